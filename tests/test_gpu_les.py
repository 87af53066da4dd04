"""GPU SFS sampler (les.jl recipe) against the numpy restatement and a
closed form.  FP64."""
import math

import numpy as np
import pytest

import distributednavierstokes_amd as dns
from distributednavierstokes_amd.les import lessampler
from distributednavierstokes_amd.problem import SpectralProblem
from oracle.oracle import OracleSpectral, les_cube, lesfreq, sfs_sample_oracle

pytestmark = pytest.mark.gpu


def compact_from_coarse(a, M):
    """Coarse rfft array (M/2+1, M, M, c) → LES compact (kc+1, 2kc+1, 2kc+1, c)."""
    kc = M // 3
    m = 2 * kc + 1
    f = np.array([j if j <= kc else j - m for j in range(m)])
    idx = np.where(f >= 0, f, M + f)
    return a[: kc + 1][:, idx][:, :, idx]


def test_sfs_matches_numpy_recipe():
    n, M = 48, 24
    s = dns.spectral_setup(n=n, visc=5e-3)
    o = OracleSpectral(SpectralProblem(n=n, visc=5e-3))
    uh = o.randomfield(seed=2, kpeak=3)
    sam = lessampler(s, M=M, delta_fac=2.0)
    sam.sample(uh)
    ub, tau = sam.spectral()
    sig = o.products(uh)
    ubo, tauo = sfs_sample_oracle(uh, sig, o.kcut, M, o.prob.l, 2.0, 2)
    ubc, tauc = compact_from_coarse(ubo, M), compact_from_coarse(tauo, M)
    scale = np.abs(tauc).max()
    assert np.abs(ub - ubc).max() < 1e-13 * max(1.0, np.abs(ubc).max())
    assert np.abs(tau - tauc).max() < 1e-12 * scale
    assert np.abs(tau[..., 0] + tau[..., 1] + tau[..., 2]).max() < 1e-15 * scale + 1e-30
    ubp, taup = sam.physical()
    assert ubp.shape == (M, M, M, 3) and taup.shape == (M, M, M, 6)
    # the physical fields transform back to the compact spectra
    w = dns.specvelocity(sam.les)
    back = dns.phys_to_spec_(w, ubp, sam.les)
    assert np.abs(back - ub).max() < 1e-13


def test_sfs_closed_form_shear_wave():
    """u = (cos 4z, 0, 0): τ_xx(±8 ez) = ⅔·¼(G(8) − G(4)²), τ_yy = τ_zz = −½ τ_xx.

    The DNS samples sit at cell centres z_j = (j + ½)h, h = 2π/n, and the
    spectra are index-based DFTs, so mode k carries the phase e^{ikh/2}."""
    n, M = 48, 24
    s = dns.spectral_setup(n=n, visc=1e-3)
    uh = dns.spectral_velocityfield_(dns.specvelocity(s), s, lambda x, y, z: math.cos(4 * z),
                                     lambda x, y, z: 0.0, lambda x, y, z: 0.0)
    sam = lessampler(s, M=M, delta_fac=2.0).sample(uh)
    ub, tau = sam.spectral()
    d = 2.0 * 2 * math.pi / M
    G = lambda k: math.exp(-d * d * k * k / 24)  # noqa: E731
    want = (2 / 3) * 0.25 * (G(8) - G(4) ** 2)
    m = sam.les.m
    h = 2 * math.pi / n
    for kz, k in ((8, 8), (m - 8, -8)):
        ph = complex(math.cos(k * h / 2), math.sin(k * h / 2))
        assert abs(tau[0, 0, kz, 0] - want * ph) < 1e-14
        assert abs(tau[0, 0, kz, 1] + want * ph / 2) < 1e-14
        assert abs(tau[0, 0, kz, 2] + want * ph / 2) < 1e-14
    assert abs(ub[0, 0, 4, 0] - 0.5 * G(4) * complex(math.cos(2 * h), math.sin(2 * h))) < 1e-15
    # nothing else is nonzero: xx, yy, zz at kz = ±8 carry (1 + ½ + ½)|want| each
    assert abs(np.abs(tau).sum() - 4 * abs(want)) < 1e-12


def test_sfswriter_dataset(tmp_path):
    """sfswriter (src/spectral/les.jl:463-511) on one Gaussian cell: the
    archives hold the reference's keys, the samples at the requested times
    (the last one equal to the sampler's current ū, τ̂), the filter Reynolds
    numbers and spectra; the physical pairs read back from disk equal the
    GPU's physical fields and load into the network's dataset."""
    from distributednavierstokes_amd.cnn import SFSNet
    from distributednavierstokes_amd.les import read_sfs, sfs_physical_pairs, sfs_to_network, sfswriter

    n, M = 48, 16
    s = dns.spectral_setup(n=n, visc=5e-3)
    uh = dns.specvelocity(s)
    dns.spectral_randomfield_(uh, s, totalenergy=0.4, kpeak=3, seed=5)
    times = [0.02, 0.04, 0.06]
    w = sfswriter(s, str(tmp_path), times, M=M, delta_fac=2.0)
    dns.spectral_solve_(uh=uh, setup=s, tlims=(0.0, 0.06), dt=0.01, tstops=times, processors={"sfs": w})
    d = read_sfs(str(tmp_path), 2.0)
    meta, f, lm = d["dns_meta"], d["fields"], d["les_meta"]
    assert set(meta) == {"times", "spectra_dns", "statistics_dns", "statistics_dns_keys", "t_int", "walltime"}
    assert set(f) == {"inputs.x", "inputs.y", "inputs.z", "outputs.xx", "outputs.yy", "outputs.zz", "outputs.xy",
                      "outputs.yz", "outputs.zx", "redelta", "Δ", "Δ_factor", "visc", "kernel", "M", "delta_eta"}
    assert set(lm) == {"spectra_les", "redelta_mean"}
    np.testing.assert_allclose(meta["times"], times, rtol=0, atol=1e-12)
    assert meta["spectra_dns"].shape == (s.kcut, 3) and meta["statistics_dns"].shape == (12, 3)
    assert f["inputs.x"].shape == (M // 2 + 1, M, M, 3) and f["outputs.zx"].shape == (M // 2 + 1, M, M, 3)
    assert float(f["Δ"]) == 2.0 * s.l / M and int(f["M"]) == M and str(f["kernel"]) == "gaussian"
    assert lm["spectra_les"].shape == (M // 3, 3)
    sam = w.sampler
    ub, tau = sam.spectral()
    assert np.array_equal(compact_from_coarse(np.stack([f[f"inputs.{c}"][..., 2] for c in "xyz"], -1), M), ub)
    names = ("xx", "yy", "zz", "xy", "yz", "zx")
    assert np.array_equal(compact_from_coarse(np.stack([f[f"outputs.{c}"][..., 2] for c in names], -1), M), tau)
    assert f["redelta"][2] == sam.filter_reynolds()
    assert np.array_equal(lm["spectra_les"][:, 2], sam.spectrum())
    assert math.isclose(float(lm["redelta_mean"]), float(np.mean(f["redelta"])), rel_tol=1e-15)
    pairs = list(sfs_physical_pairs(str(tmp_path), 2.0))
    ubp, taup = sam.physical()
    assert np.abs(pairs[-1][0] - ubp).max() < 1e-12 * np.abs(ubp).max()
    assert np.abs(pairs[-1][1] - taup).max() < 1e-12 * np.abs(taup).max()
    net = SFSNet(M=M, batch=2, seed=0)
    assert sfs_to_network(net, str(tmp_path), 2.0) == 3 and net.nsamples == 3
