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
