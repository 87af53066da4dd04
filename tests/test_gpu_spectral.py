"""GPU parity of the pseudo-spectral path against the FP64 CPU oracle and
numpy's FFT, through the C ABI.  FP64 tolerances as in
test/spectralmpitests.jl (1e-13 transforms, 1e-12 trajectories)."""
import math

import numpy as np
import pytest

import distributednavierstokes_amd as dns
from distributednavierstokes_amd.problem import SpectralProblem
from oracle.oracle import OracleSpectral
from test_oracle_spectral import fx, fy, fz, grid_field, rfft3

pytestmark = pytest.mark.gpu


def pair(n, visc=1e-3, kcut=None, T=np.float64):
    s = dns.spectral_setup(n=n, visc=visc, kcut=kcut, T=T)
    o = OracleSpectral(SpectralProblem(n=n, visc=visc, kcut=kcut))
    return s, o


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("N", [12, 18, 24, 64])
def test_pipeline_against_numpy_and_oracle(N):
    s, o = pair(N)
    kcut, m = s.kcut, s.m
    u = grid_field(N)
    uh = dns.phys_to_spec_(dns.specvelocity(s), u, s)
    F = rfft3(u) / N ** 3
    fine = lambda j: j if j <= kcut else j + (N - m)  # noqa: E731
    idx = [fine(j) for j in range(m)]
    ref = F[:kcut + 1][:, idx][:, :, idx]
    assert np.abs(uh - ref).max() < 1e-13
    assert np.abs(uh - o.phys_to_spec(u)).max() < 1e-13
    v = dns.spec_to_phys_(np.zeros((N,) * 3 + (3,), order="F"), uh, s)
    assert np.abs(v - o.spec_to_phys(uh)).max() < 1e-13
    dns.spectral_project_(uh, s)
    assert dns.spectral_maxdiv(uh, s) < 1e-14


@pytest.mark.parametrize("N", [12, 24])
def test_rhs_parity(N):
    s, o = pair(N)
    uh = o.velocityfield(fx, fy, fz)
    du = dns.spectral_rhs_(dns.specvelocity(s), uh, s)
    duo, vmax = o.rhs(uh)
    assert rel(du, duo) < 1e-12
    out = np.zeros(1)
    s.upload(uh)
    assert math.isclose(s.scalar(dns._lib.lib().dns_spec_vsummax), vmax, rel_tol=1e-13)


@pytest.mark.parametrize("xpass", [{}, {"DNS_SP_XRHS": "1"}, {"DNS_SP_XRHS": "1", "DNS_SP_XLPI": "3"},
                                   {"DNS_SP_XRHS": "0"}])
@pytest.mark.parametrize("N", [12, 18, 64])
def test_step_parity(N, xpass, monkeypatch):
    """Three IF-RK3 steps and their CFL velocity bound against the oracle,
    with the persistent RHS x pass (16 or 3 lines per block: a partial last
    block) and with the one-line-per-block kernel."""
    for key, val in xpass.items():
        monkeypatch.setenv(key, val)
    s, o = pair(N, visc=5e-3)
    uh = o.velocityfield(fx, fy, fz)
    ug = uh.copy(order="F")
    for _ in range(3):
        vo = o.step(uh, 0.01)
        vg = dns.spectral_step_(ug, s, 0.01)
        assert math.isclose(vg, vo, rel_tol=1e-12)
    assert rel(ug, uh) < 1e-12
    assert dns.spectral_maxdiv(ug, s) < 1e-13


def test_taylor_green_2d_exact_decay():
    visc = 0.01
    s = dns.spectral_setup(n=12, visc=visc)
    uh = dns.spectral_velocityfield_(dns.specvelocity(s), s, lambda x, y, z: math.sin(x) * math.cos(y),
                                     lambda x, y, z: -math.cos(x) * math.sin(y), lambda x, y, z: 0.0)
    dns.spectral_solve_(uh=uh, setup=s, tlims=(0.0, 1.0), dt=0.05)
    v = dns.spec_to_phys_(None, uh, s)
    d = math.exp(-2 * visc)
    ex = grid_field(12, (lambda x, y, z: d * np.sin(x) * np.cos(y), lambda x, y, z: -d * np.cos(x) * np.sin(y),
                         lambda x, y, z: 0 * x))
    assert np.abs(v - ex).max() < 1e-12


def test_inviscid_energy_conservation():
    s = dns.spectral_setup(n=18, visc=0.0)
    uh = dns.taylorgreen_(dns.specvelocity(s), s)
    e0 = dns.spectral_energy(uh, s)
    dns.spectral_solve_(uh=uh, setup=s, tlims=(0.0, 0.1), dt=0.002)
    assert abs(dns.spectral_energy(uh, s) - e0) / e0 < 1e-10
    assert dns.spectral_maxdiv(uh, s) < 1e-13


def test_noise_parity():
    s, o = pair(12)
    from distributednavierstokes_amd._lib import check, lib
    check(lib().dns_spec_noise(s.handle, 7))
    v = s.download_phys()
    assert np.abs(v - o.noise(7)).max() < 1e-12


def test_random_ic_forcing_trajectory_parity():
    """test/spectralmpitests.jl:128-161 against the oracle step by step."""
    s, o = pair(18, visc=5e-3)
    uh0 = dns.spectral_randomfield_(dns.specvelocity(s), s, totalenergy=0.4, kpeak=2, seed=7)
    uo = o.randomfield(totalenergy=0.4, kpeak=2, seed=7)
    assert rel(uh0, uo) < 1e-12
    assert math.isclose(dns.spectral_energy(uh0, s), 0.4, rel_tol=1e-12)
    assert dns.spectral_maxdiv(uh0, s) < 1e-14
    sp = dns.spectral_spectrum(uh0, s)
    kdiag = int(math.floor(math.sqrt(3) * s.kcut))
    prof = np.array([k ** 4 * math.exp(-2 * (k / 2) ** 2) for k in range(kdiag + 1)])
    target = 0.4 * prof / prof.sum()
    assert np.abs(sp.E - target[:s.kcut + 1]).max() < 1e-13
    ug = uh0.copy(order="F")
    f = dns.shellforcing(ug, s, shells=(1, 2))
    st = dns.spectral_solve_(uh=ug, setup=s, tlims=(0.0, 0.1), cfl=0.4, forcing=f)
    fo = o.shellforcing(uo, shells=(1, 2))
    sto = o.solve(uo, (0.0, 0.1), cfl=0.4, forcing=fo)
    assert st.n == sto["n"]
    assert rel(ug, uo) < 1e-11
    assert np.allclose(dns.shell_energies(ug, s, (1, 2)), dns.shell_energies(uh0, s, (1, 2)), rtol=1e-12)


def test_energy_dissipation_parity():
    s, o = pair(24, visc=0.01)
    uh = o.randomfield(seed=3)
    assert math.isclose(dns.spectral_energy(uh, s), o.energy(uh), rel_tol=1e-13)
    assert math.isclose(dns.spectral_dissipation(uh, s), o.dissipation(uh), rel_tol=1e-13)
    st = dns.spectral_stats(uh, s)
    assert st.e > 0 and st.diss > 0 and st.Re_tay > 0


@pytest.mark.parametrize("xrhs", ["", "0"])
def test_mixed_precision_close_to_fp64(xrhs, monkeypatch):
    """FP32 transforms (persistent or one-line RHS x pass) stay within 1e-5 of FP64."""
    if xrhs:
        monkeypatch.setenv("DNS_SP_XRHS", xrhs)
    s64 = dns.spectral_setup(n=32, visc=5e-3)
    s32 = dns.spectral_setup(n=32, visc=5e-3, T=np.float32)
    uh = dns.spectral_randomfield_(dns.specvelocity(s64), s64, seed=1)
    u32 = uh.copy(order="F")
    dns.spectral_solve_(uh=uh, setup=s64, tlims=(0.0, 0.05), dt=0.005)
    dns.spectral_solve_(uh=u32, setup=s32, tlims=(0.0, 0.05), dt=0.005)
    err = np.linalg.norm(u32 - uh) / np.linalg.norm(uh)
    assert err < 1e-5
    assert abs(dns.spectral_energy(u32, s64) - dns.spectral_energy(uh, s64)) / dns.spectral_energy(uh, s64) < 1e-5
