"""Pins the FP64 CPU oracle of the pseudo-spectral path with the reference's
own tests (test/spectralmpitests.jl), numpy's FFT (pocketfft) and closed
forms.  CPU only."""
import json
import math
import os

import numpy as np
import pytest

from distributednavierstokes_amd.problem import SpectralProblem
from oracle.oracle import OracleSpectral, splitmix64

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fx(x, y, z):
    return np.sin(x) * np.cos(y) * np.cos(z) + 0.3 * np.cos(2 * y) * np.sin(z)


def fy(x, y, z):
    return -np.cos(x) * np.sin(y) * np.cos(z) + 0.2 * np.sin(2 * z) * np.cos(x)


def fz(x, y, z):
    return 0.1 * np.sin(x + 2 * z) * np.cos(2 * y) + 0.25 * np.cos(3 * x)


def grid_field(N, fs=(fx, fy, fz), l=2 * math.pi):
    x = l / N * (np.arange(1, N + 1) - 0.5)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    return np.asfortranarray(np.stack([f(X, Y, Z) for f in fs], -1))


def rfft3(u):
    """numpy rfft over dims 1:3 with dim 1 halved (the reference's rfft(u, 1:3))."""
    return np.stack([np.transpose(np.fft.rfftn(np.transpose(u[..., c]))) for c in range(u.shape[-1])], -1)


def test_splitmix64_known_answers():
    kat = json.load(open(os.path.join(GOLDEN, "splitmix64_kat.json")))
    for x, want in kat["pairs"]:
        assert splitmix64(int(x)) == int(want, 16)


@pytest.mark.parametrize("N", [12, 18])
def test_pipeline_matches_numpy(N):
    """test/spectralmpitests.jl:30-71."""
    o = OracleSpectral(SpectralProblem(n=N))
    kcut, m = o.kcut, o.m
    u = grid_field(N)
    F = rfft3(u) / N ** 3
    fine = lambda j: j if j <= kcut else j + (N - m)  # noqa: E731
    uh = o.velocityfield(fx, fy, fz, project=False)
    err = max(abs(uh[i, j, k, c] - F[i, fine(j), fine(k), c])
              for c in range(3) for k in range(m) for j in range(m) for i in range(kcut + 1))
    assert err < 1e-13
    ii, jj, kk = np.meshgrid(np.arange(N // 2 + 1), np.arange(N), np.arange(N), indexing="ij")
    mask = (ii <= kcut) & (np.minimum(jj, N - jj) <= kcut) & (np.minimum(kk, N - kk) <= kcut)
    udeal = np.stack([np.transpose(np.fft.irfftn(np.transpose(rfft3(u)[..., c] * mask), s=(N, N, N), axes=(0, 1, 2)))
                      for c in range(3)], -1)
    v = o.spec_to_phys(uh)
    assert np.abs(v - udeal).max() < 1e-13
    o.project(uh)
    assert o.maxdiv(uh) < 1e-14


def test_taylor_green_2d_exact_decay():
    """test/spectralmpitests.jl:73-97."""
    visc = 0.01
    o = OracleSpectral(SpectralProblem(n=12, visc=visc))
    uh = o.velocityfield(lambda x, y, z: np.sin(x) * np.cos(y), lambda x, y, z: -np.cos(x) * np.sin(y),
                         lambda x, y, z: 0 * x)
    o.solve(uh, (0.0, 1.0), dt=0.05)
    v = o.spec_to_phys(uh)
    d = math.exp(-2 * visc)
    ex = grid_field(12, (lambda x, y, z: d * np.sin(x) * np.cos(y), lambda x, y, z: -d * np.cos(x) * np.sin(y),
                         lambda x, y, z: 0 * x))
    assert np.abs(v - ex).max() < 1e-12


def test_integrating_factor_telescopes_without_nonlinearity():
    """A shear wave u = (sin 2y, 0, 0) has zero nonlinear term: the three IF
    stages must multiply to exp(−ν k² Δt) exactly (src/spectral/step.jl:12-13)."""
    visc = 0.03
    o = OracleSpectral(SpectralProblem(n=12, visc=visc))
    uh = o.velocityfield(lambda x, y, z: np.sin(2 * y), lambda x, y, z: 0 * x, lambda x, y, z: 0 * x)
    o.solve(uh, (0.0, 0.7), dt=0.1)
    v = o.spec_to_phys(uh)
    ex = grid_field(12, (lambda x, y, z: math.exp(-4 * visc * 0.7) * np.sin(2 * y), lambda x, y, z: 0 * x,
                         lambda x, y, z: 0 * x))
    assert np.abs(v - ex).max() < 1e-13


def test_inviscid_energy_conservation():
    """test/spectralmpitests.jl:118-126."""
    o = OracleSpectral(SpectralProblem(n=18, visc=0.0))
    uh = o.velocityfield(lambda x, y, z: np.sin(x) * np.cos(y) * np.cos(z),
                         lambda x, y, z: -np.cos(x) * np.sin(y) * np.cos(z), lambda x, y, z: 0 * x)
    e0 = o.energy(uh)
    o.solve(uh, (0.0, 0.1), dt=0.002)
    assert abs(o.energy(uh) - e0) / e0 < 1e-10
    assert o.maxdiv(uh) < 1e-13


def test_random_ic_and_forcing():
    """test/spectralmpitests.jl:128-161."""
    prob = SpectralProblem(n=18, visc=5e-3)
    o = OracleSpectral(prob)
    uh0 = o.randomfield(totalenergy=0.4, kpeak=2, seed=7)
    assert math.isclose(o.energy(uh0), 0.4, rel_tol=1e-12)
    assert o.maxdiv(uh0) < 1e-14
    E = o.shell_energies(uh0, range(0, o.kcut + 1))
    kdiag = int(math.floor(math.sqrt(3) * o.kcut))
    prof = np.array([k ** 4 * math.exp(-2 * (k / 2) ** 2) for k in range(kdiag + 1)])
    target = 0.4 * prof / prof.sum()
    assert np.abs(E - target[:o.kcut + 1]).max() < 1e-13
    uh = uh0.copy(order="F")
    f = o.shellforcing(uh, shells=(1, 2))
    st = o.solve(uh, (0.0, 0.1), cfl=0.4, forcing=f)
    assert st["n"] > 1
    assert np.allclose(o.shell_energies(uh, (1, 2)), o.shell_energies(uh0, (1, 2)), rtol=1e-12)


def test_dissipation_closed_form():
    """Taylor–Green: ε = ν Σ k²|û|² = 3ν·2e for the k² = 3 mode set."""
    o = OracleSpectral(SpectralProblem(n=12, visc=0.01))
    uh = o.velocityfield(lambda x, y, z: np.sin(x) * np.cos(y) * np.cos(z),
                         lambda x, y, z: -np.cos(x) * np.sin(y) * np.cos(z), lambda x, y, z: 0 * x)
    e = o.energy(uh)
    assert math.isclose(e, 0.125, rel_tol=1e-12)
    assert math.isclose(o.dissipation(uh), 0.01 * 3 * 2 * e, rel_tol=1e-12)


def test_noise_known_answers():
    """noise_kernel! (src/spectral/ic.jl:38-48) at fixed (n, seed, i, j, k, c)
    against values computed from the published recipe with Python integers
    (tests/golden/make_golden.py): pins the key order ((c·n + k)·n + j)·n + i,
    the 2·key / 2·key + 1 counters, the 53-bit uniforms and Box–Muller."""
    kat = json.load(open(os.path.join(GOLDEN, "noise_kat.json")))
    fields = {}
    for n, seed, i, j, k, c, want in kat["points"]:
        seed, want = int(seed), float(want)
        if (n, seed) not in fields:
            fields[(n, seed)] = OracleSpectral(SpectralProblem(n=n)).noise(seed)
        got = fields[(n, seed)][i, j, k, c]
        assert abs(got - want) <= 2e-15 * max(1.0, abs(want)), (n, seed, i, j, k, c, got, want)


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_convection_carries_a_shear_wave_downstream(direction):
    """A shear wave in a uniform mean flow U along `direction` (ν = 0): the
    only nonzero term of du_i = −ik_j F[v_i v_j] (src/spectral/operators.jl:60-79)
    is −ik U û for the wave component, so its k = 1 mode turns by exactly
    −U t (Wray RK3 adds O((UΔt)⁵) per step).  Pins the sign and size of the
    convection term along each axis (the mean mode is untouched by the
    projection, src/spectral/operators.jl:9)."""
    U, t_end, dt = 0.8, 0.6, 0.01
    o = OracleSpectral(SpectralProblem(n=12, visc=0.0))
    wave = (direction + 2) % 3
    fs = [lambda x, y, z: 0 * x] * 3
    fs[direction] = lambda x, y, z: U + 0 * x
    fs[wave] = lambda x, y, z, d=direction: 0.5 * np.sin((x, y, z)[d])
    uh = o.velocityfield(*fs)
    idx = [0, 0, 0]
    idx[direction] = 1
    c0 = uh[idx[0], idx[1], idx[2], wave]
    o.solve(uh, (0.0, t_end), dt=dt)
    c1 = uh[idx[0], idx[1], idx[2], wave]
    shift = -np.angle(c1 / c0)
    assert abs(shift - U * t_end) < 1e-6 * U * t_end, (shift, U * t_end)
    assert abs(abs(c1 / c0) - 1) < 1e-6
    assert abs(uh[0, 0, 0, direction] - U) < 1e-14


def test_rhs_equals_advective_form():
    """The divergence form du_i = −ik_j F[v_i v_j] (src/spectral/operators.jl:60-79)
    equals the advective form −F[(v·∇)v_i] for a divergence-free band-limited
    v, computed here independently with numpy (spectral derivatives of v,
    products in physical space).  With n = 12, kcut = 3 no product mode
    aliases onto a retained one (n − 2kcut > kcut), so the two agree to
    roundoff.  Pins every term of the nonlinearity, diagonal (v_i v_i) and
    off-diagonal, with its sign."""
    N, kcut = 12, 3
    o = OracleSpectral(SpectralProblem(n=N, kcut=kcut, visc=0.0))
    uh = o.randomfield(totalenergy=1.0, kpeak=2, seed=3)
    du, _ = o.rhs(uh)
    m = o.m
    fine = lambda j: j if j <= kcut else j + (N - m)  # noqa: E731
    Uf = np.zeros((N // 2 + 1, N, N, 3), dtype=complex)
    for k in range(m):
        for j in range(m):
            Uf[:kcut + 1, fine(j), fine(k), :] = uh[:, j, k, :] * N ** 3
    kx = np.arange(N // 2 + 1)[:, None, None]
    ky = np.fft.fftfreq(N, 1.0 / N)[None, :, None]
    kz = np.fft.fftfreq(N, 1.0 / N)[None, None, :]
    kvec = (kx, ky, kz)

    def irfft3(a):
        return np.transpose(np.fft.irfftn(np.transpose(a), s=(N, N, N), axes=(0, 1, 2)))

    v = [irfft3(Uf[..., c]) for c in range(3)]
    adv = np.stack([sum(v[d] * irfft3(1j * kvec[d] * Uf[..., c]) for d in range(3)) for c in range(3)], -1)
    A = rfft3(adv) / N ** 3
    want = np.zeros_like(du)
    for k in range(m):
        for j in range(m):
            want[:, j, k, :] = -A[:kcut + 1, fine(j), fine(k), :]
    assert np.abs(du - want).max() < 1e-13 * np.abs(want).max()
