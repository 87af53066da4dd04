"""Pins the FP64 CPU oracle of the finite-volume path with the reference's
own closed-form and invariant tests (test/serialtests.jl, test/mpitests.jl)
plus dense brute-force checks.  CPU only."""
import math

import numpy as np
import pytest

from fvhelpers import (caseproblem, channelproblem, fill_velocity, randomfield, vzeros, weighted_dot, zeros)
from distributednavierstokes_amd.problem import FVProblem, tanhstretch
from oracle.oracle import OracleFV


def _cases():
    return (("periodic", FVProblem(n=(16, 12, 8), visc=0.0)), ("channel", channelproblem(visc=0.0)))


@pytest.mark.parametrize("case", ["periodic", "channel"])
def test_projection_and_skew_symmetry(case):
    """test/serialtests.jl:43-59."""
    prob = dict(_cases())[case]
    o = OracleFV(prob)
    u = randomfield(prob, halo=o.exchange_halo_u)
    p = zeros(prob)
    o.project(u, p)
    assert o.maxdiv(u) < 1e-12
    q = vzeros(prob)
    o.momentum(q, u, 0.0, 1.0)
    d, n2 = weighted_dot(prob, o.griddata(), u, q)
    assert abs(d) < 1e-12 * n2


def test_skew_symmetry_catches_planted_error():
    """A dropped Δy weight in the y-momentum transporters breaks the
    skew-symmetry identity, so the test above can fail."""
    prob = channelproblem(visc=0.0)
    o = OracleFV(prob)
    u = randomfield(prob, halo=o.exchange_halo_u)
    o.project(u, zeros(prob))
    q = vzeros(prob)
    o.momentum(q, u, 0.0, 1.0)
    g = o.griddata()
    # perturb q.y by an asymmetric term that a plausible bug would add
    qbad = [a.copy() for a in q]
    qbad[1][1:-1, 1:-2, 1:-1] += 1e-3 * u[0][1:-1, 1:-2, 1:-1] * g["dy"][None, 1:-2, None]
    d, n2 = weighted_dot(prob, g, u, qbad)
    assert abs(d) > 1e-8 * n2


def test_inviscid_energy_conservation():
    """test/serialtests.jl:61-72."""
    prob = FVProblem(n=(16, 16, 16), visc=0.0)
    o = OracleFV(prob)
    u = randomfield(prob, halo=o.exchange_halo_u)
    o.project(u, zeros(prob))
    e0 = o.kinetic_energy(u)
    o.solve(u, (0.0, 0.1), dt=0.002)
    e1 = o.kinetic_energy(u)
    assert abs(e1 - e0) / e0 < 1e-7
    assert o.maxdiv(u) < 1e-12


def test_taylor_green_accuracy():
    """test/serialtests.jl:74-101: exp(−2νt) decay, second order in space."""
    nu, tend = 0.01, 0.5
    errs = []
    for n in (16, 32):
        prob = FVProblem(n=(n, n, 4), visc=nu)
        o = OracleFV(prob)
        u = vzeros(prob)
        fill_velocity(prob, u, lambda x, y, z: math.sin(x) * math.cos(y), lambda x, y, z: -math.cos(x) * math.sin(y))
        o.exchange_halo_u(u)
        o.solve(u, (0.0, tend), dt=1e-3)
        co = prob.coordinates()
        ex = np.sin(co["xf"])[:, None, None] * np.cos(co["yc"])[None, :, None] * math.exp(-2 * nu * tend)
        errs.append(float(np.abs(u[0][1:-1, 1:-1, 1:-1] - ex).max()))
    assert errs[1] < 1e-3
    assert 3 < errs[0] / errs[1] < 5


def test_energy_spectrum_parseval_and_taylor_green():
    """test/serialtests.jl:103-125."""
    prob = FVProblem(n=(16, 16, 16), visc=0.0)
    o = OracleFV(prob)
    u = randomfield(prob, halo=o.exchange_halo_u)
    o.project(u, zeros(prob))
    nbins = math.ceil(math.sqrt(3 * 8 ** 2)) + 1
    E = np.zeros(nbins)
    o.spectrum_sample(u, E)
    V = (2 * math.pi) ** 3
    assert math.isclose(E.sum(), o.kinetic_energy(u) / V, rel_tol=1e-12)
    utg = vzeros(prob)
    fill_velocity(prob, utg, lambda x, y, z: math.sin(x) * math.cos(y), lambda x, y, z: -math.cos(x) * math.sin(y))
    o.exchange_halo_u(utg)
    E2 = np.zeros(nbins)
    o.spectrum_sample(utg, E2)
    assert math.isclose(E2[1] / E2.sum(), 1.0, rel_tol=1e-12)


def test_channel_statistics_exact_laminar():
    """test/serialtests.jl:127-139."""
    prob = channelproblem()
    o = OracleFV(prob)
    u = vzeros(prob)
    fill_velocity(prob, u, lambda x, y, z: 1 - y ** 2)
    o.exchange_halo_u(u)
    ny, nf = prob.n[1], prob.n[1] - 1
    acc = np.zeros(4 * ny + 4 * nf)
    o.channelstats_sample(u, acc)
    norm = prob.n[0] * prob.n[2]
    U = acc[:ny] / norm
    uu = acc[2 * ny:3 * ny] / norm - U ** 2
    V, Uf, uvf = acc[4 * ny:4 * ny + nf] / norm, acc[4 * ny + 2 * nf:4 * ny + 3 * nf] / norm, acc[4 * ny + 3 * nf:] / norm
    y = o.griddata()["ycg"]
    assert np.abs(U - (1 - y ** 2)).max() < 1e-13
    assert np.abs(uu).max() < 1e-13
    assert np.abs(uvf - Uf * V).max() < 1e-13


def _run_cn(yd, dt):
    prob = channelproblem(visc=0.05, bodyforce=(1.0, 0.0, 0.0), ydiffusion=yd)
    o = OracleFV(prob)
    u = vzeros(prob)
    fill_velocity(prob, u, lambda x, y, z: (1 - y ** 2) * (1 + 0.1 * math.sin(2 * x) * math.cos(z)))
    o.exchange_halo_u(u)
    o.solve(u, (0.0, 0.2), dt=dt)
    return u, o


def test_implicit_y_diffusion_second_order_and_stable():
    """test/serialtests.jl:141-174."""
    d = []
    for dt in (0.004, 0.002):
        ue, _ = _run_cn("explicit", dt)
        ui, oi = _run_cn("implicit", dt)
        assert oi.maxdiv(ui) < 1e-12
        d.append(max(float(np.abs(ue[c] - ui[c]).max()) for c in range(3)))
    assert d[0] < 1e-6
    assert 3 < d[0] / d[1] < 5
    prob = channelproblem(n=(8, 48, 8), stretch=tanhstretch(2.5), visc=0.2, bodyforce=(1.0, 0.0, 0.0),
                          ydiffusion="implicit")
    o = OracleFV(prob)
    u = vzeros(prob)
    fill_velocity(prob, u, lambda x, y, z: (1 - y ** 2) * (1 + 0.1 * math.sin(2 * x)))
    o.exchange_halo_u(u)
    o.solve(u, (0.0, 1.0), dt=5e-3)
    assert all(np.isfinite(a).all() for a in u)
    assert o.kinetic_energy(u) < 100
    assert o.maxdiv(u) < 1e-12


def test_channel_flow_sanity():
    """test/serialtests.jl:176-189 (CFL-adaptive)."""
    prob = channelproblem(visc=0.05, bodyforce=(1.0, 0.0, 0.0))
    o = OracleFV(prob)
    u = vzeros(prob)
    st = o.solve(u, (0.0, 1.0))
    assert math.isclose(st["t"], 1.0)
    assert o.maxdiv(u) < 1e-12
    assert 0 < o.kinetic_energy(u) < 100
    prof = u[0][1:-1, 1:-1, 1:-1].sum(axis=(0, 2))
    assert (prof > 0).all()
    assert np.allclose(prof, prof[::-1], rtol=1e-8)


def _expectedvalue(stag, g, gi, gj, gk, n, wall):
    """test/mpitests.jl:86-106 (global 1-based indices)."""
    ny = n[1]
    if wall:
        if stag == 1:
            if gj in (0, ny):
                return 0.0
            if gj < 0:
                return -_expectedvalue(stag, g, gi, -gj, gk, n, wall)
            if gj > ny:
                return -_expectedvalue(stag, g, gi, 2 * ny - gj, gk, n, wall)
        else:
            sgn = 1.0 if stag == 3 else -1.0
            if gj < 1:
                return sgn * _expectedvalue(stag, g, gi, 1 - gj, gk, n, wall)
            if gj > ny:
                return sgn * _expectedvalue(stag, g, gi, 2 * ny + 1 - gj, gk, n, wall)
    m1 = lambda a, b: (a - 1) % b + 1  # noqa: E731
    return g(m1(gi, n[0]), m1(gj, n[1]), m1(gk, n[2]))


@pytest.mark.parametrize("bcy", ["periodic", "wall"])
@pytest.mark.parametrize("stag", [0, 1, 2, 3])
def test_halo_expected_values(bcy, stag):
    """test/mpitests.jl:108-129, single rank: ghosts are exact (==)."""
    prob = caseproblem(bcy)
    o = OracleFV(prob)
    g = lambda i, j, k: math.sin(0.7 * i) + math.cos(0.3 * j) * math.sin(0.2 * k)  # noqa: E731
    phi = zeros(prob)
    r = prob.active_ranges(stag)
    for k in r[2]:
        for j in r[1]:
            for i in r[0]:
                phi[i, j, k] = g(i, j, k)
    o.exchange_halo(phi, stag)
    n = prob.n
    for k in range(n[2] + 2):
        for j in range(n[1] + 2):
            for i in range(n[0] + 2):
                assert phi[i, j, k] == _expectedvalue(stag, g, i, j, k, n, prob.wall)


@pytest.mark.parametrize("bcy", ["periodic", "wall"])
def test_poisson_manufactured(bcy):
    """test/mpitests.jl:133-154: u = −∇pex, r = div u, the solve returns −pex + const."""
    prob = caseproblem(bcy)
    o = OracleFV(prob)
    pex = zeros(prob)
    r = prob.active_ranges(3)
    for k in r[2]:
        for j in r[1]:
            for i in r[0]:
                pex[i, j, k] = math.sin(0.5 * i) + math.cos(0.3 * j) * math.sin(0.2 * k)
    o.exchange_halo(pex, 3)
    u = vzeros(prob)
    for c in range(3):
        o.pgrad(u[c], pex, c, 1.0)
    o.exchange_halo_u(u)
    rr = o.divergence(u)
    p2 = zeros(prob)
    o.poisson_from(rr, p2)
    d = p2[1:-1, 1:-1, 1:-1] + pex[1:-1, 1:-1, 1:-1]
    assert np.abs(d - d.mean()).max() < 1e-8


@pytest.mark.parametrize("bcy", ["periodic", "wall"])
def test_poisson_inverts_dense_div_grad(bcy):
    """Brute force: assemble L = div∘grad column by column at n = (8, 6, 4)
    and check L p = r for the solver's p (r of zero mean)."""
    prob = FVProblem(n=(8, 6, 4), lims=((0, 2 * math.pi), (-1, 1) if bcy == "wall" else (0, 2 * math.pi), (0, 3.0)),
                     bc=("periodic", bcy, "periodic"), stretch=tanhstretch(1.4) if bcy == "wall" else (lambda t: t))
    o = OracleFV(prob)
    n = prob.n
    Ncell = int(np.prod(n))
    Lmat = np.zeros((Ncell, Ncell))
    for m in range(Ncell):
        e = np.zeros(n, order="F")
        e.ravel(order="F")[m] = 1.0
        p = zeros(prob)
        p[1:-1, 1:-1, 1:-1] = e
        o.exchange_halo(p, 3)
        u = vzeros(prob)
        for c in range(3):
            o.pgrad(u[c], p, c, -1.0)
        o.exchange_halo_u(u)
        Lmat[:, m] = o.divergence(u).ravel(order="F")
    rng = np.random.default_rng(1)
    r = rng.standard_normal(Ncell)
    r -= r.mean() if bcy == "periodic" else 0.0
    if bcy == "wall":
        # compatibility: r orthogonal to the null space (constants) in the Δy-weighted sense
        dyw = np.tile(o.griddata()["dy"][1:-1][None, :, None], (n[0], 1, n[2])).ravel(order="F")
        r -= (r * dyw).sum() / dyw.sum()
    p = zeros(prob)
    o.poisson_from(r.reshape(n, order="F"), p)
    pv = p[1:-1, 1:-1, 1:-1].ravel(order="F")
    assert np.abs(Lmat @ pv - r).max() < 1e-10 * np.abs(r).max()


def _mode1_phase(line):
    """Phase of the k = 1 Fourier coefficient of a periodic sample line."""
    n = len(line)
    return np.angle(np.sum(line * np.exp(-2j * np.pi * np.arange(n) / n)))


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_convection_carries_a_shear_wave_downstream(direction):
    """A shear wave in a uniform flow U along `direction` (periodic, ν = 0):
    the wave component c = (direction + 2) % 3 depends only on that
    coordinate, so the only nonzero term is its convection, which the
    staggered skew-symmetric scheme (src/operators.jl:44-145) discretises as
    the central flux U (w_{i+1} − w_{i−1}) / 2Δ.  The k = 1 mode therefore
    moves downstream with U_eff = U sin(Δ)/Δ: its phase at time t is
    −U_eff t.  Pins the sign and the size of each component's convection
    term (a flipped sign advects upstream; a one-sided or doubled flux
    changes U_eff)."""
    n, U, t_end, dt = 16, 1.0, 0.5, 0.01
    prob = FVProblem(n=(n, n, n), visc=0.0)
    o = OracleFV(prob)
    u = vzeros(prob)
    wave = (direction + 2) % 3
    carrier = [None, None, None]
    carrier[direction] = lambda x, y, z: U
    carrier[wave] = lambda x, y, z, d=direction: 0.5 * math.sin((x, y, z)[d])
    fill_velocity(prob, u, *carrier)
    o.exchange_halo_u(u)

    def line(a):
        a = a[1:-1, 1:-1, 1:-1]
        return a.mean(axis=tuple(ax for ax in range(3) if ax != direction))

    ph0 = _mode1_phase(line(u[wave]))
    o.solve(u, (0.0, t_end), dt=dt)
    shift = ph0 - _mode1_phase(line(u[wave]))
    delta = 2 * math.pi / n
    expected = U * math.sin(delta) / delta * t_end
    assert abs(shift - expected) < 1e-4 * expected, (shift, expected)
    assert np.abs(u[direction][1:-1, 1:-1, 1:-1] - U).max() < 1e-13  # the carrier is untouched


def _hand_dymin(prob):
    ny = prob.n[1]
    y0, y1 = prob.lims[1]
    yf = [y0 + (y1 - y0) * prob.stretch(j / ny) for j in range(ny + 1)]
    return min(yf[j + 1] - yf[j] for j in range(ny))


@pytest.mark.parametrize("ydiffusion", ["explicit", "implicit"])
def test_cfl_timestep_closed_forms(ydiffusion):
    """cfl_timestep (src/timestep.jl:69-80) from hand-computed metrics:
    a uniform flow along each axis hits the convective bound cfl·Δ/U (Δymin
    for y, the smallest stretched cell); a fluid at rest hits the diffusive
    bound cfl/(2ν(1/Δx² + 1/Δymin² + 1/Δz²)), without the 1/Δymin² term
    when y-diffusion is implicit."""
    cfl, eps = 0.45, 2.220446049250313e-16
    lims = ((0.0, 4 * math.pi), (-1.0, 1.0), (0.0, 4 * math.pi / 3))
    stretch = tanhstretch(1.5)
    prob_c = FVProblem(n=(16, 12, 8), lims=lims, bc=("periodic", "wall", "periodic"), stretch=stretch, visc=1e-12,
                       ydiffusion=ydiffusion)
    dx, dz, dymin = 4 * math.pi / 16, 4 * math.pi / 3 / 8, _hand_dymin(prob_c)
    o = OracleFV(prob_c)
    for c, (d, U) in enumerate(((dx, 3.0), (dymin, 2.5), (dz, 1.5))):
        u = vzeros(prob_c)
        f = [None, None, None]
        f[c] = lambda x, y, z, U=U: U
        fill_velocity(prob_c, u, *f)
        o.exchange_halo_u(u)
        assert math.isclose(o.cfl_timestep(u, cfl), cfl * d / (U + eps), rel_tol=1e-14), c
    nu = 0.01
    prob_d = FVProblem(n=(16, 12, 8), lims=lims, bc=("periodic", "wall", "periodic"), stretch=stretch, visc=nu,
                       ydiffusion=ydiffusion)
    o = OracleFV(prob_d)
    invdy2 = 0.0 if ydiffusion == "implicit" else 1 / dymin ** 2
    want = cfl / (2 * nu * (1 / dx ** 2 + invdy2 + 1 / dz ** 2) + eps)
    assert math.isclose(o.cfl_timestep(vzeros(prob_d), cfl), want, rel_tol=1e-14)
