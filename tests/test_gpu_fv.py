"""GPU parity of the finite-volume path against the FP64 CPU oracle on the
same seeded inputs, through the C ABI (via the package's ctypes binding).
FP64 tolerance: 1e-11 relative (trajectories), 1e-12 (single operators)."""
import ctypes as ct
import dataclasses
import math

import numpy as np
import pytest

import distributednavierstokes_amd as dns
from distributednavierstokes_amd.fv import Setup, ROCmBackend
from distributednavierstokes_amd.problem import FVProblem, tanhstretch
from fvhelpers import (caseproblem, channelproblem, fill_velocity, randomfield, relmax, vzeros, weighted_dot, zeros)
from oracle.oracle import OracleFV

pytestmark = pytest.mark.gpu


def pair(prob):
    return Setup(prob, ROCmBackend()), OracleFV(prob)


def problems():
    return {
        "periodic": FVProblem(n=(16, 12, 8), visc=0.01, bodyforce=(0.3, 0.1, -0.2)),
        "channel": channelproblem(visc=0.02, bodyforce=(1.0, 0.0, 0.0)),
        "channel_cn": channelproblem(visc=0.02, bodyforce=(1.0, 0.0, 0.0), ydiffusion="implicit"),
        "case_wall": caseproblem("wall", visc=0.02),
        "case_periodic": caseproblem("periodic", visc=0.02),
    }


@pytest.mark.parametrize("name", ["periodic", "channel", "case_wall", "case_periodic"])
def test_ghost_fill_matches_oracle(name):
    prob = problems()[name]
    s, o = pair(prob)
    u = randomfield(prob, seed=3)
    ref = [a.copy(order="F") for a in u]
    o.exchange_halo_u(ref)
    dns.exchange_halo_(u, s)
    for c in range(3):
        assert np.array_equal(u[c], ref[c])


@pytest.mark.parametrize("name", ["periodic", "channel", "channel_cn"])
@pytest.mark.parametrize("A", [0.0, -5 / 9])
def test_momentum_parity(name, A):
    prob = problems()[name]
    s, o = pair(prob)
    u = randomfield(prob, seed=5, halo=o.exchange_halo_u)
    q0 = randomfield(prob, seed=6)
    qg = [a.copy(order="F") for a in q0]
    qo = [a.copy(order="F") for a in q0]
    o.momentum(qo, u, A, 0.01)
    dns.momentum_(qg, u, A, 0.01, s)
    for c in range(3):
        assert relmax(qg[c], qo[c]) < 1e-12, c


@pytest.mark.parametrize("kind", ["slab", "lz", "direct"])
@pytest.mark.parametrize("shape,wall", [((16, 12, 8), True), ((96, 20, 6), True), ((128, 16, 4), False),
                                        ((64, 9, 3), False)])
@pytest.mark.parametrize("A", [0.0, -5 / 9])
def test_momentum_kernels_parity(kind, shape, wall, A, monkeypatch):
    """Every momentum stencil kernel against the oracle, on tiles that are
    partial in x and y, with walls and periodic y."""
    monkeypatch.setenv("DNS_MOMENTUM", kind)
    if wall:
        prob = channelproblem(n=shape, visc=0.02, bodyforce=(1.0, 0.0, 0.0))
    else:
        prob = FVProblem(n=shape, visc=0.01, bodyforce=(0.3, 0.1, -0.2))
    s, o = pair(prob)
    u = randomfield(prob, seed=15, halo=o.exchange_halo_u)
    q0 = randomfield(prob, seed=16)
    qg = [a.copy(order="F") for a in q0]
    qo = [a.copy(order="F") for a in q0]
    o.momentum(qo, u, A, 0.01)
    dns.momentum_(qg, u, A, 0.01, s)
    for c in range(3):
        assert relmax(qg[c], qo[c]) < 1e-12, c


@pytest.mark.parametrize("n", [(32, 128, 8), (16, 256, 4), (16, 24, 4)])
@pytest.mark.parametrize("variant", ["", "tile", "sequential"])
def test_thomas_variants_projection_parity(n, variant, monkeypatch):
    """Every wall-normal Thomas kernel of the pressure solve (equal segments,
    general segments, one thread per line) against the oracle."""
    if variant:
        monkeypatch.setenv("DNS_THOMAS", variant)
    prob = channelproblem(n=n, visc=0.02)
    s, o = pair(prob)
    u = randomfield(prob, seed=17, halo=o.exchange_halo_u)
    ug = [a.copy(order="F") for a in u]
    po, pg = zeros(prob), zeros(prob)
    o.project(u, po)
    dns.project_(ug, pg, s)
    for c in range(3):
        assert relmax(ug[c], u[c]) < 1e-12
    assert relmax(pg, po) < 1e-11
    # the divergence left is rounding of the same size as the oracle's own
    assert dns.maxdiv(ug, s) < max(1e-12, 4 * o.maxdiv(u))


@pytest.mark.parametrize("name", ["periodic", "channel", "case_wall", "case_periodic"])
def test_projection_parity_and_divergence(name):
    prob = problems()[name]
    s, o = pair(prob)
    u = randomfield(prob, seed=7, halo=o.exchange_halo_u)
    ug = [a.copy(order="F") for a in u]
    po, pg = zeros(prob), zeros(prob)
    o.project(u, po)
    dns.project_(ug, pg, s)
    for c in range(3):
        assert relmax(ug[c], u[c]) < 1e-12
    assert relmax(pg, po) < 1e-11
    assert dns.maxdiv(ug, s) < 1e-12
    assert abs(dns.maxdiv(ug, s) - o.maxdiv(ug)) < 1e-13


@pytest.mark.parametrize("name", ["periodic", "channel", "channel_cn"])
def test_step_parity(name):
    prob = problems()[name]
    s, o = pair(prob)
    u = randomfield(prob, seed=9, halo=o.exchange_halo_u)
    o.project(u, zeros(prob))
    ug = [a.copy(order="F") for a in u]
    po, pg = zeros(prob), zeros(prob)
    qo, qg = vzeros(prob), vzeros(prob)
    do, dg = (zeros(prob), zeros(prob)) if prob.implicit else (po, pg)
    for _ in range(3):
        o.step(u, po, qo, do, 0.004)
        dns.step_(ug, pg, qg, dg, s, 0.004)
    for c in range(3):
        assert relmax(ug[c], u[c]) < 1e-11, c
    assert relmax(pg, po) < 1e-11
    assert dns.maxdiv(ug, s) < 1e-12


@pytest.mark.parametrize("pgrad", ["", "scalar"])
@pytest.mark.parametrize("name", ["periodic", "channel", "channel_cn"])
def test_pgrad_kernels_step_parity(name, pgrad, monkeypatch):
    """The vector (16-byte) and one-cell pressure-gradient kernels inside the
    step, including the CFL bound their partial maxima feed."""
    if pgrad:
        monkeypatch.setenv("DNS_PGRAD", pgrad)
    prob = problems()[name]
    s, o = pair(prob)
    u = randomfield(prob, seed=19, halo=o.exchange_halo_u)
    o.project(u, zeros(prob))
    ug = [a.copy(order="F") for a in u]
    po, pg = zeros(prob), zeros(prob)
    qo, qg = vzeros(prob), vzeros(prob)
    do, dg = (zeros(prob), zeros(prob)) if prob.implicit else (po, pg)
    for _ in range(2):
        o.step(u, po, qo, do, 0.004)
        dns.step_(ug, pg, qg, dg, s, 0.004)
    for c in range(3):
        assert relmax(ug[c], u[c]) < 1e-11, c
    assert relmax(pg, po) < 1e-11
    assert math.isclose(dns.cfl_timestep(ug, s, 0.5), o.cfl_timestep(ug, 0.5), rel_tol=1e-14)


@pytest.mark.parametrize("n", [(32, 128, 8), (16, 256, 4), (16, 24, 4)])
@pytest.mark.parametrize("variant", [{}, {"DNS_CN_NL": "2"}, {"DNS_CN": "tile"}, {"DNS_CN": "sequential"}])
def test_cn_variants_step_parity(n, variant, monkeypatch):
    """Every Crank–Nicolson y-diffusion kernel (equal segments with two or one
    x-lines per thread, general segments, one thread per line) against the
    oracle, on line lengths that
    split into equal register segments (128, 256 rows) and one that does not."""
    for key, val in variant.items():
        monkeypatch.setenv(key, val)
    prob = channelproblem(n=n, visc=0.02, bodyforce=(1.0, 0.0, 0.0), ydiffusion="implicit")
    s, o = pair(prob)
    u = randomfield(prob, seed=21, halo=o.exchange_halo_u)
    o.project(u, zeros(prob))
    ug = [a.copy(order="F") for a in u]
    po, pg = zeros(prob), zeros(prob)
    qo, qg = vzeros(prob), vzeros(prob)
    do, dg = zeros(prob), zeros(prob)
    for _ in range(2):
        o.step(u, po, qo, do, 0.004)
        dns.step_(ug, pg, qg, dg, s, 0.004)
    for c in range(3):
        assert relmax(ug[c], u[c]) < 1e-11, c
    assert relmax(pg, po) < 1e-11


@pytest.mark.parametrize("graph", ["", "0"])
@pytest.mark.parametrize("name", ["periodic", "channel", "channel_cn"])
def test_fixed_dt_graph_replay_parity(name, graph, monkeypatch):
    """Steps with a repeated Δt replay captured HIP graphs (one per u / u′
    buffer parity); a new Δt recaptures.  Same trajectory as the oracle with
    replay on and off (DNS_FV_GRAPH=0)."""
    if graph:
        monkeypatch.setenv("DNS_FV_GRAPH", graph)
    prob = problems()[name]
    s, o = pair(prob)
    u = randomfield(prob, seed=25, halo=o.exchange_halo_u)
    o.project(u, zeros(prob))
    ug = [a.copy(order="F") for a in u]
    po, pg = zeros(prob), zeros(prob)
    qo, qg = vzeros(prob), vzeros(prob)
    do, dg = (zeros(prob), zeros(prob)) if prob.implicit else (po, pg)
    for dt in (0.004, 0.004, 0.004, 0.003, 0.003, 0.004, 0.004, 0.004):
        o.step(u, po, qo, do, dt)
    s.upload_u(ug)
    s.upload("p", pg)
    for c, nm in enumerate(("qx", "qy", "qz")):
        s.upload(nm, qg[c])
    if prob.implicit:
        s.upload("delta", dg)
    L = dns._lib.lib()
    for dt in (0.004, 0.004, 0.004, 0.003, 0.003, 0.004, 0.004, 0.004):
        dns._lib.check(L.dns_fv_step(s.handle, dt))
    # CFL from the partial maxima the last (replayed) pressure-gradient pass left
    out = np.zeros(1)
    dns._lib.check(L.dns_fv_cfl_timestep(s.handle, 0.5, dns._lib.dptr(out)))
    s.download_u(ug)
    s.download("p", pg)
    for c in range(3):
        assert relmax(ug[c], u[c]) < 1e-11, c
    assert relmax(pg, po) < 1e-11
    assert math.isclose(float(out[0]), o.cfl_timestep(u, 0.5), rel_tol=1e-13)


@pytest.mark.parametrize("name", ["periodic", "channel", "channel_cn"])
def test_plain_step_parity(name):
    """The reference-structured kernel sequence (the measurement baseline)
    against the FP64 oracle, as the fused step."""
    prob = problems()[name]
    o = OracleFV(prob)
    plainprob = dataclasses.replace(prob, precision="plain")
    s = Setup(plainprob, ROCmBackend())
    u = randomfield(prob, seed=9, halo=o.exchange_halo_u)
    o.project(u, zeros(prob))
    ug = [a.copy(order="F") for a in u]
    po, pg = zeros(prob), zeros(prob)
    qo, qg = vzeros(prob), vzeros(prob)
    do, dg = (zeros(prob), zeros(prob)) if prob.implicit else (po, pg)
    for _ in range(3):
        o.step(u, po, qo, do, 0.004)
        dns.step_(ug, pg, qg, dg, s, 0.004)
    for c in range(3):
        assert relmax(ug[c], u[c]) < 1e-11, c
    assert relmax(pg, po) < 1e-11
    assert dns.maxdiv(ug, s) < 1e-12


@pytest.mark.parametrize("bcy,yd", [("periodic", "explicit"), ("wall", "explicit"), ("wall", "implicit")])
def test_solve_with_statistics_parity(bcy, yd):
    """runcase of test/mpitests.jl:158-172: trajectories and statistics
    agree with the oracle to 1e-11."""
    prob = caseproblem(bcy, visc=0.02, bodyforce=(0.5, 0.0, 0.0), ydiffusion=yd)
    s, o = pair(prob)
    fx = lambda x, y, z: math.sin(x) * math.cos(y) * math.cos(z) + 0.2 * math.cos(2 * z)  # noqa: E731
    fy = lambda x, y, z: (1 - y ** 2) * math.sin(2 * x)  # noqa: E731
    fz = lambda x, y, z: math.cos(x) * math.sin(z) + 0.1 * math.sin(2 * y)  # noqa: E731
    u = vzeros(prob)
    fill_velocity(prob, u, fx, fy, fz)
    o.exchange_halo_u(u)
    ug = [a.copy(order="F") for a in u]
    if bcy == "wall":
        ny, nf = prob.n[1], prob.n[1] - 1
        acc = np.zeros(4 * ny + 4 * nf)
        procs = [lambda uu, p, t, n: (n % 2 == 0) and o.channelstats_sample(uu, acc)]
        cs = dns.channelstats(s, nupdate=2)
        gp = {"cs": cs.sample}
    else:
        nb = math.ceil(math.sqrt(sum((v // 2) ** 2 for v in prob.n))) + 1
        E = np.zeros(nb)
        procs = [lambda uu, p, t, n: (n % 2 == 0) and o.spectrum_sample(uu, E)]
        sp = dns.spectrumstats(s, nupdate=2)
        gp = {"sp": sp.sample}
    o.solve(u, (0.0, 0.05), dt=0.01, processors=procs)
    dns.solve_(u=ug, setup=s, tlims=(0.0, 0.05), dt=0.01, processors=gp)
    for c in range(3):
        assert np.abs(ug[c][1:-1, 1:-1, 1:-1] - u[c][1:-1, 1:-1, 1:-1]).max() < 1e-11
    assert dns.maxdiv(ug, s) < 1e-12
    if bcy == "wall":
        assert np.abs(cs.acc - acc).max() < 1e-11 * max(1.0, np.abs(acc).max())
    else:
        assert np.abs(sp.E - E).max() < 1e-11 * max(1.0, np.abs(E).max())


def test_cfl_and_kinetic_energy_parity():
    prob = channelproblem(visc=0.05, bodyforce=(1.0, 0.0, 0.0))
    s, o = pair(prob)
    u = randomfield(prob, seed=11, halo=o.exchange_halo_u)
    assert math.isclose(dns.cfl_timestep(u, s, 0.5), o.cfl_timestep(u, 0.5), rel_tol=1e-15)
    assert math.isclose(dns.kinetic_energy(u, s), o.kinetic_energy(u), rel_tol=1e-13)


def test_skew_symmetry_on_device():
    """⟨u, C(u)⟩ = 0 for divergence-free u (test/serialtests.jl:52-58)."""
    prob = channelproblem(n=(32, 24, 16), visc=0.0)
    s, o = pair(prob)
    u = randomfield(prob, seed=12, halo=o.exchange_halo_u)
    dns.project_(u, zeros(prob), s)
    q = vzeros(prob)
    dns.momentum_(q, u, 0.0, 1.0, s)
    d, n2 = weighted_dot(prob, o.griddata(), u, q)
    assert abs(d) < 1e-12 * n2


def test_taylor_green_64_parity():
    """The 64³ periodic Taylor–Green box (FP64, fixed Δt = 1e-3, Fourier y):
    20 steps against the oracle."""
    prob = FVProblem(n=(64, 64, 64), visc=1 / 1600)
    s, o = pair(prob)
    u = vzeros(prob)
    fill_velocity(prob, u, lambda x, y, z: math.sin(x) * math.cos(y) * math.cos(z),
                  lambda x, y, z: -math.cos(x) * math.sin(y) * math.cos(z))
    o.exchange_halo_u(u)
    ug = [a.copy(order="F") for a in u]
    o.solve(u, (0.0, 0.02), dt=1e-3)
    dns.solve_(u=ug, setup=s, tlims=(0.0, 0.02), dt=1e-3)
    for c in range(3):
        assert relmax(ug[c], u[c]) < 1e-11


def test_taylor_green_64_full_config_parity():
    """The 64³ periodic Taylor–Green configuration in full: ν = 1/1600, fixed
    Δt = 1e-3, 200 Williamson RK3 steps (steps 2–200 replay a captured HIP
    graph), against the oracle: velocity and kinetic energy."""
    prob = FVProblem(n=(64, 64, 64), visc=1 / 1600)
    s, o = pair(prob)
    u = vzeros(prob)
    fill_velocity(prob, u, lambda x, y, z: math.sin(x) * math.cos(y) * math.cos(z),
                  lambda x, y, z: -math.cos(x) * math.sin(y) * math.cos(z))
    o.exchange_halo_u(u)
    ug = [a.copy(order="F") for a in u]
    o.solve(u, (0.0, 0.2), dt=1e-3)
    dns.solve_(u=ug, setup=s, tlims=(0.0, 0.2), dt=1e-3)
    for c in range(3):
        assert relmax(ug[c], u[c]) < 1e-11
    assert math.isclose(dns.kinetic_energy(ug, s), o.kinetic_energy(u), rel_tol=1e-12)


@pytest.mark.parametrize("name,prec", [("channel", "fp64"), ("periodic", "fp64"), ("channel", "mixed"),
                                       ("channel_cn", "fp64")])
def test_advance_cfl_matches_host_loop(name, prec):
    """dns_fv_advance_cfl (device Δt and replayed graphs in the explicit form)
    reproduces the host loop of cfl_timestep + step! bit for bit: same Δt
    sequence, same fields, and the tend clamp of solve!."""
    prob = dataclasses.replace(problems()[name], precision=prec)
    o = OracleFV(problems()[name])
    u0 = randomfield(prob, seed=27, halo=o.exchange_halo_u)
    o.project(u0, zeros(prob))
    L = dns._lib.lib()
    nsteps, cfl = 7, 0.5

    def fresh():
        s = Setup(prob, ROCmBackend())
        s.upload_u([a.copy(order="F") for a in u0])
        for nm in ("p", "qx", "qy", "qz"):
            s.upload(nm, zeros(prob))
        if prob.implicit:
            s.upload("delta", zeros(prob))
        return s

    # host loop with the tend clamp on the last step
    sh = fresh()
    out = np.zeros(1)
    t, dts_host = 0.0, []
    for n in range(nsteps):
        dns._lib.check(L.dns_fv_cfl_timestep(sh.handle, cfl, dns._lib.dptr(out)))
        dts_host.append(float(out[0]))
        t += dts_host[-1]
    tend = sum(dts_host[:-1]) + 0.5 * dts_host[-1]
    sh = fresh()
    t, dts_host = 0.0, []
    for n in range(nsteps):
        dns._lib.check(L.dns_fv_cfl_timestep(sh.handle, cfl, dns._lib.dptr(out)))
        dt = min(float(out[0]), tend - t)
        dns._lib.check(L.dns_fv_step(sh.handle, dt))
        dts_host.append(dt)
        t += dt
    uh = vzeros(prob)
    sh.download_u(uh)
    # device loop
    sd = fresh()
    tdev = np.zeros(1)
    dts = np.zeros(nsteps)
    nd = ct.c_int32(-1)
    dns._lib.check(L.dns_fv_advance_cfl(sd.handle, cfl, nsteps, tend, dns._lib.dptr(tdev), dns._lib.dptr(dts),
                                        ct.byref(nd)))
    ud = vzeros(prob)
    sd.download_u(ud)
    assert nd.value == nsteps
    assert np.array_equal(dts, np.array(dts_host))
    assert tdev[0] == t
    for c in range(3):
        assert np.array_equal(ud[c], uh[c]), c


@pytest.mark.parametrize("name,prec", [("channel", "fp64"), ("channel_cn", "fp64"), ("channel", "mixed")])
def test_advance_cfl_stops_at_tend(name, prec):
    """More steps requested than solve! takes (src/timestep.jl:102): the loop
    stops at t >= tend − 1e-10(|tend| + 1), reports the steps taken, and the
    state (u and p) is the host loop's bit for bit; a later call with a
    larger tend continues from there."""
    prob = dataclasses.replace(problems()[name], precision=prec)
    o = OracleFV(problems()[name])
    u0 = randomfield(prob, seed=31, halo=o.exchange_halo_u)
    o.project(u0, zeros(prob))
    L = dns._lib.lib()
    cfl = 0.5

    def fresh():
        s = Setup(prob, ROCmBackend())
        s.upload_u([a.copy(order="F") for a in u0])
        for nm in ("p", "qx", "qy", "qz"):
            s.upload(nm, zeros(prob))
        if prob.implicit:
            s.upload("delta", zeros(prob))
        return s

    out = np.zeros(1)
    sh = fresh()
    dns._lib.check(L.dns_fv_cfl_timestep(sh.handle, cfl, dns._lib.dptr(out)))
    tend = 3.5 * float(out[0])
    t, dts_host = 0.0, []
    while t < tend - 1e-10 * (abs(tend) + 1):
        dns._lib.check(L.dns_fv_cfl_timestep(sh.handle, cfl, dns._lib.dptr(out)))
        dt = min(float(out[0]), tend - t)
        dns._lib.check(L.dns_fv_step(sh.handle, dt))
        dts_host.append(dt)
        t += dt
    uh, ph = vzeros(prob), zeros(prob)
    sh.download_u(uh)
    sh.download("p", ph)
    sd = fresh()
    nmax = len(dts_host) + 6
    tdev = np.zeros(1)
    dts = np.full(nmax, -1.0)
    nd = ct.c_int32(-1)
    dns._lib.check(L.dns_fv_advance_cfl(sd.handle, cfl, nmax, tend, dns._lib.dptr(tdev), dns._lib.dptr(dts),
                                        ct.byref(nd)))
    ud, pd = vzeros(prob), zeros(prob)
    sd.download_u(ud)
    sd.download("p", pd)
    assert nd.value == len(dts_host)
    assert np.array_equal(dts[:nd.value], np.array(dts_host))
    assert np.all(dts[nd.value:] == -1.0)  # entries past the steps taken are untouched
    assert tdev[0] == t
    for c in range(3):
        assert np.array_equal(ud[c], uh[c]), c
    assert np.array_equal(pd, ph)
    # a call that starts at or past tend takes no step
    t2, d2 = dns.advance_cfl_(sd, cfl, 3, tend, t=float(tdev[0]))
    assert len(d2) == 0 and t2 == tdev[0]
    # and the loop continues to a later end time like the host loop
    tend2 = tend + 2.5 * dts_host[0]
    t3, d3 = dns.advance_cfl_(sd, cfl, 10, tend2, t=t2)
    while t < tend2 - 1e-10 * (abs(tend2) + 1):
        dns._lib.check(L.dns_fv_cfl_timestep(sh.handle, cfl, dns._lib.dptr(out)))
        dt = min(float(out[0]), tend2 - t)
        dns._lib.check(L.dns_fv_step(sh.handle, dt))
        t += dt
    sd.download_u(ud)
    sh.download_u(uh)
    assert t3 == t and len(d3) >= 3
    for c in range(3):
        assert np.array_equal(ud[c], uh[c]), c


def test_step_rejects_nonpositive_dt_with_cn():
    """Δt = 0 would divide the lagged-pressure update by zero (src/timestep.jl:51)."""
    prob = problems()["channel_cn"]
    s = Setup(prob, ROCmBackend())
    L = dns._lib.lib()
    assert L.dns_fv_step(s.handle, 0.0) == -1  # DNS_EINVAL
    assert L.dns_fv_step(s.handle, -1e-3) != 0
    assert L.dns_fv_step(s.handle, float("nan")) != 0


def test_advance_cfl_after_spectrum_sample():
    """A spectrum sample between two device CFL loops (it sizes its own bin
    buffers) leaves the captured CFL graphs valid: same Δt and fields as the
    host loop, bit for bit."""
    prob = FVProblem(n=(64, 16, 64), visc=0.01)
    o = OracleFV(prob)
    u0 = randomfield(prob, seed=41, halo=o.exchange_halo_u)
    o.project(u0, zeros(prob))
    L = dns._lib.lib()

    def fresh():
        s = Setup(prob, ROCmBackend())
        s.upload_u([a.copy(order="F") for a in u0])
        for nm in ("p", "qx", "qy", "qz"):
            s.upload(nm, zeros(prob))
        return s

    sd, sh = fresh(), fresh()
    tA, dA = dns.advance_cfl_(sd, 0.5, 4)
    E = np.zeros(4096)
    dns._lib.check(L.dns_fv_spectrum_sample(sd.handle, dns._lib.dptr(E), 4096))
    tB, dB = dns.advance_cfl_(sd, 0.5, 4, t=tA)
    out = np.zeros(1)
    dts = []
    for _ in range(8):
        dns._lib.check(L.dns_fv_cfl_timestep(sh.handle, 0.5, dns._lib.dptr(out)))
        dns._lib.check(L.dns_fv_step(sh.handle, float(out[0])))
        dts.append(float(out[0]))
    assert np.array_equal(np.concatenate([dA, dB]), np.array(dts))
    ud, uh = vzeros(prob), vzeros(prob)
    sd.download_u(ud)
    sh.download_u(uh)
    for c in range(3):
        assert np.array_equal(ud[c], uh[c]), c
    assert E.sum() > 0


def test_channel_cfl_trajectory_parity():
    """Channel with walls, tanh stretching and CFL-adaptive Δt: 10 steps."""
    prob = channelproblem(n=(32, 24, 16), lims=((0, 4 * math.pi), (-1, 1), (0, 4 * math.pi / 3)),
                          visc=1 / 180, bodyforce=(1.0, 0.0, 0.0))
    s, o = pair(prob)
    u = randomfield(prob, seed=13)
    fill_velocity(prob, u, lambda x, y, z: 1 - y ** 2)
    o.exchange_halo_u(u)
    ug = [a.copy(order="F") for a in u]
    st = o.solve(u, (0.0, 0.3), cfl=0.5)
    gst = dns.solve_(u=ug, setup=s, tlims=(0.0, 0.3), cfl=0.5)
    assert gst.n == st["n"]
    err = max(relmax(ug[c], u[c]) for c in range(3))
    print(f"channel CFL trajectory: rel max err {err:.3e}", flush=True)
    assert err < 1e-11


def test_advance_cfl_helper_matches_oracle():
    """advance_cfl_ (Δt from cfl_timestep computed on the device, clamped to
    tend: the loop of src/timestep.jl:92-111 after the initial projection)
    against the oracle's cfl_timestep + step! loop: same Δt sequence, the
    half-length last step, the end time and the fields."""
    prob = channelproblem(n=(32, 24, 16), lims=((0, 4 * math.pi), (-1, 1), (0, 4 * math.pi / 3)),
                          visc=1 / 180, bodyforce=(1.0, 0.0, 0.0))
    s, o = pair(prob)
    u = randomfield(prob, seed=13)
    fill_velocity(prob, u, lambda x, y, z: 1 - y ** 2)
    o.exchange_halo_u(u)
    o.project(u, zeros(prob))
    ug = [a.copy(order="F") for a in u]
    p, q = o.scalarfield(), o.vectorfield()
    dt0 = o.cfl_timestep(u, 0.5)
    t, tend, dts = 0.0, 4.5 * dt0, []
    while t < tend - 1e-10 * (abs(tend) + 1):
        dt = min(o.cfl_timestep(u, 0.5), tend - t)
        o.step(u, p, q, p, dt)
        dts.append(dt)
        t += dt
    assert len(dts) >= 4 and dts[-1] < 0.9 * dts[-2]  # the last step is clamped
    s.upload_u(ug)
    for nm in ("p", "qx", "qy", "qz"):
        s.upload(nm, zeros(prob))
    tg, dtg = dns.advance_cfl_(s, 0.5, len(dts), tend)
    ud = vzeros(prob)
    s.download_u(ud)
    np.testing.assert_allclose(dtg, dts, rtol=1e-12, atol=0)
    assert math.isclose(tg, t, rel_tol=1e-12)
    err = max(relmax(ud[c], u[c]) for c in range(3))
    print(f"advance_cfl_ vs oracle: rel max err {err:.3e}", flush=True)
    assert err < 1e-11


@pytest.mark.parametrize("n", [(32, 24, 16), (32, 128, 8)])
def test_mixed_precision_error_is_small(n):
    prob = channelproblem(n=n, visc=0.02, bodyforce=(1.0, 0.0, 0.0), ydiffusion="implicit")
    prob32 = channelproblem(n=n, visc=0.02, bodyforce=(1.0, 0.0, 0.0), ydiffusion="implicit",
                            precision="mixed")
    s64, o = pair(prob)
    s32 = Setup(prob32, ROCmBackend())
    u = randomfield(prob, seed=14)
    fill_velocity(prob, u, lambda x, y, z: 1 - y ** 2)
    o.exchange_halo_u(u)
    u64 = [a.copy(order="F") for a in u]
    u32 = [a.copy(order="F") for a in u]
    dns.solve_(u=u64, setup=s64, tlims=(0.0, 0.05), dt=0.005)
    dns.solve_(u=u32, setup=s32, tlims=(0.0, 0.05), dt=0.005)
    num = math.sqrt(sum(float(((u32[c] - u64[c])[1:-1, 1:-1, 1:-1] ** 2).sum()) for c in range(3)))
    den = math.sqrt(sum(float((u64[c][1:-1, 1:-1, 1:-1] ** 2).sum()) for c in range(3)))
    assert num / den < 1e-5
    assert dns.maxdiv(u32, s64) < 1e-5


@pytest.mark.parametrize("n", [(16, 12, 8), (64, 64, 64)])
def test_mixed_periodic_box_matches_fp64(n):
    """The mixed form of a periodic box (FP32 storage and transforms, the
    FP32 periodic-y Poisson tiles k_fourier_y<float, 16, ·>) against the FP64
    form on the same field: relative L2 velocity difference after ten steps
    (below 1e-5) and the divergence at the FP32 storage level (a few FP32
    ulps of max|u| per grid spacing)."""
    prob = FVProblem(n=n, visc=0.01, bodyforce=(0.3, 0.1, -0.2))
    prob32 = dataclasses.replace(prob, precision="mixed")
    s64, s32 = Setup(prob, ROCmBackend()), Setup(prob32, ROCmBackend())
    u = randomfield(prob, seed=21)
    OracleFV(prob).exchange_halo_u(u)
    u64 = [a.copy(order="F") for a in u]
    u32 = [a.copy(order="F") for a in u]
    dns.solve_(u=u64, setup=s64, tlims=(0.0, 0.05), dt=0.005)
    dns.solve_(u=u32, setup=s32, tlims=(0.0, 0.05), dt=0.005)
    num = math.sqrt(sum(float(((u32[c] - u64[c])[1:-1, 1:-1, 1:-1] ** 2).sum()) for c in range(3)))
    den = math.sqrt(sum(float((u64[c][1:-1, 1:-1, 1:-1] ** 2).sum()) for c in range(3)))
    div = dns.maxdiv(u32, s64)
    umax = max(float(np.abs(u64[c][1:-1, 1:-1, 1:-1]).max()) for c in range(3))
    h = min(2 * math.pi / m for m in n)
    print(f"mixed periodic {n}: rel L2 {num / den:.3e}, maxdiv {div:.3e}", flush=True)
    assert num / den < 1e-5
    assert div < 50 * float(np.finfo(np.float32).eps) * umax / h


def test_errors_raise():
    with pytest.raises(dns.DNSError):
        dns.setup(n=(8, 8, 8), bc=("periodic", "periodic", "periodic"), ydiffusion="implicit")
    with pytest.raises(dns.DNSError):
        dns.setup(n=(14, 8, 8))  # nx/2 = 7: unsupported radix


def test_channelprofiles_laminar_and_moments():
    """channelprofiles (src/stats.jl:65-98): the laminar profile U = 1 − y²
    comes back exactly with zero fluctuations and the reference's lengths
    (test/serialtests.jl:127-139); for a random field every profile equals
    the plane averages and central moments computed with numpy from the
    downloaded field (u interpolated to the y faces for uv, src/stats.jl:52-54)."""
    prob = channelproblem(n=(16, 12, 8), visc=0.02)
    s = Setup(prob, ROCmBackend())
    u = vzeros(prob)
    fill_velocity(prob, u, lambda x, y, z: 1 - y ** 2)
    cs = dns.channelstats(s)
    s.upload_u(u)
    L = dns._lib.lib()
    for _ in range(3):
        dns._lib.check(L.dns_fv_channelstats_sample(s.handle, dns._lib.dptr(cs.acc)))
        cs.nsamples += 1
    pr = dns.channelprofiles(cs, s)
    ny = prob.n[1]
    assert len(pr.U) == len(pr.W) == len(pr.uu) == len(pr.ww) == ny
    assert len(pr.V) == len(pr.vv) == len(pr.uv) == len(pr.yf) == ny - 1
    assert np.abs(pr.U - (1 - pr.y ** 2)).max() < 1e-13
    for a in (pr.V, pr.W, pr.uu, pr.vv, pr.ww, pr.uv):
        assert np.abs(a).max() < 1e-13
    assert pr.nsamples == 3
    # random field: numpy plane statistics of the same field
    ur = randomfield(prob, seed=61)
    OracleFV(prob).exchange_halo_u(ur)
    cs2 = dns.channelstats(s)
    s.upload_u(ur)
    dns._lib.check(L.dns_fv_channelstats_sample(s.handle, dns._lib.dptr(cs2.acc)))
    cs2.nsamples = 1
    pr2 = dns.channelprofiles(cs2, s)
    ux, uy, uz = (a[1:-1, 1:-1, 1:-1] for a in ur)
    U, W = ux.mean(axis=(0, 2)), uz.mean(axis=(0, 2))
    V = uy[:, :ny - 1, :].mean(axis=(0, 2))
    uf = (ur[0][1:-1, 1:ny, 1:-1] + ur[0][1:-1, 2:ny + 1, 1:-1]) / 2      # u at (xf, yf, zc)
    vx = (ur[1][1:-1, 1:ny, 1:-1] + ur[1][2:, 1:ny, 1:-1]) / 2             # v at (xf, yf, zc)
    Uf = uf.mean(axis=(0, 2))
    np.testing.assert_allclose(pr2.U, U, rtol=0, atol=1e-14)
    np.testing.assert_allclose(pr2.W, W, rtol=0, atol=1e-14)
    np.testing.assert_allclose(pr2.V, V, rtol=0, atol=1e-14)
    np.testing.assert_allclose(pr2.uu, (ux ** 2).mean(axis=(0, 2)) - U ** 2, rtol=0, atol=1e-13)
    np.testing.assert_allclose(pr2.ww, (uz ** 2).mean(axis=(0, 2)) - W ** 2, rtol=0, atol=1e-13)
    np.testing.assert_allclose(pr2.vv, (uy[:, :ny - 1, :] ** 2).mean(axis=(0, 2)) - V ** 2, rtol=0, atol=1e-13)
    np.testing.assert_allclose(pr2.uv, (uf * vx).mean(axis=(0, 2)) - Uf * V, rtol=0, atol=1e-13)
