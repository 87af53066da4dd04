"""Pseudo-spectral HIT API on the MI355X library: ``spectral_setup`` /
``spectral_solve!`` / ``shellforcing`` / ``spectral_randomfield!`` and the
diagnostics (src/DistributedNavierStokes.jl:39-48) over the C ABI of
libdnsrocm.

Host states are numpy complex arrays of shape ``(kcut+1, m, m, 3)``
(``order="F"``), exactly the reference's ``specvelocity``; physical fields
are real ``(n, n, n, 3)``.  Inside ``spectral_solve_`` the state stays on the
GPU and processors/forcing receive a device-backed state.
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np

from . import _lib
from ._lib import check, dptr, lib
from .problem import DNSError, SpectralProblem

__all__ = [
    "spectral_setup", "specvelocity", "spectral_velocityfield_", "taylorgreen_", "spectral_randomfield_",
    "spec_to_phys_", "phys_to_spec_", "spectral_project_", "spectral_rhs_", "spectral_step_", "spectral_solve_",
    "shell_energies", "shellforcing", "spectral_energy", "spectral_dissipation", "spectral_maxdiv",
    "spectral_stats", "spectral_spectrum", "wray3",
]

# Wray RK3 with integrating factor (src/spectral/step.jl:16-20)
wray3 = SimpleNamespace(a=(8 / 15, 5 / 12, 3 / 4), b=(1 / 4, 0.0, 0.0), dc=(8 / 15, 2 / 15, 1 / 3))


class SpecSetup:
    def __init__(self, prob: SpectralProblem, backend=None):
        self.prob = prob
        self.n, self.l, self.visc, self.kcut, self.m = prob.n, prob.l, prob.visc, prob.kcut, prob.m
        self.K = prob.K
        self.T = np.float32 if prob.precision == "mixed" else np.float64
        self.backend = backend
        par = _lib.SpecParams()
        par.n, par.kcut, par.l, par.visc = prob.n, prob.kcut, prob.l, prob.visc
        par.precision = _lib.DNS_MIXED if prob.precision == "mixed" else _lib.DNS_FP64
        h = _lib.ct.c_void_p()
        check(lib().dns_spec_create(_lib.ct.byref(par), _lib.ct.byref(h)))
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib._LIB is not None:
            _lib._LIB.dns_spec_destroy(h)
            self.handle = None

    @property
    def state_shape(self):
        return self.prob.state_shape

    def upload(self, uh, which=0):
        if isinstance(uh, DeviceState) and uh.setup is self and which == 0:
            return
        a = np.asfortranarray(uh, dtype=np.complex128)
        if a.shape != self.state_shape:
            raise DNSError(f"state shape must be {self.state_shape}, got {a.shape}")
        check(lib().dns_spec_set_state(self.handle, which, dptr(a)))

    def download(self, out=None, which=0):
        buf = np.zeros(self.state_shape, dtype=np.complex128, order="F")
        check(lib().dns_spec_get_state(self.handle, which, dptr(buf)))
        if out is None:
            return buf
        out[...] = buf
        return out

    def upload_phys(self, w):
        a = np.asfortranarray(w, dtype=np.float64)
        if a.shape != (self.n,) * 3 + (3,):
            raise DNSError("physical field shape must be (n, n, n, 3)")
        check(lib().dns_spec_set_physical(self.handle, dptr(a)))

    def download_phys(self, out=None):
        buf = np.zeros((self.n,) * 3 + (3,), dtype=np.float64, order="F")
        check(lib().dns_spec_get_physical(self.handle, dptr(buf)))
        if out is None:
            return buf
        out[...] = buf
        return out

    def scalar(self, fn):
        out = np.zeros(1)
        check(fn(self.handle, dptr(out)))
        return float(out[0])


class DeviceState:
    """The spectral state resident on the GPU (inside ``spectral_solve_``)."""

    def __init__(self, setup):
        self.setup = setup

    def __array__(self, dtype=None, copy=None):
        a = self.setup.download()
        return a if dtype is None else a.astype(dtype)


def spectral_setup(*, n, l=2 * math.pi, visc=1e-3, kcut=None, procgrid=None, mpibuf="auto", backend=None,
                   T=np.float64, comm=None):
    """``spectral_setup(; n, kwargs...)`` -- src/spectral/setup.jl:33-158, on
    one GPU.  ``T = np.float32`` selects FP32 state and FFTs with FP64
    reductions."""
    if procgrid not in (None, (1, 1)):
        raise DNSError("this backend runs on one GPU: procgrid must be (1, 1)")
    if mpibuf not in ("auto", "device", "host", "nccl", ":auto", ":device", ":host", ":nccl"):
        raise DNSError("mpibuf must be :auto, :device, :host, or :nccl")
    prob = SpectralProblem(n=n, l=l, visc=visc, kcut=kcut,
                           precision="mixed" if np.dtype(T) == np.float32 else "fp64")
    return SpecSetup(prob, backend)


def specvelocity(s):
    """Zero spectral state (src/spectral/setup.jl:161-162)."""
    return np.zeros(s.state_shape, dtype=np.complex128, order="F")


def _finish(uh, s):
    if not isinstance(uh, DeviceState):
        s.download(uh)
    return uh


def spec_to_phys_(v, uh, s):
    """``spec_to_phys!(v, uh, s)`` -- src/spectral/fft.jl:200-214."""
    s.upload(uh)
    check(lib().dns_spec_spec_to_phys(s.handle))
    return s.download_phys(v)


def phys_to_spec_(uh, w, s):
    """``phys_to_spec!(uh, w, s)`` -- src/spectral/fft.jl:221-232."""
    s.upload_phys(w)
    check(lib().dns_spec_phys_to_spec(s.handle))
    return _finish(uh, s)


def spectral_project_(uh, s):
    """``spectral_project!(uh, s)`` -- src/spectral/operators.jl:25-29."""
    s.upload(uh)
    check(lib().dns_spec_project(s.handle))
    return _finish(uh, s)


def spectral_rhs_(du, uh, s):
    """``spectral_rhs!(du, uh, s)`` -- src/spectral/operators.jl:60-79."""
    s.upload(uh)
    check(lib().dns_spec_rhs(s.handle))
    return s.download(du, which=2)


def spectral_step_(uh, s, dt, cache=None):
    """``spectral_step!(uh, s, Δt, cache)`` -- src/spectral/step.jl:68-92;
    returns vsummax of the first stage."""
    s.upload(uh)
    out = np.zeros(1)
    check(lib().dns_spec_step(s.handle, float(dt), dptr(out)))
    _finish(uh, s)
    return float(out[0])


def spectral_velocityfield_(uh, s, x, y, z, project=True):
    """``spectral_velocityfield!`` (src/spectral/ic.jl:13-25): closures of the
    cell-centre coordinates, forward transformed (dealiased)."""
    n, l = s.n, s.l
    c = l / n * (np.arange(1, n + 1) - 0.5)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    w = np.zeros((n, n, n, 3), order="F")
    for k, f in enumerate((x, y, z)):
        w[..., k] = np.broadcast_to(np.asarray(np.vectorize(f)(X, Y, Z), dtype=np.float64), (n, n, n))
    s.upload_phys(w)
    check(lib().dns_spec_phys_to_spec(s.handle))
    if project:
        check(lib().dns_spec_project(s.handle))
    return _finish(uh, s)


def taylorgreen_(uh, s, V0=1.0):
    """``taylorgreen!`` (src/spectral/ic.jl:90-96)."""
    L = s.l
    return spectral_velocityfield_(
        uh, s,
        x=lambda x, y, z: V0 * math.sin(2 * math.pi * x / L) * math.cos(2 * math.pi * y / L) * math.cos(2 * math.pi * z / L),
        y=lambda x, y, z: -V0 * math.cos(2 * math.pi * x / L) * math.sin(2 * math.pi * y / L) * math.cos(2 * math.pi * z / L),
        z=lambda x, y, z: 0.0)


def _shells(s, shells):
    shells = [int(v) for v in shells]
    s0, s1 = min(shells), max(shells)
    E = np.zeros(s1 - s0 + 1)
    check(lib().dns_spec_shell_energies(s.handle, s0, s1 - s0 + 1, dptr(E)))
    return np.array([E[v - s0] for v in shells])


def _rescale(s, factors: dict):
    s0, s1 = min(factors), max(factors)
    f = np.ones(s1 - s0 + 1)
    for sh, v in factors.items():
        f[sh - s0] = v
    check(lib().dns_spec_rescale_shells(s.handle, s0, s1 - s0 + 1, dptr(f)))


def shell_energies(uh, s, shells):
    """``shell_energies(uh, s, shells)`` -- src/spectral/forcing.jl:12-24."""
    s.upload(uh)
    return _shells(s, shells)


def shellforcing(uh, s, shells=(1, 2), eref=None):
    """``shellforcing(uh, s; shells, eref)`` -- src/spectral/forcing.jl:47-56."""
    shells = list(shells)
    eref = list(shell_energies(uh, s, shells) if eref is None else eref)

    def forcing(uh2, s2):
        s2.upload(uh2)
        e = _shells(s2, shells)
        _rescale(s2, {sh: (math.sqrt(e0 / e1) if e1 > 0 else 1.0) for sh, e0, e1 in zip(shells, eref, e)})
        _finish(uh2, s2)
        return None

    forcing.eref = eref
    return forcing


def spectral_randomfield_(uh, s, totalenergy=1.0, kpeak=4.0, seed=0, profile=None):
    """``spectral_randomfield!`` (src/spectral/ic.jl:59-84): counter-based
    noise, forward transform, projection, shell rescaling to the profile."""
    profile = profile or (lambda k: k ** 4 * math.exp(-2 * (k / kpeak) ** 2))
    L = lib()
    check(L.dns_spec_noise(s.handle, int(seed) & 0xFFFFFFFFFFFFFFFF))
    check(L.dns_spec_phys_to_spec(s.handle))
    check(L.dns_spec_project(s.handle))
    kdiag = int(math.floor(math.sqrt(3) * s.kcut))
    E = _shells(s, range(0, kdiag + 1))
    tot = sum(profile(k) for k in range(0, kdiag + 1))
    factors = {}
    for sh, e in zip(range(0, kdiag + 1), E):
        e0 = totalenergy * profile(sh) / tot
        factors[sh] = math.sqrt(e0 / e) if e > 0 else 0.0
    _rescale(s, factors)
    return _finish(uh, s)


def spectral_energy(uh, s):
    """``spectral_energy`` -- src/spectral/operators.jl:91-92."""
    s.upload(uh)
    return s.scalar(lib().dns_spec_energy)


def spectral_dissipation(uh, s):
    """``spectral_dissipation`` -- src/spectral/operators.jl:95-101."""
    s.upload(uh)
    return s.scalar(lib().dns_spec_dissipation)


def spectral_maxdiv(uh, s):
    """``spectral_maxdiv`` -- src/spectral/operators.jl:104-110."""
    s.upload(uh)
    return s.scalar(lib().dns_spec_maxdiv)


def spectral_stats(uh, s):
    """``spectral_stats`` -- src/spectral/stats.jl:11-26 (K41 quantities)."""
    visc, n, l = s.visc, s.n, s.l
    e = spectral_energy(uh, s)
    diss = spectral_dissipation(uh, s)
    uavg = math.sqrt(2 * e / 3)
    l_kol = (visc ** 3 / diss) ** 0.25
    l_tay = math.sqrt(15 * visc / diss) * uavg
    l_int = uavg ** 3 / diss
    return SimpleNamespace(e=e, uavg=uavg, diss=diss, l_int=l_int, l_tay=l_tay, l_kol=l_kol, t_int=l_int / uavg,
                           t_tay=l_tay / uavg, t_kol=math.sqrt(visc / diss), Re_int=l_int * uavg / visc,
                           Re_tay=l_tay * uavg / visc, kmax_eta=(math.pi * n / l) * l_kol)


def spectral_spectrum(uh, s):
    """``spectral_spectrum`` -- src/spectral/stats.jl:33-36."""
    E = shell_energies(uh, s, range(0, s.kcut + 1))
    return SimpleNamespace(kappa=np.arange(s.kcut + 1), E=E)


class SpecState:
    """``state = (; uh, t, n)`` passed to processors (src/spectral/step.jl:151,166)."""

    def __init__(self, s, t, n):
        self.setup, self.t, self.n = s, t, n
        self.uh = DeviceState(s)


def spectral_solve_(*, uh, setup, tlims, dt=None, cfl=0.85, forcing=None, tstops=(), nstart=0, processors=None,
                    closure=None):
    """``spectral_solve!`` -- src/spectral/step.jl:118-170.  The state stays
    on the GPU; a processor returning ``"stop"`` ends the solve after the
    current step.  ``uh`` is updated in place at the end.  ``closure`` (an
    ``SFSNet`` on this grid) adds its learned sub-filter stress to every
    stage (an LES step); None is the no-model solve."""
    s = setup
    L = lib()
    processors = dict(processors or {})
    s.upload(uh)
    check(L.dns_spec_project(s.handle))
    kmax = 2 * math.pi * s.kcut / s.l
    eps = float(np.finfo(s.T).eps)
    out = np.zeros(1)
    vmax = 0.0
    if dt is None:
        check(L.dns_spec_vsummax(s.handle, dptr(out)))
        vmax = float(out[0])
    t, tend = float(tlims[0]), float(tlims[1])
    tol = 1e-10 * (abs(tend) + 1)
    stops = sorted(x for x in tstops if t + tol < x < tend - tol)
    istop = 0
    n = nstart
    stop = False
    dev = DeviceState(s)

    def call_all(state):
        nonlocal stop
        for proc in processors.values():
            if proc(state, s) in ("stop", ":stop"):
                stop = True

    call_all(SpecState(s, t, n))
    while not stop and t < tend - tol:
        while istop < len(stops) and stops[istop] <= t + tol:
            istop += 1
        tnext = stops[istop] if istop < len(stops) else tend
        dtn = min(dt if dt is not None else cfl * math.sqrt(3) / (kmax * (vmax + eps)), tnext - t)
        if closure is None:
            check(L.dns_spec_step(s.handle, float(dtn), dptr(out) if dt is None else None))
        else:
            check(L.dns_les_closure_step(s.handle, closure.handle, float(dtn), dptr(out) if dt is None else None))
        if forcing is not None:
            forcing(dev, s)
        if dt is None:
            vmax = float(out[0])
        t += dtn
        n += 1
        call_all(SpecState(s, t, n))
    check(L.dns_spec_sync(s.handle))
    if not isinstance(uh, DeviceState):
        s.download(uh)
    return SpecState(s, t, n)
