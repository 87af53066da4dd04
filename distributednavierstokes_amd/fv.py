"""Finite-volume API on the MI355X library: the reference's ``setup`` /
``vectorfield`` / ``project!`` / ``step!`` / ``solve!`` / ``channelstats`` /
``spectrumstats`` surface (src/DistributedNavierStokes.jl:39-48) over the
C ABI of libdnsrocm.

Julia's mutating ``name!`` functions are spelled ``name_`` here.  Host
fields are the reference's ghosted column-major arrays (numpy, ``order="F"``,
shape ``(nx+2, ny+2, nz+2)``); inside ``solve_`` the state lives on the GPU
and processors receive a device-backed state that downloads lazily.
"""
from __future__ import annotations

import ctypes as ct
import math
from types import SimpleNamespace

import numpy as np

from . import _lib
from ._lib import FIELD, LAYOUT_GHOSTED, check, dptr, lib
from .problem import DNSError, FVProblem

__all__ = [
    "ROCmBackend", "setup", "scalarfield", "vectorfield", "velocityfield_", "exchange_halo_",
    "project_", "momentum_", "step_", "cfl_timestep", "solve_", "maxdiv", "kinetic_energy",
    "channelstats", "channelprofiles", "spectrumstats", "energyspectrum", "rk3",
]

# Williamson low-storage RK3 (src/timestep.jl:10-14)
rk3 = SimpleNamespace(A=(0.0, -5 / 9, -153 / 128), B=(1 / 3, 15 / 16, 8 / 15), alpha=(1 / 3, 5 / 12, 1 / 4))


class ROCmBackend:
    """Marker backend selecting the gfx950 library (the role of the
    reference's ``backend`` keyword, src/setup.jl:45)."""

    def __repr__(self):
        return "ROCmBackend()"


class VectorField:
    """``(; x, y, z)`` of ghosted host arrays (src/fields.jl:8-9)."""

    __slots__ = ("x", "y", "z")

    def __init__(self, x, y, z):
        self.x, self.y, self.z = x, y, z

    def __getitem__(self, c):
        return (self.x, self.y, self.z)[c] if isinstance(c, int) else getattr(self, c)

    def __iter__(self):
        return iter((self.x, self.y, self.z))


class Setup:
    """The problem NamedTuple of ``setup`` plus the device handle."""

    def __init__(self, prob: FVProblem, backend):
        self.prob = prob
        self.backend = backend
        self.n = prob.n
        self.lims = prob.lims
        self.bc = prob.bc
        self.visc = prob.visc
        self.bodyforce = prob.bodyforce
        self.imexy = prob.implicit
        self.T = np.float32 if prob.precision == "mixed" else np.float64
        self.w = 1
        self.rk = rk3
        self.grid = SimpleNamespace(**prob.coordinates())
        nx, ny, nz = prob.n
        Ly = prob.lims[1][1] - prob.lims[1][0]
        self.grid.dx = (prob.lims[0][1] - prob.lims[0][0]) / nx
        self.grid.dz = (prob.lims[2][1] - prob.lims[2][0]) / nz
        yfg = prob.yfaces
        self.grid.yfg = yfg.copy()
        self.grid.ycg = (yfg[:-1] + yfg[1:]) / 2
        del Ly
        self.ranges = SimpleNamespace(x=prob.active_ranges(0), y=prob.active_ranges(1),
                                      z=prob.active_ranges(2), p=prob.active_ranges(3))
        L = lib()
        par = _lib.FVParams()
        par.n[:] = list(prob.n)
        par.lims[:] = [v for ab in prob.lims for v in ab]
        par.ywall = int(prob.wall)
        self._yf = np.ascontiguousarray(prob.yfaces, dtype=np.float64)
        par.yfaces = dptr(self._yf)
        par.visc = prob.visc
        par.bodyforce[:] = list(prob.bodyforce)
        par.ydiffusion = int(prob.implicit)
        par.precision = {"mixed": _lib.DNS_MIXED, "plain": _lib.DNS_PLAIN_FP64}.get(prob.precision, _lib.DNS_FP64)
        h = _lib.ct.c_void_p()
        check(L.dns_fv_create(_lib.ct.byref(par), _lib.ct.byref(h)))
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib._LIB is not None:
            _lib._LIB.dns_fv_destroy(h)
            self.handle = None

    def describe(self):
        """Kernel instances and launch shapes of this handle's step
        (``dns_fv_describe``) as a dict of ``key -> text``."""
        buf = ct.create_string_buffer(512)
        check(lib().dns_fv_describe(self.handle, buf, 512))
        return dict(part.strip().split("=", 1) for part in buf.value.decode().split(";"))

    # host <-> device ------------------------------------------------------
    def upload(self, name, a):
        a = np.asarray(a, dtype=np.float64)
        if a.shape != self.prob.ghosted_shape:
            raise DNSError(f"field {name}: expected ghosted shape {self.prob.ghosted_shape}, got {a.shape}")
        a = np.asfortranarray(a)
        check(lib().dns_fv_set_field(self.handle, FIELD[name], dptr(a), LAYOUT_GHOSTED))

    def download(self, name, out=None):
        buf = np.zeros(self.prob.ghosted_shape, dtype=np.float64, order="F")
        check(lib().dns_fv_get_field(self.handle, FIELD[name], dptr(buf), LAYOUT_GHOSTED))
        if out is None:
            return buf
        out[...] = buf
        return out

    def upload_u(self, u):
        if isinstance(u, DeviceVelocity) and u.setup is self:
            return
        for c, name in enumerate(("ux", "uy", "uz")):
            self.upload(name, u[c])

    def download_u(self, u):
        for c, name in enumerate(("ux", "uy", "uz")):
            self.download(name, u[c])


class DeviceVelocity:
    """Velocity resident on the GPU (inside ``solve_``); components
    download on access."""

    def __init__(self, setup):
        self.setup = setup

    def __getitem__(self, c):
        c = "xyz".index(c) if isinstance(c, str) else c
        return self.setup.download(("ux", "uy", "uz")[c])

    @property
    def x(self):
        return self[0]

    @property
    def y(self):
        return self[1]

    @property
    def z(self):
        return self[2]

    def __iter__(self):
        return iter((self[0], self[1], self[2]))


def setup(*, n, lims=((0.0, 2 * math.pi),) * 3, bc=("periodic", "periodic", "periodic"), stretch=lambda t: t,
          visc=1e-3, bodyforce=(0.0, 0.0, 0.0), procgrid=None, ydiffusion="explicit", mpibuf="auto", order=2,
          backend=None, T=np.float64, comm=None):
    """``setup(; n, kwargs...)`` -- src/setup.jl:34-113, on one GPU.

    ``T = np.float64`` is the FP64 form; ``T = np.float32`` selects the mixed
    form (FP32 storage, stencils and FFTs; FP64 Thomas solves and
    reductions).  ``procgrid`` must be ``None`` or ``(1, 1)``: this backend
    runs one device without ranks."""
    if procgrid not in (None, (1, 1)):
        raise DNSError("this backend runs on one GPU: procgrid must be (1, 1)")
    if mpibuf not in ("auto", "device", "host", ":auto", ":device", ":host"):
        raise DNSError("mpibuf must be :auto, :device, or :host")
    precision = "mixed" if np.dtype(T) == np.float32 else "fp64"
    prob = FVProblem(n=n, lims=lims, bc=bc, stretch=stretch, visc=visc, bodyforce=bodyforce,
                     ydiffusion=ydiffusion, order=order, precision=precision)
    return Setup(prob, backend if backend is not None else ROCmBackend())


def scalarfield(s):
    """Ghosted scalar field of zeros (src/fields.jl:4-5)."""
    return np.zeros(s.prob.ghosted_shape, dtype=np.float64, order="F")


def vectorfield(s):
    """Ghosted velocity field of zeros (src/fields.jl:8-9)."""
    return VectorField(scalarfield(s), scalarfield(s), scalarfield(s))


def exchange_halo_(u, s):
    """Fills the ghost layers of a host velocity field as exchange_halo!
    does on one rank (src/halo.jl:81-133), via a device round trip."""
    s.upload_u(u)
    s.download_u(u)
    return u


def velocityfield_(u, s, x=None, y=None, z=None):
    """``velocityfield!`` (src/fields.jl:28-35): evaluates the closures at
    each component's staggered points on the active DOFs, then fills halos."""
    g = s.grid
    coords = ((g.xf, g.yc, g.zc), (g.xc, g.yf, g.zc), (g.xc, g.yc, g.zf))
    for c, f in enumerate((x, y, z)):
        if f is None:
            continue
        r = s.prob.active_ranges(c)
        cx, cy, cz = coords[c]
        phi = u[c]
        for k in r[2]:
            for j in r[1]:
                for i in r[0]:
                    phi[i, j, k] = f(cx[i - 1], cy[j - 1], cz[k - 1])
    return exchange_halo_(u, s)


def project_(u, psi, s):
    """``project!(u, ψ, s)`` -- src/timestep.jl:25-31."""
    s.upload_u(u)
    check(lib().dns_fv_project(s.handle, FIELD["p"]))
    if not isinstance(u, DeviceVelocity):
        s.download_u(u)
    if psi is not None:
        s.download("p", psi)
    return u


def momentum_(q, u, A, dt, s):
    """``momentum!(q, u, A, Δt, s)`` -- src/operators.jl:155-178."""
    s.upload_u(u)
    for c, name in enumerate(("qx", "qy", "qz")):
        s.upload(name, q[c])
    check(lib().dns_fv_momentum(s.handle, float(A), float(dt)))
    for c, name in enumerate(("qx", "qy", "qz")):
        s.download(name, q[c])
    return q


def step_(u, p, q, delta, s, dt):
    """``step!(u, p, q, δ, s, Δt)`` -- src/timestep.jl:41-63."""
    s.upload_u(u)
    s.upload("p", p)
    for c, name in enumerate(("qx", "qy", "qz")):
        s.upload(name, q[c])
    if s.imexy and delta is not None and delta is not p:
        s.upload("delta", delta)
    check(lib().dns_fv_step(s.handle, float(dt)))
    s.download_u(u)
    s.download("p", p)
    for c, name in enumerate(("qx", "qy", "qz")):
        s.download(name, q[c])
    if s.imexy and delta is not None and delta is not p:
        s.download("delta", delta)
    return u


def cfl_timestep(u, s, cfl):
    """``cfl_timestep(u, s; cfl)`` -- src/timestep.jl:69-80."""
    s.upload_u(u)
    out = np.zeros(1)
    check(lib().dns_fv_cfl_timestep(s.handle, float(cfl), dptr(out)))
    return float(out[0])


def advance_cfl_(s, cfl, nsteps, tend=math.inf, t=0.0):
    """At most ``nsteps`` CFL steps of ``solve!`` (src/timestep.jl:92-111) on
    the device-resident state: while t < tend − 1e-10(|tend| + 1),
    Δt = min(cfl_timestep, tend − t), computed on the device in the explicit
    form.  Returns (t, Δt of each step taken)."""
    tt = np.array([float(t)])
    dts = np.zeros(max(1, int(nsteps)))
    nd = ct.c_int32(0)
    check(lib().dns_fv_advance_cfl(s.handle, float(cfl), int(nsteps), float(tend), dptr(tt), dptr(dts),
                                   ct.byref(nd)))
    return float(tt[0]), dts[:nd.value].copy()


def maxdiv(u, s):
    """``maxdiv(u, s)`` -- src/fields.jl:38-42 (FP64 residual)."""
    s.upload_u(u)
    out = np.zeros(1)
    check(lib().dns_fv_maxdiv(s.handle, dptr(out)))
    return float(out[0])


def kinetic_energy(u, s):
    """``kinetic_energy(u, s)`` -- src/fields.jl:45-60 (compensated FP64)."""
    s.upload_u(u)
    out = np.zeros(1)
    check(lib().dns_fv_kinetic_energy(s.handle, dptr(out)))
    return float(out[0])


class State:
    """``state = (; u, p, t, n)`` passed to processors (src/timestep.jl:100)."""

    def __init__(self, s, t, n):
        self.setup, self.t, self.n = s, t, n
        self.u = DeviceVelocity(s)

    @property
    def p(self):
        return self.setup.download("p")


def solve_(*, u, setup, tlims, dt=None, cfl=0.5, processors=None):
    """``solve!(; u, setup, tlims, Δt, cfl, processors)`` -- src/timestep.jl:92-111.

    The state stays on the GPU for the whole loop; ``u`` is updated in place
    at the end and the final state is returned."""
    s = setup
    processors = dict(processors or {})
    s.upload_u(u)
    zero = scalarfield(s)
    s.upload("p", zero)
    for name in ("qx", "qy", "qz"):
        s.upload(name, zero)
    if s.imexy:
        s.upload("delta", zero)
    L = lib()
    check(L.dns_fv_project(s.handle, FIELD["delta"] if s.imexy else FIELD["p"]))
    t, tend = float(tlims[0]), float(tlims[1])
    n = 0
    out = np.zeros(1)
    for proc in processors.values():
        proc(State(s, t, n), s)
    while t < tend - 1e-10 * (abs(tend) + 1):
        if dt is None:
            check(L.dns_fv_cfl_timestep(s.handle, float(cfl), dptr(out)))
            dtn = float(out[0])
        else:
            dtn = float(dt)
        dtn = min(dtn, tend - t)
        check(L.dns_fv_step(s.handle, dtn))
        t += dtn
        n += 1
        for proc in processors.values():
            proc(State(s, t, n), s)
    check(L.dns_fv_sync(s.handle))
    if not isinstance(u, DeviceVelocity):
        s.download_u(u)
    st = State(s, t, n)
    return st


# ---------------------------------------------------------------- statistics


class ChannelStats(SimpleNamespace):
    pass


def channelstats(s, nupdate=1):
    """``channelstats(setup; nupdate)`` -- src/stats.jl:15-37."""
    ny = s.n[1]
    nf = ny - 1 if s.prob.wall else ny
    acc = np.zeros(4 * ny + 4 * nf)
    cs = ChannelStats(nsamples=0, acc=acc, ny=ny, nf=nf)

    def sample(state, s2):
        if state.n % nupdate == 0:
            s2.upload_u(state.u)
            check(lib().dns_fv_channelstats_sample(s2.handle, dptr(cs.acc)))
            cs.nsamples += 1
        return None

    cs.sample = sample
    return cs


def channelprofiles(cs, s):
    """``channelprofiles(cs, setup)`` -- src/stats.jl:65-98."""
    ny, nf = cs.ny, cs.nf
    norm = max(cs.nsamples, 1) * s.n[0] * s.n[2]
    a = cs.acc
    Uc, Wc, uuc, wwc = (a[i * ny:(i + 1) * ny] / norm for i in range(4))
    Vf, vvf, Uf, uvf = (a[4 * ny + i * nf:4 * ny + (i + 1) * nf] / norm for i in range(4))
    yfg = s.grid.yfg
    return SimpleNamespace(
        y=s.grid.ycg.copy(),
        yf=yfg[1:ny] if s.prob.wall else yfg[1:ny + 1],
        U=Uc, V=Vf, W=Wc, uu=uuc - Uc ** 2, vv=vvf - Vf ** 2, ww=wwc - Wc ** 2, uv=uvf - Uf * Vf,
        nsamples=cs.nsamples,
    )


def spectrumstats(s, nupdate=1):
    """``spectrumstats(setup; nupdate)`` -- src/stats.jl:108-118."""
    if s.bc != ("periodic", "periodic", "periodic"):
        raise DNSError("spectra require a fully periodic box")
    nbins = math.ceil(math.sqrt(sum((v // 2) ** 2 for v in s.n))) + 1
    sp = SimpleNamespace(nsamples=0, E=np.zeros(nbins))

    def sample(state, s2):
        if state.n % nupdate == 0:
            s2.upload_u(state.u)
            check(lib().dns_fv_spectrum_sample(s2.handle, dptr(sp.E), len(sp.E)))
            sp.nsamples += 1
        return None

    sp.sample = sample
    return sp


def energyspectrum(sp, s):
    """``energyspectrum(sp, setup)`` -- src/stats.jl:164-167."""
    E = sp.E / max(sp.nsamples, 1)
    return SimpleNamespace(kappa=np.arange(len(E)), E=E)
