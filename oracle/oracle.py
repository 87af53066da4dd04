"""ctypes wrapper of the FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py import this module; it checks the HIP library and is never part of
the product path.  Arrays are the reference's ghosted column-major fields
(numpy ``order="F"``, shape ``(nx+2, ny+2, nz+2)``).
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ct.POINTER(ct.c_double)


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ct.CDLL(path)
        L.ofv_create.restype = ct.c_int
        L.ofv_create.argtypes = [ct.POINTER(ct.c_int), _dp, ct.c_int, _dp, ct.c_double, _dp, ct.c_int,
                                 ct.POINTER(ct.c_void_p)]
        L.ofv_destroy.argtypes = [ct.c_void_p]
        for name in ("ofv_cfl_timestep", "ofv_maxdiv", "ofv_kinetic_energy"):
            getattr(L, name).restype = ct.c_double
        L.ofv_rbuf.restype = _dp
        L.oracle_set_threads.restype = ct.c_int
        L.oracle_set_threads.argtypes = [ct.c_int]
        L.ospec_splitmix64.restype = ct.c_uint64
        L.ospec_splitmix64.argtypes = [ct.c_uint64]
        _LIB = L
    return _LIB


def set_threads(n):
    """OpenMP threads of later oracle calls (n <= 0: unchanged); returns the
    count in effect."""
    return int(lib().oracle_set_threads(int(n)))


def _p(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1), "need float64 F-order"
    return a.ctypes.data_as(_dp)


class OracleFV:
    """The reference's FV single-rank CPU path on ghosted numpy arrays."""

    def __init__(self, prob):
        self.prob = prob
        L = lib()
        n = (ct.c_int * 3)(*prob.n)
        lims = np.array([v for ab in prob.lims for v in ab], dtype=np.float64)
        force = np.array(prob.bodyforce, dtype=np.float64)
        h = ct.c_void_p()
        rc = L.ofv_create(n, _p(lims), int(prob.wall), _p(prob.yfaces), prob.visc, _p(force),
                          int(prob.implicit), ct.byref(h))
        if rc != 0:
            raise RuntimeError(f"oracle create failed ({rc})")
        self.h = h
        self.L = L

    def __del__(self):
        try:
            self.L.ofv_destroy(self.h)
        except Exception:
            pass

    # fields --------------------------------------------------------------
    def scalarfield(self):
        return np.zeros(self.prob.ghosted_shape, dtype=np.float64, order="F")

    def vectorfield(self):
        return [self.scalarfield() for _ in range(3)]

    def griddata(self):
        ny = self.prob.n[1]
        dy, dyu = np.zeros(ny + 2), np.zeros(ny + 2)
        yf, yc = np.zeros(ny + 1), np.zeros(ny)
        dymin = ct.c_double()
        self.L.ofv_griddata(self.h, _p(dy), _p(dyu), ct.byref(dymin), _p(yf), _p(yc))
        return dict(dy=dy, dyu=dyu, dymin=dymin.value, yfg=yf, ycg=yc)

    def exchange_halo(self, phi, stag):
        self.L.ofv_exchange_halo(self.h, _p(phi), ct.c_int(stag))

    def exchange_halo_u(self, u):
        self.L.ofv_exchange_halo_u(self.h, *[_p(a) for a in u])

    def divergence(self, u):
        r = np.zeros(self.prob.n, dtype=np.float64, order="F")
        self.L.ofv_divergence(self.h, _p(r), *[_p(a) for a in u])
        return r

    def momentum(self, q, u, A, dt):
        self.L.ofv_momentum(self.h, *[_p(a) for a in q], *[_p(a) for a in u], ct.c_double(A), ct.c_double(dt))

    def pgrad(self, phi, p, c, s=1.0):
        self.L.ofv_pgrad(self.h, _p(phi), _p(p), ct.c_int(c), ct.c_double(s))

    def poisson_from(self, r, p):
        """poisson!(p) with the divergence buffer set to r (interior shape)."""
        buf = self.L.ofv_rbuf(self.h)
        N = int(np.prod(self.prob.n))
        np.ctypeslib.as_array(buf, shape=(N,))[:] = np.asarray(r, dtype=np.float64).ravel(order="F")
        self.L.ofv_poisson(self.h, _p(p))

    def project(self, u, psi):
        self.L.ofv_project(self.h, *[_p(a) for a in u], _p(psi))

    def step(self, u, p, q, delta, dt):
        self.L.ofv_step(self.h, *[_p(a) for a in u], _p(p), *[_p(a) for a in q], _p(delta), ct.c_double(dt))

    def cfl_timestep(self, u, cfl):
        return self.L.ofv_cfl_timestep(self.h, *[_p(a) for a in u], ct.c_double(cfl))

    def maxdiv(self, u):
        return self.L.ofv_maxdiv(self.h, *[_p(a) for a in u])

    def kinetic_energy(self, u):
        return self.L.ofv_kinetic_energy(self.h, *[_p(a) for a in u])

    def channelstats_sample(self, u, acc):
        self.L.ofv_channelstats_sample(self.h, *[_p(a) for a in u], _p(acc))

    def spectrum_sample(self, u, E):
        self.L.ofv_spectrum_sample(self.h, *[_p(a) for a in u], _p(E), ct.c_int(len(E)))

    # host loop: solve! (src/timestep.jl:92-111) ----------------------------
    def solve(self, u, tlims, dt=None, cfl=0.5, processors=(), p=None):
        p = self.scalarfield() if p is None else p
        q = self.vectorfield()
        delta = self.scalarfield() if self.prob.implicit else p
        self.exchange_halo_u(u)
        self.project(u, delta)
        t, tend = float(tlims[0]), float(tlims[1])
        n = 0
        dts = []
        for proc in processors:
            proc(u, p, t, n)
        while t < tend - 1e-10 * (abs(tend) + 1):
            dtn = min(dt if dt is not None else self.cfl_timestep(u, cfl), tend - t)
            self.step(u, p, q, delta, dtn)
            dts.append(dtn)
            t += dtn
            n += 1
            for proc in processors:
                proc(u, p, t, n)
        return dict(u=u, p=p, t=t, n=n, dts=dts)


def splitmix64(x: int) -> int:
    return int(lib().ospec_splitmix64(ct.c_uint64(x & 0xFFFFFFFFFFFFFFFF)))


# ----------------------------------------------------------------- spectral

_cp = ct.POINTER(ct.c_double)


def _pc(a):
    assert a.dtype == np.complex128 and a.flags.f_contiguous, "need complex128 F-order"
    return a.ctypes.data_as(_cp)


def _spec_sigs(L):
    L.ospec_create.restype = ct.c_void_p
    L.ospec_create.argtypes = [ct.c_int, ct.c_int, ct.c_double, ct.c_double]
    L.ospec_destroy.argtypes = [ct.c_void_p]
    for name in ("ospec_rhs", "ospec_step", "ospec_vsummax", "ospec_energy", "ospec_dissipation", "ospec_maxdiv"):
        getattr(L, name).restype = ct.c_double
    L.ospec_step.argtypes = [ct.c_void_p, _cp, ct.c_double, _cp, _cp]


class OracleSpectral:
    """The reference's single-rank pseudo-spectral CPU path (FP64).

    State arrays: complex (K, m, m, 3) F-order; physical: real (n, n, n, 3)."""

    def __init__(self, prob):
        self.prob = prob
        self.L = lib()
        _spec_sigs(self.L)
        self.h = ct.c_void_p(self.L.ospec_create(prob.n, prob.kcut, prob.l, prob.visc))
        self.n, self.kcut, self.K, self.m = prob.n, prob.kcut, prob.K, prob.m

    def __del__(self):
        try:
            self.L.ospec_destroy(self.h)
        except Exception:
            pass

    def specvelocity(self):
        return np.zeros((self.K, self.m, self.m, 3), dtype=np.complex128, order="F")

    def physfield(self):
        return np.zeros((self.n, self.n, self.n, 3), dtype=np.float64, order="F")

    def spec_to_phys(self, uh, v=None):
        v = self.physfield() if v is None else v
        self.L.ospec_spec_to_phys(self.h, _pc(uh), _p(v))
        return v

    def phys_to_spec(self, w, uh=None):
        uh = self.specvelocity() if uh is None else uh
        self.L.ospec_phys_to_spec(self.h, _p(np.asfortranarray(w)), _pc(uh))
        return uh

    def project(self, uh):
        self.L.ospec_project(self.h, _pc(uh))
        return uh

    def rhs(self, uh):
        du = self.specvelocity()
        vmax = self.L.ospec_rhs(self.h, _pc(du), _pc(uh))
        return du, vmax

    def step(self, uh, dt, cache=None):
        cache = cache or (self.specvelocity(), self.specvelocity())
        return self.L.ospec_step(self.h, _pc(uh), ct.c_double(dt), _pc(cache[0]), _pc(cache[1]))

    def shell_energies(self, uh, shells):
        shells = list(shells)
        s0, ns = min(shells), max(shells) - min(shells) + 1
        E = np.zeros(ns)
        self.L.ospec_shell_energies(self.h, _pc(uh), ct.c_int(s0), ct.c_int(ns), _p(E))
        return np.array([E[s - s0] for s in shells])

    def rescale_shell(self, uh, sh, factor):
        self.L.ospec_rescale_shell(self.h, _pc(uh), ct.c_int(sh), ct.c_double(factor))

    def energy(self, uh):
        return self.L.ospec_energy(self.h, _pc(uh))

    def dissipation(self, uh):
        return self.L.ospec_dissipation(self.h, _pc(uh))

    def maxdiv(self, uh):
        return self.L.ospec_maxdiv(self.h, _pc(uh))

    def noise(self, seed):
        v = self.physfield()
        self.L.ospec_noise(self.h, ct.c_uint64(seed), _p(v))
        return v

    # host logic of the reference ---------------------------------------
    def velocityfield(self, fx, fy, fz, project=True):
        """spectral_velocityfield! (src/spectral/ic.jl:13-25)."""
        n, l = self.n, self.prob.l
        x = l / n * (np.arange(1, n + 1) - 0.5)
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        w = self.physfield()
        for c, f in enumerate((fx, fy, fz)):
            w[..., c] = np.vectorize(f)(X, Y, Z)
        uh = self.phys_to_spec(w)
        if project:
            self.project(uh)
        return uh

    def randomfield(self, totalenergy=1.0, kpeak=4.0, seed=0, profile=None):
        """spectral_randomfield! (src/spectral/ic.jl:59-84)."""
        profile = profile or (lambda k: k ** 4 * np.exp(-2 * (k / kpeak) ** 2))
        uh = self.phys_to_spec(self.noise(seed))
        self.project(uh)
        kdiag = int(np.floor(np.sqrt(3) * self.kcut))
        E = self.shell_energies(uh, range(0, kdiag + 1))
        tot = sum(profile(k) for k in range(0, kdiag + 1))
        for sh, e in zip(range(0, kdiag + 1), E):
            e0 = totalenergy * profile(sh) / tot
            self.rescale_shell(uh, sh, np.sqrt(e0 / e) if e > 0 else 0.0)
        return uh

    def shellforcing(self, uh, shells=(1, 2), eref=None):
        """shellforcing (src/spectral/forcing.jl:47-56)."""
        shells = list(shells)
        eref = list(self.shell_energies(uh, shells) if eref is None else eref)

        def forcing(uh2):
            e = self.shell_energies(uh2, shells)
            for sh, e0, e1 in zip(shells, eref, e):
                self.rescale_shell(uh2, sh, np.sqrt(e0 / e1) if e1 > 0 else 1.0)

        return forcing

    def solve(self, uh, tlims, dt=None, cfl=0.85, forcing=None):
        """spectral_solve! host loop (src/spectral/step.jl:118-170), no tstops."""
        self.project(uh)
        kmax = 2 * np.pi * self.kcut / self.prob.l
        eps = 2.220446049250313e-16
        vmax = 0.0
        if dt is None:
            v = self.spec_to_phys(uh)
            vmax = self.L.ospec_vsummax(self.h, _p(v))
        t, tend = float(tlims[0]), float(tlims[1])
        tol = 1e-10 * (abs(tend) + 1)
        n = 0
        cache = (self.specvelocity(), self.specvelocity())
        dts = []
        while t < tend - tol:
            dtn = min(dt if dt is not None else cfl * np.sqrt(3) / (kmax * (vmax + eps)), tend - t)
            vloc = self.step(uh, dtn, cache)
            if forcing is not None:
                forcing(uh)
            if dt is None:
                vmax = vloc
            t += dtn
            n += 1
            dts.append(dtn)
        return dict(uh=uh, t=t, n=n, dts=dts)


def _spec_products(self, uh):
    """σ̂ (K, m, m, 6) of the state's velocity: xx, yy, zz, xy, yz, zx."""
    sig = np.zeros((self.K, self.m, self.m, 6), dtype=np.complex128, order="F")
    self.L.ospec_products(self.h, _pc(uh), _pc(sig))
    return sig


OracleSpectral.products = _spec_products


# ------------------------------------------------------- SFS data recipe
# Restatement of the rank-0 recipe of sfs_sample! (src/spectral/les.jl:313-382)
# on coarse rfft arrays (M/2+1, M, M) with wrapped negative frequencies, in
# numpy: lesfreq (les.jl:41), les_twothirds! (46-54), les_filterkernel!
# (69-86, Gaussian kernel), les_nonlinearity! (89-103), les_tracefree!
# (134-143), les_spectrum (119-131), les_filter_reynolds (106-116).

def lesfreq(M):
    q = np.arange(M)
    return np.where(q < (M + 1) // 2, q, q - M)


def les_cube(state, kcut, M):
    """The LES cube |k_i| <= M/2 of a compact DNS state (K, m, m, nc) as coarse
    rfft arrays (M/2+1, M, M, nc) (les_gather! + lescoarse, les.jl:149-160)."""
    m = 2 * kcut + 1
    f = lesfreq(M)
    jd = np.where(f >= 0, f, m + f)
    return state[: M // 2 + 1][:, jd][:, :, jd]


def les_twothirds(a, M):
    kc = M // 3
    k1 = np.arange(a.shape[0])[:, None, None]
    f = lesfreq(M)
    k2, k3 = f[None, :, None], f[None, None, :]
    keep = (k1 <= kc) & (np.abs(k2) <= kc) & (np.abs(k3) <= kc) & ~((k1 == 0) & (k2 == 0) & (k3 == 0))
    return a * keep[..., None] if a.ndim == 4 else a * keep


def les_gaussian(a, M, l, delta, protectshells):
    kf = 2 * np.pi / l
    f = lesfreq(M)
    k1 = kf * np.arange(a.shape[0])[:, None, None]
    k2, k3 = kf * f[None, :, None], kf * f[None, None, :]
    kk = k1 ** 2 + k2 ** 2 + k3 ** 2
    w = np.where(kk < (kf * (protectshells + 1)) ** 2, 1.0, np.exp(-delta ** 2 * kk / 24))
    return a * w[..., None] if a.ndim == 4 else a * w


def les_nonlinearity(ub, M):
    """σ2 = F[v_i v_j]/M³ from dealiased ub (M/2+1, M, M, 3)."""
    vs = [np.fft.irfftn(les_twothirds(ub[..., i], M).transpose(2, 1, 0), s=(M, M, M), axes=(0, 1, 2)) * M ** 3
          for i in range(3)]
    out = []
    for i, j in ((0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)):
        out.append(np.fft.rfftn(vs[i] * vs[j], axes=(0, 1, 2)).transpose(2, 1, 0) / M ** 3)
    return np.stack(out, -1)


def sfs_sample_oracle(uh, sig, kcut, M, l, delta_fac=2.0, protectshells=2):
    """ū and trace-free τ on the coarse rfft grid (Gaussian kernel)."""
    delta = delta_fac * l / M
    ub = les_cube(uh, kcut, M)
    sb = les_cube(sig, kcut, M)
    ubf = les_twothirds(les_gaussian(ub, M, l, delta, protectshells), M)
    tau = les_gaussian(sb, M, l, delta, protectshells)
    tau = les_twothirds(tau - les_nonlinearity(ubf, M), M)
    tr = (tau[..., 0] + tau[..., 1] + tau[..., 2]) / 3
    for c in range(3):
        tau[..., c] = tau[..., c] - tr
    return ubf, tau


def les_spectrum(ub, M):
    kc = M // 3
    f = lesfreq(M)
    s2 = np.arange(ub.shape[0])[:, None, None] ** 2 + f[None, :, None] ** 2 + f[None, None, :] ** 2
    sh = np.floor(np.sqrt(s2)).astype(int)
    # isqrt on integers
    sh = np.array([[[int(np.sqrt(v)) if int(np.sqrt(v)) ** 2 <= v else int(np.sqrt(v)) - 1 for v in row]
                    for row in plane] for plane in s2])
    w = np.where((np.arange(ub.shape[0]) > 0) & (np.arange(ub.shape[0]) < ub.shape[0] - 1), 2.0, 1.0)[:, None, None]
    u2 = (np.abs(ub) ** 2).sum(-1)
    E = np.zeros(kc)
    for s in range(1, kc + 1):
        E[s - 1] = (w * u2 * (sh == s)).sum() / 2
    return E
