"""Shared helpers for the finite-volume tests (problem builders mirroring
test/serialtests.jl and test/mpitests.jl of the reference)."""
import math

import numpy as np

from distributednavierstokes_amd.problem import FVProblem, tanhstretch

TWO_PI = 2 * math.pi


def channelproblem(n=(16, 12, 8), **kw):
    """channelsetup (test/serialtests.jl:6-13)."""
    kw.setdefault("lims", ((0.0, TWO_PI), (-1.0, 1.0), (0.0, math.pi)))
    kw.setdefault("bc", ("periodic", "wall", "periodic"))
    kw.setdefault("stretch", tanhstretch(1.5))
    return FVProblem(n=n, **kw)


def caseproblem(bcy, **kw):
    """casesetup (test/mpitests.jl:22-28)."""
    wall = bcy == "wall"
    return FVProblem(n=(12, 10, 8), lims=((0.0, TWO_PI), (-1.0, 1.0) if wall else (0.0, TWO_PI), (0.0, math.pi)),
                     bc=("periodic", bcy, "periodic"), stretch=tanhstretch(1.3) if wall else (lambda t: t), **kw)


def zeros(prob):
    return np.zeros(prob.ghosted_shape, dtype=np.float64, order="F")


def vzeros(prob):
    return [zeros(prob) for _ in range(3)]


def randomfield(prob, seed=42, halo=None):
    """Random values on all active DOFs (test/serialtests.jl:16-23); the
    caller's `halo(u)` fills the ghosts."""
    rng = np.random.default_rng(seed)
    u = vzeros(prob)
    for c in range(3):
        r = prob.active_ranges(c)
        sl = tuple(slice(rr.start, rr.stop) for rr in r)
        u[c][sl] = rng.standard_normal(tuple(len(rr) for rr in r))
    if halo is not None:
        halo(u)
    return u


def fill_velocity(prob, u, fx=None, fy=None, fz=None):
    """velocityfield! on the host (src/fields.jl:11-35), without halos."""
    co = prob.coordinates()
    coords = ((co["xf"], co["yc"], co["zc"]), (co["xc"], co["yf"], co["zc"]), (co["xc"], co["yc"], co["zf"]))
    for c, f in enumerate((fx, fy, fz)):
        if f is None:
            continue
        r = prob.active_ranges(c)
        X, Y, Z = coords[c]
        for k in r[2]:
            for j in r[1]:
                for i in r[0]:
                    u[c][i, j, k] = f(X[i - 1], Y[j - 1], Z[k - 1])
    return u


def weighted_dot(prob, grid, u, q):
    """Volume-weighted ⟨u, q⟩ and ⟨u, u⟩ over active DOFs (test/serialtests.jl:26-41)."""
    dx = (prob.lims[0][1] - prob.lims[0][0]) / prob.n[0]
    dz = (prob.lims[2][1] - prob.lims[2][0]) / prob.n[2]
    d = n2 = 0.0
    for c in range(3):
        r = prob.active_ranges(c)
        D2 = grid["dyu"] if c == 1 else grid["dy"]
        sl = tuple(slice(rr.start, rr.stop) for rr in r)
        vol = dx * dz * D2[sl[1]][None, :, None]
        d += float((vol * u[c][sl] * q[c][sl]).sum())
        n2 += float((vol * u[c][sl] ** 2).sum())
    return d, n2


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


def relmax(a, b):
    """max|a - b| / max|b| over interiors."""
    ia, ib = interior(np.asarray(a)), interior(np.asarray(b))
    return float(np.abs(ia - ib).max() / max(np.abs(ib).max(), 1e-300))
