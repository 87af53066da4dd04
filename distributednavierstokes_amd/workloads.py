"""The benchmark configurations as problem builders and initial conditions.

- Channel (the reference's channel example, examples/channel.jl:33-53, and
  README.md:175-183): 4π × 2 × 4π/3, walls in y with tanh stretching 1.5,
  ν = 1/180, body force (1, 0, 0) so that u_τ = 1 and Re_τ = 180; IC the
  laminar profile U = 1 − y² plus a 10 % uniform perturbation per component
  on the active degrees of freedom, 0.1·(2r − 1) with r the 53-bit uniform of
  splitmix64(seed + key), key = ((c·nz + k)·ny + j)·nx + i (0-based) -- the
  counter-based noise of src/spectral/ic.jl:31-47 keyed per staggered
  component, so every run and the oracle start from the same field.
- Taylor–Green box 64³ on [0, 2π]³, ν = 1/1600, u = sin x cos y cos z,
  v = −cos x sin y cos z, w = 0 at the staggered points (src/fields.jl:28-35).
"""
import math

import numpy as np

from .problem import FVProblem, tanhstretch


def splitmix64(x):
    """splitmix64 finaliser (src/spectral/ic.jl:31-36) on uint64 arrays."""
    x = (x + np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)).astype(np.uint64)
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)).astype(np.uint64)
    return x ^ (x >> np.uint64(31))


def channel_ic(prob, seed=0, amp=0.1):
    """Laminar 1 − y² plus amp·(2r − 1) on active DOFs (ghosted arrays)."""
    nx, ny, nz = prob.n
    co = prob.coordinates()
    u = [np.zeros(prob.ghosted_shape, order="F") for _ in range(3)]
    i = np.arange(nx, dtype=np.uint64)[:, None, None]
    j = np.arange(ny, dtype=np.uint64)[None, :, None]
    with np.errstate(over="ignore"):
        for c in range(3):
            for k0 in range(0, nz, 16):  # k-slabs bound the temporaries
                k = np.arange(k0, min(nz, k0 + 16), dtype=np.uint64)[None, None, :]
                key = ((np.uint64(c) * np.uint64(nz) + k) * np.uint64(ny) + j) * np.uint64(nx) + i
                r = ((splitmix64(np.uint64(seed) + key) >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
                u[c][1:-1, 1:-1, 1 + k0:1 + k0 + k.shape[2]] = amp * (2 * r - 1)
    u[0][1:-1, 1:-1, 1:-1] += (1 - co["yc"] ** 2)[None, :, None]
    if prob.wall:
        u[1][:, -2, :] = 0.0  # top wall face
    return u


def channel_problem(n, ydiffusion="explicit", precision="fp64"):
    """The channel of the benchmark configurations at n = (nx, ny, nz)."""
    return FVProblem(n=n, lims=((0.0, 4 * math.pi), (-1.0, 1.0), (0.0, 4 * math.pi / 3)),
                     bc=("periodic", "wall", "periodic"), stretch=tanhstretch(1.5), visc=1 / 180,
                     bodyforce=(1.0, 0.0, 0.0), ydiffusion=ydiffusion, precision=precision)


def tgv_problem(n=64):
    """Fully periodic Taylor–Green box n³ on [0, 2π]³, ν = 1/1600."""
    return FVProblem(n=(n, n, n), visc=1 / 1600)


def tgv_ic(prob):
    """Taylor–Green velocity at the staggered points (ghosted arrays)."""
    co = prob.coordinates()
    u = [np.zeros(prob.ghosted_shape, order="F") for _ in range(3)]
    u[0][1:-1, 1:-1, 1:-1] = (np.sin(co["xf"])[:, None, None] * np.cos(co["yc"])[None, :, None]
                              * np.cos(co["zc"])[None, None, :])
    u[1][1:-1, 1:-1, 1:-1] = (-np.cos(co["xc"])[:, None, None] * np.sin(co["yf"])[None, :, None]
                              * np.cos(co["zc"])[None, None, :])
    return u
