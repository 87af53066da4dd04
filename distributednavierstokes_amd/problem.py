"""Problem descriptions shared by the device backend and by host-side helpers.

Pure Python/numpy: validates the keyword arguments exactly as the
reference's ``setup`` does (src/setup.jl:34-113) and evaluates everything
that cannot cross a C ABI -- the ``stretch`` closure is evaluated here into
the global y faces ``yf[0..ny]`` (src/setup.jl:155-156), which the C library
then turns into the ghost-extended metrics of ``griddata``
(src/setup.jl:148-193).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

__all__ = ["FVProblem", "SpectralProblem", "tanhstretch", "DNSError"]


class DNSError(RuntimeError):
    """Error raised for invalid problems and failed library calls."""


def tanhstretch(beta):
    """tanh y-face clustering towards both walls (test/serialtests.jl:4)."""
    tb = math.tanh(beta)
    return lambda t: (math.tanh(beta * (2 * t - 1)) / tb + 1) / 2


def _identity(t):
    return t


@dataclass
class FVProblem:
    """Keyword arguments of ``setup`` after validation (src/setup.jl:34-62)."""

    n: tuple
    lims: tuple = ((0.0, 2 * math.pi),) * 3
    bc: tuple = ("periodic", "periodic", "periodic")
    stretch: object = _identity
    visc: float = 1e-3
    bodyforce: tuple = (0.0, 0.0, 0.0)
    ydiffusion: str = "explicit"
    order: int = 2
    precision: str = "fp64"
    yfaces: np.ndarray = field(init=False)

    def __post_init__(self):
        self.n = tuple(int(v) for v in self.n)
        self.lims = tuple((float(a), float(b)) for a, b in self.lims)
        self.bc = tuple(str(b).lstrip(":") for b in self.bc)
        self.bodyforce = tuple(float(f) for f in self.bodyforce)
        self.visc = float(self.visc)
        self.ydiffusion = str(self.ydiffusion).lstrip(":")
        if not (self.bc[0] == self.bc[2] == "periodic"):
            raise DNSError("x and z must be periodic")
        if self.bc[1] not in ("periodic", "wall"):
            raise DNSError("bc[2] must be :periodic or :wall")
        if self.ydiffusion not in ("explicit", "implicit"):
            raise DNSError("ydiffusion must be :explicit or :implicit")
        if self.order != 2:
            raise DNSError("only order 2 is implemented so far")
        if self.precision not in ("fp64", "mixed", "plain"):
            raise DNSError("precision must be fp64, mixed or plain")
        if min(self.n) < 2:
            raise DNSError("local block smaller than ghost width; use fewer ranks")
        if self.ydiffusion == "implicit" and self.bc[1] != "wall":
            raise DNSError("implicit y-diffusion requires bc[2] == :wall")
        ny = self.n[1]
        y0, y1 = self.lims[1]
        ly = y1 - y0
        # yface(j) = lims[2][1] + Ly * stretch(j / ny)   (src/setup.jl:155)
        self.yfaces = np.array([y0 + ly * self.stretch(j / ny) for j in range(ny + 1)], dtype=np.float64)
        if self.bc[1] == "periodic":
            dyi = np.diff(self.yfaces)
            if not (dyi.max() - dyi.min() < 1e-12 * dyi.max()):
                raise DNSError("periodic y requires a uniform grid (stretch = identity)")

    @property
    def wall(self) -> bool:
        return self.bc[1] == "wall"

    @property
    def implicit(self) -> bool:
        return self.ydiffusion == "implicit"

    @property
    def ghosted_shape(self):
        return tuple(v + 2 for v in self.n)

    def coordinates(self):
        """Face/centre coordinates of the (single-rank) block (src/setup.jl:183-188)."""
        (nx, ny, nz), (lx, _, lz) = self.n, self.lims
        Lx, Lz = lx[1] - lx[0], lz[1] - lz[0]
        yf = self.yfaces
        xf = np.array([lx[0] + i * Lx / nx for i in range(1, nx + 1)])
        xc = np.array([lx[0] + (i - 0.5) * Lx / nx for i in range(1, nx + 1)])
        zf = np.array([lz[0] + k * Lz / nz for k in range(1, nz + 1)])
        zc = np.array([lz[0] + (k - 0.5) * Lz / nz for k in range(1, nz + 1)])
        yff = yf[1:]                      # faces 1..ny (face j tops cell j)
        yc = (yf[:-1] + yf[1:]) / 2       # centres 1..ny
        return dict(xf=xf, xc=xc, yf=yff, yc=yc, zf=zf, zc=zc)

    def active_ranges(self, c):
        """0-based ghosted index ranges of the active DOFs of component c
        (0 = x, 1 = y, 2 = z, 3 = p), src/setup.jl:132-138."""
        nx, ny, nz = self.n
        jhi = ny - 1 if (c == 1 and self.wall) else ny
        return (range(1, nx + 1), range(1, jhi + 1), range(1, nz + 1))


@dataclass
class SpectralProblem:
    """Keyword arguments of ``spectral_setup`` after validation
    (src/spectral/setup.jl:33-48)."""

    n: int
    l: float = 2 * math.pi
    visc: float = 1e-3
    kcut: int | None = None
    precision: str = "fp64"

    def __post_init__(self):
        self.n = int(self.n)
        if self.kcut is None:
            self.kcut = self.n // 3
        self.kcut = int(self.kcut)
        self.l = float(self.l)
        self.visc = float(self.visc)
        if self.n % 2 != 0:
            raise DNSError("n must be even")
        if not 1 <= self.kcut:
            raise DNSError("kcut must be at least 1")
        if not 3 * self.kcut <= self.n:
            raise DNSError("aliasing: 3kcut ≤ n is required (2/3 rule)")
        if self.precision not in ("fp64", "mixed"):
            raise DNSError("precision must be fp64 or mixed")

    @property
    def m(self) -> int:
        return 2 * self.kcut + 1

    @property
    def K(self) -> int:
        return self.kcut + 1

    @property
    def state_shape(self):
        return (self.K, self.m, self.m, 3)
