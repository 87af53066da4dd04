"""GPU parity at the kernel instances and line lengths the benchmark
configurations run:

- channel 256×192×256 explicit FP64 (C2): Thomas segments of 6 rows, the
  32-plane momentum march, 128-point half-length x lines, 256-point z lines;
- channel 512×384×512 with Crank–Nicolson y-diffusion (C3) in FP64 and in
  the mixed form: Thomas segments of 12 (FP64) / 24 (FP32 pencil) rows, CN
  segments of 24 rows, 256-point half-length x lines, 512-point z lines
  (the FP32 16-line z tile);
- HIT 512³, kcut 170 (C4): one IF-RK3 step of the production pipeline, and
  the same step at 128³ and 256³ (their lane-group pencil FFT lengths).

The FV cases are thin slabs that keep one production dimension each, so the
FP64 oracle finishes in seconds; `dns_fv_describe` shows that a slab runs the
same kernel instance as the full-size handle.  The spectral case uses a
band-limited field (|k_i| <= 2): over one RK3 step its stages stay within
|k_i| <= 16 with no aliasing at n = 48, so the 512³ step must equal the
oracle's 48³ step mode by mode.  Full-size FP64 runs of the channel
configurations (a few steps) close the file.

FP64 tolerance 1e-11 relative (trajectories), as everywhere; the mixed form
against the FP64 oracle within 2e-5 relative (FP32 storage and FFTs, FP64
Thomas solves: a few hundred FP32 ulps after two steps)."""
import math

import numpy as np
import pytest

import distributednavierstokes_amd as dns
from distributednavierstokes_amd.fv import ROCmBackend, Setup
from distributednavierstokes_amd.problem import SpectralProblem
from distributednavierstokes_amd.workloads import channel_ic, channel_problem
from fvhelpers import relmax, vzeros, zeros
from oracle.oracle import OracleFV, OracleSpectral

pytestmark = pytest.mark.gpu

MIXED_TOL = 2e-5


def _run_pair(n, ydiffusion, precision, nsteps=2):
    """nsteps CFL-sized steps of the GPU handle and the oracle from the
    channel IC (projected); returns (gpu u, oracle u, gpu p, oracle p, setup)."""
    prob = channel_problem(n, ydiffusion, "fp64")
    gprob = channel_problem(n, ydiffusion, precision)
    o = OracleFV(prob)
    s = Setup(gprob, ROCmBackend())
    u = channel_ic(prob)
    o.exchange_halo_u(u)
    o.project(u, zeros(prob))
    ug = [a.copy(order="F") for a in u]
    po, pg = zeros(prob), zeros(prob)
    qo, qg = vzeros(prob), vzeros(prob)
    implicit = ydiffusion == "implicit"
    do, dg = (zeros(prob), zeros(prob)) if implicit else (po, pg)
    dt = o.cfl_timestep(u, 0.5)
    for _ in range(nsteps):
        o.step(u, po, qo, do, dt)
        dns.step_(ug, pg, qg, dg, s, dt)
    return ug, u, pg, po, s


def _check(ug, u, pg, po, precision):
    errs = [relmax(ug[c], u[c]) for c in range(3)]
    perr = relmax(pg, po)
    print(f"rel max err u {max(errs):.3e} p {perr:.3e}", flush=True)
    if precision == "fp64":
        assert max(errs) < 1e-11, errs
        assert perr < 1e-10
    else:
        assert max(errs) < MIXED_TOL, errs


def _same_instances(a, full, keys):
    """The slab's description `a` names the same kernel instances as the
    full-size handle's for `keys`."""
    b = full.describe()
    for k in keys:
        assert a[k] == b[k], (k, a[k], b[k])


@pytest.fixture(scope="module")
def full_handles():
    """Full-size handles of the benchmark channels (fields stay zero)."""
    out = {"c2": Setup(channel_problem((256, 192, 256), "explicit", "fp64"), ROCmBackend())}
    for prec in ("fp64", "mixed"):
        out["c3_" + prec] = Setup(channel_problem((512, 384, 512), "implicit", prec), ROCmBackend())
    return out


def test_c2_y_solve_and_momentum_march(full_handles, monkeypatch):
    """(256, 192, 64) with the 32-plane march of the full grid: k_thomas_seg
    with 6-row segments, k_momentum_lz over 32 planes, 128-point x lines."""
    monkeypatch.setenv("DNS_MOM_ZC", "32")
    ug, u, pg, po, s = _run_pair((256, 192, 64), "explicit", "fp64")
    d = s.describe()
    monkeypatch.delenv("DNS_MOM_ZC")
    assert d["ysolve"].startswith("seg R=6")
    assert d["momentum"] == "lz zc=32"
    _same_instances(d, full_handles["c2"], ["ysolve", "xpass", "pgrad", "momentum"])
    _check(ug, u, pg, po, "fp64")


def test_c2_z_lines():
    """(256, 8, 256): the full grid's 256-point z lines and x lines."""
    ug, u, pg, po, s = _run_pair((256, 8, 256), "explicit", "fp64")
    full = Setup(channel_problem((256, 192, 256), "explicit", "fp64"), ROCmBackend())
    _same_instances(s.describe(), full, ["zpass", "xpass"])
    _check(ug, u, pg, po, "fp64")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c3_y_solves(precision, full_handles, monkeypatch):
    """(64, 384, 64) CN with the 32-plane march: the full grid's Thomas
    segments (12 rows FP64, 24 rows for the FP32 pencil) and CN segments of
    24 rows."""
    monkeypatch.setenv("DNS_MOM_ZC", "32")
    ug, u, pg, po, s = _run_pair((64, 384, 64), "implicit", precision)
    d = s.describe()
    monkeypatch.delenv("DNS_MOM_ZC")
    assert d["ysolve"].startswith("seg R=12" if precision == "fp64" else "seg R=24")
    # FP64: the three components per launch by default (k_cn_seg3, pressure in LDS)
    assert d["cn"].startswith("seg3 R=24" if precision == "fp64" else "seg R=24")
    _same_instances(d, full_handles["c3_" + precision], ["ysolve", "cn", "momentum"])
    _check(ug, u, pg, po, precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("fused", ["0", "1", "2"])
def test_c3_cn_launch_forms(precision, fused, monkeypatch):
    """The same CN segments one component per launch (k_cn_seg,
    DNS_CN_FUSED=0) and with the three velocity components in one launch
    (k_cn_seg3, DNS_CN_FUSED=1: pressure rows re-read per component; 2: staged
    once in LDS)."""
    monkeypatch.setenv("DNS_CN_FUSED", fused)
    ug, u, pg, po, s = _run_pair((64, 384, 64), "implicit", precision)
    d = s.describe()
    monkeypatch.delenv("DNS_CN_FUSED")
    if fused == "0":
        assert d["cn"].startswith("seg R=24")
    else:
        assert d["cn"].startswith("seg3 R=24") and d["cn"].endswith("p=lds" if fused == "2" else "p=l2")
    _check(ug, u, pg, po, precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c3_x_and_z_lines(precision, full_handles):
    """(512, 8, 512) CN: 256-point half-length x lines and 512-point z lines
    (in the mixed form the 16-line FP32 z tile)."""
    ug, u, pg, po, s = _run_pair((512, 8, 512), "implicit", precision)
    _same_instances(s.describe(), full_handles["c3_" + precision], ["zpass", "xpass"])
    _check(ug, u, pg, po, precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n", [128, 256, 512])
def test_hit_step_matches_band_limited_oracle(n, precision):
    """One IF-RK3 step (src/spectral/step.jl:68-92) at n = 128, 256, 512 with
    the 2/3 cut kcut = n/3 on a field with |k_i| <= 2, against the oracle's
    step of the same modes at n = 48, kcut = 16 (no truncation or aliasing
    differs between the two).  n = 512 is the production pipeline (the
    four-step 256-point x lines, the 512-point lane-group pencils); 128 and
    256 run the 128- and 256-point lane-group pencil FFTs."""
    visc, dt = 0.01, 0.02
    small = OracleSpectral(SpectralProblem(n=48, kcut=16, visc=visc))
    uh = small.velocityfield(lambda x, y, z: np.sin(x + 2 * y) * np.cos(z) + 0.4 * np.cos(2 * z - y),
                             lambda x, y, z: np.cos(2 * x) * np.sin(z - y) + 0.3 * np.sin(x),
                             lambda x, y, z: np.sin(y) * np.cos(2 * x + z) - 0.2 * np.cos(2 * y))
    assert np.abs(uh).max() > 0.1
    big = dns.spectral_setup(n=n, kcut=n // 3, visc=visc, T=np.float32 if precision == "mixed" else np.float64)
    K, m = big.K, big.m

    def to_big(a):
        out = np.zeros((K, m, m, 3), dtype=np.complex128, order="F")
        for k in range(small.m):
            kk = k if k <= small.kcut else k - small.m + m
            for j in range(small.m):
                jj = j if j <= small.kcut else j - small.m + m
                out[:small.K, jj, kk, :] = a[:, j, k, :]
        return out

    def from_big(a):
        out = small.specvelocity()
        for k in range(small.m):
            kk = k if k <= small.kcut else k - small.m + m
            for j in range(small.m):
                jj = j if j <= small.kcut else j - small.m + m
                out[:, j, k, :] = a[:small.K, jj, kk, :]
        return out

    ug = to_big(uh)
    small.step(uh, dt)
    dns.spectral_step_(ug, big, dt)
    got = from_big(ug)
    err = float(np.abs(got - uh).max() / np.abs(uh).max())
    rest = ug.copy()
    kept = list(range(small.kcut + 1)) + list(range(m - small.kcut, m))
    rest[np.ix_(range(small.K), kept, kept)] = 0
    leak = float(np.abs(rest).max() / np.abs(uh).max())
    print(f"HIT {n} {precision}: rel err {err:.3e}, modes outside |k| <= 16 {leak:.3e}", flush=True)
    tol = 1e-12 if precision == "fp64" else MIXED_TOL
    assert err < tol
    assert leak < tol


def test_c2_full_grid_five_steps():
    """The full 256×192×256 explicit FP64 channel, 5 CFL steps, against the
    oracle (the reference-structured CPU restatement)."""
    ug, u, pg, po, s = _run_pair((256, 192, 256), "explicit", "fp64", nsteps=5)
    _check(ug, u, pg, po, "fp64")


def test_c3_full_grid_two_steps():
    """The full 512×384×512 CN channel in FP64, 2 CFL steps, against the oracle."""
    ug, u, pg, po, s = _run_pair((512, 384, 512), "implicit", "fp64", nsteps=2)
    _check(ug, u, pg, po, "fp64")
    assert math.isfinite(pg[5, 5, 5])
