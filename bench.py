#!/usr/bin/env python3
"""Throughput benchmark of the single-GPU time step.

Default workload: the FV channel 256×192×256 on 4π × 2 × 4π/3, walls in y
with tanh stretching 1.5, Re_τ ≈ 180 (ν = 1/180, body force (1, 0, 0)),
explicit diffusion, FP64, CFL 0.5 (each step includes its CFL reduction).
IC: laminar U = 1 − y² plus a 10 % uniform perturbation per component from
splitmix64(seed + key), key = ((c·nz + k)·ny + j)·nx + i, seed 0.

Prints one JSON line: RK3 steps/s (value), Mcell-updates/s, the HBM
roofline of the dominant kernel family (HIP events on the library stream
over the timed region) and of the whole step (minimal-pass byte model), and
the FP64 CPU oracle timed on a bounded sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FAMILIES = {0: "momentum_rk", 1: "x_pass_fft", 2: "z_pass_fft", 3: "y_solve", 4: "pressure_gradient", 5: "cn_ydiffusion"}


def splitmix64(x):
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
    k = np.arange(nz, dtype=np.uint64)[None, None, :]
    with np.errstate(over="ignore"):
        for c in range(3):
            key = ((np.uint64(c) * np.uint64(nz) + k) * np.uint64(ny) + j) * np.uint64(nx) + i
            r = ((splitmix64(np.uint64(seed) + key) >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
            u[c][1:-1, 1:-1, 1:-1] = amp * (2 * r - 1)
    u[0][1:-1, 1:-1, 1:-1] += (1 - co["yc"] ** 2)[None, :, None]
    if prob.wall:
        u[1][:, -2, :] = 0.0  # top wall face
    return u


def channel_problem(n, ydiffusion="explicit", precision="fp64"):
    from distributednavierstokes_amd.problem import FVProblem, tanhstretch

    return FVProblem(n=n, lims=((0.0, 4 * math.pi), (-1.0, 1.0), (0.0, 4 * math.pi / 3)),
                     bc=("periodic", "wall", "periodic"), stretch=tanhstretch(1.5), visc=1 / 180,
                     bodyforce=(1.0, 0.0, 0.0), ydiffusion=ydiffusion, precision=precision)


def mmin_units_per_step(prob, cfl=True):
    """Minimal-pass model: units (one real field of N values) per RK3 step."""
    h = (prob.n[0] + 2) / prob.n[0]
    if prob.implicit:
        units = 3 * (34 + 8 * h) - 3
    else:
        units = 3 * (25 + 8 * h) - 3
    return units + (3 if cfl else 0)


def family_bytes_per_launch(prob, fam, esize):
    """Algorithmic bytes of one launch of a kernel family (per launch average)."""
    nx, ny, nz = prob.n
    N = nx * ny * nz
    NC = (nx // 2 + 1) * ny * nz
    if fam == 0:  # momentum + RK: u(3) + q(3, stages 2-3) read, q(3) + v(3) written
        units = (9 + 12 + 12) / 3 if not prob.implicit else (6 + 9 + 9) / 3
        return units * N * esize
    if fam == 1:  # x pass: (div: u read 3N, A write) and (c2r: A read, p write), averaged
        return ((3 * N * esize + NC * 2 * esize) + (NC * 2 * esize + N * esize)) / 2
    if fam == 2:  # z pass: A read + write
        return 2 * NC * 2 * esize
    if fam == 3:  # y solve: A read + write
        return 2 * NC * 2 * esize
    if fam == 4:  # pgrad: p + u read, u written (CN: + p read/write)
        return (7 if not prob.implicit else 9) * N * esize
    if fam == 5:  # CN, 3 components: u, q, p read, u written per component
        return 10 * N * esize
    return 0.0


def run_gpu(prob, steps, warmup, cfl=0.5):
    import ctypes as ct

    from distributednavierstokes_amd import _lib
    from distributednavierstokes_amd.fv import ROCmBackend, Setup

    L = _lib.lib()
    s = Setup(prob, ROCmBackend())
    u = channel_ic(prob)
    s.upload_u(u)
    zero = np.zeros(prob.ghosted_shape, order="F")
    for name in ("p", "qx", "qy", "qz"):
        s.upload(name, zero)
    if prob.implicit:
        s.upload("delta", zero)
    _lib.check(L.dns_fv_project(s.handle, _lib.FIELD["delta"] if prob.implicit else _lib.FIELD["p"]))
    out = np.zeros(1)
    t = 0.0

    def one_step():
        nonlocal t
        _lib.check(L.dns_fv_cfl_timestep(s.handle, cfl, _lib.dptr(out)))
        _lib.check(L.dns_fv_step(s.handle, float(out[0])))
        t += float(out[0])

    for _ in range(warmup):
        one_step()
    _lib.check(L.dns_fv_sync(s.handle))
    _lib.check(L.dns_fv_timing(s.handle, 1))
    t0 = time.perf_counter()
    for _ in range(steps):
        one_step()
    _lib.check(L.dns_fv_sync(s.handle))
    t1 = time.perf_counter()
    fam = {}
    for f in FAMILIES:
        ms = ct.c_double()
        cnt = ct.c_int64()
        _lib.check(L.dns_fv_timing_read(s.handle, f, ct.byref(ms), ct.byref(cnt)))
        if cnt.value:
            fam[f] = (ms.value, cnt.value)
    _lib.check(L.dns_fv_timing(s.handle, 0))
    md = np.zeros(1)
    _lib.check(L.dns_fv_maxdiv(s.handle, _lib.dptr(md)))
    ke = np.zeros(1)
    _lib.check(L.dns_fv_kinetic_energy(s.handle, _lib.dptr(ke)))
    return dict(seconds=t1 - t0, fam=fam, maxdiv=float(md[0]), ke=float(ke[0]), t=t)


def run_cpu_baseline(prob, steps=2, warmup=1, cfl=0.5):
    from oracle.oracle import OracleFV

    o = OracleFV(prob)
    u = channel_ic(prob)
    o.exchange_halo_u(u)
    p = o.scalarfield()
    q = o.vectorfield()
    d = o.scalarfield() if prob.implicit else p
    o.project(u, d)
    for _ in range(warmup):
        o.step(u, p, q, d, o.cfl_timestep(u, cfl))
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step(u, p, q, d, o.cfl_timestep(u, cfl))
    return steps / (time.perf_counter() - t0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, nargs=3, default=(256, 192, 256))
    ap.add_argument("--ydiffusion", default="explicit")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (scales the FETCH/WRITE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        a = torch.ones(2 ** 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 2 ** 30, "copy_bytes_written": 2 ** 30, "launches": 3}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    prob = channel_problem(tuple(args.n), args.ydiffusion, args.precision)
    N = int(np.prod(prob.n))
    esize = 4 if prob.precision == "mixed" else 8
    r = run_gpu(prob, args.steps, args.warmup)
    sps = args.steps / r["seconds"]
    ms_step = 1e3 * r["seconds"] / args.steps
    units = mmin_units_per_step(prob)
    step_bytes = units * N * esize
    # dominant kernel family by total event time
    dom = max(r["fam"], key=lambda f: r["fam"][f][0])
    ms_tot, cnt = r["fam"][dom]
    avg_s = ms_tot / cnt / 1e3
    fam_bytes = family_bytes_per_launch(prob, dom, esize)
    achieved = fam_bytes / avg_s / 1e9
    fam_summary = {FAMILIES[f]: {"ms_per_step": v[0] / args.steps, "launches_per_step": v[1] / args.steps,
                                 "avg_us": 1e3 * v[0] / v[1],
                                 "gbs": family_bytes_per_launch(prob, f, esize) / (v[0] / v[1] / 1e3) / 1e9}
                   for f, v in r["fam"].items()}
    cpu = None
    if not args.no_cpu_baseline:
        cps = run_cpu_baseline(prob, steps=args.cpu_steps)
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": cps, "unit": "RK3 steps/s", "cores": cores, "kind": "port",
               "sample": f"{args.cpu_steps} CFL steps of the same {prob.n[0]}x{prob.n[1]}x{prob.n[2]} "
                         f"workload after 1 warm-up step (FP64 C oracle, reference-literal kernel sequence, "
                         f"own FFT instead of FFTW, OpenMP)"}
    line = {
        "metric": "RK3 steps/s & Mcells/s, FP64 vs mixed at stated err; %HBM roofline; LES err",
        "value": sps,
        "unit": "RK3 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if esize == 8 else "f32 storage / f64 Thomas+reductions",
        "data": "synthetic (laminar profile + seeded splitmix64 perturbation)",
        "config": {"workload": f"FV channel {prob.n[0]}x{prob.n[1]}x{prob.n[2]}, walls in y (tanh 1.5), "
                               f"Re_tau~180, {prob.ydiffusion} y-diffusion, {prob.precision}, CFL 0.5",
                   "n": list(prob.n), "ydiffusion": prob.ydiffusion, "precision": prob.precision},
        "mcell_updates_per_s": N * sps / 1e6,
        "roofline": {"bound": "hbm", "kernel": FAMILIES[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "step_roofline": {"model": "minimal-pass", "units_per_step": units, "bytes_per_step": step_bytes,
                          "achieved_gbs": step_bytes * sps / 1e9, "frac": step_bytes * sps / 1e9 / HBM_PEAK_GBS},
        "kernels": fam_summary,
        "checks": {"maxdiv": r["maxdiv"], "kinetic_energy": r["ke"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
