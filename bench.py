#!/usr/bin/env python3
"""Throughput benchmark of the single-GPU time step.

Default workload: the FV channel 512×384×512 on 4π × 2 × 4π/3, walls in y
with tanh stretching 1.5, Re_τ ≈ 180 (ν = 1/180, body force (1, 0, 0)),
semi-implicit Crank–Nicolson y-diffusion, FP64, CFL 0.5 (each step includes
its CFL reduction).  IC: laminar U = 1 − y² plus a 10 % uniform perturbation
per component from splitmix64(seed + key), key = ((c·nz + k)·ny + j)·nx + i,
seed 0.

Prints one JSON line: RK3 steps/s of that workload (value), Mcell-updates/s,
the HBM roofline of the dominant kernel family (HIP events on the library
stream over the timed region) and of the whole step (minimal-pass byte
model), the measured device-copy bandwidth, the mixed and plain FP64 forms of
the same workload with their speed ratios, the mixed-vs-FP64 error over 100
steps driven by the FP64 Δt sequence, the latest learned-closure LES error,
the FP64 CPU oracle timed on the host (all cores on the workload, one thread
on a slab of it), and the other configurations (`forms`).
"""
import argparse
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FAMILIES = {0: "momentum_rk", 1: "x_pass_fft", 2: "z_pass_fft", 3: "y_solve", 4: "pressure_gradient", 5: "cn_ydiffusion"}


from distributednavierstokes_amd.workloads import channel_ic, channel_problem, tgv_ic, tgv_problem  # noqa: E402


def spectral_mmin_bytes(n, kcut, esize):
    """Minimal-pass model of one IF-RK3 step (complex elements × 2·esize):
    per stage y back 3Km² + 3Knm, z back 3Knm + 3Kn², fused x 3Kn² + 6Kn²,
    z forward 6Kn² + 6Knm, y forward + RHS + IF + projection 6Knm + 9Km²
    (+3Km² in stage 1 for the register copy)."""
    K, m = kcut + 1, 2 * kcut + 1
    Km2, Knm, Kn2 = K * m * m, K * n * m, K * n * n
    elems = 39 * Km2 + 54 * Knm + 54 * Kn2
    return elems * 2 * esize


def run_spectral(n, kcut, precision, steps, warmup, cfl=0.5):
    import ctypes as ct

    from distributednavierstokes_amd import _lib
    import distributednavierstokes_amd as dns

    s = dns.spectral_setup(n=n, kcut=kcut, visc=2e-4, T=np.float32 if precision == "mixed" else np.float64)
    L = _lib.lib()
    dns.spectral_randomfield_(dns.spectral.DeviceState(s), s, totalenergy=1.0, kpeak=4, seed=0)
    forcing = dns.shellforcing(dns.spectral.DeviceState(s), s, shells=(1, 2))
    out = np.zeros(1)
    _lib.check(L.dns_spec_vsummax(s.handle, _lib.dptr(out)))
    vmax = float(out[0])
    kmax = 2 * math.pi * kcut / s.l

    def one_step():
        nonlocal vmax
        dt = cfl * math.sqrt(3) / (kmax * (vmax + 1e-16))
        _lib.check(L.dns_spec_step(s.handle, dt, _lib.dptr(out)))
        forcing(dns.spectral.DeviceState(s), s)
        vmax = float(out[0])

    for _ in range(warmup):
        one_step()
    _lib.check(L.dns_spec_sync(s.handle))
    _lib.check(L.dns_spec_timing(s.handle, 1))
    t0 = time.perf_counter()
    for _ in range(steps):
        one_step()
    _lib.check(L.dns_spec_sync(s.handle))
    t1 = time.perf_counter()
    fam = {}
    names = {0: "y_pass", 1: "z_pass", 2: "x_pass_products", 3: "y_pass_if_stage", 4: "forcing"}
    for f, nm in names.items():
        ms, cnt = ct.c_double(), ct.c_int64()
        _lib.check(L.dns_spec_timing_read(s.handle, f, ct.byref(ms), ct.byref(cnt)))
        if cnt.value:
            fam[nm] = {"ms_per_step": ms.value / steps, "launches_per_step": cnt.value / steps}
    _lib.check(L.dns_spec_timing(s.handle, 0))
    e = s.scalar(L.dns_spec_energy)
    md = s.scalar(L.dns_spec_maxdiv)
    esize = 4 if precision == "mixed" else 8
    sps = steps / (t1 - t0)
    b = spectral_mmin_bytes(n, kcut, esize)
    return {"workload": f"HIT {n}^3 kcut={kcut} {precision}, shell forcing k<=2, nu=2e-4, CFL {cfl}",
            "steps_per_s": sps, "ms_per_step": 1e3 / sps, "mcell_updates_per_s": n ** 3 * sps / 1e6,
            "bytes_per_step": b, "achieved_gbs": b * sps / 1e9, "frac": b * sps / 1e9 / HBM_PEAK_GBS,
            "kernels": fam, "checks": {"energy": e, "maxdiv": md}}


CNN_MACS_PER_VOXEL = 27 * (3 * 64 + 64 * 64 + 64 * 64 + 64 * 6)   # 236,736
BF16_PEAK_TFLOPS = 2500.0


def run_cnn_train(steps, warmup, M=64, batch=16, nsamples=16):
    """Adam steps of the closure CNN on batches of 16 × 64³ (synthetic
    pairs); FLOPs per step = 3 × 2 × MACs (forward, data and weight grads)."""
    from distributednavierstokes_amd.cnn import SFSNet

    net = SFSNet(M=M, batch=batch, seed=0)
    rng = np.random.default_rng(0)
    for _ in range(nsamples):
        net.add_sample(rng.standard_normal((M, M, M, 3)).astype(np.float32),
                       (0.1 * rng.standard_normal((M, M, M, 6))).astype(np.float32))
    net.fit_scales()
    net.train(warmup, lr=1e-3, seed=1)
    net.timing(True)
    t0 = time.perf_counter()
    losses = net.train(steps, lr=1e-3, seed=2)
    t1 = time.perf_counter()
    fam = {nm: net.timing_read(f) for f, nm in ((0, "forward"), (1, "backward"), (2, "adam_pack"))}
    net.timing(False)
    sps = steps / (t1 - t0)
    flops = 3 * 2 * CNN_MACS_PER_VOXEL * batch * M ** 3
    fwd_ms = fam["forward"][0] / max(1, fam["forward"][1])
    bwd_ms = fam["backward"][0] / max(1, fam["backward"][1])
    return {"workload": f"CNN 3-64-64-64-6, 3x3x3 periodic, batch {batch} x {M}^3, BF16 MFMA / FP32 accumulate, Adam",
            "adam_steps_per_s": sps, "ms_per_step": 1e3 / sps, "flop_per_step": flops,
            "achieved_tflops": flops * sps / 1e12, "frac_bf16_peak": flops * sps / 1e12 / BF16_PEAK_TFLOPS,
            "forward_ms": fwd_ms, "forward_tflops": flops / 3 / (fwd_ms * 1e-3) / 1e12 if fwd_ms > 0 else None,
            "backward_ms": bwd_ms, "backward_tflops": 2 * flops / 3 / (bwd_ms * 1e-3) / 1e12 if bwd_ms > 0 else None,
            "loss_first_last": [float(losses[0]), float(losses[-1])]}


def run_les_closure(steps, warmup, M=64):
    """64³ spectral LES steps with the learned closure (CNN on every stage)
    and without it (no-model fallback), random IC, fixed Δt."""
    from distributednavierstokes_amd import _lib
    from distributednavierstokes_amd.cnn import SFSNet
    import distributednavierstokes_amd as dns

    s = dns.spectral_setup(n=M, kcut=M // 3, visc=2e-4)
    dns.spectral_randomfield_(dns.spectral.DeviceState(s), s, totalenergy=1.0, kpeak=4, seed=0)
    net = SFSNet(M=M, batch=1, seed=0)
    net.scales = (1.0, 0.01)
    L = _lib.lib()
    out = np.zeros(1)
    res = {}
    for name, step in (("no_model", lambda: L.dns_spec_step(s.handle, 1e-3, _lib.dptr(out))),
                       ("closure", lambda: L.dns_les_closure_step(s.handle, net.handle, 1e-3, _lib.dptr(out)))):
        for _ in range(warmup):
            _lib.check(step())
        _lib.check(L.dns_spec_sync(s.handle))
        t0 = time.perf_counter()
        for _ in range(steps):
            _lib.check(step())
        _lib.check(L.dns_spec_sync(s.handle))
        res[name] = steps / (time.perf_counter() - t0)
    return {"workload": f"spectral LES {M}^3 kcut={M // 3}, FP64, learned closure vs no model",
            "closure_steps_per_s": res["closure"], "no_model_steps_per_s": res["no_model"]}


def run_closure_experiment(n=512, M=64, nsnap=64, every=40, spinup_t=2.0, train_steps=2000, batch=16,
                           nheld=8, window_tint=3.0, nwin=12, dns_precision="mixed", cfl=0.5, log=print):
    """Learned-closure experiment: forced HIT DNS (random IC k^4 exp(-2(k/4)^2),
    E = 1, seed 0, shells 1-2 clamped, nu = 2e-4) spun up to t = spinup_t,
    then nsnap SFS samples (Gaussian, filter width 2 l/M) every `every` steps
    on the M^3 grid; the CNN trains on the first nsnap - nheld (batch 16,
    Adam lr 1e-3, seed 0); a-priori score on the held-out samples; then M^3
    LES (FP64) with and without the closure from the filtered DNS at the
    last sample, over `window_tint` integral times (t_int of the DNS at the
    last sample), against filtered-DNS spectra at nwin equally spaced times:
    per-time and time-averaged shell errors, over all shells and over the
    unforced shells 3..kcut."""
    from distributednavierstokes_amd import closure as cl
    from distributednavierstokes_amd.cnn import SFSNet
    from distributednavierstokes_amd.spectral import DeviceState, _shells
    import distributednavierstokes_amd as dns

    T = np.float32 if dns_precision == "mixed" else np.float64
    s = dns.spectral_setup(n=n, visc=2e-4, T=T)
    dev = DeviceState(s)
    dns.spectral_randomfield_(dev, s, totalenergy=1.0, kpeak=4, seed=0)
    forcing = dns.shellforcing(dev, s, shells=(1, 2))
    t_start = time.perf_counter()

    def progress(state, s_):
        if state.n % 200 == 0:
            log(f"dns spin-up step {state.n} t={state.t:.3f} ({time.perf_counter() - t_start:.0f} s)")

    st = dns.spectral_solve_(uh=dev, setup=s, tlims=(0.0, spinup_t), cfl=cfl, forcing=forcing,
                             processors={"log": progress})
    log(f"spin-up done: {st.n} steps to t={st.t:.3f}")
    net = SFSNet(M=M, batch=batch, seed=0)
    kept_idx = tuple(range(nsnap - nheld, nsnap))
    sam, st2, times, kept = cl.collect_pairs(s, dev, net, M=M, nsnap=nsnap, every=every, spinup=0, cfl=cfl,
                                             forcing=forcing, keep_physical=kept_idx, t0=st.t)
    log(f"{net.nsamples} samples at t={times[0]:.3f}..{times[-1]:.3f} ({time.perf_counter() - t_start:.0f} s)")
    su, stau = net.fit_scales()
    # LES initial condition: filtered DNS at the last sample
    ub0, _ = sam.spectral()
    # continue the DNS over the window, filtered spectra at the stop times
    t_int = float(dns.spectral_stats(dev, s).t_int)
    window = window_tint * t_int
    stops = [times[-1] + window * (q + 1) / nwin for q in range(nwin)]
    E_dns = []

    def dns_proc(state, s_):
        if any(abs(state.t - x) < 1e-9 for x in stops):
            sam.sample(None)
            E_dns.append(np.asarray(sam.spectrum()))

    dns.spectral_solve_(uh=dev, setup=s, tlims=(times[-1], stops[-1]), cfl=cfl, forcing=forcing, tstops=stops,
                        processors={"spec": dns_proc})
    log(f"DNS window done ({time.perf_counter() - t_start:.0f} s)")
    # a-priori: untrained vs trained on held-out samples, no model = mean square of normalised tau
    nofit = float(np.mean([np.mean((kept[q][1] / stau) ** 2) for q in kept_idx]))
    t0 = time.perf_counter()
    losses = net.train(train_steps, lr=1e-3, seed=0, ntrain=nsnap - nheld)
    train_s = time.perf_counter() - t0
    held = net.evaluate(nsnap - nheld, nheld)
    corr = float(np.mean([cl.stress_correlation(net.predict(q), kept[q][1]) for q in kept_idx]))
    log(f"trained {train_steps} steps in {train_s:.1f} s: held-out MSE {held:.4f} (no model {nofit:.4f}), corr {corr:.3f}")

    def les(netarg):
        les_s = cl.les_setup_like(sam, T=np.float64)
        E = []
        uh = np.array(ub0, order="F")
        fr = dns.shellforcing(uh, les_s, shells=(1, 2))

        def proc(state, s_):
            if any(abs(state.t - x) < 1e-9 for x in stops):
                E.append(np.asarray(_shells(s_, range(1, les_s.kcut + 1))))

        dns.spectral_solve_(uh=uh, setup=les_s, tlims=(times[-1], stops[-1]), cfl=cfl, forcing=fr, tstops=stops,
                            processors={"spec": proc}, closure=netarg)
        return E

    E_nm, E_cl = les(None), les(net)
    Ed, Enm, Ecl = np.mean(E_dns, axis=0), np.mean(E_nm, axis=0), np.mean(E_cl, axis=0)
    err_nm = np.abs(Enm / Ed - 1.0)
    err_cl = np.abs(Ecl / Ed - 1.0)
    # per stop time over the unforced shells 3..kcut (index 2.. of shells 1..kcut)
    per_time = {"t": stops,
                "max_closure": [float(np.max(np.abs(np.asarray(c)[2:] / np.asarray(d)[2:] - 1))) for c, d in zip(E_cl, E_dns)],
                "mean_closure": [float(np.mean(np.abs(np.asarray(c)[2:] / np.asarray(d)[2:] - 1))) for c, d in zip(E_cl, E_dns)],
                "max_no_model": [float(np.max(np.abs(np.asarray(c)[2:] / np.asarray(d)[2:] - 1))) for c, d in zip(E_nm, E_dns)],
                "mean_no_model": [float(np.mean(np.abs(np.asarray(c)[2:] / np.asarray(d)[2:] - 1))) for c, d in zip(E_nm, E_dns)]}
    log(f"LES done ({time.perf_counter() - t_start:.0f} s)")
    return {"workload": f"HIT {n}^3 DNS ({dns_precision}) -> {nsnap} SFS samples on {M}^3 (Gaussian, 2l/M) -> "
                        f"CNN {train_steps} Adam steps batch {batch} -> {M}^3 LES closure vs no model",
            "sample_times": [times[0], times[-1]], "scales": [su, stau],
            "train_seconds": train_s, "train_loss_first_last": [float(losses[0]), float(losses[-1])],
            "a_priori": {"held_out_mse_trained": held, "held_out_mse_no_model": nofit, "tau_correlation": corr},
            "a_posteriori": {"window": [times[-1], stops[-1]], "times": nwin, "t_int": t_int,
                             "window_integral_times": window_tint,
                             "unforced_shells": [3, len(Ed)],
                             "unforced_rel_err_mean_closure": float(err_cl[2:].mean()),
                             "unforced_rel_err_max_closure": float(err_cl[2:].max()),
                             "unforced_rel_err_mean_no_model": float(err_nm[2:].mean()),
                             "unforced_rel_err_max_no_model": float(err_nm[2:].max()),
                             "per_time_unforced": per_time,
                             "spectrum_rel_err_max_closure": float(err_cl.max()),
                             "spectrum_rel_err_max_no_model": float(err_nm.max()),
                             "spectrum_rel_err_mean_closure": float(err_cl.mean()),
                             "spectrum_rel_err_mean_no_model": float(err_nm.mean()),
                             "E_filtered_dns": Ed.tolist(), "E_les_closure": Ecl.tolist(),
                             "E_les_no_model": Enm.tolist()}}


def run_mixed_accuracy(n=(512, 384, 512), steps=1000, checkpoints=(10, 100, 1000), cfl=0.5, log=print,
                       ydiffusion="implicit"):
    """Mixed (FP32 storage/stencils/FFTs, FP64 Thomas and reductions) against
    FP64 on the CN channel from the same IC, the mixed run driven by the FP64
    run's Δt sequence.  Errors at the checkpoints: relative L2 of the
    velocity, kinetic energy, mean streamwise profile U(y) and Re_τ
    (u_τ² = ν dU/dy from the first cell rows at both walls)."""
    from distributednavierstokes_amd import _lib
    from distributednavierstokes_amd.fv import ROCmBackend, Setup

    L = _lib.lib()
    res = {"workload": f"FV channel {n[0]}x{n[1]}x{n[2]}, {ydiffusion} y-diffusion, CFL {cfl}, {steps} steps; "
                       f"mixed driven by the FP64 dt sequence", "checkpoints": {}}
    snaps, dts = {}, []
    for prec in ("fp64", "mixed"):
        prob = channel_problem(n, ydiffusion, prec)
        s = Setup(prob, ROCmBackend())
        s.upload_u(channel_ic(prob))
        zero = np.zeros(prob.ghosted_shape, order="F")
        for name in ("p", "qx", "qy", "qz") + (("delta",) if prob.implicit else ()):
            s.upload(name, zero)
        _lib.check(L.dns_fv_project(s.handle, _lib.FIELD["delta" if prob.implicit else "p"]))
        out = np.zeros(1)
        t0 = time.perf_counter()
        for it in range(1, steps + 1):
            if prec == "fp64":
                _lib.check(L.dns_fv_cfl_timestep(s.handle, cfl, _lib.dptr(out)))
                dts.append(float(out[0]))
            _lib.check(L.dns_fv_step(s.handle, dts[it - 1]))
            if it % 100 == 0:
                _lib.check(L.dns_fv_sync(s.handle))
                log(f"mixed-vs-fp64 check: {prec} step {it} ({time.perf_counter() - t0:.0f} s)")
            if it in checkpoints:
                u = [s.download(c)[1:-1, 1:-1, 1:-1].copy() for c in ("ux", "uy", "uz")]
                ke = np.zeros(1)
                _lib.check(L.dns_fv_kinetic_energy(s.handle, _lib.dptr(ke)))
                md = np.zeros(1)
                _lib.check(L.dns_fv_maxdiv(s.handle, _lib.dptr(md)))
                if prec == "fp64":
                    snaps[it] = (u, float(ke[0]), float(md[0]))
                else:
                    ref, ke_ref, md_ref = snaps.pop(it)
                    num = sum(float(np.sum((a - b) ** 2)) for a, b in zip(u, ref))
                    den = sum(float(np.sum(b ** 2)) for b in ref)
                    U, Ur = u[0].mean(axis=(0, 2)), ref[0].mean(axis=(0, 2))
                    yc = 0.5 * (prob.yfaces[:-1] + prob.yfaces[1:])

                    def retau(prof):
                        dUdy = 0.5 * (prof[0] / (yc[0] - prob.yfaces[0]) + prof[-1] / (prob.yfaces[-1] - yc[-1]))
                        return math.sqrt(prob.visc * dUdy) / prob.visc

                    res["checkpoints"][str(it)] = {
                        "t": float(sum(dts[:it])),
                        "velocity_rel_l2": math.sqrt(num / den),
                        "kinetic_energy_rel": abs(float(ke[0]) - ke_ref) / ke_ref,
                        "mean_profile_rel": float(np.max(np.abs(U - Ur)) / np.max(np.abs(Ur))),
                        "re_tau_rel": abs(retau(U) - retau(Ur)) / retau(Ur),
                        "maxdiv_mixed": float(md[0]), "maxdiv_fp64": md_ref}
                    log(f"checkpoint {it}: {res['checkpoints'][str(it)]}")
        _lib.check(L.dns_fv_sync(s.handle))
        res[f"{prec}_seconds"] = time.perf_counter() - t0
        del s
    return res


def fv_form(prob, steps, warmup, fixed_dt=None, ic=None):
    r = run_gpu(prob, steps, warmup, fixed_dt=fixed_dt, ic=ic)
    N = int(np.prod(prob.n))
    esize = 4 if prob.precision == "mixed" else 8
    sps = steps / r["seconds"]
    sps_events = None
    if prob.precision != "plain":
        # without per-launch events the steps replay HIP graphs (fixed Δt) or
        # run the device-Δt CFL loop (explicit form); the rate is taken from
        # that run, the kernel times from the run with events
        sps_events = sps
        sps = steps / run_gpu(prob, steps, warmup, fixed_dt=fixed_dt, ic=ic, events=False)["seconds"]
    units = mmin_units_per_step(prob, cfl=fixed_dt is None)
    b = units * N * esize
    out = {"workload": f"FV {prob.n[0]}x{prob.n[1]}x{prob.n[2]} {prob.bc[1]} y, {prob.ydiffusion}, {prob.precision}, "
                       + ("fixed dt" if fixed_dt else "CFL 0.5"),
           "steps_per_s": sps, "ms_per_step": 1e3 / sps, "mcell_updates_per_s": N * sps / 1e6,
           "units_per_step": units, "bytes_per_step": b, "achieved_gbs": b * sps / 1e9,
           "frac": b * sps / 1e9 / HBM_PEAK_GBS,
           "kernels": {FAMILIES[f]: v[0] / steps for f, v in r["fam"].items()},
           "checks": {"maxdiv": r["maxdiv"], "kinetic_energy": r["ke"]}}
    if sps_events is not None:
        out["steps_per_s_with_launch_events"] = sps_events
        out["timing"] = "wall clock of a run without per-launch events (captured HIP graphs / device-side CFL)"
    if prob.precision == "plain":
        rl = refliteral_units_per_step(prob, cfl=fixed_dt is None)
        out["refliteral_units_per_step"] = rl
        out["refliteral_bytes_per_step"] = rl * N * 8
        out["refliteral_frac"] = rl * N * 8 * sps / 1e9 / HBM_PEAK_GBS
    return out


def refliteral_units_per_step(prob, cfl=True):
    """Reference-literal byte model of the plain kernel sequence: 54 units per
    explicit stage, 129 per CN stage (the reference's separate broadcasts,
    copies/transposes and HBM Thomas arrays), plus the 3 CFL reductions."""
    return 3 * (129 if prob.implicit else 54) + (3 if cfl else 0)


def mmin_units_per_step(prob, cfl=True):
    """Minimal-pass model: units (one real field of N values) per RK3 step."""
    h = (prob.n[0] + 2) / prob.n[0]
    if prob.implicit:
        units = 3 * (34 + 8 * h) - 3
    else:
        units = 3 * (25 + 8 * h) - 3
    return units + (3 if cfl else 0)


def family_bytes_per_launch(prob, fam, esize, cn3=False):
    """Algorithmic bytes of one launch of a kernel family (per launch average).
    cn3: the CN family runs the three components per launch (k_cn_seg3)."""
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
    if fam == 5:  # CN per component: u_c, q_c, p read, u_c written; three per launch: p read once
        return (10 if cn3 else 4) * N * esize
    return 0.0


def run_gpu(prob, steps, warmup, cfl=0.5, fixed_dt=None, ic=None, events=True):
    """Steps of the FV workload on the GPU; `events` records HIP events around
    every launch for the per-family times (a fixed-Δt run without them replays
    captured HIP graphs of the step)."""
    import ctypes as ct

    from distributednavierstokes_amd import _lib
    from distributednavierstokes_amd.fv import ROCmBackend, Setup

    L = _lib.lib()
    s = Setup(prob, ROCmBackend())
    u = (ic or channel_ic)(prob)
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
        if fixed_dt is None:
            _lib.check(L.dns_fv_cfl_timestep(s.handle, cfl, _lib.dptr(out)))
            dt = float(out[0])
        else:
            dt = fixed_dt
        _lib.check(L.dns_fv_step(s.handle, dt))
        t += dt

    def run(n):
        # CFL steps without per-launch events: the library's device-Δt loop
        # (dns_fv_advance_cfl), which needs no host round trip per step
        nonlocal t
        if fixed_dt is None and not events:
            tt = np.array([t])
            _lib.check(L.dns_fv_advance_cfl(s.handle, cfl, n, math.inf, _lib.dptr(tt), None, None))
            t = float(tt[0])
        else:
            for _ in range(n):
                one_step()

    run(warmup)
    _lib.check(L.dns_fv_sync(s.handle))
    _lib.check(L.dns_fv_timing(s.handle, 1 if events else 0))
    t0 = time.perf_counter()
    run(steps)
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
    return dict(seconds=t1 - t0, fam=fam, maxdiv=float(md[0]), ke=float(ke[0]), t=t, describe=s.describe())


def cpu_model():
    """Model name of the host CPU (/proc/cpuinfo, as lscpu reports it)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def run_cpu_baseline(prob, steps=2, warmup=1, cfl=0.5, threads=0):
    """The FP64 C oracle (reference-literal kernel sequence, own FFT) on the
    host: `warmup` CFL steps, then `steps` individually timed steps
    (scaling.jl:90-109 protocol).  threads <= 0 keeps the OpenMP default."""
    from oracle.oracle import OracleFV, set_threads

    nthreads = set_threads(threads)
    o = OracleFV(prob)
    u = channel_ic(prob)
    o.exchange_halo_u(u)
    p = o.scalarfield()
    q = o.vectorfield()
    d = o.scalarfield() if prob.implicit else p
    o.project(u, d)
    for _ in range(warmup):
        o.step(u, p, q, d, o.cfl_timestep(u, cfl))
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        o.step(u, p, q, d, o.cfl_timestep(u, cfl))
        ts.append(time.perf_counter() - t0)
    del o, u, p, q, d
    return {"threads": nthreads, "step_seconds": {"min": min(ts), "median": float(np.median(ts)),
                                                  "mean": float(np.mean(ts))}, "steps": steps, "warmup": warmup}


def cpu_baseline_block(prob, steps=2, slab_nz=16):
    """All host cores on the full workload, and one thread on a z-slab of it
    (same nx, ny; nz = slab_nz) scaled to the full grid by nz / slab_nz."""
    full = run_cpu_baseline(prob, steps=steps, warmup=1, threads=0)
    med = full["step_seconds"]["median"]
    out = {"value": 1.0 / med, "unit": "RK3 steps/s", "cores": full["threads"], "kind": "port",
           "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
           "sample": f"{steps} individually timed CFL steps after 1 warm-up step of the same "
                     f"{prob.n[0]}x{prob.n[1]}x{prob.n[2]} workload on {full['threads']} OpenMP threads; value = "
                     f"1 / median step time (FP64 C oracle: the reference's kernel sequence with its copies and "
                     f"HBM Thomas arrays, own FFT instead of FFTW)",
           "step_seconds": full["step_seconds"]}
    nz = prob.n[2]
    if slab_nz and slab_nz < nz:
        import dataclasses

        slab = dataclasses.replace(prob, n=(prob.n[0], prob.n[1], slab_nz))
        one = run_cpu_baseline(slab, steps=steps, warmup=1, threads=1)
        scale = nz / slab_nz
        out["single_thread"] = {
            "value": 1.0 / (one["step_seconds"]["median"] * scale), "unit": "RK3 steps/s", "cores": 1,
            "sample": f"{steps} timed steps after 1 warm-up on a {prob.n[0]}x{prob.n[1]}x{slab_nz} slab of the "
                      f"workload, one thread, step time scaled by {scale:g} to the full grid (its z FFTs are "
                      f"shorter: an upper bound on the full-grid rate)",
            "slab_step_seconds": one["step_seconds"]}
        set_threads_default()
    return out


def set_threads_default():
    from oracle.oracle import set_threads

    set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))


def stream_copy_gbs(nbytes=2 ** 31, reps=10):
    """Device-to-device copy of nbytes (read + write counted) timed with HIP
    events on torch's stream: the attainable HBM bandwidth next to the 8 TB/s
    peak.  Runs in a child process: torch's bundled HIP runtime does not
    share a process with the library's."""
    import subprocess

    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--stream-copy", str(nbytes), str(reps)],
                         capture_output=True, text=True, check=True)
    return float(json.loads(out.stdout.strip().splitlines()[-1])["copy_gbs"])


def _stream_copy_child(nbytes, reps):
    import torch

    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) / 1e3 / reps
    return 2 * nbytes / sec / 1e9


def measured_traffic(prob, family, form=None):
    """DRAM bytes per launch of `family` on this workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (copy-calibrated), or None when
    no profile of this exact workload, kernel family and kernel form (the
    first word of the handle's description of the family, when given) is
    committed."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"{family}_traffic_channel_{prob.n[0]}x{prob.n[1]}x{prob.n[2]}_{prob.precision}.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if (list(rec.get("n", [])) != list(prob.n) or rec.get("ydiffusion") != prob.ydiffusion
            or rec.get("precision") != prob.precision or rec.get("family") != family
            or (form is not None and rec.get("form") != form)):
        return None
    return rec["bytes_per_launch"]


def closure_error_record():
    """The latest learned-closure experiment committed under profiles/ (written
    by `bench.py --closure-experiment`): LES spectrum errors against the
    filtered DNS, labelled with their source."""
    path = os.path.join(ROOT, "profiles", "closure_experiment_hit512_les64.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    ap = rec.get("a_posteriori", {})
    return {"source": "profiles/closure_experiment_hit512_les64.json (bench.py --closure-experiment; "
                      "not re-run by the default bench)",
            "workload": rec.get("workload"),
            "spectrum_rel_err_mean_closure": ap.get("spectrum_rel_err_mean_closure"),
            "spectrum_rel_err_max_closure": ap.get("spectrum_rel_err_max_closure"),
            "spectrum_rel_err_mean_no_model": ap.get("spectrum_rel_err_mean_no_model"),
            "spectrum_rel_err_max_no_model": ap.get("spectrum_rel_err_max_no_model"),
            "window": ap.get("window"), "window_integral_times": ap.get("window_integral_times"),
            "unforced_shells": ap.get("unforced_shells"),
            "unforced_rel_err_mean_closure": ap.get("unforced_rel_err_mean_closure"),
            "unforced_rel_err_max_closure": ap.get("unforced_rel_err_max_closure"),
            "unforced_rel_err_mean_no_model": ap.get("unforced_rel_err_mean_no_model"),
            "unforced_rel_err_max_no_model": ap.get("unforced_rel_err_max_no_model"),
            "tau_correlation": rec.get("a_priori", {}).get("tau_correlation")}


def main():
    if len(sys.argv) == 4 and sys.argv[1] == "--stream-copy":
        print(json.dumps({"copy_gbs": _stream_copy_child(int(sys.argv[2]), int(sys.argv[3]))}), flush=True)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, nargs=3, default=(512, 384, 512))
    ap.add_argument("--ydiffusion", default="implicit")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the other configurations")
    ap.add_argument("--no-stream-copy", action="store_true",
                    help="skip the device-copy bandwidth measurement (a child process; for profiler runs)")
    ap.add_argument("--mixed-check-steps", type=int, default=100,
                    help="steps of the mixed-vs-FP64 comparison in the line (0: skip)")
    ap.add_argument("--spectral-n", type=int, default=0,
                    help="run only the pseudo-spectral HIT form at this n (kcut = n/3; --precision fp64|mixed)")
    ap.add_argument("--cnn", action="store_true",
                    help="run only the closure-CNN training step (batch 16 x 64^3, --steps Adam steps)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (scales the FETCH/WRITE counters)")
    ap.add_argument("--closure-experiment", action="store_true",
                    help="run only the learned-closure experiment (DNS -> SFS data -> training -> LES)")
    ap.add_argument("--dns-n", type=int, default=512)
    ap.add_argument("--les-m", type=int, default=64)
    ap.add_argument("--nsnap", type=int, default=64)
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--mixed-accuracy", action="store_true",
                    help="run only the mixed-vs-FP64 CN channel comparison (same dt sequence)")
    ap.add_argument("--accuracy-steps", type=int, default=1000)
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if args.mixed_accuracy:
        cps = tuple(c for c in (10, 100, 1000) if c <= args.accuracy_steps) or (args.accuracy_steps,)
        res = run_mixed_accuracy(tuple(args.n), steps=args.accuracy_steps, checkpoints=cps, log=log,
                                 ydiffusion=args.ydiffusion)
        print(json.dumps(res), flush=True)
        return
    if args.closure_experiment:
        res = run_closure_experiment(n=args.dns_n, M=args.les_m, nsnap=args.nsnap, train_steps=args.train_steps,
                                     log=log)
        print(json.dumps(res), flush=True)
        return
    if args.spectral_n:
        print(json.dumps(run_spectral(args.spectral_n, args.spectral_n // 3, args.precision, args.steps, args.warmup)),
              flush=True)
        return
    if args.cnn:
        print(json.dumps(run_cnn_train(args.steps, args.warmup)), flush=True)
        return
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

    # first, while this process has not touched the GPU: a plain device copy
    # in a child process (torch), the attainable HBM bandwidth
    copy_gbs = None
    if not args.no_stream_copy:
        log("step 0: device-to-device copy (attainable HBM bandwidth)")
        copy_gbs = stream_copy_gbs()
    prob = channel_problem(tuple(args.n), args.ydiffusion, args.precision)
    N = int(np.prod(prob.n))
    esize = 4 if prob.precision == "mixed" else 8
    # kernel times from a run with HIP events around every launch; the step
    # rate from a run without them (device Δt / replayed graphs where the form has them)
    log(f"step 1: {prob.n} {prob.ydiffusion} {prob.precision}, {args.steps} steps with launch events")
    r = run_gpu(prob, args.steps, args.warmup)
    log("step 2: the same steps without launch events (the rate)")
    r_rate = run_gpu(prob, args.steps, args.warmup, events=False)
    sps = args.steps / r_rate["seconds"]
    ms_step = 1e3 * r_rate["seconds"] / args.steps
    units = mmin_units_per_step(prob)
    step_bytes = units * N * esize
    # dominant kernel family by total event time
    dom = max(r["fam"], key=lambda f: r["fam"][f][0])
    ms_tot, cnt = r["fam"][dom]
    avg_s = ms_tot / cnt / 1e3
    cn_form = r["describe"].get("cn", "").split(" ")[0] or None
    cn3 = cn_form == "seg3"
    fam_bytes = family_bytes_per_launch(prob, dom, esize, cn3)
    achieved = fam_bytes / avg_s / 1e9
    fam_summary = {FAMILIES[f]: {"ms_per_step": v[0] / args.steps, "launches_per_step": v[1] / args.steps,
                                 "avg_us": 1e3 * v[0] / v[1],
                                 "gbs": family_bytes_per_launch(prob, f, esize, cn3) / (v[0] / v[1] / 1e3) / 1e9}
                   for f, v in r["fam"].items()}
    # the other precision forms of the same workload and their speed ratios
    precision_forms = {prob.precision: {"steps_per_s": sps, "frac": step_bytes * sps / 1e9 / HBM_PEAK_GBS}}
    if not args.no_forms:
        for prec in ("fp64", "mixed", "plain"):
            if prec != prob.precision:
                log(f"step 3: the {prec} form of the same workload")
                f = fv_form(channel_problem(tuple(args.n), args.ydiffusion, prec), 10, 2)
                precision_forms[prec] = {"steps_per_s": f["steps_per_s"], "frac": f["frac"]}
                if prec == "plain":
                    precision_forms[prec]["refliteral_frac"] = f["refliteral_frac"]
        rate = {k: v["steps_per_s"] for k, v in precision_forms.items()}
        precision_forms["ratios"] = {"mixed_over_plain_fp64": rate["mixed"] / rate["plain"],
                                     "fp64_over_plain_fp64": rate["fp64"] / rate["plain"],
                                     "mixed_over_fp64": rate["mixed"] / rate["fp64"]}
    mixed_err = None
    if args.mixed_check_steps > 0:
        ns = args.mixed_check_steps
        log(f"step 4: mixed vs FP64 over {ns} steps (the mixed run driven by the FP64 dt sequence)")
        mixed_err = run_mixed_accuracy(tuple(args.n), steps=ns, checkpoints=tuple(c for c in (10, ns) if c <= ns),
                                       log=log, ydiffusion=args.ydiffusion)
    cpu = None
    if not args.no_cpu_baseline:
        log("step 5: CPU baseline (FP64 oracle on the host)")
        cpu = cpu_baseline_block(prob, steps=args.cpu_steps)
    forms = None
    if not args.no_forms:
        log("step 6: the other configurations")
        forms = {}
        forms["channel_256x192x256_fp64_explicit"] = fv_form(channel_problem((256, 192, 256), "explicit", "fp64"),
                                                             50, 5)
        forms["channel_256x192x256_plain"] = fv_form(channel_problem((256, 192, 256), "explicit", "plain"), 10, 2)
        forms["tgv64_fp64"] = fv_form(tgv_problem(64), 50, 5, fixed_dt=1e-3, ic=tgv_ic)
        for prec in ("fp64", "mixed"):
            forms[f"hit512_{prec}"] = run_spectral(512, 170, prec, 5, 1)
        forms["cnn_train_16x64cube"] = run_cnn_train(10, 2)
        forms["les64_closure"] = run_les_closure(20, 2)
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
        "dtype": "f64" if prob.precision != "mixed" else "f32 storage / f64 Thomas+reductions",
        "data": "synthetic (laminar profile + seeded splitmix64 perturbation)",
        "config": {"workload": f"FV channel {prob.n[0]}x{prob.n[1]}x{prob.n[2]} on 4pi x 2 x 4pi/3, walls in y "
                               f"(tanh 1.5), Re_tau~180, {prob.ydiffusion} y-diffusion"
                               + (" (Crank-Nicolson)" if prob.implicit else "") + f", {prob.precision}, CFL 0.5",
                   "n": list(prob.n), "ydiffusion": prob.ydiffusion, "precision": prob.precision},
        "mcell_updates_per_s": N * sps / 1e6,
        "roofline": {"bound": "hbm", "kernel": FAMILIES[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(prob, FAMILIES[dom], cn_form if dom == 5 else None),
                     "algorithmic_bytes": fam_bytes, "attainable_gbs": copy_gbs},
        "step_roofline": {"model": "minimal-pass", "units_per_step": units, "bytes_per_step": step_bytes,
                          "achieved_gbs": step_bytes * sps / 1e9, "frac": step_bytes * sps / 1e9 / HBM_PEAK_GBS,
                          "frac_of_attainable": step_bytes * sps / 1e9 / copy_gbs if copy_gbs else None},
        "kernels": fam_summary,
        "precision_forms": precision_forms,
        "mixed_vs_fp64": mixed_err,
        "les_error": closure_error_record(),
        "stream_copy_gbs": copy_gbs,
        "checks": {"maxdiv": r["maxdiv"], "kinetic_energy": r["ke"]},
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
