"""Closure LES on the GPU: the learned-stress step against the no-model step
and against an FP64 restatement (oracle RHS + numpy CNN), the device dataset
path, and what a trained network gains over no model on held-out samples."""
import math

import numpy as np
import pytest

import distributednavierstokes_amd as dns
from distributednavierstokes_amd import closure
from distributednavierstokes_amd.cnn import SFSNet
from distributednavierstokes_amd.les import lessampler
from distributednavierstokes_amd.problem import SpectralProblem
from oracle import oracle_cnn as oc
from oracle.oracle import OracleSpectral

pytestmark = pytest.mark.gpu

WRAY_A = (8 / 15, 5 / 12, 3 / 4)
WRAY_B = (1 / 4, 0.0, 0.0)
WRAY_DC = (8 / 15, 2 / 15, 1 / 3)


def les_case(n=16, visc=2e-3, seed=1):
    prob = SpectralProblem(n=n, kcut=n // 3, visc=visc)
    o = OracleSpectral(prob)
    uh = o.randomfield(seed=seed, kpeak=2)
    s = dns.spectral_setup(n=n, kcut=n // 3, visc=visc)
    return o, s, uh


def gpu_step(s, uh, dt, net=None):
    s.upload(uh)
    out = np.zeros(1)
    lib = dns._lib.lib()
    if net is None:
        dns._lib.check(lib.dns_spec_step(s.handle, dt, dns._lib.dptr(out)))
    else:
        dns._lib.check(lib.dns_les_closure_step(s.handle, net.handle, dt, dns._lib.dptr(out)))
    return s.download(), float(out[0])


def test_zero_network_step_is_the_plain_step():
    o, s, uh = les_case()
    net = SFSNet(M=16, batch=1, seed=0)
    net.set_params(np.zeros(net.nparams, dtype=np.float32))
    u_plain, v_plain = gpu_step(s, uh, 0.01)
    u_cl, v_cl = gpu_step(s, uh, 0.01, net)
    assert np.array_equal(u_plain, u_cl)
    assert v_plain == v_cl


def oracle_closure_step(o, uh, dt, p, su, stau, visc):
    """Wray IF-RK3 with du = oracle RHS − ik_j F[τ(ū)] (τ from the FP64 CNN on
    BF16-rounded weights and inputs), stage structure of step.jl:22-45."""
    n, K, m = o.n, o.K, o.m
    kf = 2 * math.pi / o.prob.l
    kx = kf * np.arange(K)[:, None, None]
    f = np.array([j if j <= o.kcut else j - m for j in range(m)])
    ky = kf * f[None, :, None]
    kz = kf * f[None, None, :]
    kk = kx ** 2 + ky ** 2 + kz ** 2
    pb = oc.bf16_round(p)

    def closure_du(u):
        v = o.spec_to_phys(u)
        x = oc.bf16_round((v / su).transpose(2, 1, 0, 3)[None])
        y, _, _ = oc.forward(pb, x)
        tau = stau * y[0].transpose(2, 1, 0, 3)
        t1 = o.phys_to_spec(np.asfortranarray(tau[..., 0:3]))
        t2 = o.phys_to_spec(np.asfortranarray(tau[..., 3:6]))
        t = [t1[..., 0], t1[..., 1], t1[..., 2], t2[..., 0], t2[..., 1], t2[..., 2]]  # xx yy zz xy yz zx
        mi = lambda k, z: -1j * k * z  # noqa: E731
        du = np.zeros_like(u)
        du[..., 0] = mi(kx, t[0]) + mi(ky, t[3]) + mi(kz, t[5])
        du[..., 1] = mi(ky, t[1]) + mi(kx, t[3]) + mi(kz, t[4])
        du[..., 2] = mi(kz, t[2]) + mi(ky, t[4]) + mi(kx, t[5])
        return du

    u = uh.copy(order="F")
    ustart = uh.copy(order="F")
    for i in range(3):
        du, _ = o.rhs(u)
        du = du + closure_du(u)
        phi = np.exp(-visc * WRAY_DC[i] * dt * kk)[..., None]
        unew = phi * (ustart + WRAY_A[i] * dt * du)
        ustart = phi * (ustart + WRAY_B[i] * dt * du)
        u = o.project(np.asfortranarray(unew))
    return u


def test_oracle_stage_restatement_is_the_oracle_step():
    """With a zero network the restated stage loop reproduces the oracle step."""
    o, _, uh = les_case()
    p = np.zeros(oc.nparams())
    u1 = oracle_closure_step(o, uh, 0.01, p, 1.0, 1.0, 2e-3)
    u2 = uh.copy(order="F")
    o.step(u2, 0.01)
    assert np.abs(u1 - u2).max() < 1e-14 * np.abs(u2).max()


def test_closure_increment_matches_fp64_restatement():
    o, s, uh = les_case()
    rng = np.random.default_rng(5)
    net = SFSNet(M=16, batch=1, seed=0)
    p = (rng.standard_normal(net.nparams) * 0.05).astype(np.float32)
    net.set_params(p)
    su, stau = 0.7, 0.05
    net.scales = (su, stau)
    dt = 0.01
    u_plain, _ = gpu_step(s, uh, dt)
    u_cl, _ = gpu_step(s, uh, dt, net)
    ref = oracle_closure_step(o, uh, dt, p.astype(np.float64), su, stau, 2e-3)
    ref_plain = uh.copy(order="F")
    o.step(ref_plain, dt)
    inc, inc_ref = u_cl - u_plain, ref - ref_plain
    assert np.linalg.norm(inc_ref) > 1e-3 * np.linalg.norm(ref_plain)  # the closure does something
    assert np.linalg.norm(inc - inc_ref) < 3e-2 * np.linalg.norm(inc_ref)


def test_device_dataset_matches_host_samples():
    n, M = 48, 16
    s = dns.spectral_setup(n=n, visc=5e-3)
    o = OracleSpectral(SpectralProblem(n=n, visc=5e-3))
    uh = o.randomfield(seed=3, kpeak=3)
    sam = lessampler(s, M=M).sample(uh)
    a, b = SFSNet(M=M, batch=1, seed=0), SFSNet(M=M, batch=1, seed=0)
    sam.append_to(a)
    ub, tau = sam.physical()
    b.add_sample(ub, tau)
    assert a.nsamples == b.nsamples == 1
    sa, sb = a.fit_scales(), b.fit_scales()
    assert sa == sb
    assert abs(sa[0] - math.sqrt(np.mean(ub ** 2))) < 1e-6 * sa[0]
    assert abs(sa[1] - math.sqrt(np.mean(tau ** 2))) < 1e-6 * sa[1]
    assert np.array_equal(a.predict(0), b.predict(0))


def test_trained_network_beats_no_model_on_held_out_samples():
    """DNS 64³ (forced, random IC) → 24 SFS samples on 16³ → train on 20 →
    held-out MSE and τ correlation against the no-model closure (τ = 0)."""
    n, M = 64, 16
    s = dns.spectral_setup(n=n, visc=2e-3)
    uh = dns.spectral_randomfield_(dns.specvelocity(s), s, kpeak=3, seed=0)
    forcing = dns.shellforcing(uh, s, shells=(1, 2))
    net = SFSNet(M=M, batch=4, seed=0)
    sam, _, times, kept = closure.collect_pairs(s, uh, net, M=M, nsnap=24, every=8, spinup=40, forcing=forcing,
                                                keep_physical=(20, 21, 22, 23))
    assert net.nsamples == 24 and len(times) == 24
    su, stau = net.fit_scales()
    base = np.mean([np.mean((kept[q][1] / stau) ** 2) for q in range(20, 24)])
    before = net.evaluate(20, 4)
    losses = net.train(300, lr=1e-3, seed=0, ntrain=20)
    after = net.evaluate(20, 4)
    corr = np.mean([closure.stress_correlation(net.predict(q), kept[q][1]) for q in range(20, 24)])
    print(f"held-out MSE: no model {base:.4f}, untrained {before:.4f}, trained {after:.4f}; "
          f"train loss {losses[0]:.4f} -> {losses[-1]:.4f}; τ correlation {corr:.3f}")
    assert losses[-10:].mean() < 0.7 * losses[:10].mean()
    assert after < 0.7 * base
    assert corr > 0.5
