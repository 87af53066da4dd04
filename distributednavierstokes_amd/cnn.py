"""Learned sub-filter-stress closure: the 4-layer 3×3×3 periodic CNN
(3 → 64 → 64 → 64 → 6 channels, erf-GELU) on the GPU via the dns_cnn C ABI.
BF16 weights and activations with FP32 accumulation; Adam on FP32 masters.
Batches are numpy float32 arrays [b, z, y, x, c]."""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from ._lib import check, dptr, fptr, lib
from .problem import DNSError


class SFSNet:
    def __init__(self, M=64, batch=16, seed=0):
        self.M, self.batch = int(M), int(batch)
        h = ct.c_void_p()
        check(lib().dns_cnn_create(self.M, self.batch, int(seed), ct.byref(h)))
        self.handle = h
        n = ct.c_int64()
        check(lib().dns_cnn_num_params(h, ct.byref(n)))
        self.nparams = int(n.value)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib._LIB is not None:
            _lib._LIB.dns_cnn_destroy(h)
            self.handle = None

    def _check_batch(self, x, c):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.shape != (self.batch, self.M, self.M, self.M, c):
            raise DNSError(f"batch shape must be {(self.batch, self.M, self.M, self.M, c)}")
        return x

    def params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        check(lib().dns_cnn_get_params(self.handle, fptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.nparams:
            raise DNSError("wrong parameter count")
        check(lib().dns_cnn_set_params(self.handle, fptr(p)))

    def grads(self):
        g = np.zeros(self.nparams, dtype=np.float32)
        check(lib().dns_cnn_get_grads(self.handle, fptr(g)))
        return g

    def forward(self, x):
        x = self._check_batch(x, 3)
        y = np.zeros(x.shape[:4] + (6,), dtype=np.float32)
        check(lib().dns_cnn_forward(self.handle, fptr(x), fptr(y)))
        return y

    def loss_grad(self, x, t):
        x, t = self._check_batch(x, 3), self._check_batch(t, 6)
        out = np.zeros(1)
        check(lib().dns_cnn_loss_grad(self.handle, fptr(x), fptr(t), dptr(out)))
        return float(out[0])

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        check(lib().dns_cnn_adam_step(self.handle, lr, beta1, beta2, eps))

    # ------------------------------------------------------------- dataset
    def add_sample(self, ubar, tau):
        """Adds a physical pair ū (M, M, M, 3), τ (M, M, M, 6) (x fastest)."""
        M = self.M
        if ubar.shape != (M, M, M, 3) or tau.shape != (M, M, M, 6):
            raise DNSError("sample shapes must be (M, M, M, 3) and (M, M, M, 6)")
        u = np.ascontiguousarray(np.asarray(ubar, dtype=np.float32).ravel(order="F"))
        t = np.ascontiguousarray(np.asarray(tau, dtype=np.float32).ravel(order="F"))
        check(lib().dns_cnn_add_sample(self.handle, fptr(u), fptr(t)))

    @property
    def nsamples(self):
        n = ct.c_int32()
        check(lib().dns_cnn_dataset_size(self.handle, ct.byref(n)))
        return int(n.value)

    def fit_scales(self):
        su, st = np.zeros(1), np.zeros(1)
        check(lib().dns_cnn_fit_scales(self.handle, dptr(su), dptr(st)))
        return float(su[0]), float(st[0])

    @property
    def scales(self):
        su, st = np.zeros(1), np.zeros(1)
        check(lib().dns_cnn_get_scales(self.handle, dptr(su), dptr(st)))
        return float(su[0]), float(st[0])

    @scales.setter
    def scales(self, v):
        check(lib().dns_cnn_set_scales(self.handle, float(v[0]), float(v[1])))

    def train(self, nsteps, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, seed=0, ntrain=0):
        """nsteps Adam steps on shuffled batches of the first ``ntrain``
        dataset samples (all if 0); returns the normalised loss before each step."""
        losses = np.zeros(max(1, int(nsteps)))
        check(lib().dns_cnn_train(self.handle, int(nsteps), lr, beta1, beta2, eps, int(seed), int(ntrain),
                                  dptr(losses)))
        return losses[: int(nsteps)]

    def evaluate(self, first=0, count=None):
        """Normalised MSE over dataset samples [first, first + count); the
        no-model closure (τ = 0) scores the mean square of the normalised τ."""
        count = self.nsamples - first if count is None else count
        out = np.zeros(1)
        check(lib().dns_cnn_eval(self.handle, int(first), int(count), dptr(out)))
        return float(out[0])

    def predict(self, sample):
        """Predicted τ (M, M, M, 6) of dataset sample ``sample`` (physical units)."""
        M = self.M
        out = np.zeros(6 * M ** 3, dtype=np.float32)
        check(lib().dns_cnn_predict(self.handle, int(sample), fptr(out)))
        return out.reshape((M, M, M, 6), order="F")

    def timing(self, enable=True):
        check(lib().dns_cnn_timing(self.handle, 1 if enable else 0))

    def timing_read(self, family):
        ms, cnt = np.zeros(1), ct.c_int64()
        check(lib().dns_cnn_timing_read(self.handle, int(family), dptr(ms), ct.byref(cnt)))
        return float(ms[0]), int(cnt.value)
