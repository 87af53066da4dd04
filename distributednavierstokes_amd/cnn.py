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
