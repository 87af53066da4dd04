"""Filtered-DNS / sub-filter-stress sampling on the GPU (the recipe of
``sfs_sample!``, src/spectral/les.jl:313-382, Gaussian bank cell).

``lessampler(dns_setup, M)`` builds an LES spectral handle on the coarse grid
(n = M, kcut = M ÷ 3: exactly the modes the 2/3 truncation keeps); each
``sample(uh)`` computes ū and the trace-free τ on the GPU from the DNS state
and keeps them on the LES handle.  Diagnostics follow les_spectrum
(les.jl:119-131) and les_filter_reynolds (les.jl:106-116) through GPU
reductions on the LES handle.
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np

from ._lib import check, dptr, lib
from .problem import DNSError
from .spectral import DeviceState, _shells, spectral_setup


class LESSampler:
    def __init__(self, dns, M=64, delta_fac=2.0, protectshells=2):
        if M % 2 or M // 2 > dns.kcut:
            raise DNSError("LES grid must be even with M/2 <= the DNS kcut")
        self.dns, self.M, self.delta_fac, self.protectshells = dns, M, float(delta_fac), int(protectshells)
        self.delta = self.delta_fac * dns.l / M
        self.les = spectral_setup(n=M, kcut=M // 3, l=dns.l, visc=dns.visc, T=dns.T)
        self.nsamples = 0

    def sample(self, uh=None):
        """ū, τ of the DNS state ``uh`` (host array or the device state)."""
        if uh is not None:
            self.dns.upload(uh)
        check(lib().dns_sfs_sample(self.dns.handle, self.les.handle, self.delta_fac, self.protectshells))
        self.nsamples += 1
        return self

    def append_to(self, net):
        """Adds the last sample's physical (ū, τ) to the device dataset of the
        closure network ``net`` (an ``SFSNet`` with M equal to this grid)."""
        check(lib().dns_sfs_append(self.les.handle, net.handle))
        return self

    def spectral(self):
        """ū (K, m, m, 3) and τ̂ (K, m, m, 6) on the LES compact grid."""
        s = self.les
        ub = np.zeros(s.state_shape, dtype=np.complex128, order="F")
        tau = np.zeros(s.state_shape[:3] + (6,), dtype=np.complex128, order="F")
        check(lib().dns_sfs_get(s.handle, dptr(ub), dptr(tau)))
        return ub, tau

    def physical(self):
        """ū (M, M, M, 3) and τ (M, M, M, 6) in physical space."""
        M = self.M
        ub = np.zeros((M, M, M, 3), order="F")
        tau = np.zeros((M, M, M, 6), order="F")
        check(lib().dns_sfs_physical(self.les.handle, dptr(ub), dptr(tau)))
        return ub, tau

    def spectrum(self):
        """les_spectrum of ū: shells 1..M÷3 (les.jl:119-131)."""
        return _shells(self.les, range(1, self.M // 3 + 1))

    def filter_reynolds(self):
        """Re_Δ = Δ² √⟨|∇ū|²⟩ / ν (les.jl:106-116)."""
        s = self.les
        out = np.zeros(1)
        check(lib().dns_spec_dissipation(s.handle, dptr(out)))
        a2 = out[0] / s.visc
        return self.delta ** 2 * math.sqrt(a2) / self.dns.visc


def lessampler(dns, M=64, delta_fac=2.0, protectshells=2):
    return LESSampler(dns, M, delta_fac, protectshells)
