"""Filtered-DNS / sub-filter-stress sampling on the GPU (the recipe of
``sfs_sample!``, src/spectral/les.jl:313-382, Gaussian bank cell).

``lessampler(dns_setup, M)`` builds an LES spectral handle on the coarse grid
(n = M, kcut = M ÷ 3: exactly the modes the 2/3 truncation keeps); each
``sample(uh)`` computes ū and the trace-free τ on the GPU from the DNS state
and keeps them on the LES handle.  Diagnostics follow les_spectrum
(les.jl:119-131) and les_filter_reynolds (les.jl:106-116) through GPU
reductions on the LES handle.  ``sfswriter`` keeps the samples of a run as an
on-disk dataset with the reference's key names (below); ``sfs_to_network``
loads one into the closure network's device dataset.
"""
from __future__ import annotations

import math
import os
import time as _time
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


# ------------------------------------------------------------ SFS dataset
# The artifact schemas of sfswriter (src/spectral/les.jl:383-511) as NumPy
# .npz archives instead of JLD2 (Julia-specific), with the reference's key
# names and directory layout (a single Gaussian bank cell: `delta=<Δf>/`):
#
#   <dir>/dns_meta.npz      times (nsnap), spectra_dns (kcut, nsnap): shells
#                           1..kcut of the DNS, statistics_dns (12, nsnap) with
#                           statistics_dns_keys (the K41 fields of
#                           spectral_stats), t_int, walltime
#   <dir>/delta=<Δf>/fields.npz
#                           inputs.x|y|z  ū, outputs.xx|yy|zz|xy|yz|zx  τ
#                           (trace-free): complex (M/2+1, M, M, nsnap) in
#                           column-major order, the LES rfft layout with wrapped
#                           ky, kz and zeros outside the retained 2/3 modes,
#                           û = F[u]/M³; redelta (nsnap), Δ, Δ_factor, visc,
#                           kernel ("gaussian"), M, delta_eta (NaN: not set)
#   <dir>/delta=<Δf>/les_meta.npz
#                           spectra_les (M÷3, nsnap), redelta_mean
#
# Every archive is written to a temporary name and renamed when complete;
# no array needs pickling (load with allow_pickle=False).

def _expand_rfft(a, M):
    """Compact (K, m, m, C) LES spectra -> (M/2+1, M, M, C), wrapped ky/kz."""
    K, m = a.shape[0], a.shape[1]
    kc = K - 1
    out = np.zeros((M // 2 + 1, M, M, a.shape[3]), dtype=np.complex128, order="F")
    idx = np.array([j if j <= kc else j - m + M for j in range(m)])
    out[np.ix_(range(K), idx, idx)] = a
    return out


def _compact_rfft(a, M):
    """Inverse of _expand_rfft for the retained modes."""
    kc = M // 3
    m = 2 * kc + 1
    idx = np.array([j if j <= kc else j - m + M for j in range(m)])
    return np.asfortranarray(a[np.ix_(range(kc + 1), idx, idx)])


def _savez_atomic(path, **arrays):
    tmp = path + ".tmp.npz"
    np.savez(tmp, **arrays)
    os.replace(tmp, path)


def sfswriter(dns, dir, times, M=64, delta_fac=2.0, protectshells=2):
    """``sfswriter(s; dir, times, nles, filters, protectshells)`` --
    src/spectral/les.jl:463-511 for one Gaussian bank cell: a processor that
    samples (ū, τ) on the GPU (``dns_sfs_sample``) each time t reaches an
    entry of ``times`` (pass them as ``tstops``) and, after the last, writes
    the archives described above.  Returns the processor; its ``sampler``
    attribute is the ``LESSampler``."""
    from .spectral import spectral_spectrum, spectral_stats
    from .spectral_io import STATS_KEYS

    sam = LESSampler(dns, M, delta_fac, protectshells)
    times = sorted(float(t) for t in times)
    rec = dict(t=[], spec_dns=[], stats=[], ub=[], tau=[], redelta=[], spec_les=[])
    inext = [0]
    tstart = _time.time()

    def finish():
        os.makedirs(dir, exist_ok=True)
        st = np.array([[getattr(s, k) for k in STATS_KEYS] for s in rec["stats"]]).T
        _savez_atomic(os.path.join(dir, "dns_meta.npz"), times=np.array(rec["t"]),
                      spectra_dns=np.asfortranarray(np.array(rec["spec_dns"]).T), statistics_dns=np.asfortranarray(st),
                      statistics_dns_keys=np.array(STATS_KEYS), t_int=np.float64(rec["stats"][0].t_int),
                      walltime=np.float64(_time.time() - tstart))
        cdir = os.path.join(dir, f"delta={sam.delta_fac}")
        os.makedirs(cdir, exist_ok=True)
        fields = {}
        for c, name in enumerate("xyz"):
            fields[f"inputs.{name}"] = np.asfortranarray(np.stack([u[..., c] for u in rec["ub"]], -1))
        for c, name in enumerate(("xx", "yy", "zz", "xy", "yz", "zx")):
            fields[f"outputs.{name}"] = np.asfortranarray(np.stack([t[..., c] for t in rec["tau"]], -1))
        _savez_atomic(os.path.join(cdir, "fields.npz"), **fields, redelta=np.array(rec["redelta"]),
                      **{"Δ": np.float64(sam.delta), "Δ_factor": np.float64(sam.delta_fac)},
                      visc=np.float64(dns.visc), kernel=np.array("gaussian"), M=np.int64(M),
                      delta_eta=np.float64(np.nan))
        _savez_atomic(os.path.join(cdir, "les_meta.npz"),
                      spectra_les=np.asfortranarray(np.array(rec["spec_les"]).T),
                      redelta_mean=np.float64(np.mean(rec["redelta"])))

    def proc(state, s):
        while inext[0] < len(times) and state.t >= times[inext[0]] - 1e-10 * (abs(times[inext[0]]) + 1):
            rec["t"].append(state.t)
            rec["stats"].append(spectral_stats(state.uh, s))
            rec["spec_dns"].append(np.asarray(spectral_spectrum(state.uh, s).E[1:], dtype=np.float64))
            sam.sample(None)
            ub, tau = sam.spectral()
            rec["ub"].append(_expand_rfft(ub, M))
            rec["tau"].append(_expand_rfft(tau, M))
            rec["redelta"].append(sam.filter_reynolds())
            rec["spec_les"].append(np.asarray(sam.spectrum(), dtype=np.float64))
            inext[0] += 1
            if inext[0] == len(times):
                finish()
        return None

    proc.sampler = sam
    return proc


def read_sfs(dir, delta_fac=2.0):
    """The archives written by ``sfswriter`` as dicts of arrays
    (``dns_meta``, ``fields``, ``les_meta``)."""
    cdir = os.path.join(dir, f"delta={float(delta_fac)}")
    out = {}
    for key, path in (("dns_meta", os.path.join(dir, "dns_meta.npz")), ("fields", os.path.join(cdir, "fields.npz")),
                      ("les_meta", os.path.join(cdir, "les_meta.npz"))):
        with np.load(path, allow_pickle=False) as z:
            out[key] = {k: z[k] for k in z.files}
    return out


def sfs_physical_pairs(dir, delta_fac=2.0):
    """The (ū, τ) pairs of an ``sfswriter`` dataset as physical fields on the
    M³ grid, (M, M, M, 3) and (M, M, M, 6) per snapshot: u = M³ · irfft(û)."""
    f = read_sfs(dir, delta_fac)["fields"]
    M = int(f["M"])

    def phys(a):
        return np.transpose(np.fft.irfftn(np.transpose(a), s=(M, M, M), axes=(0, 1, 2))) * M ** 3

    for k in range(f["inputs.x"].shape[-1]):
        ub = np.stack([phys(f[f"inputs.{c}"][..., k]) for c in "xyz"], -1)
        tau = np.stack([phys(f[f"outputs.{c}"][..., k]) for c in ("xx", "yy", "zz", "xy", "yz", "zx")], -1)
        yield ub, tau


def sfs_to_network(net, dir, delta_fac=2.0):
    """Loads the (ū, τ) pairs of an ``sfswriter`` dataset into the device
    dataset of the closure network ``net``, so training can be redone without
    the DNS; returns the number of pairs."""
    n = 0
    for ub, tau in sfs_physical_pairs(dir, delta_fac):
        net.add_sample(ub, tau)
        n += 1
    return n
