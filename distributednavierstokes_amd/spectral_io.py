"""Snapshot / restart I/O and run-archive processors of the pseudo-spectral
solver (src/spectral/io.jl), on one GPU.

- ``spectral_save`` / ``spectral_load_`` write and read ``<prefix>.bin`` (the
  (kcut+1, m, m, 3) complex state, column-major in global index order,
  component after component, Complex{Float64} or Complex{Float32} in the
  mixed form) with the ``<prefix>.toml`` sidecar, through the library's
  ``dns_spec_save`` / ``dns_spec_load`` (atomic: both files appear under
  their final names only when complete).  Byte layout and sidecar keys are
  those of src/spectral/io.jl:29-104, so a file written here restarts the
  reference and the reverse.
- ``checkpointlist``, ``spectral_latest``, ``checkpointer``, ``snapshotsaver``
  and ``statswriter`` are the reference's processors and helpers
  (src/spectral/io.jl:107-262), host logic over those two calls and the
  device diagnostics.
"""
from __future__ import annotations

import math
import os
import re
import time as _time

from ._lib import check, lib
from .problem import DNSError
from .spectral import DeviceState, spectral_stats

__all__ = ["spectral_save", "spectral_load_", "checkpointlist", "spectral_latest", "checkpointer",
           "snapshotsaver", "statswriter"]


def _toml_value(v):
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, int):
        return str(int(v))
    if isinstance(v, float):
        if math.isnan(v):
            return "nan"
        if math.isinf(v):
            return "inf" if v > 0 else "-inf"
        r = repr(float(v))
        return r if any(c in r for c in ".en") else r + ".0"
    if isinstance(v, str):
        return '"' + v.replace("\\", "\\\\").replace('"', '\\"') + '"'
    if isinstance(v, (list, tuple)):
        return "[" + ", ".join(_toml_value(x) for x in v) + "]"
    if hasattr(v, "tolist"):
        return _toml_value(v.tolist())
    raise DNSError(f"meta value {v!r} has no TOML form")


def _meta_toml(meta):
    if not meta:
        return None
    return "".join(f"{k} = {_toml_value(meta[k])}\n" for k in sorted(meta))


def spectral_save(prefix, uh, s, time=0.0, step=0, meta=None):
    """``spectral_save(prefix, uh, s; time, step, meta)`` -- src/spectral/io.jl:29-75.
    Returns ``prefix``."""
    s.upload(uh)
    m = _meta_toml(meta)
    check(lib().dns_spec_save(s.handle, os.fsencode(prefix), float(time), int(step),
                              m.encode() if m is not None else None))
    return prefix


def spectral_load_(uh, prefix, s):
    """``spectral_load!(uh, prefix, s)`` -- src/spectral/io.jl:77-104: the
    state is loaded into ``uh`` (a host array, or the device state); returns
    ``(time, step, meta)``."""
    import ctypes as ct

    import numpy as np

    t = np.zeros(1)
    st = ct.c_int64(0)
    check(lib().dns_spec_load(s.handle, os.fsencode(prefix), t.ctypes.data_as(ct.POINTER(ct.c_double)),
                              ct.byref(st)))
    try:
        import tomli

        with open(f"{prefix}.toml", "rb") as f:
            meta = tomli.load(f).get("meta", {})
    except ImportError:  # the meta table needs a TOML reader
        meta = {}
    if not isinstance(uh, DeviceState):
        s.download(uh)
    return float(t[0]), int(st.value), meta


def checkpointlist(prefix):
    """Checkpoint prefixes ``<prefix>_<step>`` with complete file pairs,
    ascending by step (src/spectral/io.jl:107-121)."""
    d = os.path.dirname(prefix) or "."
    base = os.path.basename(prefix)
    if not os.path.isdir(d):
        return []
    pat = re.compile("^" + re.escape(base) + r"_(\d+)\.toml$")
    steps = []
    for f in os.listdir(d):
        mm = pat.match(f)
        if mm and os.path.isfile(os.path.join(d, f"{base}_{mm.group(1)}.bin")):
            steps.append(int(mm.group(1)))
    return [os.path.join(d, "%s_%09d" % (base, st)) for st in sorted(steps)]


def spectral_latest(prefix, s=None):
    """``spectral_latest(prefix, s)`` -- src/spectral/io.jl:123-133: the newest
    complete checkpoint, or None."""
    lst = checkpointlist(prefix)
    return lst[-1] if lst else None


def checkpointer(prefix, interval=900.0, stopfile=None, keep=2, meta=None):
    """``checkpointer(prefix; interval, stopfile, keep, meta)`` --
    src/spectral/io.jl:135-180: a processor writing ``<prefix>_<step>`` every
    ``interval`` wall-clock seconds and when ``stopfile`` exists (then it
    returns ``"stop"``); only the newest ``keep`` checkpoints are kept."""
    tlast = [_time.time()]

    def proc(state, s):
        due = _time.time() - tlast[0] >= interval
        stop = stopfile is not None and os.path.isfile(stopfile)
        if not (due or stop):
            return None
        spectral_save("%s_%09d" % (prefix, state.n), state.uh, s, time=state.t, step=state.n, meta=meta)
        tlast[0] = _time.time()
        old = checkpointlist(prefix)
        for p in old[:len(old) - min(keep, len(old))]:
            for ext in (".bin", ".toml"):
                try:
                    os.remove(p + ext)
                except FileNotFoundError:
                    pass
        return "stop" if stop else None

    return proc


def snapshotsaver(prefix, times, meta=None, stats=True, tstart=-math.inf):
    """``snapshotsaver(prefix; times, meta, stats, tstart)`` --
    src/spectral/io.jl:182-229: saves ``<prefix>_0001``, ``_0002``, ... the
    first time t reaches each entry of ``times`` (K41 numbers in the meta when
    ``stats``); entries before ``tstart`` are skipped."""
    times = sorted(float(t) for t in times)
    i0 = next((i for i, t in enumerate(times) if t >= tstart - 1e-10 * (abs(t) + 1)), len(times))
    inext = [i0]

    def proc(state, s):
        while inext[0] < len(times) and state.t >= times[inext[0]] - 1e-10 * (abs(times[inext[0]]) + 1):
            m = dict(meta or {})
            if stats:
                st = spectral_stats(state.uh, s)
                m.update(eta=st.l_kol, t_int=st.t_int, e=st.e, diss=st.diss)
            spectral_save("%s_%04d" % (prefix, inext[0] + 1), state.uh, s, time=state.t, step=state.n, meta=m)
            inext[0] += 1
        return None

    return proc


STATS_KEYS = ("e", "uavg", "diss", "l_int", "l_tay", "l_kol", "t_int", "t_tay", "t_kol", "Re_int", "Re_tay",
              "kmax_eta")


def statswriter(file, nupdate=10):
    """``statswriter(; file, nupdate)`` -- src/spectral/io.jl:231-262: appends
    one CSV row ``t, step, <K41 statistics>, walltime`` every ``nupdate``
    steps; the header only when the file is new; flushed per row."""
    io = [None]

    def proc(state, s):
        if state.n % nupdate != 0:
            return None
        st = spectral_stats(state.uh, s)
        if io[0] is None:
            d = os.path.dirname(os.path.abspath(file))
            os.makedirs(d, exist_ok=True)
            fresh = not os.path.isfile(file) or os.path.getsize(file) == 0
            io[0] = open(file, "a")
            if fresh:
                io[0].write(",".join(("t", "step") + STATS_KEYS + ("walltime",)) + "\n")
        vals = [repr(float(state.t)), str(int(state.n))] + [repr(float(getattr(st, k))) for k in STATS_KEYS]
        io[0].write(",".join(vals + [repr(_time.time())]) + "\n")
        io[0].flush()
        return None

    proc.close = lambda: io[0] is not None and io[0].close()
    return proc
