// Host-side helpers shared by the FV and spectral handles: twiddle tables,
// device uploads/allocations and per-family HIP-event launch timing.
#pragma once

#include <cmath>
#include <vector>

#include "common.hpp"

namespace dns {

// tw[t] = exp(-2πi t/n), t = 0..len-1, rounded from long double.
template <typename T> inline std::vector<cplx<T>> twiddles(int n, int len)
{
    const long double twopi = 6.283185307179586476925286766559005768L;
    std::vector<cplx<T>> tw(len);
    for (int t = 0; t < len; ++t) {
        long double a = twopi * (long double)t / (long double)n;
        tw[t] = mk<T>((T)cosl(a), (T)(-sinl(a)));
    }
    return tw;
}

// Device copy of a host table.  hipMemcpy from pageable memory may return
// once the data is staged, before the DMA lands, and the handles' streams are
// non-blocking: wait for it before any kernel can read the table.
template <typename V> inline int dupload(V **dptr, const std::vector<V> &h)
{
    *dptr = nullptr;
    size_t bytes = std::max<size_t>(1, h.size()) * sizeof(V);
    DNS_HIP_CHECK(hipMalloc((void **)dptr, bytes));
    if (!h.empty()) DNS_HIP_CHECK(hipMemcpy(*dptr, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice));
    DNS_HIP_CHECK(hipDeviceSynchronize());
    return DNS_OK;
}

// Zero-initialised device allocation.  The zero fill runs on the null stream,
// which has no implicit ordering with the handles' non-blocking streams, so
// it is waited for here: otherwise a copy or kernel enqueued right after on a
// handle's stream could run before (and be overwritten by) the fill.
template <typename V> inline int dalloc(V **dptr, size_t count)
{
    *dptr = nullptr;
    DNS_HIP_CHECK(hipMalloc((void **)dptr, std::max<size_t>(1, count) * sizeof(V)));
    DNS_HIP_CHECK(hipMemset(*dptr, 0, std::max<size_t>(1, count) * sizeof(V)));
    DNS_HIP_CHECK(hipDeviceSynchronize());
    return DNS_OK;
}

// A temporary device buffer released on every return path (a DNS_TRY or
// DNS_HIP_CHECK returning early included).  The release synchronises the
// device first, so work still queued on a stream never reads freed memory.
template <typename V> struct DevTemp {
    V *p = nullptr;
    DevTemp() = default;
    DevTemp(const DevTemp &) = delete;
    DevTemp &operator=(const DevTemp &) = delete;
    int alloc(size_t count) { return dalloc(&p, count); }
    ~DevTemp()
    {
        if (p) {
            hipDeviceSynchronize();
            hipFree(p);
        }
    }
};

// Per-family HIP-event timing of kernel launches on the handle's stream.
struct LaunchTimer {
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[8];
    double ms[8] = {0};
    long long cnt[8] = {0};
    hipStream_t st = nullptr;
    hipEvent_t cur = nullptr;
    void begin(int fam)
    {
        (void)fam;
        if (!on) return;
        hipEventCreate(&cur);
        hipEventRecord(cur, st);
    }
    void end(int fam)
    {
        if (!on) return;
        hipEvent_t e;
        hipEventCreate(&e);
        hipEventRecord(e, st);
        ev[fam].push_back({cur, e});
        if (ev[fam].size() > 4096) flush();
    }
    void flush()
    {
        hipStreamSynchronize(st);
        for (int f = 0; f < 8; ++f) {
            for (auto &p : ev[f]) {
                float t = 0;
                hipEventElapsedTime(&t, p.first, p.second);
                ms[f] += t;
                cnt[f] += 1;
                hipEventDestroy(p.first);
                hipEventDestroy(p.second);
            }
            ev[f].clear();
        }
    }
};

}  // namespace dns
