// Shared host/device helpers for the gfx950 kernels of libdnsrocm.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/dns_rocm.h"

namespace dns {

// ---------------------------------------------------------------- errors

void set_error(const char *fmt, ...);

#define DNS_HIP_CHECK(call)                                                             \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::dns::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                             __FILE__, __LINE__);                                       \
            return DNS_EHIP;                                                            \
        }                                                                               \
    } while (0)

#define DNS_TRY(call)                  \
    do {                               \
        int rc_ = (call);              \
        if (rc_ != DNS_OK) return rc_; \
    } while (0)

// --------------------------------------------------------- complex types

template <typename T> struct c2t;
template <> struct c2t<double> { using type = double2; };
template <> struct c2t<float> { using type = float2; };
template <typename T> using cplx = typename c2t<T>::type;

template <typename T> __host__ __device__ inline cplx<T> mk(T a, T b)
{
    cplx<T> r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename T> __device__ inline cplx<T> cadd(cplx<T> a, cplx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ inline cplx<T> csub(cplx<T> a, cplx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b)
{
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T> __device__ inline cplx<T> cmulc(cplx<T> a, cplx<T> b)
{
    return mk<T>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T> __device__ inline cplx<T> cscale(cplx<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
template <typename T> __device__ inline cplx<T> cconj(cplx<T> a) { return mk<T>(a.x, -a.y); }
// i * a
template <typename T> __device__ inline cplx<T> cmuli(cplx<T> a) { return mk<T>(-a.y, a.x); }

// ---------------------------------------------- compensated FP64 sums

// Error-free transformation: s + e == a + b exactly.
__device__ inline void two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Double-double accumulator (hi, lo).
struct dd {
    double hi, lo;
};
__device__ inline dd dd_add(dd a, dd b)
{
    double s, e;
    two_sum(a.hi, b.hi, s, e);
    e += a.lo + b.lo;
    double hi = s + e;
    double lo = e - (hi - s);
    return dd{hi, lo};
}
__device__ inline dd dd_add1(dd a, double b)
{
    double s, e;
    two_sum(a.hi, b, s, e);
    e += a.lo;
    double hi = s + e;
    double lo = e - (hi - s);
    return dd{hi, lo};
}

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// Workgroups are dispatched round-robin over the 8 XCDs of the chip, each
// with its own L2.  This bijection of linear block ids gives every XCD one
// contiguous range of the domain, so the y/z stencil neighbours a block
// reads were mostly fetched into the same L2 by the blocks before it.
__device__ inline int xcd_block(int bid, int nb)
{
    const int X = 8;
    const int q = nb / X, r = nb - q * X;
    const int xcd = bid % X, slot = bid / X;
    return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// 3D block coordinates after the XCD remap of a (gx, gy, gz) grid.
__device__ inline void xcd_block3(int &bx, int &by, int &bz)
{
    const int gx = gridDim.x, gy = gridDim.y;
    const int nb = gx * gy * gridDim.z;
    const int lb = xcd_block(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), nb);
    bx = lb % gx;
    by = (lb / gx) % gy;
    bz = lb / (gx * gy);
}
// 2D block coordinates after the XCD remap of a (gx, gy) grid: blocks that
// are consecutive in (x fastest) order land on the same XCD, so tiles that
// share 128-byte lines or neighbour rows share its L2.
__device__ inline void xcd_block2(int &bx, int &by)
{
    const int gx = gridDim.x;
    const int lb = xcd_block(blockIdx.x + gx * blockIdx.y, gx * gridDim.y);
    bx = lb % gx;
    by = lb / gx;
}

// Block-strided loop over [0, total) that keeps U independent loads in
// flight per thread: ld(idx) for U indices into registers, then st(idx, v)
// for them.  A plain loop whose body loads from HBM and stores to LDS waits
// out each load's latency in turn.
template <int U, typename V, typename LD, typename ST>
__device__ __forceinline__ void batched_for(int total, LD ld, ST st)
{
    for (int base = threadIdx.x; base < total; base += U * (int)blockDim.x) {
        V r[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int idx = base + q * (int)blockDim.x;
            if (idx < total) r[q] = ld(idx);
        }
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int idx = base + q * (int)blockDim.x;
            if (idx < total) st(idx, r[q]);
        }
    }
}

// Division by a runtime-invariant divisor d ≥ 1 as a multiply-high, add and
// shift (Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, round-up variant): with s = ceil(log2 d) and
// m = floor(2^32 (2^s − d) / d) + 1, n / d = (mulhi(n, m) + n) >> s for
// 0 ≤ n < 2^31 (the sum stays below 2^32).  A 32-bit integer division by a
// runtime value costs some twenty instructions on the GPU.
struct FastDiv {
    uint32_t m;
    int s;
};
inline FastDiv make_fastdiv(int d)
{
    FastDiv f{1u, 0};
    if (d <= 1) return f;
    while ((1u << f.s) < (uint32_t)d) ++f.s;
    f.m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << f.s) - (uint64_t)d)) / (uint64_t)d + 1);
    return f;
}
__device__ inline int fdiv(int n, FastDiv f) { return (int)((__umulhi((uint32_t)n, f.m) + (uint32_t)n) >> f.s); }

// Buffer-resource access: a wave-uniform base and byte range, a per-lane
// 32-bit byte offset (one VGPR for every row a thread reads) and a
// wave-uniform byte offset per access (an SGPR).  The hardware range check
// makes reads at or beyond `bytes` return 0 and drops such writes.
// Word 3 of the descriptor (0x00020000) is the raw 32-bit data format of
// the gfx9 family.
using bufres = __amdgpu_buffer_rsrc_t;
__device__ inline bufres make_bufres(const void *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}
template <typename T> __device__ inline T bload(bufres r, int voff, int soff);
template <> __device__ inline float bload<float>(bufres r, int voff, int soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <> __device__ inline double bload<double>(bufres r, int voff, int soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <> __device__ inline float2 bload<float2>(bufres r, int voff, int soff)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <> __device__ inline double2 bload<double2>(bufres r, int voff, int soff)
{
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
__device__ inline void bstore(bufres r, int voff, int soff, float2 v)
{
    using u2 = __attribute__((ext_vector_type(2))) unsigned int;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, soff, 0);
}
__device__ inline void bstore(bufres r, int voff, int soff, double2 v)
{
    using u4 = __attribute__((ext_vector_type(4))) unsigned int;
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), r, voff, soff, 0);
}
__device__ inline void bstore(bufres r, int voff, int soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
}
__device__ inline void bstore(bufres r, int voff, int soff, double v)
{
    using u2 = __attribute__((ext_vector_type(2))) unsigned int;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, voff, soff, 0);
}

}  // namespace dns
