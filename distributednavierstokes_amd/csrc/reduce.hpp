// Block-level reductions shared by the FV and spectral kernels: FP64 maxima,
// compensated (double-double) sums and binned sums, each finished by a
// deterministic second pass over per-block partials.
#pragma once

#include "common.hpp"

namespace dns {

// First level of a two-level max: block b reduces partials
// [b·chunk, (b+1)·chunk) of each slot into mid[b·nslots + slot].
static __global__ void __launch_bounds__(256) k_max_mid(const double *part, int nblocks, int nslots, double *mid)
{
    __shared__ double s[256];
    const int chunk = (nblocks + gridDim.x - 1) / gridDim.x;
    const int b0 = blockIdx.x * chunk, b1 = min(nblocks, b0 + chunk);
    for (int slot = 0; slot < nslots; ++slot) {
        double m = 0;
        for (int b = b0 + threadIdx.x; b < b1; b += blockDim.x) m = fmax(m, part[(size_t)b * nslots + slot]);
        s[threadIdx.x] = m;
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
            __syncthreads();
        }
        if (threadIdx.x == 0) mid[blockIdx.x * nslots + slot] = s[0];
        __syncthreads();
    }
}

// Block max of `cnt` partial maxima per slot; writes out[slot].
static __global__ void k_max_final(const double *part, int nblocks, int nslots, double *out)
{
    __shared__ double s[256];
    for (int slot = 0; slot < nslots; ++slot) {
        double m = 0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) m = fmax(m, part[b * nslots + slot]);
        s[threadIdx.x] = m;
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
            __syncthreads();
        }
        if (threadIdx.x == 0) out[slot] = s[0];
        __syncthreads();
    }
}

// Tree-reduces a per-thread double-double in LDS; thread 0 returns the total.
__device__ inline dd block_dd_sum(dd v, dd *s)
{
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = dd_add(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    dd r = s[0];
    __syncthreads();
    return r;
}

// Sums nparts double-doubles per slot (slot-major: part[slot * nparts + b]).
static __global__ void k_dd_final(const dd *part, int nparts, int nslots, double *out)
{
    __shared__ dd s[256];
    for (int slot = blockIdx.x; slot < nslots; slot += gridDim.x) {
        dd acc{0.0, 0.0};
        for (int b = threadIdx.x; b < nparts; b += blockDim.x) acc = dd_add(acc, part[(size_t)slot * nparts + b]);
        dd r = block_dd_sum(acc, s);
        if (threadIdx.x == 0) out[slot] = r.hi + r.lo;
    }
}

// Sums per-block bins in block order (deterministic final pass).
// Adds each lane's value v to bins[key] (lanes with key < 0 add nothing),
// deterministically and without atomics: the wavefront walks its distinct
// keys (lowest pending lane first), sums the lanes of each key with a fixed
// shuffle tree (other lanes contribute an exact 0) and lane 0 adds the sum.
// bins is this wavefront's own array, so no other wavefront writes it.
__device__ inline void wave_bin_add(double *bins, int key, double v)
{
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(key >= 0);
    while (pending) {
        const int first = __ffsll(pending) - 1;
        const int k = __shfl(key, first, 64);
        const unsigned long long same = __ballot(key == k);
        double x = key == k ? v : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) bins[k] += x;
        pending &= ~same;
    }
}

static __global__ void k_bins_final(const double *part, int nblocks, int nbins, double *out)
{
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nbins; b += gridDim.x * blockDim.x) {
        double s = 0;
        for (int q = 0; q < nblocks; ++q) s += part[(size_t)q * nbins + b];
        out[b] = s;
    }
}

}  // namespace dns
