// BF16 MFMA kernels of the sub-filter-stress CNN closure (gfx950): 3×3×3
// periodic convolutions as implicit GEMMs, erf-GELU, MSE loss and Adam.
//
// Activations are channels-last BF16 arrays act[v][c], v = x + M(y + M(z + M b)),
// with the channel count padded to a multiple of 8.  Weights are BF16
// W[o][n][k] (o = 0..26 the offset (dx, dy, dz) = (o%3−1, (o/3)%3−1, o/9−1),
// n = output channel padded to a multiple of 32, k = input channel, contiguous)
// for the forward pass, and the transposed copy WT[o][k][n] for the data
// gradient.  FP32 master weights and Adam moments live beside them.
//
// MFMA: v_mfma_f32_32x32x8_bf16 (wave64): C[32×32] += A[32×8]·B[8×32]; lane l
// supplies A[l%32][4(l/32)..4(l/32)+3] and B[4(l/32)..+3][l%32]; it receives
// C[8(r/4) + 4(l/32) + r%4][l%32] in accumulator r = 0..15.
#pragma once

#include "common.hpp"

namespace dns {

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
// 16-byte register vector for staged loads: a native vector type, so arrays
// of them stay in registers (HIP's uint4 is a struct whose copies are
// memcpys, which keeps such arrays in scratch memory)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));

__device__ inline unsigned short f2bf(float f)
{
    // round to nearest even
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16) | ((u & 0xffff) ? 0x40 : 0);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

// Normal CDF Φ(x) = ½(1 + erf(x/√2)) with erf from Abramowitz & Stegun
// 7.1.26 (|error| ≤ 1.5e-7): erf(u) = 1 − t(a1 + t(a2 + t(a3 + t(a4 + t a5))))
// e^{−u²}, t = 1/(1 + p u), u ≥ 0, on the hardware reciprocal and exp2;
// the tail ½ t(…) e^{−u²} is used directly for x < 0 (no cancellation).
// ex returns e^{−x²/2}, which the GELU derivative shares.  The activations
// are rounded to BF16 (2^−8) after every layer, far above this error.
__device__ inline float phi_cdf(float x, float &ex)
{
    const float u = fabsf(x) * 0.70710678118654752f;
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, u, 1.0f));
    ex = __expf(-0.5f * x * x);
    const float poly =
        t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, 1.061405429f, -1.453152027f), 1.421413741f), -0.284496736f), 0.254829592f);
    const float tail = 0.5f * poly * ex; // ½ erfc(u)
    return x >= 0.0f ? 1.0f - tail : tail;
}
__device__ inline float gelu_f(float x)
{
    float ex;
    return x * phi_cdf(x, ex);
}
__device__ inline float gelu_d(float x)
{
    float ex;
    const float c = phi_cdf(x, ex);
    return c + x * 0.39894228040143268f * ex;
}

struct ConvGeom {
    int M;     // grid points per direction
    int B;     // batch
    long V;    // B M³ voxels
};

// Neighbour voxel of v at offset o (sign = +1: v + d, −1: v − d), periodic.
__device__ inline long conv_nb(long v, int o, int sign, const ConvGeom &g)
{
    const int M = g.M;
    int x = (int)(v % M);
    long r = v / M;
    int y = (int)(r % M);
    r /= M;
    int z = (int)(r % M);
    long b = r / M;
    const int dx = (o % 3 - 1) * sign, dy = ((o / 3) % 3 - 1) * sign, dz = (o / 9 - 1) * sign;
    x += dx;
    y += dy;
    z += dz;
    x = x < 0 ? x + M : (x >= M ? x - M : x);
    y = y < 0 ? y + M : (y >= M ? y - M : y);
    z = z < 0 ? z + M : (z >= M ? z - M : z);
    return x + (long)M * (y + (long)M * (z + (long)M * b));
}

// out[v][n] = Σ_o Σ_k W[o][n][k] · in[v + sign·d_o][k]   (+ epilogue)
//   EPI 0: forward hidden layer: z = acc + bias → zbuf (BF16), a = gelu(z) → out (BF16)
//   EPI 1: forward output layer: y = acc + bias → yout (FP32, nout channels)
//   EPI 2: data gradient: d = acc · gelu'(zprev[v][n]) → out (BF16)
// Block: 4 waves; tile (32·WM voxels) × (32·WN channels), WM·WN = 4.
// KIN = channels of `in` (multiple of 8), NP = padded output channels
// (multiple of 32) = row count of W, NOUT = channels stored.
struct ConvArgs {
    const unsigned short *in;
    const unsigned short *W;
    const float *bias;
    const unsigned short *zprev;
    unsigned short *out;
    unsigned short *zbuf;
    float *yout;
    int kin, np, nout, ldo; // ldo: channel stride of out/zbuf/yout/zprev
    int sign;
};

template <int WM, int WN, int EPI>
__global__ void __launch_bounds__(256) k_conv(ConvArgs a, ConvGeom g)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const long v0 = (long)blockIdx.x * (32 * WM) + 32 * wm;
    const int n0 = blockIdx.y * (32 * WN) + 32 * wn;
    const int row = lane & 31, kh = 4 * (lane >> 5);
    const long v = v0 + row;
    const bool vok = v < g.V;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    if (n0 < a.np) {
        for (int o = 0; o < 27; ++o) {
            const long nb = vok ? conv_nb(v, o, a.sign, g) : 0;
            const unsigned short *ap = a.in + nb * a.kin + kh;
            const unsigned short *bp = a.W + ((size_t)o * a.np + n0 + row) * a.kin + kh;
            for (int k0 = 0; k0 < a.kin; k0 += 8) {
                bf16x4 av, bv;
                if (vok) {
                    const uint2 raw = *reinterpret_cast<const uint2 *>(ap + k0);
                    av = __builtin_bit_cast(bf16x4, raw);
                } else {
                    av = bf16x4{0, 0, 0, 0};
                }
                const uint2 rawb = *reinterpret_cast<const uint2 *>(bp + k0);
                bv = __builtin_bit_cast(bf16x4, rawb);
                acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
            }
        }
    }
    // epilogue: accumulator r holds row 8(r/4) + 4(lane/32) + r%4, column lane%32
    const int col = n0 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int rr = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const long vv = v0 + rr;
        if (vv >= g.V || col >= a.nout) continue;
        const float c = acc[r];
        if (EPI == 0) {
            const float z = c + a.bias[col];
            a.zbuf[vv * a.ldo + col] = f2bf(z);
            a.out[vv * a.ldo + col] = f2bf(gelu_f(z));
        } else if (EPI == 1) {
            a.yout[vv * a.ldo + col] = c + a.bias[col];
        } else {
            const float zp = bf2f(a.zprev[vv * a.ldo + col]);
            a.out[vv * a.ldo + col] = f2bf(c * gelu_d(zp));
        }
    }
}

// Same convolution on v_mfma_f32_32x32x16_bf16 (gfx950): each lane feeds 8
// consecutive channels (one 16-byte load per operand), and each wave owns a
// (32·MB voxels) × (32·NB channels) tile so every A load feeds NB MFMAs and
// every B load MB.  Block: 4 waves stacked along the voxels.  KIN must be a
// multiple of 16 and NP a multiple of 32·NB (one block column per 32·NB).
// The k order inside an MFMA is the same permutation for A and B, so the
// products do not depend on how the hardware assigns k to lanes.
template <int MB, int NB, int EPI>
__global__ void __launch_bounds__(256) k_conv16(ConvArgs a, ConvGeom g)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long vbase = ((long)blockIdx.x * 4 + wave) * (32 * MB);
    const int n0 = blockIdx.y * (32 * NB);
    const int row = lane & 31, kh = 8 * (lane >> 5);
    long vrow[MB];
    bool vok[MB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        vrow[mb] = vbase + 32 * mb + row;
        vok[mb] = vrow[mb] < g.V;
    }
    f32x16 acc[MB][NB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[mb][nb][r] = 0.0f;
    const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
    for (int o = 0; o < 27; ++o) {
        const unsigned short *ap[MB];
        const unsigned short *bp[NB];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb)
            ap[mb] = a.in + (vok[mb] ? conv_nb(vrow[mb], o, a.sign, g) : 0) * a.kin + kh;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) bp[nb] = a.W + ((size_t)o * a.np + n0 + 32 * nb + row) * a.kin + kh;
        for (int k0 = 0; k0 < a.kin; k0 += 16) {
            bf16x8 av[MB], bv[NB];
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
                av[mb] = __builtin_bit_cast(bf16x8, vok[mb] ? *reinterpret_cast<const uint4 *>(ap[mb] + k0) : zero4);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) bv[nb] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(bp[nb] + k0));
#pragma unroll
            for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    acc[mb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[mb], bv[nb], acc[mb][nb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int col = n0 + 32 * nb + (lane & 31);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const long vv = vbase + 32 * mb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (vv >= g.V || col >= a.nout) continue;
                const float c = acc[mb][nb][r];
                if (EPI == 0) {
                    const float z = c + a.bias[col];
                    a.zbuf[vv * a.ldo + col] = f2bf(z);
                    a.out[vv * a.ldo + col] = f2bf(gelu_f(z));
                } else if (EPI == 1) {
                    a.yout[vv * a.ldo + col] = c + a.bias[col];
                } else {
                    const float zp = bf2f(a.zprev[vv * a.ldo + col]);
                    a.out[vv * a.ldo + col] = f2bf(c * gelu_d(zp));
                }
            }
        }
}

// Row-staged convolution: a block owns 4 consecutive x-rows (y0..y0+3 at
// fixed z, b) of M = 32·MB voxels, one row per wave; for each of the three
// z offsets it stages the input rows y0−1..y0+4 of plane z + sign·dz in LDS
// (voxel stride kin + 8 halves, so the 16-byte fragment reads of one wave
// spread over the banks), then runs the nine (dx, dy) offsets from LDS with
// the x wrap applied per lane.  Weights stream from L2; each wave keeps a
// (32·MB voxels) × (32·NB channels) accumulator, so every fragment feeds MB
// or NB MFMAs.  Epilogues as k_conv.  Grid (M/4, M, B) blocks of 256 threads;
// LDS 6·M·(kin + 8) halves.  Needs M = 32·MB and kin % 16 == 0.
template <int MB, int NB, int EPI>
__global__ void __launch_bounds__(256) k_conv_rows(ConvArgs a, ConvGeom g)
{
    extern __shared__ unsigned char smem_raw[];
    unsigned short *sa = reinterpret_cast<unsigned short *>(smem_raw);
    constexpr int M = 32 * MB;
    const int KS = a.kin + 8; // voxel stride in LDS (halves)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 31, h = lane >> 5;
    const int y0 = blockIdx.x * 4, z = blockIdx.y, b = blockIdx.z;
    const int n0 = 0;
    f32x16 acc[MB][NB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[mb][nb][r] = 0.0f;
    const int kc8 = a.kin / 8;
    for (int dzi = 0; dzi < 3; ++dzi) {
        const int zs = (z + a.sign * (dzi - 1) + M) % M;
        __syncthreads();
        // stage rows y0-1 .. y0+4 of plane zs
        for (int e = threadIdx.x; e < 6 * M * kc8; e += blockDim.x) {
            const int c8 = e % kc8, t = e / kc8, x = t % M, rr = t / M;
            const int ys = (y0 - 1 + rr + M) % M;
            const long v = x + (long)M * (ys + (long)M * (zs + (long)M * b));
            *reinterpret_cast<uint4 *>(sa + (rr * M + x) * KS + 8 * c8) =
                *reinterpret_cast<const uint4 *>(a.in + v * a.kin + 8 * c8);
        }
        __syncthreads();
        for (int dyi = 0; dyi < 3; ++dyi) {
            const int rr = wave + 1 + a.sign * (dyi - 1); // staged row of this wave's neighbour
            for (int dxi = 0; dxi < 3; ++dxi) {
                const int o = dxi + 3 * dyi + 9 * dzi;
                const int dx = a.sign * (dxi - 1);
                const unsigned short *ap[MB];
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) {
                    const int xs = (32 * mb + row + dx + M) % M;
                    ap[mb] = sa + (rr * M + xs) * KS + 8 * h;
                }
                const unsigned short *bp[NB];
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) bp[nb] = a.W + ((size_t)o * a.np + n0 + 32 * nb + row) * a.kin + 8 * h;
                for (int k0 = 0; k0 < a.kin; k0 += 16) {
                    bf16x8 av[MB], bv[NB];
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        bv[nb] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(bp[nb] + k0));
#pragma unroll
                    for (int mb = 0; mb < MB; ++mb)
                        av[mb] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(ap[mb] + k0));
#pragma unroll
                    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                        for (int nb = 0; nb < NB; ++nb)
                            acc[mb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[mb], bv[nb], acc[mb][nb], 0, 0, 0);
                }
            }
        }
    }
    const long vrow = (long)M * ((y0 + wave) + (long)M * (z + (long)M * b));
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int col = n0 + 32 * nb + (lane & 31);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const long vv = vrow + 32 * mb + 8 * (r >> 2) + 4 * h + (r & 3);
                if (col >= a.nout) continue;
                const float c = acc[mb][nb][r];
                if (EPI == 0) {
                    const float zz = c + a.bias[col];
                    a.zbuf[vv * a.ldo + col] = f2bf(zz);
                    a.out[vv * a.ldo + col] = f2bf(gelu_f(zz));
                } else if (EPI == 1) {
                    a.yout[vv * a.ldo + col] = c + a.bias[col];
                } else {
                    const float zp = bf2f(a.zprev[vv * a.ldo + col]);
                    a.out[vv * a.ldo + col] = f2bf(c * gelu_d(zp));
                }
            }
        }
}

// k_conv_rows with the input channel count KIN (16 or 64) fixed at compile
// time and the next dz plane's input rows loaded into registers while the
// current plane is computed.  Weights: with WLDS (the default) each offset's
// NP × KIN block is loaded once per block, an offset ahead, into a
// double-buffered LDS copy that the four wavefronts share (one barrier per
// offset); without it every wavefront streams its own fragments from L1/L2
// into registers an offset ahead — four times the L1 traffic, which bound the
// kernel (forward 5.2 → 3.0 ms per training step with WLDS).
template <int MB, int NB, int KIN, int EPI, bool WLDS = false>
__global__ void __launch_bounds__(256) k_conv_rows_pf(ConvArgs a, ConvGeom g)
{
    extern __shared__ unsigned char smem_raw[];
    unsigned short *sa = reinterpret_cast<unsigned short *>(smem_raw);
    constexpr int M = 32 * MB, KS = KIN + 8, KST = KIN / 16, KC8 = KIN / 8;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 31, h = lane >> 5;
    const int y0 = blockIdx.x * 4, z = blockIdx.y, b = blockIdx.z;
    f32x16 acc[MB][NB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[mb][nb][r] = 0.0f;
    // kernel arguments as locals: the lambdas below capture by reference,
    // and a captured by-value argument struct would be kept in scratch
    const unsigned short *const gin = a.in;
    const int sign = a.sign, np = a.np;
    u32x4 bf[2][NB][KST];
    const unsigned short *wrow = a.W + (size_t)row * KIN + 8 * h;
    auto loadb = [&](int o, u32x4 (&dst)[NB][KST]) __attribute__((always_inline)) {
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int ks = 0; ks < KST; ++ks)
                dst[nb][ks] = *reinterpret_cast<const u32x4 *>(wrow + ((size_t)o * np + 32 * nb) * KIN + 16 * ks);
    };
    // input rows y0-1 .. y0+4 of plane z + sign·dz: NIT 16-byte pieces per
    // thread, loaded for the next dz while the current one is computed
    constexpr int NPC = 6 * M * KC8, NIT = (NPC + 255) / 256;
    u32x4 st[NIT];
    auto fetch_rows = [&](int dzi) __attribute__((always_inline)) {
        const int zs = (z + sign * (dzi - 1) + M) % M;
#pragma unroll
        for (int i = 0; i < NIT; ++i) {
            const int e = threadIdx.x + 256 * i;
            if (NPC % 256 == 0 || e < NPC) {
                const int c8 = e % KC8, t = e / KC8, x = t % M, rr = t / M;
                const int ys = (y0 - 1 + rr + M) % M;
                const long v = x + (long)M * (ys + (long)M * (zs + (long)M * b));
                st[i] = *reinterpret_cast<const u32x4 *>(gin + v * KIN + 8 * c8);
            }
        }
    };
    // WLDS: the block's four wavefronts share each offset's weights through
    // a double-buffered LDS copy [2][NP][KIN + 8] (two 16-byte pieces per
    // thread, loaded an offset ahead), instead of every wavefront streaming
    // its own copy from L1/L2
    constexpr int NP = 32 * NB, KW = KIN + 8, NPW = NP * KIN / 8, WPI = (NPW + 255) / 256;
    unsigned short *sw = sa + 6 * M * KS;
    u32x4 wr[WPI];
    const unsigned short *const gw = a.W;
    auto loadw = [&](int o) __attribute__((always_inline)) {
#pragma unroll
        for (int i = 0; i < WPI; ++i) {
            const int e = threadIdx.x + 256 * i;
            if (NPW % 256 == 0 || e < NPW)
                wr[i] = *reinterpret_cast<const u32x4 *>(gw + ((size_t)o * np + e / (KIN / 8)) * KIN + 8 * (e % (KIN / 8)));
        }
    };
    auto storew = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int i = 0; i < WPI; ++i) {
            const int e = threadIdx.x + 256 * i;
            if (NPW % 256 == 0 || e < NPW)
                *reinterpret_cast<u32x4 *>(sw + (buf * NP + e / (KIN / 8)) * KW + 8 * (e % (KIN / 8))) = wr[i];
        }
    };
    if constexpr (WLDS) {
        loadw(0);
        storew(0);
        loadw(1);
    } else {
        loadb(0, bf[0]);
    }
    fetch_rows(0);
#pragma unroll
    for (int dzi = 0; dzi < 3; ++dzi) {
        if (!WLDS || dzi == 0) __syncthreads(); // (WLDS: the last offset's barrier)
#pragma unroll
        for (int i = 0; i < NIT; ++i) {
            const int e = threadIdx.x + 256 * i;
            if (NPC % 256 == 0 || e < NPC) {
                const int c8 = e % KC8, t = e / KC8, x = t % M, rr = t / M;
                *reinterpret_cast<u32x4 *>(sa + (rr * M + x) * KS + 8 * c8) = st[i];
            }
        }
        __syncthreads();
        if (dzi < 2) fetch_rows(dzi + 1);
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const int o = q + 9 * dzi, cur = o & 1;
            if (!WLDS && o + 1 < 27) loadb(o + 1, bf[cur ^ 1]);
            const int dyi = q / 3, dxi = q % 3;
            const int rr = wave + 1 + a.sign * (dyi - 1);
            const int dx = a.sign * (dxi - 1);
            const unsigned short *ap[MB];
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                const int xs = (32 * mb + row + dx + M) % M;
                ap[mb] = sa + (rr * M + xs) * KS + 8 * h;
            }
#pragma unroll
            for (int ks = 0; ks < KST; ++ks) {
                bf16x8 av[MB];
#pragma unroll
                for (int mb = 0; mb < MB; ++mb) av[mb] = *reinterpret_cast<const bf16x8 *>(ap[mb] + 16 * ks);
                bf16x8 wv[NB];
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    if constexpr (WLDS)
                        wv[nb] = *reinterpret_cast<const bf16x8 *>(sw + (cur * NP + 32 * nb + row) * KW + 16 * ks + 8 * h);
                    else
                        wv[nb] = __builtin_bit_cast(bf16x8, bf[cur][nb][ks]);
                }
#pragma unroll
                for (int mb = 0; mb < MB; ++mb)
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        acc[mb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wv[nb], av[mb], acc[mb][nb], 0, 0, 0);
            }
            if constexpr (WLDS) {
                // the buffer written here was read in offset o − 1, which every
                // wavefront has finished (the barrier at its end)
                if (o + 1 < 27) storew(cur ^ 1);
                if (o + 2 < 27) loadw(o + 2);
                __syncthreads();
            }
        }
    }
    // The MFMAs take the weights as A and the voxels as B, so the accumulators
    // hold C^T: lane l has voxel 32·mb + l % 32 and, in accumulators 4g..4g+3,
    // the four consecutive channels 32·nb + 8g + 4(l / 32) + 0..3, which
    // leave as one 8-byte store per group (channels-last), converted to BF16
    // in pairs by the hardware (round to nearest even).
    const long vrow = (long)M * ((y0 + wave) + (long)M * (z + (long)M * b));
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const long vv = vrow + 32 * mb + (lane & 31);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int gq = 0; gq < 4; ++gq) {
                const int c0 = 32 * nb + 8 * gq + 4 * h;
                float v[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) v[k] = acc[mb][nb][4 * gq + k];
                if (EPI == 1) {
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        if (c0 + k < a.nout) a.yout[vv * a.ldo + c0 + k] = v[k] + a.bias[c0 + k];
                    continue;
                }
                if (c0 >= a.nout) continue; // nout is a multiple of 4 for the BF16 outputs
                const size_t off = (size_t)vv * a.ldo + c0;
                if (EPI == 0) {
                    float zz[4], gg[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        zz[k] = v[k] + a.bias[c0 + k];
                        gg[k] = gelu_f(zz[k]);
                    }
                    const bf16x4v zb = {(__bf16)zz[0], (__bf16)zz[1], (__bf16)zz[2], (__bf16)zz[3]};
                    const bf16x4v gb = {(__bf16)gg[0], (__bf16)gg[1], (__bf16)gg[2], (__bf16)gg[3]};
                    *reinterpret_cast<bf16x4v *>(a.zbuf + off) = zb;
                    *reinterpret_cast<bf16x4v *>(a.out + off) = gb;
                } else {
                    const bf16x4v zp = *reinterpret_cast<const bf16x4v *>(a.zprev + off);
                    float d[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) d[k] = v[k] * gelu_d((float)zp[k]);
                    const bf16x4v db = {(__bf16)d[0], (__bf16)d[1], (__bf16)d[2], (__bf16)d[3]};
                    *reinterpret_cast<bf16x4v *>(a.out + off) = db;
                }
            }
    }

}

// Weight gradient of one offset per block-row: for o = blockIdx.y,
// dW[o][n][k] += Σ_{v in chunk} dout[v][n] · in[v + d_o][k], tiles of 64
// voxels staged (transposed) through LDS.  Each block covers a range of
// voxels (blockIdx.x) and the full (NP × KP) output (NP, KP <= 64, multiples
// of 32 after padding), writing FP32 partials part[blockIdx.x][o][n][k].
struct WGradArgs {
    const unsigned short *dout; // [V][ldd]
    const unsigned short *in;   // [V][kin]
    float *part;
    int ldd, nout, kin, np, kp; // np, kp: padded extents of the partial tile
    long chunk;                 // voxels per block
};

__global__ void __launch_bounds__(256) k_conv_wgrad(WGradArgs a, ConvGeom g)
{
    __shared__ unsigned short sd[64][64 + 4]; // [channel n][voxel]
    __shared__ unsigned short si[64][64 + 4]; // [channel k][voxel]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wn = wave & 1, wk = wave >> 1; // 2 × 2 waves of 32 × 32
    const int o = blockIdx.y;
    const long vbeg = (long)blockIdx.x * a.chunk;
    const long vend = min(vbeg + a.chunk, g.V);
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    const int row = lane & 31, kh = 4 * (lane >> 5);
    for (long vt = vbeg; vt < vend; vt += 64) {
        // stage dout[vt..vt+63][0..np) and in[nb(v)][0..kp) transposed
        for (int e = threadIdx.x; e < 64 * 64; e += 256) {
            const int vl = e >> 6, c = e & 63;
            const long v = vt + vl;
            unsigned short dval = 0, ival = 0;
            if (v < vend) {
                if (c < a.nout) dval = a.dout[v * a.ldd + c];
                if (c < a.kin) ival = a.in[conv_nb(v, o, 1, g) * a.kin + c];
            }
            sd[c][vl] = dval;
            si[c][vl] = ival;
        }
        __syncthreads();
        if (32 * wn < a.np && 32 * wk < a.kp) {
#pragma unroll
            for (int k0 = 0; k0 < 64; k0 += 8) {
                bf16x4 av, bv;
                const unsigned short *pa = &sd[32 * wn + row][k0 + kh];
                const unsigned short *pb = &si[32 * wk + row][k0 + kh];
                av = bf16x4{(short)pa[0], (short)pa[1], (short)pa[2], (short)pa[3]};
                bv = bf16x4{(short)pb[0], (short)pb[1], (short)pb[2], (short)pb[3]};
                acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    if (32 * wn < a.np && 32 * wk < a.kp) {
        const int col = 32 * wk + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int n = 32 * wn + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            a.part[(((size_t)blockIdx.x * 27 + o) * a.np + n) * a.kp + col] = acc[r];
        }
    }
}

// Weight gradient on the K = 16 MFMA, organised by x-rows: a block owns a
// range of rows (y, z, b) and one (dy, dz) pair (blockIdx.y), stages each
// row's output gradient dout[x][n] transposed as sd[n][x] and the matching
// input row at (y + dy, z + dz) as three x-shifted copies si[dx][k][x], so the
// three dx offsets are aligned 16-byte LDS reads; a wave owns up to three of
// the (n-tile, k-tile, dx) 32×32 accumulators.  Partials
// part[blockIdx.x][o][n][k] as k_conv_wgrad.  Needs M % 16 == 0, np and kp
// in {32, 64}, ldd and kin multiples of 8.  LDS 4·64·(M + 8) halves.
// Halves q of four 16-byte vectors (8 BF16 channels of 4 voxels) packed
// into one 8-byte word in voxel order: the 4×8 → 8×4 register transpose of
// the weight-gradient staging.
__device__ inline uint2 pack4_bf16(const u32x4 (&v)[4], int q)
{
    auto half = [&](const u32x4 &w) -> unsigned int {
        const unsigned int word = (q >> 1) == 0 ? w.x : (q >> 1) == 1 ? w.y : (q >> 1) == 2 ? w.z : w.w;
        return (q & 1) ? (word >> 16) : (word & 0xffffu);
    };
    return make_uint2(half(v[0]) | (half(v[1]) << 16), half(v[2]) | (half(v[3]) << 16));
}

struct WGrad16Args {
    const unsigned short *dout; // [V][ldd]
    const unsigned short *in;   // [V][kin]
    float *part;
    int ldd, nout, kin, np, kp;
    int rows_per_block;
};

__global__ void __launch_bounds__(256) k_conv_wgrad16(WGrad16Args a, ConvGeom g)
{
    extern __shared__ unsigned char smem_raw[];
    unsigned short *lds = reinterpret_cast<unsigned short *>(smem_raw);
    const int M = g.M, LD = M + 8;
    unsigned short *sd = lds;           // [64][LD]
    unsigned short *si = lds + 64 * LD; // [3][64][LD]
    const int grp = blockIdx.y, dy = grp % 3 - 1, dz = grp / 3 - 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 31, h = lane >> 5;
    const int NT = a.np / 32, KT = a.kp / 32, ntiles = NT * KT * 3;
    const int nc8 = a.np / 8, kc8 = a.kp / 8;
    f32x16 acc[3];
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;
    const long nrows = (long)g.B * M * M;
    const long r0 = (long)blockIdx.x * a.rows_per_block;
    const long r1 = min(nrows, r0 + a.rows_per_block);
    const u32x4 zero4 = {0u, 0u, 0u, 0u};
    for (long r = r0; r < r1; ++r) {
        const int y = (int)(r % M);
        const long t = r / M;
        const int z = (int)(t % M);
        const long b = t / M;
        const int ys = y + dy < 0 ? y + dy + M : (y + dy >= M ? y + dy - M : y + dy);
        const int zs = z + dz < 0 ? z + dz + M : (z + dz >= M ? z + dz - M : z + dz);
        const long vrow = r * M, vsrc = (long)M * (ys + (long)M * (zs + (long)M * b));
        __syncthreads();
        // groups of 4 voxels × 8 channels are loaded as four 16-byte vectors and
        // written transposed as eight 8-byte LDS stores
        for (int e = threadIdx.x; e < (M / 4) * nc8; e += blockDim.x) {
            const int xg = e / nc8, c8 = (e - xg * nc8) * 8, x = 4 * xg;
            u32x4 v[4];
#pragma unroll
            for (int vi = 0; vi < 4; ++vi)
                v[vi] = c8 + 8 <= a.ldd ? *reinterpret_cast<const u32x4 *>(a.dout + (vrow + x + vi) * a.ldd + c8) : zero4;
#pragma unroll
            for (int q = 0; q < 8; ++q)
                *reinterpret_cast<uint2 *>(sd + (c8 + q) * LD + x) = pack4_bf16(v, q);
        }
        for (int e = threadIdx.x; e < 3 * (M / 4) * kc8; e += blockDim.x) {
            const int dxi = e / ((M / 4) * kc8), f = e - dxi * ((M / 4) * kc8);
            const int xg = f / kc8, c8 = (f - xg * kc8) * 8, x = 4 * xg;
            u32x4 v[4];
#pragma unroll
            for (int vi = 0; vi < 4; ++vi) {
                int xs = x + vi + dxi - 1; // si[dx][k][x] = in[(x + dx) mod M][k]
                xs = xs < 0 ? xs + M : (xs >= M ? xs - M : xs);
                v[vi] = c8 + 8 <= a.kin ? *reinterpret_cast<const u32x4 *>(a.in + (vsrc + xs) * a.kin + c8) : zero4;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q)
                *reinterpret_cast<uint2 *>(si + (dxi * 64 + c8 + q) * LD + x) = pack4_bf16(v, q);
        }
        __syncthreads();
#pragma unroll
        for (int ti = 0; ti < 3; ++ti) {
            const int tile = wave + 4 * ti;
            if (tile >= ntiles) break;
            const int dxi = tile % 3, rest = tile / 3, nt = rest % NT, kt = rest / NT;
            const unsigned short *pa = sd + (32 * nt + row) * LD + 8 * h;
            const unsigned short *pb = si + (dxi * 64 + 32 * kt + row) * LD + 8 * h;
            for (int x0 = 0; x0 < M; x0 += 16) {
                const bf16x8 av = *reinterpret_cast<const bf16x8 *>(pa + x0);
                const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(pb + x0);
                acc[ti] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[ti], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int ti = 0; ti < 3; ++ti) {
        const int tile = wave + 4 * ti;
        if (tile >= ntiles) break;
        const int dxi = tile % 3, rest = tile / 3, nt = rest % NT, kt = rest / NT;
        const int o = dxi + 3 * grp;
        const int k = 32 * kt + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int n = 32 * nt + 8 * (r >> 2) + 4 * h + (r & 3);
            a.part[(((size_t)blockIdx.x * 27 + o) * a.np + n) * a.kp + k] = acc[ti][r];
        }
    }
}

// Weight gradient by x-rows with one staged input row for all three dx
// offsets.  A block owns a range of rows (y, z, b) and one (dy, dz) pair
// (blockIdx.y); per row it stages the output gradient transposed,
// sd[n][x], and the input row at (y + dy, z + dz) transposed with an
// 8-voxel periodic halo on each side, si[k][x + 8].  Wave w takes the
// 16-voxel reduction slices x0 = 16 s (s ≡ w mod 4) and accumulates every
// (n-tile, k-tile, dx) tile of the block, NT·KT·3 ≤ 12: each slice reads one
// A fragment per n-tile and, per k-tile, the aligned 16-byte chunk at x0 plus
// the neighbouring dwords, from which the dx = ±1 operands are built by
// 16-bit funnel shifts (the reduction index pairs dout[x] with in[x + dx]).
// The global loads of the next row are issued into registers before the
// current row's MFMAs, so their latency overlaps the arithmetic.  Staging
// writes 8 voxels × 8 channels per thread as eight 16-byte transposed rows;
// row strides are ≡ 4 dwords (mod 8), which keeps both the transposed
// writes (lanes along x) and the fragment reads (lanes along channels) free
// of bank conflicts.  At the end the waves' tiles are summed in LDS in wave
// order (deterministic) and written as part[blockIdx.x][o][n][k].
// Needs M % 16 == 0, M <= 64, NT = np/32 and KT = kp/32 in {1, 2}, ldd and
// kin multiples of 8.
struct WGradRowArgs {
    const unsigned short *dout; // [V][ldd]
    const unsigned short *in;   // [V][kin]
    float *part;
    int ldd, kin;
    int rows_per_block;
    int lds_d, lds_i; // row strides of sd and si in halves
};

__host__ __device__ constexpr int wgrad_row_stride(int halves)
{
    // dwords rounded up to ≡ 4 (mod 8), returned in halves
    return 2 * ((halves / 2 + 3) / 8 * 8 + 4);
}

__device__ inline unsigned int bf_pair(unsigned int a, unsigned int b, int q)
{
    // element q (0..1 inside the dword) of a and of b as (a_q | b_q << 16)
    return q ? ((a >> 16) | (b & 0xffff0000u)) : ((a & 0xffffu) | (b << 16));
}

template <int NT, int KT>
__global__ void __launch_bounds__(256) k_conv_wgrad_rows(WGradRowArgs a, ConvGeom g)
{
    extern __shared__ unsigned char smem_raw[];
    unsigned short *lds = reinterpret_cast<unsigned short *>(smem_raw);
    constexpr int NP = 32 * NT, KP = 32 * KT, NTILE = NT * KT * 3;
    const int M = g.M, LDD = a.lds_d, LDI = a.lds_i;
    unsigned short *sd = lds;            // [NP][LDD]
    unsigned short *si = lds + NP * LDD; // [KP][LDI], column x + 8
    const int grp = blockIdx.y, dy = grp % 3 - 1, dz = grp / 3 - 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 31, h = lane >> 5;
    const int MG = M / 8;                        // 8-voxel groups per row
    const int nd = MG * (NP / 8);                // staging items of dout
    const int ni = (MG + 2) * (KP / 8);          // staging items of the input (with halo)
    const long nrows = (long)g.B * M * M;
    const long r0 = (long)blockIdx.x * a.rows_per_block;
    const long r1 = min(nrows, r0 + a.rows_per_block);
    const u32x4 zero4 = {0u, 0u, 0u, 0u};

    f32x16 acc[NTILE];
#pragma unroll
    for (int t = 0; t < NTILE; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;

    // this thread's staging item (at most one of the 2 kinds; nd + ni <= 256)
    const int e = threadIdx.x;
    const bool isd = e < nd, isi = !isd && e < nd + ni;
    int xg = 0, c8 = 0;
    if (isd) {
        xg = e % MG;
        c8 = 8 * (e / MG);
    } else if (isi) {
        xg = (e - nd) % (MG + 2);
        c8 = 8 * ((e - nd) / (MG + 2));
    }
    u32x4 pf[8];
    auto fetch = [&](long r) __attribute__((always_inline)) {
        const int y = (int)(r % M);
        const long t = r / M;
        const int z = (int)(t % M);
        const long b = t / M;
        if (isd) {
            const long v0 = r * M + 8 * xg;
            const bool ok = c8 + 8 <= a.ldd;
#pragma unroll
            for (int i = 0; i < 8; ++i)
                pf[i] = ok ? *reinterpret_cast<const u32x4 *>(a.dout + (v0 + i) * a.ldd + c8) : zero4;
        } else if (isi) {
            const int ys = y + dy < 0 ? y + dy + M : (y + dy >= M ? y + dy - M : y + dy);
            const int zs = z + dz < 0 ? z + dz + M : (z + dz >= M ? z + dz - M : z + dz);
            const long vrow = (long)M * (ys + (long)M * (zs + (long)M * b));
            int xs = 8 * xg - 8; // halo groups wrap periodically
            xs = xs < 0 ? xs + M : (xs >= M ? xs - M : xs);
            const bool ok = c8 + 8 <= a.kin;
#pragma unroll
            for (int i = 0; i < 8; ++i)
                pf[i] = ok ? *reinterpret_cast<const u32x4 *>(a.in + (vrow + xs + i) * a.kin + c8) : zero4;
        }
    };
    auto stage = [&]() __attribute__((always_inline)) {
        if (!isd && !isi) return;
        unsigned short *dst = isd ? sd + c8 * LDD + 8 * xg : si + c8 * LDI + 8 * xg;
        const int ld = isd ? LDD : LDI;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int wq = q >> 1, hq = q & 1;
            auto w = [&](const u32x4 &v) { return wq == 0 ? v.x : wq == 1 ? v.y : wq == 2 ? v.z : v.w; };
            u32x4 o;
            o.x = bf_pair(w(pf[0]), w(pf[1]), hq);
            o.y = bf_pair(w(pf[2]), w(pf[3]), hq);
            o.z = bf_pair(w(pf[4]), w(pf[5]), hq);
            o.w = bf_pair(w(pf[6]), w(pf[7]), hq);
            *reinterpret_cast<u32x4 *>(dst + q * ld) = o;
        }
    };

    if (r0 < r1) fetch(r0);
    for (long r = r0; r < r1; ++r) {
        __syncthreads(); // the previous row's fragments are read
        stage();
        __syncthreads();
        if (r + 1 < r1) fetch(r + 1);
        for (int s = wave; s < M / 16; s += 4) {
            const int x0 = 16 * s;
            bf16x8 av[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                av[nt] = *reinterpret_cast<const bf16x8 *>(sd + (32 * nt + row) * LDD + x0 + 8 * h);
#pragma unroll
            for (int kt = 0; kt < KT; ++kt) {
                const unsigned short *pk = si + (32 * kt + row) * LDI + 8 + x0 + 8 * h;
                const u32x4 c = *reinterpret_cast<const u32x4 *>(pk);
                const unsigned int prev = *reinterpret_cast<const unsigned int *>(pk - 2); // x - 2, x - 1
                const unsigned int next = *reinterpret_cast<const unsigned int *>(pk + 8); // x + 8, x + 9
                const unsigned int s01 = __builtin_amdgcn_alignbit(c.y, c.x, 16);
                const unsigned int s12 = __builtin_amdgcn_alignbit(c.z, c.y, 16);
                const unsigned int s23 = __builtin_amdgcn_alignbit(c.w, c.z, 16);
                const u32x4 bm = {__builtin_amdgcn_alignbit(c.x, prev, 16), s01, s12, s23}; // in[x − 1]
                const u32x4 bp = {s01, s12, s23, __builtin_amdgcn_alignbit(next, c.w, 16)}; // in[x + 1]
                const bf16x8 bv[3] = {__builtin_bit_cast(bf16x8, bm), __builtin_bit_cast(bf16x8, c),
                                      __builtin_bit_cast(bf16x8, bp)};
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                    for (int dxi = 0; dxi < 3; ++dxi) {
                        f32x16 &ac = acc[(nt * KT + kt) * 3 + dxi];
                        ac = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[nt], bv[dxi], ac, 0, 0, 0);
                    }
            }
        }
    }
    // ordered sum of the waves' tiles in LDS, then the block's partials
    float *red = reinterpret_cast<float *>(smem_raw);
    __syncthreads();
    for (int w = 0; w < 4; ++w) {
        if (wave == w) {
#pragma unroll
            for (int t = 0; t < NTILE; ++t)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int idx = t * 1024 + (8 * (r >> 2) + 4 * h + (r & 3)) * 32 + row;
                    red[idx] = w == 0 ? acc[t][r] : red[idx] + acc[t][r];
                }
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < NTILE * 1024; idx += blockDim.x) {
        const int t = idx >> 10, i = (idx >> 5) & 31, j = idx & 31;
        const int dxi = t % 3, kt = (t / 3) % KT, nt = t / (3 * KT);
        const int o = dxi + 3 * grp;
        a.part[(((size_t)blockIdx.x * 27 + o) * NP + 32 * nt + i) * KP + 32 * kt + j] = red[idx];
    }
}

// Weight gradient by (z, b) planes, one dz per block (blockIdx.y): the block
// walks the x-rows y = 0..M-1 of its planes and accumulates the nine (dy, dx)
// offsets of its dz at once, so each output-gradient row is staged once per
// dz (three times in all, not nine) and each input row once per plane: the
// input rows y−1, y, y+1 of plane z + dz sit in a ring of three LDS slots and
// advancing y stages only row y+2 (slot (y' + 1) mod 3 holds row y').  Each
// unit of work stages at most one 8 × 8 item per thread (the output-gradient
// row and the next input row, or the first two input rows of a plane), whose
// global loads are issued into registers before the previous unit's MFMAs.
// Wave w takes the (n-tile, k-tile) pair w mod NT·KT and the 16-voxel
// reduction slices s ≡ w / (NT·KT) (mod 4 / (NT·KT)), accumulating its nine
// (dy, dx) tiles; waves that share a pair are summed in LDS in wave order.
// Operands, halo and funnel shifts as k_conv_wgrad_rows.  Partials
// part[blockIdx.x][o][n][k] for the block's nine offsets.
// Needs M % 16 == 0, M <= 64, NT = np/32 and KT = kp/32 in {1, 2}.
struct WGradPlaneArgs {
    const unsigned short *dout; // [V][ldd]
    const unsigned short *in;   // [V][kin]
    float *part;
    float *bpart; // bias-gradient partials bpart[c · gridDim.x + blockIdx.x] (dz = 0 blocks), or null
    int nout;
    int ldd, kin;
    int planes_per_block;
    int lds_d, lds_i; // row strides of sd and si in halves
};

template <int NT, int KT>
__global__ void __launch_bounds__(256, 2) k_conv_wgrad_planes(WGradPlaneArgs a, ConvGeom g)
{
    extern __shared__ unsigned char smem_raw[];
    unsigned short *lds = reinterpret_cast<unsigned short *>(smem_raw);
    constexpr int NP = 32 * NT, KP = 32 * KT, NC = NT * KT, WPC = 4 / NC;
    const int M = g.M, LDD = a.lds_d, LDI = a.lds_i;
    unsigned short *sd = lds;            // [NP][LDD]
    unsigned short *si = lds + NP * LDD; // [3][KP][LDI], column x + 8
    const int dzi = blockIdx.y, dz = dzi - 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 31, h = lane >> 5;
    const int combo = wave % NC, sg = wave / NC;
    const int nt = combo % NT, kt = combo / NT;
    const int MG = M / 8;
    const int nd = MG * (NP / 8), ni = (MG + 2) * (KP / 8);
    const long nplanes = (long)g.B * M;
    const long p0 = (long)blockIdx.x * a.planes_per_block;
    const long p1 = min(nplanes, p0 + a.planes_per_block);
    const long nunits = (p1 - p0) * (M + 1); // per plane: a start unit, then rows 0..M-1
    const u32x4 zero4 = {0u, 0u, 0u, 0u};

    f32x16 acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[t][r] = 0.0f;

    // what this thread stages in unit u: kind 0 none, 1 dout row, 2 input row
    u32x4 pf[8];
    int kind = 0, xg = 0, c8 = 0, slot = 0;
    auto fetch = [&](long u) __attribute__((always_inline)) {
        const long p = p0 + u / (M + 1);
        const int yu = (int)(u % (M + 1)) - 1; // -1: plane start
        const int z = (int)(p % M);
        const long b = p / M;
        const int e = threadIdx.x;
        int yin = 0;
        kind = 0;
        if (yu < 0) {
            if (e < 2 * ni) {
                kind = 2;
                const int q = e / ni, f = e - q * ni;
                yin = q - 1; // rows -1 and 0
                xg = f % (MG + 2);
                c8 = 8 * (f / (MG + 2));
            }
        } else if (e < nd) {
            kind = 1;
            xg = e % MG;
            c8 = 8 * (e / MG);
            const long v0 = (long)M * (yu + (long)M * p) + 8 * xg;
            const bool ok = c8 + 8 <= a.ldd;
#pragma unroll
            for (int i = 0; i < 8; ++i)
                pf[i] = ok ? *reinterpret_cast<const u32x4 *>(a.dout + (v0 + i) * a.ldd + c8) : zero4;
        } else if (e < nd + ni) {
            kind = 2;
            yin = yu + 1;
            const int f = e - nd;
            xg = f % (MG + 2);
            c8 = 8 * (f / (MG + 2));
        }
        if (kind == 2) {
            slot = (yin + 1) % 3;
            const int ys = yin < 0 ? yin + M : (yin >= M ? yin - M : yin);
            const int zs = z + dz < 0 ? z + dz + M : (z + dz >= M ? z + dz - M : z + dz);
            const long vrow = (long)M * (ys + (long)M * (zs + (long)M * b));
            int xs = 8 * xg - 8;
            xs = xs < 0 ? xs + M : (xs >= M ? xs - M : xs);
            const bool ok = c8 + 8 <= a.kin;
#pragma unroll
            for (int i = 0; i < 8; ++i)
                pf[i] = ok ? *reinterpret_cast<const u32x4 *>(a.in + (vrow + xs + i) * a.kin + c8) : zero4;
        }
    };
    // bias gradient (Σ_v dout[v][c]) from the output-gradient rows the dz = 0
    // blocks stage anyway: 8 channels per staging thread, FP32 per block
    float bsum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const bool dobias = a.bpart != nullptr && dz == 0;
    auto stage = [&]() __attribute__((always_inline)) {
        if (kind == 0) return;
        if (kind == 1 && dobias) {
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const unsigned int wd = q >> 1 == 0 ? pf[i].x : q >> 1 == 1 ? pf[i].y : q >> 1 == 2 ? pf[i].z : pf[i].w;
                    bsum[q] += bf2f((unsigned short)((q & 1) ? (wd >> 16) : (wd & 0xffffu)));
                }
        }
        unsigned short *dst = kind == 1 ? sd + c8 * LDD + 8 * xg : si + (slot * KP + c8) * LDI + 8 * xg;
        const int ld = kind == 1 ? LDD : LDI;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int wq = q >> 1, hq = q & 1;
            auto w = [&](const u32x4 &v) { return wq == 0 ? v.x : wq == 1 ? v.y : wq == 2 ? v.z : v.w; };
            u32x4 o;
            o.x = bf_pair(w(pf[0]), w(pf[1]), hq);
            o.y = bf_pair(w(pf[2]), w(pf[3]), hq);
            o.z = bf_pair(w(pf[4]), w(pf[5]), hq);
            o.w = bf_pair(w(pf[6]), w(pf[7]), hq);
            *reinterpret_cast<u32x4 *>(dst + q * ld) = o;
        }
    };

    if (nunits > 0) fetch(0);
    for (long u = 0; u < nunits; ++u) {
        __syncthreads(); // the previous unit's fragments are read
        stage();
        __syncthreads();
        if (u + 1 < nunits) fetch(u + 1);
        const int y = (int)(u % (M + 1)) - 1;
        if (y < 0) continue;
        for (int s = sg; s < M / 16; s += WPC) {
            const int x0 = 16 * s;
            const bf16x8 av = *reinterpret_cast<const bf16x8 *>(sd + (32 * nt + row) * LDD + x0 + 8 * h);
#pragma unroll
            for (int dyi = 0; dyi < 3; ++dyi) {
                const int sl = (y + dyi) % 3; // input row y + dyi − 1
                const unsigned short *pk = si + (sl * KP + 32 * kt + row) * LDI + 8 + x0 + 8 * h;
                const u32x4 c = *reinterpret_cast<const u32x4 *>(pk);
                const unsigned int prev = *reinterpret_cast<const unsigned int *>(pk - 2);
                const unsigned int next = *reinterpret_cast<const unsigned int *>(pk + 8);
                const unsigned int s01 = __builtin_amdgcn_alignbit(c.y, c.x, 16);
                const unsigned int s12 = __builtin_amdgcn_alignbit(c.z, c.y, 16);
                const unsigned int s23 = __builtin_amdgcn_alignbit(c.w, c.z, 16);
                const u32x4 bm = {__builtin_amdgcn_alignbit(c.x, prev, 16), s01, s12, s23};
                const u32x4 bp = {s01, s12, s23, __builtin_amdgcn_alignbit(next, c.w, 16)};
                acc[3 * dyi + 0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, __builtin_bit_cast(bf16x8, bm),
                                                                           acc[3 * dyi + 0], 0, 0, 0);
                acc[3 * dyi + 1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, __builtin_bit_cast(bf16x8, c),
                                                                           acc[3 * dyi + 1], 0, 0, 0);
                acc[3 * dyi + 2] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, __builtin_bit_cast(bf16x8, bp),
                                                                           acc[3 * dyi + 2], 0, 0, 0);
            }
        }
    }
    // waves sharing an (n, k) tile pair are summed in LDS in wave order
    float *red = reinterpret_cast<float *>(smem_raw);
    if (dobias) {
        // the block's bias partials: the dout-staging threads (x-group
        // fastest) hold 8 channels each; summed over the x-groups in order
        __syncthreads();
        if (threadIdx.x < nd)
#pragma unroll
            for (int q = 0; q < 8; ++q) red[threadIdx.x * 8 + q] = bsum[q];
        __syncthreads();
        if (threadIdx.x < NP) {
            const int c = threadIdx.x, cg = c / 8, cq = c % 8;
            float t = 0.f;
            for (int x = 0; x < MG; ++x) t += red[(x + MG * cg) * 8 + cq];
            if (c < a.nout) a.bpart[(size_t)c * gridDim.x + blockIdx.x] = t;
        }
    }
    const size_t pbase = (size_t)blockIdx.x * 27 + 9 * dzi;
    if (WPC == 1) {
#pragma unroll
        for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = 8 * (r >> 2) + 4 * h + (r & 3);
                a.part[((pbase + t) * NP + 32 * nt + i) * KP + 32 * kt + row] = acc[t][r];
            }
        return;
    }
    __syncthreads();
    for (int q = 0; q < WPC; ++q) {
        if (sg == q) {
#pragma unroll
            for (int t = 0; t < 9; ++t)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int idx = (combo * 9 + t) * 1024 + (8 * (r >> 2) + 4 * h + (r & 3)) * 32 + row;
                    red[idx] = q == 0 ? acc[t][r] : red[idx] + acc[t][r];
                }
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < NC * 9 * 1024; idx += blockDim.x) {
        const int cb = idx / (9 * 1024), t = (idx >> 10) % 9, i = (idx >> 5) & 31, j = idx & 31;
        const int ntc = cb % NT, ktc = cb / NT;
        a.part[((pbase + t) * NP + 32 * ntc + i) * KP + 32 * ktc + j] = red[idx];
    }
}

// Sums the partials over the voxel splits into the FP32 gradient
// grad[o][n][k] (n < nout, k < kin) of the unpadded layout.
__global__ void k_wgrad_reduce(const float *part, int nsplit, int np, int kp, int nout, int kin, float *grad)
{
    const int total = 27 * nout * kin;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int k = t % kin, n = (t / kin) % nout, o = t / (kin * nout);
        // four independent chains keep four loads in flight; fixed order
        const size_t stride = (size_t)27 * np * kp, base = ((size_t)o * np + n) * kp + k;
        double s4[4] = {0, 0, 0, 0};
        int q = 0;
        for (; q + 4 <= nsplit; q += 4)
#pragma unroll
            for (int e = 0; e < 4; ++e) s4[e] += part[base + (size_t)(q + e) * stride];
        for (; q < nsplit; ++q) s4[0] += part[base + (size_t)q * stride];
        grad[t] = (float)((s4[0] + s4[1]) + (s4[2] + s4[3]));
    }
}

// Bias gradient: per-channel sum of dout over all voxels (per-block
// partials part[c · gridDim + block]).  A thread reads 8 channels (16 bytes)
// of a voxel row, ldd/8 threads cover a row and a block 256·8/ldd rows per
// pass, so waves read contiguous channel-last memory, dout is read once and
// each load moves 16 bytes (ldd a multiple of 8 dividing 2048).  1D grid.
__global__ void __launch_bounds__(256) k_bias_grad(const unsigned short *dout, int ldd, int nout, long V, float *part)
{
    __shared__ float s[8][256];
    const int tpr = ldd / 8, c8 = threadIdx.x % tpr, r = threadIdx.x / tpr, rows = blockDim.x / tpr;
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 4
    for (long v = blockIdx.x * (long)rows + r; v < V; v += (long)gridDim.x * rows) {
        const uint4 q = *reinterpret_cast<const uint4 *>(dout + v * ldd + 8 * c8);
        const unsigned int w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            acc[2 * e] += bf2f((unsigned short)(w[e] & 0xffffu));
            acc[2 * e + 1] += bf2f((unsigned short)(w[e] >> 16));
        }
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) s[e][threadIdx.x] = acc[e];
    __syncthreads();
    // thread (row 0, c8) sums its 8 channels over the block's rows
    if (r == 0) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int c = 8 * c8 + e;
            if (c >= nout) continue;
            float t = 0;
            for (int q = 0; q < rows; ++q) t += s[e][q * tpr + c8];
            part[(size_t)c * gridDim.x + blockIdx.x] = t;
        }
    }
}

// One block of 256 threads per channel (grid nout): strided FP64 partial
// sums, then a tree in LDS (fixed order).
__global__ void __launch_bounds__(256) k_bias_reduce(const float *part, int nblocks, int nout, float *grad)
{
    __shared__ double s[256];
    const int c = blockIdx.x;
    double acc = 0;
    for (int q = threadIdx.x; q < nblocks; q += blockDim.x) acc += part[(size_t)c * nblocks + q];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0 && c < nout) grad[c] = (float)s[0];
}

// MSE loss over `nout` channels: dy = 2(y − t)/(V nout) (BF16, padded to
// ldd channels with zeros), per-block partial sums of (y − t)².
__global__ void __launch_bounds__(256) k_mse(const float *y, const float *t, int nout, long V, unsigned short *dy,
                                             int ldd, double *part)
{
    // one voxel per thread and pass: the ldd BF16 gradient channels (zero
    // past nout) leave as 16-byte stores of 8 channels (ldd % 8 == 0)
    __shared__ double s[256];
    const double scale = 2.0 / ((double)V * nout);
    double acc = 0;
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x) {
        for (int c0 = 0; c0 < ldd; c0 += 8) {
            unsigned int w[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                unsigned short hb[2];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const int c = c0 + 2 * q + e;
                    float d = 0.f;
                    if (c < nout) {
                        const float r = y[v * nout + c] - t[v * nout + c];
                        acc += (double)r * r;
                        d = (float)(scale * r);
                    }
                    hb[e] = f2bf(d);
                }
                w[q] = (unsigned int)hb[0] | ((unsigned int)hb[1] << 16);
            }
            *reinterpret_cast<u32x4 *>(dy + v * ldd + c0) = u32x4{w[0], w[1], w[2], w[3]};
        }
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Adam on FP32 masters; refreshes the BF16 forward (W[o][n][k], n padded to
// np) and transposed (WT[o][k][n], k padded to kpt) copies.  bias tail of
// each layer is stored after its weights in the flat parameter vector.
struct AdamArgs {
    float *w, *m, *v;
    const float *g;
    long count;
    float lr, b1, b2, eps, c1, c2; // c1 = 1/(1−b1^t), c2 = 1/(1−b2^t)
};

__global__ void __launch_bounds__(256) k_adam(AdamArgs a)
{
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < a.count; t += (long)gridDim.x * blockDim.x) {
        const float g = a.g[t];
        const float m = a.b1 * a.m[t] + (1.f - a.b1) * g;
        const float v = a.b2 * a.v[t] + (1.f - a.b2) * g * g;
        a.m[t] = m;
        a.v[t] = v;
        const float mh = m * a.c1, vh = v * a.c2;
        a.w[t] -= a.lr * mh / (sqrtf(vh) + a.eps);
    }
}

// Writes the BF16 copies of one layer's weights from its FP32 master w[o][n][k]
// (n < nout, k < kin): W[o][n][k] with n padded to np and k to kinp, and
// WT[o][k][n] with k padded to kpt (rows) and n to noutp (contiguous).
__global__ void k_pack_weights(const float *w, int nout, int kin, unsigned short *W, int np, int kinp,
                               unsigned short *WT, int kpt, int noutp)
{
    const int total = 27 * np * kinp;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int k = t % kinp, n = (t / kinp) % np, o = t / (kinp * np);
        const float val = (n < nout && k < kin) ? w[(o * nout + n) * kin + k] : 0.f;
        W[t] = f2bf(val);
    }
    if (!WT) return;
    const int totalT = 27 * kpt * noutp;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < totalT; t += gridDim.x * blockDim.x) {
        const int n = t % noutp, k = (t / noutp) % kpt, o = t / (noutp * kpt);
        const float val = (n < nout && k < kin) ? w[(o * nout + n) * kin + k] : 0.f;
        WT[t] = f2bf(val);
    }
}

// FP32 planar (V, cin) → BF16 channels-last padded to ldc channels, with a
// scale; and the inverse for FP32 outputs.
__global__ void k_to_bf16(const float *x, int cin, float scale, long V, unsigned short *out, int ldc)
{
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < V * ldc; e += (long)gridDim.x * blockDim.x) {
        const long v = e / ldc;
        const int c = (int)(e - v * ldc);
        out[e] = c < cin ? f2bf(scale * x[v * cin + c]) : (unsigned short)0;
    }
}

}  // namespace dns
