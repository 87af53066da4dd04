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

__device__ inline unsigned short f2bf(float f)
{
    // round to nearest even
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16) | ((u & 0xffff) ? 0x40 : 0);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

__device__ inline float gelu_f(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_d(float x)
{
    const float phi = 0.39894228040143268f * expf(-0.5f * x * x);
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * phi;
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
__device__ inline uint2 pack4_bf16(const uint4 (&v)[4], int q)
{
    auto half = [&](const uint4 &w) -> unsigned int {
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
    const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
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
            uint4 v[4];
#pragma unroll
            for (int vi = 0; vi < 4; ++vi)
                v[vi] = c8 + 8 <= a.ldd ? *reinterpret_cast<const uint4 *>(a.dout + (vrow + x + vi) * a.ldd + c8) : zero4;
#pragma unroll
            for (int q = 0; q < 8; ++q)
                *reinterpret_cast<uint2 *>(sd + (c8 + q) * LD + x) = pack4_bf16(v, q);
        }
        for (int e = threadIdx.x; e < 3 * (M / 4) * kc8; e += blockDim.x) {
            const int dxi = e / ((M / 4) * kc8), f = e - dxi * ((M / 4) * kc8);
            const int xg = f / kc8, c8 = (f - xg * kc8) * 8, x = 4 * xg;
            uint4 v[4];
#pragma unroll
            for (int vi = 0; vi < 4; ++vi) {
                int xs = x + vi + dxi - 1; // si[dx][k][x] = in[(x + dx) mod M][k]
                xs = xs < 0 ? xs + M : (xs >= M ? xs - M : xs);
                v[vi] = c8 + 8 <= a.kin ? *reinterpret_cast<const uint4 *>(a.in + (vsrc + xs) * a.kin + c8) : zero4;
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

// Sums the partials over the voxel splits into the FP32 gradient
// grad[o][n][k] (n < nout, k < kin) of the unpadded layout.
__global__ void k_wgrad_reduce(const float *part, int nsplit, int np, int kp, int nout, int kin, float *grad)
{
    const int total = 27 * nout * kin;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int k = t % kin, n = (t / kin) % nout, o = t / (kin * nout);
        double s = 0;
        for (int q = 0; q < nsplit; ++q) s += part[(((size_t)q * 27 + o) * np + n) * kp + k];
        grad[t] = (float)s;
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

__global__ void k_bias_reduce(const float *part, int nblocks, int nout, float *grad)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nout) return;
    double s = 0;
    for (int q = 0; q < nblocks; ++q) s += part[(size_t)c * nblocks + q];
    grad[c] = (float)s;
}

// MSE loss over `nout` channels: dy = 2(y − t)/(V nout) (BF16, padded to
// ldd channels with zeros), per-block partial sums of (y − t)².
__global__ void __launch_bounds__(256) k_mse(const float *y, const float *t, int nout, long V, unsigned short *dy,
                                             int ldd, double *part)
{
    __shared__ double s[256];
    const double scale = 2.0 / ((double)V * nout);
    double acc = 0;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < V * ldd; e += (long)gridDim.x * blockDim.x) {
        const long v = e / ldd;
        const int c = (int)(e - v * ldd);
        float d = 0.f;
        if (c < nout) {
            const float r = y[v * nout + c] - t[v * nout + c];
            acc += (double)r * r;
            d = (float)(scale * r);
        }
        dy[e] = f2bf(d);
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
