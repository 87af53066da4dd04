// Device kernels of the pseudo-spectral HIT step (gfx950).
//
// Spectral state uh: (K, m, m, 3) complex, K = kcut+1, m = 2kcut+1, first
// index fastest, compact frequency order j <= kcut ? j : j - m
// (src/spectral/operators.jl:7); û = F[u]/n³ (src/spectral/fft.jl:176-179).
// The 3D transforms run as three line passes without transposes:
//   backward:  y pass (pad m → n, inverse FFT)      uh → B1 (K, n, m, 3)
//              z pass (pad m → n, inverse FFT)      B1 → P[..,0:3] (K, n, n)
//              x pass (c2r of the K retained kx)    P  → physical, LDS only
//   forward:   x pass (r2c, keep kx < K)            physical → P (K, n, n)
//              z pass (FFT, keep retained kz)       P → Q (K, n, m)
//              y pass (FFT, keep retained ky, /n³)  Q → state
// In the time step the x pass is fused: c2r of the 3 velocity components,
// the 6 products v_i v_j (and the CFL velocity bound), r2c of each product;
// physical fields never reach HBM.  The forward y pass of the products is
// fused with the divergence-form nonlinear term −i k_j σ̂_ij, the Wray
// integrating-factor stage and the projection.
#pragma once

#include "common.hpp"

// Loads kept in flight per thread by the HBM→LDS staging loops (capped by the
// tile's elements per thread).
// (FP64 tiles keep fewer in flight: their FFT registers already fill 256 VGPRs.)
#ifndef SP_LOAD_MAX
#define SP_LOAD_MAX 16
#endif
#define SP_LOAD_U(e) ((e) < SP_LOAD_MAX / (int)(sizeof(T) / 4) ? (e) : SP_LOAD_MAX / (int)(sizeof(T) / 4))
#include "fft_lds.hpp"
#include "reduce.hpp"

namespace dns {

struct SpecGeom {
    int n, kcut, K, m;
    double kfac; // 2π / l
};

// LDS line stride (elements) of the y/z pencil tiles.  512-point lines take
// 516 (a stride of 16 banks in FP64, 8 in FP32, between the lines of a tile),
// so the tile's row-wise stores and reads (line fastest across the lanes)
// fall on distinct banks; n + 1 otherwise.
template <typename T> __host__ __device__ constexpr int sp_line_stride(int n)
{
    return n == 512 ? 516 : n + 1;
}



__device__ inline int sp_fine(int c, const SpecGeom &g) { return c <= g.kcut ? c : c + (g.n - g.m); }
__device__ inline int sp_freq(int c, const SpecGeom &g) { return c <= g.kcut ? c : c - g.m; }
// fine position → compact index, or -1 in the killed band
__device__ inline int sp_compact(int f, const SpecGeom &g)
{
    if (f <= g.kcut) return f;
    if (f >= g.n - g.kcut) return f - (g.n - g.m);
    return -1;
}

// ------------------------------------------------------------ y / z passes

// Pencil tiles (line tiles of fft_lds.hpp: line b is kx = i0 + b) whose
// HBM side is truncated: the backward passes read the compact index of each
// fine position (zero in the killed band kcut < t < n − kcut), the forward
// passes write the retained positions only.  Same walk as tile_load /
// tile_store.

// buf[b ls + t] = base[b + stride·c(t)], t < n, c(t) the compact index of t.
template <typename T, int EPT, int NT>
__device__ inline void pencil_load_compact(cplx<T> *buf, int ls, const cplx<T> *base, int stride, int nb, int tb,
                                           FastDiv ftb, const SpecGeom &g)
{
    const int n = g.n, sz = (int)sizeof(cplx<T>);
    int b, t0;
    if (tile_walk<T, NT>(tb, ftb, n, stride, b, t0)) {
        const int S = NT / tb, E = tb * n / NT;
        const bufres r = make_bufres(base, (uint32_t)(((size_t)stride * (g.m - 1) + tb) * sz));
        const int v0 = (b + stride * t0) * sz, shift = (n - g.m) * stride * sz;
        cplx<T> pre[EPT];
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            if (e < E) {
                // the whole offset in the VGPR: a shifted lane offset may be
                // negative alone
                const int t = t0 + e * S, vs = v0 + e * S * stride * sz;
                const int voff = b >= nb ? TILE_OOB : t <= g.kcut ? vs : t >= n - g.kcut ? vs - shift : TILE_OOB;
                pre[e] = bload<cplx<T>>(r, voff, 0);
            }
        }
        cplx<T> *row = buf + b * ls + t0;
#pragma unroll
        for (int e = 0; e < EPT; ++e)
            if (e < E) row[e * S] = pre[e];
        return;
    }
    batched_for<SP_LOAD_U(EPT), cplx<T>>(
        tb * n,
        [&](int idx) {
            const int t = fdiv(idx, ftb), bb = idx - t * tb, sc = sp_compact(t, g);
            cplx<T> v = mk<T>(T(0), T(0));
            if (bb < nb && sc >= 0) v = base[bb + (size_t)stride * sc];
            return v;
        },
        [&](int idx, cplx<T> v) {
            const int t = fdiv(idx, ftb), bb = idx - t * tb;
            buf[bb * ls + t] = v;
        });
}

// base[b + stride c] = sc · buf[b ls + f(c)], c < m, f(c) the fine position
// of compact index c.
template <typename T, int EPT, int NT, bool SCALE>
__device__ inline void pencil_store_fine(const cplx<T> *buf, int ls, cplx<T> *base, int stride, int nb, int tb,
                                         FastDiv ftb, const SpecGeom &g, T sc)
{
    const int sz = (int)sizeof(cplx<T>), cnt = g.m;
    int b, t0;
    if (tile_walk<T, NT>(tb, ftb, g.n, stride, b, t0)) {
        const int S = NT / tb;
        const bufres r = make_bufres(base, (uint32_t)(((size_t)stride * (cnt - 1) + tb) * sz));
        const int v0 = b < nb ? (b + stride * t0) * sz : TILE_OOB;
        const cplx<T> *row = buf + b * ls;
        // all LDS reads first (positions past the line clamped), then the
        // stores, those of indices c >= m sent out of range: no branch
        // between them, so the reads overlap
        cplx<T> v[EPT];
#pragma unroll
        for (int e = 0; e < EPT; ++e) v[e] = row[min(sp_fine(t0 + e * S, g), g.n - 1)];
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int c = t0 + e * S;
            bstore(r, c < cnt ? v0 : TILE_OOB, e * S * stride * sz, SCALE ? cscale<T>(v[e], sc) : v[e]);
        }
        return;
    }
    for (int idx = threadIdx.x; idx < tb * cnt; idx += NT) {
        const int c = fdiv(idx, ftb), bb = idx - c * tb;
        if (bb < nb) {
            cplx<T> v = buf[bb * ls + sp_fine(c, g)];
            base[bb + (size_t)stride * c] = SCALE ? cscale<T>(v, sc) : v;
        }
    }
}

// Backward y pass: B1[i, y, kz, c] = Σ_ky uh[i, ky, kz, c] e^{+i ky y}.
// grid (ceil(K/tb), m, 3); LDS tb·ls + the twiddles.
template <typename T, int EPT, int NT>
__global__ void __launch_bounds__(NT, 2) k_sp_yback(const cplx<T> *uh, cplx<T> *B1, const cplx<T> *tw, int tb, int tbf, FastDiv ftb,
                                                     SpecGeom g, FFTDesc d)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
    const int n = g.n, ls = sp_line_stride<T>(n), K = g.K, m = g.m;
    const cplx<T> *twl = stage_tile_twiddles<T>(buf + (size_t)tb * ls, tw, n, tb);
    int bxr, byr, bzr;
    xcd_block3(bxr, byr, bzr);
    const int i0 = bxr * tb, kz = byr, c = bzr, nb = min(tb, K - i0);
    pencil_load_compact<T, EPT, NT>(buf, ls, uh + i0 + (size_t)K * m * (kz + (size_t)m * c), K, nb, tb, ftb, g);
    __syncthreads();
    fft_lds_tile<T, EPT, true, 0, true>(buf, tb, ls, d, twl, tbf);
    tile_store<T, EPT, NT>(buf, ls, B1 + i0 + (size_t)K * n * (kz + (size_t)m * c), K, nb, tb, ftb, n);
}

// Backward z pass: P[i, y, z, c] = Σ_kz B1[i, y, kz, c] e^{+i kz z}.
// grid (ceil(K/tb), n, 3).
template <typename T, int EPT, int NT>
__global__ void __launch_bounds__(NT, 2) k_sp_zback(const cplx<T> *B1, cplx<T> *P, const cplx<T> *tw, int tb, int tbf, FastDiv ftb,
                                                     SpecGeom g, FFTDesc d)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
    const int n = g.n, ls = sp_line_stride<T>(n), K = g.K, m = g.m;
    const cplx<T> *twl = stage_tile_twiddles<T>(buf + (size_t)tb * ls, tw, n, tb);
    int bxr, byr, bzr;
    xcd_block3(bxr, byr, bzr);
    const int i0 = bxr * tb, y = byr, c = bzr, nb = min(tb, K - i0);
    pencil_load_compact<T, EPT, NT>(buf, ls, B1 + i0 + (size_t)K * (y + (size_t)n * m * c), K * n, nb, tb, ftb, g);
    __syncthreads();
    fft_lds_tile<T, EPT, true, 0, true>(buf, tb, ls, d, twl, tbf);
    tile_store<T, EPT, NT>(buf, ls, P + i0 + (size_t)K * (y + (size_t)n * n * c), K * n, nb, tb, ftb, n);
}

// Forward z pass: Q[i, y, kc, p] = Σ_z P[i, y, z, p] e^{−i kz z} at the
// retained kz.  grid (ceil(K/tb), n, ncomp).
template <typename T, int EPT, int NT>
__global__ void __launch_bounds__(NT, 2) k_sp_zfwd(const cplx<T> *P, cplx<T> *Q, const cplx<T> *tw, int tb, int tbf, FastDiv ftb,
                                                    SpecGeom g, FFTDesc d)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
    const int n = g.n, ls = sp_line_stride<T>(n), K = g.K, m = g.m;
    const cplx<T> *twl = stage_tile_twiddles<T>(buf + (size_t)tb * ls, tw, n, tb);
    int bxr, byr, bzr;
    xcd_block3(bxr, byr, bzr);
    const int i0 = bxr * tb, y = byr, p = bzr, nb = min(tb, K - i0);
    tile_load<T, EPT, NT, SP_LOAD_U(EPT)>(buf, ls, P + i0 + (size_t)K * (y + (size_t)n * n * p), K * n, nb, tb, ftb, n);
    __syncthreads();
    fft_lds_tile<T, EPT, false, 0, true>(buf, tb, ls, d, twl, tbf);
    pencil_store_fine<T, EPT, NT, false>(buf, ls, Q + i0 + (size_t)K * (y + (size_t)n * m * p), K * n, nb, tb, ftb, g,
                                         T(1));
}

// Forward y pass with truncation and 1/n³ (truncate_y_kernel!,
// src/spectral/fft.jl:190-193): out[i, jc, kc, p] (K, m, m, ncomp).
// grid (ceil(K/tb), m, ncomp).
template <typename T, int EPT, int NT>
__global__ void __launch_bounds__(NT, 2) k_sp_yfwd(const cplx<T> *Q, cplx<T> *out, const cplx<T> *tw, int tb, int tbf, FastDiv ftb,
                                                    SpecGeom g, FFTDesc d)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
    const int n = g.n, ls = sp_line_stride<T>(n), K = g.K, m = g.m;
    const cplx<T> *twl = stage_tile_twiddles<T>(buf + (size_t)tb * ls, tw, n, tb);
    int bxr, byr, bzr;
    xcd_block3(bxr, byr, bzr);
    const int i0 = bxr * tb, kc = byr, p = bzr, nb = min(tb, K - i0);
    tile_load<T, EPT, NT, SP_LOAD_U(EPT)>(buf, ls, Q + i0 + (size_t)K * n * (kc + (size_t)m * p), K, nb, tb, ftb, n);
    __syncthreads();
    fft_lds_tile<T, EPT, false, 0, true>(buf, tb, ls, d, twl, tbf);
    const T inv = T(1.0 / ((double)n * n * n));
    pencil_store_fine<T, EPT, NT, true>(buf, ls, out + i0 + (size_t)K * m * (kc + (size_t)m * p), K, nb, tb, ftb, g, inv);
}

// ------------------------------------------------------------- x passes

enum { XMID_RHS = 0, XMID_C2R = 1, XMID_R2C = 2, XMID_VMAX = 3 };

// x pass over (y, z) lines L = y + n z, `lpb` lines per block, d.n = n/2.
//   XMID_RHS : c2r of P[..,0:3] → v (LDS), vmax partial, six products, r2c
//              of each product into P[.., 0:6] (kx < K)
//   XMID_C2R : c2r of P[..,0:3] → phys (n, n, n, 3)
//   XMID_VMAX: c2r of P[..,0:3] → vmax partial only
//   XMID_R2C : r2c of phys (n, n, n, 3) → P[..,0:3] (kx < K)
// The three components (and each group of three products) are transformed
// together as 3·lpb lines.  LDS: v[3][lpb][n] reals, work[3·lpb][n/2]
// complex, then the n/2 transform twiddles and (FP32) the n/2+1
// real-transform twiddles.
template <typename T> struct XMidArgs {
    cplx<T> *P;
    T *phys;
    double *vpart;
    const cplx<T> *tw;  // length n/2
    const cplx<T> *twx; // length n/2 + 1
    int lpb;
    FastDiv fN, fn, fK, flpb; // division by n/2, n, K and lpb
};

template <typename T, int EPT, int MODE>
__global__ void __launch_bounds__(256) k_sp_xmid(XMidArgs<T> a, SpecGeom g, FFTDesc d)
{
    extern __shared__ unsigned char smem_raw[];
    const int n = g.n, N = n / 2, K = g.K, lpb = a.lpb;
    T *v = reinterpret_cast<T *>(smem_raw);
    cplx<T> *work = reinterpret_cast<cplx<T> *>(smem_raw + sizeof(T) * (size_t)3 * lpb * n);
    const int nlines = n * n;
    const int L0 = xcd_block(blockIdx.x, gridDim.x) * lpb;
    const size_t pstride = (size_t)K * n * n; // component stride of P
    const int nw = 3 * lpb;                   // work lines: three components (or products) × lpb
    const int maxl = max(1, EPT * (int)blockDim.x / N); // lines one fft_lds call may hold
    // twiddles staged in LDS behind the work lines (read by every butterfly pass)
    const cplx<T> *twl = stage_twiddles<T>(work + (size_t)nw * N, a.tw, N);
    // real-transform twiddles: an LDS copy in FP32 (measured faster there),
    // read from L1/L2 in FP64
    const cplx<T> *twxl = sizeof(T) == 4 ? stage_twiddles<T>(work + (size_t)(nw + 1) * N, a.twx, N + 1) : a.twx;
    if (MODE != XMID_R2C) {
        // c2r of the three components in one batched transform: work line c·lpb + b
        batched_for<SP_LOAD_U(EPT), cplx<T>>(
            nw * N,
            [&](int idx) {
                const int cl = fdiv(idx, a.fN), k = idx - cl * N, c = fdiv(cl, a.flpb), b = cl - c * lpb, L = L0 + b;
                cplx<T> Z = mk<T>(T(0), T(0));
                if (L < nlines) {
                    const cplx<T> *row = a.P + (size_t)L * K + c * pstride;
                    cplx<T> Xk = k < K ? row[k] : mk<T>(T(0), T(0));
                    cplx<T> Xn = (N - k) < K ? row[N - k] : mk<T>(T(0), T(0));
                    if (k == 0) {
                        Xk.y = T(0);
                        Xn.y = T(0);
                    }
                    cplx<T> Xc = cconj<T>(Xn);
                    cplx<T> ev = cadd<T>(Xk, Xc);
                    cplx<T> od = cmulc<T>(csub<T>(Xk, Xc), a.twx[k]); // (the LDS copy is not yet complete)
                    Z = mk<T>(ev.x - od.y, ev.y + od.x);
                }
                return Z;
            },
            [&](int idx, cplx<T> z) { work[idx] = z; });
        __syncthreads();
        for (int l0 = 0; l0 < nw; l0 += maxl) fft_lds<T, EPT, true>(work + (size_t)l0 * N, min(maxl, nw - l0), N, d, twl);
        for (int idx = threadIdx.x; idx < nw * N; idx += blockDim.x) {
            const int cl = fdiv(idx, a.fN), t = idx - cl * N;
            const cplx<T> z = work[idx];
            T *vl = v + (size_t)cl * n; // v[c][b] = line cl
            vl[2 * t] = z.x;
            vl[2 * t + 1] = z.y;
        }
        __syncthreads();
        if (MODE == XMID_RHS || MODE == XMID_VMAX) {
            __shared__ double red[256];
            double mx = 0;
            for (int idx = threadIdx.x; idx < lpb * n; idx += blockDim.x) {
                int b = fdiv(idx, a.fn);
                if (L0 + b >= nlines) continue;
                double s = fabs((double)v[idx]) + fabs((double)v[(size_t)lpb * n + idx]) +
                           fabs((double)v[(size_t)2 * lpb * n + idx]);
                mx = fmax(mx, s);
            }
            red[threadIdx.x] = mx;
            __syncthreads();
            for (int w = blockDim.x / 2; w > 0; w >>= 1) {
                if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
                __syncthreads();
            }
            if (threadIdx.x == 0) a.vpart[blockIdx.x] = red[0];
        }
        if (MODE == XMID_C2R) {
            for (int c = 0; c < 3; ++c)
                for (int idx = threadIdx.x; idx < lpb * n; idx += blockDim.x) {
                    int b = fdiv(idx, a.fn), x = idx - b * n, L = L0 + b;
                    if (L < nlines) a.phys[x + (size_t)n * L + (size_t)c * n * nlines] = v[((size_t)c * lpb + b) * n + x];
                }
        }
        if (MODE != XMID_RHS) return;
    }
    // forward r2c of the products (RHS: two groups of three) or of the three
    // physical components (R2C), each group one batched transform
    const int ngroups = MODE == XMID_RHS ? 2 : 1;
    for (int grp = 0; grp < ngroups; ++grp) {
        batched_for<SP_LOAD_U(EPT), cplx<T>>(
            nw * N,
            [&](int idx) {
                const int cl = fdiv(idx, a.fN), t = idx - cl * N, pg = fdiv(cl, a.flpb), b = cl - pg * lpb, L = L0 + b;
                const int p = 3 * grp + pg;
                cplx<T> z = mk<T>(T(0), T(0));
                if (MODE == XMID_RHS) {
                    const int ca = p < 3 ? p : (p == 3 ? 0 : p == 4 ? 1 : 2);
                    const int cb = p < 3 ? p : (p == 3 ? 1 : p == 4 ? 2 : 0);
                    const T *va = v + ((size_t)ca * lpb + b) * n, *vb = v + ((size_t)cb * lpb + b) * n;
                    z = mk<T>(va[2 * t] * vb[2 * t], va[2 * t + 1] * vb[2 * t + 1]);
                } else if (L < nlines) {
                    const T *src = a.phys + (size_t)n * L + (size_t)p * n * nlines;
                    z = mk<T>(src[2 * t], src[2 * t + 1]);
                }
                return z;
            },
            [&](int idx, cplx<T> z) { work[idx] = z; });
        __syncthreads();
        for (int l0 = 0; l0 < nw; l0 += maxl) fft_lds<T, EPT, false>(work + (size_t)l0 * N, min(maxl, nw - l0), N, d, twl);
        for (int idx = threadIdx.x; idx < nw * K; idx += blockDim.x) {
            const int cl = fdiv(idx, a.fK), kx = idx - cl * K, pg = fdiv(cl, a.flpb), b = cl - pg * lpb, L = L0 + b;
            if (L >= nlines) continue;
            const int p = 3 * grp + pg;
            cplx<T> Zk = work[cl * N + (kx == N ? 0 : kx)];
            cplx<T> Zc = cconj<T>(work[cl * N + (kx == 0 ? 0 : N - kx)]);
            cplx<T> s = cadd<T>(Zk, Zc), df = csub<T>(Zk, Zc);
            cplx<T> wd = cmul<T>(twxl[kx], df);
            a.P[kx + (size_t)K * L + p * pstride] = mk<T>(T(0.5) * (s.x + wd.y), T(0.5) * (s.y - wd.x));
        }
        __syncthreads();
    }
}

// The RHS x pass of k_sp_xmid, persistent over `lpi` consecutive (y, z)
// lines per block, one line at a time: the raw P rows of line L+1 are loaded
// into registers before line L is computed and land in LDS after it, so the
// HBM latency of a line hides behind the previous line's transforms; the
// twiddle tables are staged once per block and the CFL bound is one block
// reduction at the end.  Needs 3K ≤ XR_RP · blockDim.
// LDS: v[3][n] reals, work[3][n/2] and raw[3][K] complex, the n/2 transform
// and n/2+1 real-transform twiddles.  One vmax partial per block.
constexpr int XR_RP = 3; // raw P elements per thread per line

template <typename T, int EPT>
__global__ void __launch_bounds__(256) k_sp_xrhs(XMidArgs<T> a, SpecGeom g, FFTDesc d, int lpi)
{
    extern __shared__ unsigned char smem_raw[];
    const int n = g.n, N = n / 2, K = g.K;
    T *v = reinterpret_cast<T *>(smem_raw);
    cplx<T> *work = reinterpret_cast<cplx<T> *>(smem_raw + sizeof(T) * (size_t)3 * n);
    cplx<T> *raw = work + 3 * N;
    const cplx<T> *twl = stage_twiddles<T>(raw + 3 * K, a.tw, N);
    // real-transform twiddles: LDS copy in FP32 (read after the first barrier), L1/L2 in FP64
    const cplx<T> *twxl = sizeof(T) == 4 ? stage_twiddles<T>(raw + 3 * K + N, a.twx, N + 1) : a.twx;
    const int nlines = n * n;
    const int Lb = xcd_block(blockIdx.x, gridDim.x) * lpi, Le = min(nlines, Lb + lpi);
    const size_t pstride = (size_t)K * n * n; // component stride of P
    const int maxl = max(1, EPT * (int)blockDim.x / N);
    // this thread's raw elements: e_s = tid + s·blockDim → P offset c·pstride + k
    size_t roff[XR_RP];
    bool rok[XR_RP];
#pragma unroll
    for (int s = 0; s < XR_RP; ++s) {
        const int e = threadIdx.x + s * (int)blockDim.x;
        rok[s] = e < 3 * K;
        const int c = fdiv(rok[s] ? e : 0, a.fK), k = (rok[s] ? e : 0) - c * K;
        roff[s] = k + c * pstride;
    }
    cplx<T> r0 = mk<T>(T(0), T(0)), r1 = r0, r2 = r0; // prefetched raw elements (XR_RP = 3)
    static_assert(XR_RP == 3, "three prefetch registers");
    double mx = 0;
    if (Lb < Le) {
        const cplx<T> *src = a.P + (size_t)K * Lb;
        if (rok[0]) r0 = src[roff[0]];
        if (rok[1]) r1 = src[roff[1]];
        if (rok[2]) r2 = src[roff[2]];
    }
    for (int L = Lb; L < Le; ++L) {
        if (rok[0]) raw[threadIdx.x] = r0;
        if (rok[1]) raw[threadIdx.x + blockDim.x] = r1;
        if (rok[2]) raw[threadIdx.x + 2 * blockDim.x] = r2;
        if (L + 1 < Le) {
            const cplx<T> *src = a.P + (size_t)K * (L + 1);
            if (rok[0]) r0 = src[roff[0]];
            if (rok[1]) r1 = src[roff[1]];
            if (rok[2]) r2 = src[roff[2]];
        }
        __syncthreads();
        // c2r of the three components in one batched transform
        for (int idx = threadIdx.x; idx < 3 * N; idx += blockDim.x) {
            const int c = fdiv(idx, a.fN), k = idx - c * N;
            const cplx<T> *row = raw + c * K;
            cplx<T> Xk = k < K ? row[k] : mk<T>(T(0), T(0));
            cplx<T> Xn = (N - k) < K ? row[N - k] : mk<T>(T(0), T(0));
            if (k == 0) {
                Xk.y = T(0);
                Xn.y = T(0);
            }
            const cplx<T> Xc = cconj<T>(Xn);
            const cplx<T> ev = cadd<T>(Xk, Xc);
            const cplx<T> od = cmulc<T>(csub<T>(Xk, Xc), twxl[k]);
            work[idx] = mk<T>(ev.x - od.y, ev.y + od.x);
        }
        __syncthreads();
        for (int l0 = 0; l0 < 3; l0 += maxl) fft_lds<T, EPT, true>(work + (size_t)l0 * N, min(maxl, 3 - l0), N, d, twl);
        for (int idx = threadIdx.x; idx < 3 * N; idx += blockDim.x) {
            const int c = fdiv(idx, a.fN), t = idx - c * N;
            const cplx<T> z = work[idx];
            v[(size_t)c * n + 2 * t] = z.x;
            v[(size_t)c * n + 2 * t + 1] = z.y;
        }
        __syncthreads();
        for (int x = threadIdx.x; x < n; x += blockDim.x)
            mx = fmax(mx, fabs((double)v[x]) + fabs((double)v[n + x]) + fabs((double)v[2 * n + x]));
        // six products, two batched r2c transforms of three
        for (int grp = 0; grp < 2; ++grp) {
            for (int idx = threadIdx.x; idx < 3 * N; idx += blockDim.x) {
                const int pg = fdiv(idx, a.fN), t = idx - pg * N, p = 3 * grp + pg;
                const int ca = p < 3 ? p : (p == 3 ? 0 : p == 4 ? 1 : 2);
                const int cb = p < 3 ? p : (p == 3 ? 1 : p == 4 ? 2 : 0);
                const T *va = v + (size_t)ca * n, *vb = v + (size_t)cb * n;
                work[idx] = mk<T>(va[2 * t] * vb[2 * t], va[2 * t + 1] * vb[2 * t + 1]);
            }
            __syncthreads();
            for (int l0 = 0; l0 < 3; l0 += maxl) fft_lds<T, EPT, false>(work + (size_t)l0 * N, min(maxl, 3 - l0), N, d, twl);
            for (int idx = threadIdx.x; idx < 3 * K; idx += blockDim.x) {
                const int pg = fdiv(idx, a.fK), kx = idx - pg * K, p = 3 * grp + pg;
                const cplx<T> Zk = work[pg * N + (kx == N ? 0 : kx)];
                const cplx<T> Zc = cconj<T>(work[pg * N + (kx == 0 ? 0 : N - kx)]);
                const cplx<T> s = cadd<T>(Zk, Zc), df = csub<T>(Zk, Zc);
                const cplx<T> wd = cmul<T>(twxl[kx], df);
                a.P[kx + (size_t)K * L + p * pstride] = mk<T>(T(0.5) * (s.x + wd.y), T(0.5) * (s.y - wd.x));
            }
            __syncthreads();
        }
    }
    __shared__ double red[256];
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) a.vpart[blockIdx.x] = red[0];
}

// ----------------------------------------- y forward + RHS + IF stage

// Forward y pass of the six product spectra with the divergence-form
// nonlinear term (src/spectral/operators.jl:35-79), the Wray
// integrating-factor stage (src/spectral/step.jl:22-45) and the projection
// (src/spectral/operators.jl:10-22).  Each thread owns up to MAXO outputs
// (b, jc) of the tile; their du accumulate in registers across the six
// product passes.  MODE 0: stage update of uh/ustart; MODE 1: write du;
// MODE 2: write the six product spectra σ̂_p (K, m, m, 6) to s.du.
// grid (ceil(K/tb), m).
template <typename T> struct StageArgs {
    const cplx<T> *Q;
    cplx<T> *uh, *ustart, *du;
    const cplx<T> *extra; // (K, m, m, 6) stress added to σ̂ before the divergence, or null
    const cplx<T> *tw;
    int tb;
    FastDiv ftb; // division by tb
    int first;   // stage 1: the stage start is uh itself
    T a, b, nudcdt, dt;
};

template <typename T, int EPT, int MAXO, int MODE, int NT>
__global__ void __launch_bounds__(NT, 2) k_sp_ystage(StageArgs<T> s, SpecGeom g, FFTDesc d)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
    const int n = g.n, ls = sp_line_stride<T>(n), K = g.K, m = g.m, tb = s.tb;
    const FastDiv ftb = s.ftb;
    // 512-point lines: the per-pass twiddle table (both transform engines
    // read it); FP32 512-point lines are also stored swizzled (fft_lds O512)
    const cplx<T> *twl = n == 512 ? stage_ptw512<T>(buf + (size_t)tb * ls, s.tw)
                                  : stage_twiddles<T>(buf + (size_t)tb * ls, s.tw, n);
    auto lsw = [n](int t) { return sizeof(T) == 4 && n == 512 ? t ^ ((t >> 3) & 31) : t; };
    int bxr, byr;
    xcd_block2(bxr, byr);
    const int i0 = bxr * tb, kc = byr;
    const T inv = T(1.0 / ((double)n * n * n));
    const T kz = T(g.kfac * sp_freq(kc, g));
    cplx<T> acc[MAXO][3];
#pragma unroll
    for (int o = 0; o < MAXO; ++o)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[o][c] = mk<T>(T(0), T(0));
    // The tile of the next product is loaded into registers while the
    // current one is transformed (tb n <= EPT NT, checked by the host).
    // Buffer loads over the (K, n) slab of (kc, p) with 32-bit offsets; the
    // elements outside the tile (i >= K, idx >= tb n) get the slab size as
    // offset, which the buffer range check reads as zero.  No branch around
    // the loads, so they all stay in flight.
    const int slab = K * n * (int)sizeof(cplx<T>);
    int voff[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int idx = threadIdx.x + q * NT;
        const int t = fdiv(idx, ftb), i = i0 + (idx - t * tb);
        voff[q] = (idx < tb * n && i < K) ? (i + K * t) * (int)sizeof(cplx<T>) : slab;
    }
    cplx<T> pre[EPT];
    auto load_tile = [&](int p) {
        const bufres r = make_bufres(s.Q + (size_t)K * n * (kc + (size_t)m * p), (uint32_t)slab);
#pragma unroll
        for (int q = 0; q < EPT; ++q) pre[q] = bload<cplx<T>>(r, voff[q], 0);
    };
    load_tile(0);
    for (int p = 0; p < 6; ++p) {
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int idx = threadIdx.x + q * NT;
            if (idx < tb * n) {
                int t = fdiv(idx, ftb), b = idx - t * tb;
                buf[b * ls + lsw(t)] = pre[q];
            }
        }
        __syncthreads();
        if (p < 5) load_tile(p + 1);
        // FP32: the block-wide engine (the lane-group transforms measured
        // slower here, 3.5 -> 4.3 ms per HIT 512 step, at 16 elements per
        // thread); FP64: wavefront transforms for 512-point lines (the
        // other lane-group lengths are not instantiated here: their code
        // would raise the register count of the kernel past its budget).
        if (sizeof(T) == 8 && n == 512)
            fft_lds_groups<T, false, 512, NT, true>(buf, tb, ls, twl);
        else
            fft_lds<T, EPT, false, NT, sizeof(T) == 4>(buf, tb, ls, d, twl);
        // the output indices are recomputed for each product rather than
        // kept live across the loop (registers)
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
#pragma unroll
        for (int o = 0; o < MAXO; ++o) {
            int oi = tid + o * NT;
            if (oi < tb * m) {
                int jc = fdiv(oi, ftb), b = oi - jc * tb, i = i0 + b;
                if (i < K && MODE == 2) {
                    s.du[i + (size_t)K * (jc + (size_t)m * (kc + (size_t)m * p))] =
                        cscale<T>(buf[b * ls + lsw(sp_fine(jc, g))], inv);
                } else if (i < K) {
                    cplx<T> sg = cscale<T>(buf[b * ls + lsw(sp_fine(jc, g))], inv);
                    if (s.extra) sg = cadd<T>(sg, s.extra[i + (size_t)K * (jc + (size_t)m * (kc + (size_t)m * p))]);
                    T kx = T(g.kfac * i), ky = T(g.kfac * sp_freq(jc, g));
                    // −i k σ  =  (k σ.y, −k σ.x)
                    auto mi = [](T k, cplx<T> z) { return mk<T>(k * z.y, -k * z.x); };
                    switch (p) {
                        case 0: acc[o][0] = cadd<T>(acc[o][0], mi(kx, sg)); break;              // σxx
                        case 1: acc[o][1] = cadd<T>(acc[o][1], mi(ky, sg)); break;              // σyy
                        case 2: acc[o][2] = cadd<T>(acc[o][2], mi(kz, sg)); break;              // σzz
                        case 3:                                                                  // σxy
                            acc[o][0] = cadd<T>(acc[o][0], mi(ky, sg));
                            acc[o][1] = cadd<T>(acc[o][1], mi(kx, sg));
                            break;
                        case 4:                                                                  // σyz
                            acc[o][1] = cadd<T>(acc[o][1], mi(kz, sg));
                            acc[o][2] = cadd<T>(acc[o][2], mi(ky, sg));
                            break;
                        default:                                                                 // σzx
                            acc[o][0] = cadd<T>(acc[o][0], mi(kz, sg));
                            acc[o][2] = cadd<T>(acc[o][2], mi(kx, sg));
                            break;
                    }
                }
            }
        }
        __syncthreads();
    }
    if (MODE == 2) return;
#pragma unroll
    for (int o = 0; o < MAXO; ++o) {
        int oi = threadIdx.x + o * blockDim.x;
        if (oi >= tb * m) continue;
        int jc = fdiv(oi, ftb), b = oi - jc * tb, i = i0 + b;
        if (i >= K) continue;
        const size_t base = i + (size_t)K * (jc + (size_t)m * kc), cs = (size_t)K * m * m;
        if (MODE == 1) {
#pragma unroll
            for (int c = 0; c < 3; ++c) s.du[base + c * cs] = acc[o][c];
            continue;
        }
        T kx = T(g.kfac * i), ky = T(g.kfac * sp_freq(jc, g));
        T kk = kx * kx + ky * ky + kz * kz;
        T phi = exp(-s.nudcdt * kk);
        cplx<T> u[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            cplx<T> st = s.first ? s.uh[base + c * cs] : s.ustart[base + c * cs];
            cplx<T> dd_ = acc[o][c];
            T at = s.a * s.dt, bt = s.b * s.dt;
            u[c] = mk<T>(phi * (st.x + at * dd_.x), phi * (st.y + at * dd_.y));
            s.ustart[base + c * cs] = mk<T>(phi * (st.x + bt * dd_.x), phi * (st.y + bt * dd_.y));
        }
        // projection û ← û − k (k·û)/|k|², k = 0 untouched
        cplx<T> pr = mk<T>(T(0), T(0));
        if (kk != T(0)) {
            pr = mk<T>((kx * u[0].x + ky * u[1].x + kz * u[2].x) / kk, (kx * u[0].y + ky * u[1].y + kz * u[2].y) / kk);
        }
        s.uh[base] = mk<T>(u[0].x - kx * pr.x, u[0].y - kx * pr.y);
        s.uh[base + cs] = mk<T>(u[1].x - ky * pr.x, u[1].y - ky * pr.y);
        s.uh[base + 2 * cs] = mk<T>(u[2].x - kz * pr.x, u[2].y - kz * pr.y);
    }
}

// ------------------------------------------------------- pointwise kernels

template <typename T> __global__ void __launch_bounds__(256) k_sp_project(cplx<T> *uh, SpecGeom g)
{
    const size_t cs = (size_t)g.K * g.m * g.m;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < cs; t += (size_t)gridDim.x * blockDim.x) {
        int i = t % g.K, r = t / g.K, j = r % g.m, k = r / g.m;
        T k1 = T(g.kfac * i), k2 = T(g.kfac * sp_freq(j, g)), k3 = T(g.kfac * sp_freq(k, g));
        T kk = k1 * k1 + k2 * k2 + k3 * k3;
        cplx<T> a = uh[t], b = uh[t + cs], c = uh[t + 2 * cs];
        cplx<T> p = mk<T>(T(0), T(0));
        if (kk != T(0)) p = mk<T>((k1 * a.x + k2 * b.x + k3 * c.x) / kk, (k1 * a.y + k2 * b.y + k3 * c.y) / kk);
        uh[t] = mk<T>(a.x - k1 * p.x, a.y - k1 * p.y);
        uh[t + cs] = mk<T>(b.x - k2 * p.x, b.y - k2 * p.y);
        uh[t + 2 * cs] = mk<T>(c.x - k3 * p.x, c.y - k3 * p.y);
    }
}

__device__ inline int sp_shell(int i, int j, int k, const SpecGeom &g)
{
    long a = i, b = sp_freq(j, g), c = sp_freq(k, g);
    return (int)floor(sqrt((double)(a * a + b * b + c * c)));
}

// Enumerates the modes of the cube |k_i| <= r (all modes when r >= kcut):
// flat index → (i, j, k) state indices.
__device__ inline bool sp_cube(long t, int r, const SpecGeom &g, int &i, int &j, int &k)
{
    if (r >= g.kcut) {
        i = t % g.K;
        long q = t / g.K;
        j = q % g.m;
        k = q / g.m;
        return true;
    }
    int w = 2 * r + 1;
    i = t % (r + 1);
    long q = t / (r + 1);
    int jj = q % w, kk = q / w;
    j = jj <= r ? jj : g.m - (w - jj); // freq jj - w → compact m + (jj - w)
    k = kk <= r ? kk : g.m - (w - kk);
    return true;
}

// Per-block shell energies (src/spectral/forcing.jl:12-24) of shells
// s0..s0+ns−1 over the cube |k_i| <= smax; part[block * ns + s].
// Deterministic: the block stages 256 (shell, energy) pairs in LDS, then the
// thread that owns shell b adds that chunk's contributions to b in index
// order, so every run sums in the same order.  Dynamic LDS: ns doubles +
// 256 doubles + 256 ints.
template <typename T>
__global__ void __launch_bounds__(256) k_sp_shells(const cplx<T> *uh, int s0, int ns, int r, long count, SpecGeom g,
                                                   double *part)
{
    extern __shared__ unsigned char smem_raw[];
    double *bins = reinterpret_cast<double *>(smem_raw);
    double *val = bins + ns;
    int *key = reinterpret_cast<int *>(val + 256);
    const int tid = threadIdx.x;
    for (int b = tid; b < ns; b += 256) bins[b] = 0.0;
    const size_t cs = (size_t)g.K * g.m * g.m;
    for (long base = (long)blockIdx.x * 256; base < count; base += (long)gridDim.x * 256) {
        const long t = base + tid;
        int sh = -1;
        double e = 0;
        if (t < count) {
            int i, j, k;
            sp_cube(t, r, g, i, j, k);
            sh = sp_shell(i, j, k, g) - s0;
            if (sh >= 0 && sh < ns) {
                const size_t at = i + (size_t)g.K * (j + (size_t)g.m * k);
                for (int c = 0; c < 3; ++c) {
                    cplx<T> a = uh[at + c * cs];
                    e += (double)a.x * (double)a.x + (double)a.y * (double)a.y;
                }
                e = (i == 0 ? 1.0 : 2.0) * e / 2;
            } else {
                sh = -1;
            }
        }
        __syncthreads(); // the previous chunk's readers are done with key/val
        key[tid] = sh;
        val[tid] = e;
        __syncthreads();
        const int cnt = (int)min((long)256, count - base);
        for (int b = tid; b < ns; b += 256) {
            double acc = bins[b];
            for (int q = 0; q < cnt; ++q)
                if (key[q] == b) acc += val[q];
            bins[b] = acc;
        }
    }
    __syncthreads();
    for (int b = tid; b < ns; b += 256) part[(size_t)blockIdx.x * ns + b] = bins[b];
}

// k_sp_shells with per-wavefront bins (4 ns doubles of LDS, ns ≤ 2048):
// each wavefront bins its 64 elements per pass with wave_bin_add (a few
// distinct shells per pass) instead of every bin owner scanning the whole
// chunk; the four arrays are summed in wave order.  Deterministic.
template <typename T>
__global__ void __launch_bounds__(256) k_sp_shells_w(const cplx<T> *uh, int s0, int ns, int r, long count, SpecGeom g,
                                                     double *part)
{
    extern __shared__ unsigned char smem_raw[];
    double *bins = reinterpret_cast<double *>(smem_raw); // [4][ns]
    const int tid = threadIdx.x, wave = tid >> 6;
    for (int b = tid; b < 4 * ns; b += 256) bins[b] = 0.0;
    __syncthreads();
    const size_t cs = (size_t)g.K * g.m * g.m;
    for (long base = (long)blockIdx.x * 256; base < count; base += (long)gridDim.x * 256) {
        const long t = base + tid;
        int sh = -1;
        double e = 0;
        if (t < count) {
            int i, j, k;
            sp_cube(t, r, g, i, j, k);
            sh = sp_shell(i, j, k, g) - s0;
            if (sh >= 0 && sh < ns) {
                const size_t at = i + (size_t)g.K * (j + (size_t)g.m * k);
                for (int c = 0; c < 3; ++c) {
                    cplx<T> a = uh[at + c * cs];
                    e += (double)a.x * (double)a.x + (double)a.y * (double)a.y;
                }
                e = (i == 0 ? 1.0 : 2.0) * e / 2;
            } else {
                sh = -1;
            }
        }
        wave_bin_add(bins + wave * ns, sh, e);
    }
    __syncthreads();
    for (int b = tid; b < ns; b += 256)
        part[(size_t)blockIdx.x * ns + b] = ((bins[b] + bins[ns + b]) + bins[2 * ns + b]) + bins[3 * ns + b];
}

template <typename T>
__global__ void __launch_bounds__(256) k_sp_rescale(cplx<T> *uh, int s0, int ns, int r, long count,
                                                    const double *factors, SpecGeom g)
{
    const size_t cs = (size_t)g.K * g.m * g.m;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < count; t += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        sp_cube(t, r, g, i, j, k);
        int sh = sp_shell(i, j, k, g) - s0;
        if (sh < 0 || sh >= ns) continue;
        T f = T(factors[sh]);
        size_t base = i + (size_t)g.K * (j + (size_t)g.m * k);
        for (int c = 0; c < 3; ++c) uh[base + c * cs] = cscale<T>(uh[base + c * cs], f);
    }
}

// Counter-based Gaussian noise (src/spectral/ic.jl:31-48), physical grid.
__device__ inline uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

template <typename T> __global__ void __launch_bounds__(256) k_sp_noise(T *v, uint64_t seed, int n)
{
    const size_t n3 = (size_t)n * n * n;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < 3 * n3; t += (size_t)gridDim.x * blockDim.x) {
        size_t c = t / n3, r = t - c * n3;
        size_t i = r % n, q = r / n, j = q % n, k = q / n;
        uint64_t key = (((uint64_t)c * n + k) * n + j) * n + i;
        uint64_t r1 = splitmix64(seed + 2 * key), r2 = splitmix64(seed + 2 * key + 1);
        double u1 = ((double)(r1 >> 11) + 0.5) * 0x1p-53;
        double u2 = ((double)(r2 >> 11) + 0.5) * 0x1p-53;
        v[t] = T(sqrt(-2.0 * log(u1)) * cospi(2.0 * u2));
    }
}

// Energy (slot 0) and Σ w k²|û|² (slot 1), compensated per block.
template <typename T>
__global__ void __launch_bounds__(256) k_sp_energy(const cplx<T> *uh, SpecGeom g, dd *part)
{
    __shared__ dd s[256];
    const size_t cs = (size_t)g.K * g.m * g.m;
    dd e{0.0, 0.0}, w2{0.0, 0.0};
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < cs; t += (size_t)gridDim.x * blockDim.x) {
        int i = t % g.K, r = t / g.K, j = r % g.m, k = r / g.m;
        double a2 = 0;
        for (int c = 0; c < 3; ++c) {
            cplx<T> a = uh[t + c * cs];
            a2 += (double)a.x * (double)a.x + (double)a.y * (double)a.y;
        }
        double k1 = g.kfac * i, k2 = g.kfac * sp_freq(j, g), k3 = g.kfac * sp_freq(k, g);
        double w = i == 0 ? 1.0 : 2.0;
        e = dd_add1(e, w * a2 / 2);
        w2 = dd_add1(w2, w * a2 * (k1 * k1 + k2 * k2 + k3 * k3));
    }
    dd r0 = block_dd_sum(e, s);
    dd r1 = block_dd_sum(w2, s);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = r0;
        part[gridDim.x + blockIdx.x] = r1;
    }
}

template <typename T> __global__ void __launch_bounds__(256) k_sp_maxdiv(const cplx<T> *uh, SpecGeom g, double *part)
{
    __shared__ double s[256];
    const size_t cs = (size_t)g.K * g.m * g.m;
    double mx = 0;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < cs; t += (size_t)gridDim.x * blockDim.x) {
        int i = t % g.K, r = t / g.K, j = r % g.m, k = r / g.m;
        double k1 = g.kfac * i, k2 = g.kfac * sp_freq(j, g), k3 = g.kfac * sp_freq(k, g);
        cplx<T> a = uh[t], b = uh[t + cs], c = uh[t + 2 * cs];
        double re = k1 * a.x + k2 * b.x + k3 * c.x, im = k1 * a.y + k2 * b.y + k3 * c.y;
        mx = fmax(mx, sqrt(re * re + im * im));
    }
    s[threadIdx.x] = mx;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// ---------------------------------------------- SFS data recipe (les.jl)

// Gaussian test filter with protected low shells (les_filterkernel!,
// src/spectral/les.jl:69-86): 1 for kf²|k|² < (kf (protect+1))², else
// exp(−Δ² k²/24).
__device__ inline double les_gauss(double kk, double kbound2, double delta)
{
    return kk < kbound2 ? 1.0 : exp(-delta * delta * kk / 24.0);
}

// ū = T⅔(G(cut(û))) and G(cut(σ̂)) on the compact layout of the LES grid
// (n = M, kcut = M ÷ 3: exactly the modes T⅔ keeps), gathered from the DNS
// state and its product spectra (src/spectral/les.jl:316-359).  The mean is
// zeroed as les_twothirds! does.
template <typename T>
__global__ void __launch_bounds__(256) k_les_gather(const cplx<T> *uhD, const cplx<T> *sigD, SpecGeom gD,
                                                    cplx<T> *uhL, cplx<T> *tauL, SpecGeom gL, double kbound2,
                                                    double delta)
{
    const size_t csL = (size_t)gL.K * gL.m * gL.m, csD = (size_t)gD.K * gD.m * gD.m;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < csL; t += (size_t)gridDim.x * blockDim.x) {
        int i = t % gL.K, r = t / gL.K, j = r % gL.m, k = r / gL.m;
        int fj = sp_freq(j, gL), fk = sp_freq(k, gL);
        int jD = fj >= 0 ? fj : gD.m + fj, kD = fk >= 0 ? fk : gD.m + fk;
        size_t d = i + (size_t)gD.K * (jD + (size_t)gD.m * kD);
        double k1 = gL.kfac * i, k2 = gL.kfac * fj, k3 = gL.kfac * fk;
        double w = les_gauss(k1 * k1 + k2 * k2 + k3 * k3, kbound2, delta);
        if (i == 0 && fj == 0 && fk == 0) w = 0.0;
        for (int c = 0; c < 3; ++c) uhL[t + c * csL] = cscale<T>(uhD[d + c * csD], T(w));
        for (int p = 0; p < 6; ++p) tauL[t + p * csL] = cscale<T>(sigD[d + p * csD], T(w));
    }
}

// τ = T⅔(G σ̂ − F[ū_i ū_j]/M³), trace-free (src/spectral/les.jl:360-370).
template <typename T>
__global__ void __launch_bounds__(256) k_les_tau(cplx<T> *tauL, const cplx<T> *sig2, SpecGeom gL)
{
    const size_t cs = (size_t)gL.K * gL.m * gL.m;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < cs; t += (size_t)gridDim.x * blockDim.x) {
        cplx<T> v[6];
        const bool mean = t == 0;
        for (int p = 0; p < 6; ++p) v[p] = mean ? mk<T>(T(0), T(0)) : csub<T>(tauL[t + p * cs], sig2[t + p * cs]);
        cplx<T> tr = mk<T>((v[0].x + v[1].x + v[2].x) / T(3), (v[0].y + v[1].y + v[2].y) / T(3));
        for (int p = 0; p < 3; ++p) v[p] = csub<T>(v[p], tr);
        for (int p = 0; p < 6; ++p) tauL[t + p * cs] = v[p];
    }
}

// ------------------------------------------- wavefront x pass (n = 512)

// 256-point complex FFT by one wavefront: lane l holds a[r] = x[l + 64 r] on
// entry and X[l + 64 r] on return; unnormalised (w = exp(∓2πi/256)).  A
// four-step 16 x 16 transform: element n = l + 64 r splits as n = na + 16 nb with
// na = l mod 16 and nb = l div 16 + 4 r, and X[kb + 16 ka] =
// Σ_na w16^(na ka) w^(na kb) Σ_nb x[na + 16 nb] w16^(nb kb).  Each 16-point
// DFT (fft16_rows) is a radix-4 over the registers, the twiddles
// w16^(h k) (h = the lane's row l div 16), a transpose of the four lane rows
// with the four registers (permlane32_swap on register pairs (0, 2), (1, 3),
// then permlane16_swap on (0, 1), (2, 3): no LDS), and a radix-4 over the
// registers again; between the two DFT stages the twiddles w^(na kb) and one
// LDS exchange (element (kb, na) at sh[16 kb + (na ^ kb)], so that sixteen
// lanes of each write and read hit sixteen distinct 16-byte bank groups)
// bring the sixteen kb columns to the lanes.  One LDS exchange per
// transform (a radix-4 Stockham schedule needs three; measured 5.28 → 4.57 ms
// per HIT 512 FP64 step for the x pass).  tw4 (stage_tw4s): w^(na (q + 4 t))
// at t 64 + l (q = l div 16), then w16^(h k) at 256 + 4 h + k, both from
// tw[t] = exp(-2πi t / 256).
constexpr int TW4S_LEN = 256 + 16;
template <typename T> __device__ inline void stage_tw4s(cplx<T> *tw4, const cplx<T> *tw)
{
    for (int e = threadIdx.x; e < TW4S_LEN; e += blockDim.x) {
        if (e < 256) {
            const int t = e >> 6, l = e & 63;
            tw4[e] = tw[(l & 15) * ((l >> 4) + 4 * t)];
        } else {
            const int h = (e - 256) >> 2, k = (e - 256) & 3;
            tw4[e] = tw[16 * h * k];
        }
    }
}

// Swaps of 32-bit words between the lane halves (X = 32) or between the odd
// rows of a and the even rows of b (X = 16), word by word.
template <int X, typename T> __device__ __forceinline__ void permlane_swap(cplx<T> &a, cplx<T> &b)
{
    constexpr int W = (int)(sizeof(cplx<T>) / 4);
    using UV = unsigned __attribute__((ext_vector_type(W)));
    UV ua = __builtin_bit_cast(UV, a), ub = __builtin_bit_cast(UV, b);
#pragma unroll
    for (int i = 0; i < W; ++i) {
        auto r = X == 32 ? __builtin_amdgcn_permlane32_swap(ua[i], ub[i], false, false)
                         : __builtin_amdgcn_permlane16_swap(ua[i], ub[i], false, false);
        ua[i] = r[0];
        ub[i] = r[1];
    }
    a = __builtin_bit_cast(cplx<T>, ua);
    b = __builtin_bit_cast(cplx<T>, ub);
}

// 16-point DFTs down the columns of the wavefront: column c = l mod 16 holds
// element h + 4 s in lane row h = l div 16, register s; on return lane row q,
// register t holds X_c[q + 4 t].
template <typename T, bool INV>
__device__ __forceinline__ void fft16_rows(cplx<T> (&a)[4], const cplx<T> *tw16, int h)
{
    Dft<T, 4, INV>::run(a);
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        const cplx<T> w = tw16[4 * h + k];
        a[k] = INV ? cmulc<T>(a[k], w) : cmul<T>(a[k], w);
    }
    permlane_swap<32, T>(a[0], a[2]);
    permlane_swap<32, T>(a[1], a[3]);
    permlane_swap<16, T>(a[0], a[1]);
    permlane_swap<16, T>(a[2], a[3]);
    Dft<T, 4, INV>::run(a);
}

template <typename T, bool INV>
__device__ __forceinline__ void wave_fft256(cplx<T> (&a)[4], cplx<T> *sh, const cplx<T> *tw4, int lane_)
{
    int lane = lane_;
    if (sizeof(T) == 8) asm volatile("" : "+v"(lane));
    const int c = lane & 15, h = lane >> 4;
    fft16_rows<T, INV>(a, tw4 + 256, h);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const cplx<T> w = tw4[64 * t + lane];
        a[t] = INV ? cmulc<T>(a[t], w) : cmul<T>(a[t], w);
    }
    // lane (na = c, row q = h) register t holds column kb = h + 4 t, row na
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int kb = h + 4 * t;
        sh[16 * kb + (c ^ kb)] = a[t];
    }
    wave_lds_sync();
    // lane (kb = c, row h) register s takes na = h + 4 s
#pragma unroll
    for (int s = 0; s < 4; ++s) a[s] = sh[16 * c + ((h + 4 * s) ^ c)];
    wave_lds_sync();
    fft16_rows<T, INV>(a, tw4 + 256, h);
}

// The RHS x pass (k_sp_xmid, XMID_RHS) for n = 512: one wavefront per
// (y, z) line L, four lines per block.  The wavefront loads the three
// velocity rows P[L, 0:K, c], forms the half-length complex sequence of
// each c2r (Z_k = E_k + i O_k from X_k and conj(X_{N-k})), transforms it
// with wave_fft256 and keeps the 512 physical values of each component in
// registers (8 per lane); the velocity bound max Σ_c |v_c| is a wavefront
// reduction; each of the six products v_a v_b is formed in registers,
// packed as z_t = (p_2t, p_2t+1), transformed forward and unpacked
// (P_k = ½[(Z_k + conj Z_{N-k}) − i w^k (Z_k − conj Z_{N-k})], the partner
// Z_{N-k} through LDS) into P[L, 0:K, p].  Same arithmetic as k_sp_xmid with
// a different factorisation of the 256-point transforms (four-step 16 x 16
// instead of 8·8·4).  No workgroup barrier after the twiddle staging.  One
// vmax partial per block.  LDS: the transform twiddles (stage_tw4s, 272), the
// 257 real-transform twiddles, 4 × 256 complex.
template <typename T>
__global__ void __launch_bounds__(256) k_sp_xrhs256(XMidArgs<T> a, SpecGeom g)
{
    constexpr int N = 256;
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *tw = reinterpret_cast<cplx<T> *>(smem_raw);
    cplx<T> *twx = tw + TW4S_LEN;
    cplx<T> *sh = twx + (N + 1) + (threadIdx.x >> 6) * N;
    __shared__ double red[4];
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    stage_tw4s<T>(tw, a.tw);
    for (int t = threadIdx.x; t <= N; t += blockDim.x) twx[t] = a.twx[t];
    __syncthreads();
    const int K = g.K, n = g.n, nlines = n * n;
    const int L = xcd_block(blockIdx.x, gridDim.x) * 4 + wv;
    const bool live = L < nlines;
    const size_t pstride = (size_t)K * n * n;
    // Rows through buffer resources of K modes (none past the last line):
    // the byte offsets of the lane's modes k = lane + 64 r and N − k are
    // fixed, those at or beyond K are out of range (read as zero, writes
    // dropped), so no branch surrounds the loads and stores.
    const int sz = (int)sizeof(cplx<T>);
    // (wave-uniform values, kept in SGPRs for the buffer resources)
    const uint32_t rowbytes = __builtin_amdgcn_readfirstlane(live ? K * sz : 0);
    const int Lrow = __builtin_amdgcn_readfirstlane(live ? L : 0);
    int ok[4], on[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int k = lane + 64 * r;
        ok[r] = k < K ? k * sz : TILE_OOB;
        on[r] = N - k < K ? (N - k) * sz : TILE_OOB;
    }
    T v[3][8];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const bufres row = make_bufres(a.P + (size_t)Lrow * K + c * pstride, rowbytes);
        cplx<T> z[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int k = lane + 64 * r;
            cplx<T> Xk = bload<cplx<T>>(row, ok[r], 0);
            cplx<T> Xn = bload<cplx<T>>(row, on[r], 0);
            if (k == 0) {
                Xk.y = T(0);
                Xn.y = T(0);
            }
            const cplx<T> Xc = cconj<T>(Xn);
            const cplx<T> ev = cadd<T>(Xk, Xc);
            const cplx<T> od = cmulc<T>(csub<T>(Xk, Xc), twx[k]);
            z[r] = mk<T>(ev.x - od.y, ev.y + od.x);
        }
        wave_fft256<T, true>(z, sh, tw, lane);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            v[c][2 * r] = z[r].x;
            v[c][2 * r + 1] = z[r].y;
        }
    }
    double mx = 0;
#pragma unroll
    for (int e = 0; e < 8; ++e)
        mx = fmax(mx, fabs((double)v[0][e]) + fabs((double)v[1][e]) + fabs((double)v[2][e]));
    if (!live) mx = 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) red[wv] = mx;
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        const int ca = p < 3 ? p : (p == 3 ? 0 : p == 4 ? 1 : 2);
        const int cb = p < 3 ? p : (p == 3 ? 1 : p == 4 ? 2 : 0);
        cplx<T> z[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) z[r] = mk<T>(v[ca][2 * r] * v[cb][2 * r], v[ca][2 * r + 1] * v[cb][2 * r + 1]);
        wave_fft256<T, false>(z, sh, tw, lane);
#pragma unroll
        for (int r = 0; r < 4; ++r) sh[lane + 64 * r] = z[r];
        wave_lds_sync();
        const bufres out = make_bufres(a.P + (size_t)Lrow * K + p * pstride, rowbytes);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int kx = lane + 64 * r;
            const cplx<T> Zk = z[r];
            const cplx<T> Zc = cconj<T>(sh[(N - kx) & (N - 1)]);
            const cplx<T> sm = cadd<T>(Zk, Zc), df = csub<T>(Zk, Zc);
            const cplx<T> wd = cmul<T>(twx[kx], df);
            bstore(out, ok[r], 0, mk<T>(T(0.5) * (sm.x + wd.y), T(0.5) * (sm.y - wd.x)));
        }
        wave_lds_sync();
    }
    __syncthreads();
    if (threadIdx.x == 0) a.vpart[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

}  // namespace dns
