// Batched line FFTs staged in LDS for gfx950.
//
// A workgroup holds `nlines` lines of length n in LDS (line stride `ls`
// complex elements) and transforms all of them in place with a mixed-radix
// Stockham autosort schedule (radices 8, 4, 2, 3, 5):
//
//   pass with radix R, current sub-transform size Ns, m = n/R:
//     for j in [0, m):  k = j mod Ns
//       v[r] = x[j + r m] * w_n^(r k n/(Ns R))          r = 0..R-1
//       V    = DFT_R(v)
//       y[(j div Ns) Ns R + k + r Ns] = V[r]
//
// Every pass first reads all of its butterfly inputs into registers, waits
// on a barrier and then writes the outputs back into the same buffer, so one
// LDS buffer suffices (no ping-pong).  The caller guarantees
// nlines * n <= EPT * blockDim.x, which bounds the butterflies per thread by
// ceil(EPT / R) and keeps every register array statically indexed.
#pragma once

#include "common.hpp"

namespace dns {

constexpr int FFT_MAXPASS = 12;

struct FFTDesc {
    int n;
    int npass;
    int radix[FFT_MAXPASS];
    FastDiv dm[FFT_MAXPASS];  // division by m = n / R of each pass
    FastDiv dns[FFT_MAXPASS]; // division by the sub-transform size Ns of each pass
};

// Factorises n into the supported radices (largest first: 8, 4, 2, then 3, 5).
// Returns false when n has a prime factor other than 2, 3 and 5.
inline bool make_fft_desc(int n, FFTDesc &d)
{
    d.n = n;
    d.npass = 0;
    int m = n;
    if (n < 1) return false;
    int p2 = 0;
    while (m % 2 == 0) {
        m /= 2;
        ++p2;
    }
    while (p2 >= 3) {
        d.radix[d.npass++] = 8;
        p2 -= 3;
    }
    if (p2 == 2) d.radix[d.npass++] = 4;
    if (p2 == 1) d.radix[d.npass++] = 2;
    while (m % 3 == 0) {
        d.radix[d.npass++] = 3;
        m /= 3;
    }
    while (m % 5 == 0) {
        d.radix[d.npass++] = 5;
        m /= 5;
    }
    if (d.npass > FFT_MAXPASS) return false;
    int ns = 1;
    for (int p = 0; p < d.npass; ++p) {
        d.dm[p] = make_fastdiv(n / d.radix[p]);
        d.dns[p] = make_fastdiv(ns);
        ns *= d.radix[p];
    }
    return m == 1;
}

// ------------------------------------------------------------ DFT kernels

template <typename T, int R, bool INV> struct Dft;

template <typename T, bool INV> struct Dft<T, 2, INV> {
    __device__ static inline void run(cplx<T> *v)
    {
        cplx<T> a = cadd<T>(v[0], v[1]), b = csub<T>(v[0], v[1]);
        v[0] = a;
        v[1] = b;
    }
};

// ±i rotation: forward multiplies by -i, inverse by +i.
template <typename T, bool INV> __device__ inline cplx<T> rot_i(cplx<T> z)
{
    return INV ? mk<T>(-z.y, z.x) : mk<T>(z.y, -z.x);
}

template <typename T, bool INV> struct Dft<T, 4, INV> {
    __device__ static inline void run(cplx<T> *v)
    {
        cplx<T> t0 = cadd<T>(v[0], v[2]), t1 = csub<T>(v[0], v[2]);
        cplx<T> t2 = cadd<T>(v[1], v[3]), t3 = rot_i<T, INV>(csub<T>(v[1], v[3]));
        v[0] = cadd<T>(t0, t2);
        v[2] = csub<T>(t0, t2);
        v[1] = cadd<T>(t1, t3);
        v[3] = csub<T>(t1, t3);
    }
};

template <typename T, bool INV> struct Dft<T, 8, INV> {
    __device__ static inline void run(cplx<T> *v)
    {
        const T c = T(0.70710678118654752440);
        cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
        cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
        Dft<T, 4, INV>::run(e);
        Dft<T, 4, INV>::run(o);
        // o[q] *= w8^q, w8 = exp(∓2πi/8)
        cplx<T> o1 = o[1], o3 = o[3];
        if (INV) {
            o[1] = mk<T>(c * (o1.x - o1.y), c * (o1.x + o1.y));
            o[3] = mk<T>(-c * (o3.x + o3.y), c * (o3.x - o3.y));
        } else {
            o[1] = mk<T>(c * (o1.x + o1.y), c * (o1.y - o1.x));
            o[3] = mk<T>(c * (o3.y - o3.x), -c * (o3.x + o3.y));
        }
        o[2] = rot_i<T, INV>(o[2]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[q] = cadd<T>(e[q], o[q]);
            v[q + 4] = csub<T>(e[q], o[q]);
        }
    }
};

template <typename T, bool INV> struct Dft<T, 3, INV> {
    __device__ static inline void run(cplx<T> *v)
    {
        const T h = T(0.86602540378443864676) * (INV ? T(1) : T(-1));
        cplx<T> t1 = cadd<T>(v[1], v[2]);
        cplx<T> t2 = mk<T>(v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y);
        cplx<T> d = csub<T>(v[1], v[2]);
        cplx<T> t3 = mk<T>(-h * d.y, h * d.x);  // i * h * (v1 - v2)
        v[0] = cadd<T>(v[0], t1);
        v[1] = cadd<T>(t2, t3);
        v[2] = csub<T>(t2, t3);
    }
};

template <typename T, bool INV> struct Dft<T, 5, INV> {
    __device__ static inline void run(cplx<T> *v)
    {
        const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
        const T sg = INV ? T(1) : T(-1);
        const T s1 = sg * T(0.95105651629515357212), s2 = sg * T(0.58778525229247312917);
        cplx<T> a1 = cadd<T>(v[1], v[4]), a2 = cadd<T>(v[2], v[3]);
        cplx<T> b1 = csub<T>(v[1], v[4]), b2 = csub<T>(v[2], v[3]);
        cplx<T> x0 = v[0];
        cplx<T> p1 = mk<T>(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
        cplx<T> p2 = mk<T>(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
        cplx<T> q1 = mk<T>(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
        cplx<T> q2 = mk<T>(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
        cplx<T> iq1 = cmuli<T>(q1), iq2 = cmuli<T>(q2);
        v[0] = mk<T>(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
        v[1] = cadd<T>(p1, iq1);
        v[4] = csub<T>(p1, iq1);
        v[2] = cadd<T>(p2, iq2);
        v[3] = csub<T>(p2, iq2);
    }
};

// --------------------------------------------------------------- one pass

// O512 (512-point lines, radix 8): element i of a line sits at
// i ^ ((i / 8) mod 32), under which the 32 lanes of each pass's stride-8
// writes hit distinct LDS banks, and tw is the per-pass table of
// stage_ptw512.
template <typename T, int R, int EPT, bool INV, bool O512 = false>
__device__ inline void fft_pass(cplx<T> *buf, int nlines, int ls, int n, int Ns, FastDiv dm, FastDiv dns,
                                const cplx<T> *__restrict__ tw, int tid, int nthr)
{
    auto x = [](int i) { return O512 ? i ^ ((i >> 3) & 31) : i; };
    constexpr int BPT = (EPT + R - 1) / R;
    const int m = n / R;
    const int total = nlines * m;
    cplx<T> v[BPT][R];
#pragma unroll
    for (int s = 0; s < BPT; ++s) {
        int t = tid + s * nthr;
        if (t < total) {
            int line = fdiv(t, dm);
            int j = t - line * m;
            const cplx<T> *src = buf + line * ls;
#pragma unroll
            for (int r = 0; r < R; ++r) v[s][r] = src[x(j + r * m)];
        }
    }
    __syncthreads();
    const int tws = n / (Ns * R);
#pragma unroll
    for (int s = 0; s < BPT; ++s) {
        int t = tid + s * nthr;
        if (t < total) {
            int line = fdiv(t, dm);
            int j = t - line * m;
            int jd = fdiv(j, dns);
            int k = j - jd * Ns;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    cplx<T> w = O512 ? tw[(Ns == 8 ? 0 : 56) + (r - 1) * Ns + k] : tw[r * k * tws];
                    v[s][r] = INV ? cmulc<T>(v[s][r], w) : cmul<T>(v[s][r], w);
                }
            }
            Dft<T, R, INV>::run(v[s]);
            cplx<T> *dst = buf + line * ls;
            const int base = jd * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[x(base + r * Ns)] = v[s][r];
        }
    }
    __syncthreads();
}

// Copies the n twiddles of a length-n transform into LDS at dst (every
// butterfly pass reads them); the first barrier after the tile load orders
// the copy before the transform.
template <typename T>
__device__ inline const cplx<T> *stage_twiddles(cplx<T> *dst, const cplx<T> *tw, int n)
{
    for (int t = threadIdx.x; t < n; t += blockDim.x) dst[t] = tw[t];
    return dst;
}

// Transforms nlines lines of d.n points held in LDS at buf[line*ls + t].
// tw[t] = exp(-2πi t / n) for t in [0, n).  Unnormalised in both directions.
// Must be called by every thread of the block (contains barriers).
// NT > 0: the block size is known at compile time (blockDim.x is then not
// read, which would be a memory load).  O512: when d.n = 512 the lines are
// stored swizzled (element i at i ^ ((i / 8) mod 32)) and tw is the per-pass
// table of stage_ptw512 (the caller stages and addresses the lines so).
template <typename T, int EPT, bool INV, int NT = 0, bool O512 = false>
__device__ inline void fft_lds(cplx<T> *buf, int nlines, int ls, const FFTDesc &d, const cplx<T> *__restrict__ tw)
{
    const int tid = threadIdx.x, nthr = NT > 0 ? NT : (int)blockDim.x;
    const bool o512 = O512 && d.n == 512;
    int Ns = 1;
    for (int p = 0; p < d.npass; ++p) {
        int R = d.radix[p];
        switch (R) {
            case 8:
                if (o512)
                    fft_pass<T, 8, EPT, INV, O512>(buf, nlines, ls, d.n, Ns, d.dm[p], d.dns[p], tw, tid, nthr);
                else
                    fft_pass<T, 8, EPT, INV>(buf, nlines, ls, d.n, Ns, d.dm[p], d.dns[p], tw, tid, nthr);
                break;
            case 4: fft_pass<T, 4, EPT, INV>(buf, nlines, ls, d.n, Ns, d.dm[p], d.dns[p], tw, tid, nthr); break;
            case 2: fft_pass<T, 2, EPT, INV>(buf, nlines, ls, d.n, Ns, d.dm[p], d.dns[p], tw, tid, nthr); break;
            case 3: fft_pass<T, 3, EPT, INV>(buf, nlines, ls, d.n, Ns, d.dm[p], d.dns[p], tw, tid, nthr); break;
            case 5: fft_pass<T, 5, EPT, INV>(buf, nlines, ls, d.n, Ns, d.dm[p], d.dns[p], tw, tid, nthr); break;
            default: break;
        }
        Ns *= R;
    }
}

// ------------------------------------------------ wavefront-group FFTs

// Orders the LDS accesses of one wavefront around an exchange (the lanes of
// a wavefront run in lockstep and LDS serves one wavefront's requests in
// order, so no workgroup barrier is needed).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// N-point complex FFT (N = 4^a or 2·4^a, 16 <= N <= 256) by a group of
// G = N/4 lanes of one wavefront (64/G lines per wavefront).  Lane t of the
// group holds a[r] = x[t + r G] on entry and X[t + r G] on return.  The same
// Stockham schedule as fft_lds (pass of radix R with sub-transform size Ns:
// inputs x[j + r N/R] twiddled by w^(r k N/(Ns R)), k = j mod Ns, outputs at
// (j div Ns) R Ns + k + r Ns): radix-4 passes, then one radix-2 pass when N
// is 2·4^a, whose two butterflies per lane (j = t, t + G) take a[0], a[2] and
// a[1], a[3].  Between passes the outputs go through the group's N-entry LDS
// line sh; the last pass's outputs land at t + r G, in the registers.
// The exchanges keep element i at sh[i ^ ((i / 4) mod 16)], a bijection of
// [0, N) under which every sixteen lanes of each exchange access sixteen
// distinct 16-byte LDS bank groups (unswizzled, the writes of the first
// passes put four lanes of every sixteen on the same banks).
// tw[t] = exp(-2πi t / N), or with PTW the per-pass table of
// stage_group_twiddles; unnormalised.
template <int N> struct GroupFFT {
    static constexpr int G = N / 4;
    static constexpr int P4 = (N == 16 || N == 32) ? 2 : (N == 64 || N == 128) ? 3 : 4; // radix-4 passes
    static constexpr bool R2 = (N == 32 || N == 128);
    // per-pass table: w^(r k N / (4 Ns)) of pass p >= 1 (Ns = 4^p) at
    // off(Ns) + (r - 1) Ns + k, then (R2) w^j, j < N/2, at r2off
    static constexpr __host__ __device__ int off(int Ns) { return Ns == 4 ? 0 : Ns == 16 ? 12 : 60; }
    static constexpr int r2off = (1 << (2 * P4)) - 4;
    static constexpr int len = r2off + (R2 ? N / 2 : 0);
};

// Stages the per-pass twiddle table of group_fft<N> into LDS from
// tw[t] = exp(-2πi t / N), t < N (the same values, rearranged so that the
// lanes of a pass read consecutive entries).
template <typename T, int N> __device__ inline const cplx<T> *stage_group_twiddles(cplx<T> *dst, const cplx<T> *tw)
{
    using F = GroupFFT<N>;
    for (int e = threadIdx.x; e < F::len; e += blockDim.x) {
        if (e >= F::r2off) {
            dst[e] = tw[e - F::r2off];
        } else {
            const int Ns = e < 12 ? 4 : e < 60 ? 16 : 64;
            const int j = e - F::off(Ns), r = j / Ns + 1, k = j - (r - 1) * Ns;
            dst[e] = tw[r * k * (N / (4 * Ns))];
        }
    }
    return dst;
}

template <typename T, bool INV, int N, bool PTW = false>
__device__ __forceinline__ void group_fft(cplx<T> (&a)[4], cplx<T> *sh, const cplx<T> *tw, int t)
{
    using F = GroupFFT<N>;
    constexpr int G = F::G, P4 = F::P4;
    constexpr bool R2 = F::R2;
    static_assert(N == 16 || N == 32 || N == 64 || N == 128 || N == 256, "group_fft: N not supported");
    auto x = [](int i) { return i ^ ((i >> 2) & 15); };
#pragma unroll
    for (int p = 0; p < P4; ++p) {
        const int Ns = 1 << (2 * p);
        const int k = t & (Ns - 1);
        if (p > 0) {
#pragma unroll
            for (int r = 1; r < 4; ++r) {
                const cplx<T> w = PTW ? tw[F::off(Ns) + (r - 1) * Ns + k] : tw[r * k * (N / (4 * Ns))];
                a[r] = INV ? cmulc<T>(a[r], w) : cmul<T>(a[r], w);
            }
        }
        Dft<T, 4, INV>::run(a);
        if (p < P4 - 1 || R2) {
            const int base = (t - k) * 4 + k;
#pragma unroll
            for (int r = 0; r < 4; ++r) sh[x(base + r * Ns)] = a[r];
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = sh[x(t + G * r)];
            wave_lds_sync();
        }
    }
    if (R2) {
        // Ns = N/2: butterflies j = t (a[0], a[2]) and j = t + G (a[1], a[3]),
        // k = j, outputs at j and j + N/2: a[r] ends as X[t + r G]
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            cplx<T> b[2] = {a[h], a[h + 2]};
            const cplx<T> w = tw[(PTW ? F::r2off : 0) + t + h * G];
            b[1] = INV ? cmulc<T>(b[1], w) : cmul<T>(b[1], w);
            Dft<T, 2, INV>::run(b);
            a[h] = b[0];
            a[h + 2] = b[1];
        }
    }
}

// 512-point complex FFT by one wavefront: lane t holds a[r] = x[t + 64 r]
// (r = 0..7) on entry and X[t + 64 r] on return; Stockham radix 8 in three
// passes (Ns = 1, 8, 64) with the exchanges through the line's LDS storage
// sh, as group_fft.  tw[t] = exp(-2πi t / 512); unnormalised.
// The exchanges keep element i at sh[i ^ ((i / 8) mod 16)], a bijection of
// [0, 512): the first pass writes element 8 t + r from lane t, which would
// put every other lane on the same LDS banks; swizzled, every sixteen lanes
// of each exchange write and read sixteen distinct 16-byte bank groups.
// PTW: tw is the per-pass table of stage_ptw512 (w^(r k 64 / Ns) at
// ptw[o_p + (r - 1) Ns + k]), read by the lanes of a pass at consecutive
// entries; indexing w^t by r k 64 / Ns puts the lanes of the Ns = 8 pass on
// one bank.
constexpr int PTW512_LEN = 7 * (8 + 64);
template <typename T> __device__ inline const cplx<T> *stage_ptw512(cplx<T> *dst, const cplx<T> *tw)
{
    for (int e = threadIdx.x; e < PTW512_LEN; e += blockDim.x) {
        const int Ns = e < 56 ? 8 : 64;
        const int j = e - (Ns == 8 ? 0 : 56), r = j / Ns + 1, k = j - (r - 1) * Ns;
        dst[e] = tw[r * k * (64 / Ns)];
    }
    return dst;
}

template <typename T, bool INV, bool PTW = false>
__device__ __forceinline__ void wave_fft512(cplx<T> (&a)[8], cplx<T> *sh, const cplx<T> *tw, int t)
{
    auto x = [](int i) { return i ^ ((i >> 3) & 15); };
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const int Ns = 1 << (3 * p);
        const int k = t & (Ns - 1);
        if (p > 0) {
#pragma unroll
            for (int r = 1; r < 8; ++r) {
                const cplx<T> w = PTW ? tw[(Ns == 8 ? 0 : 56) + (r - 1) * Ns + k] : tw[r * k * (64 / Ns)];
                a[r] = INV ? cmulc<T>(a[r], w) : cmul<T>(a[r], w);
            }
        }
        Dft<T, 8, INV>::run(a);
        if (p < 2) {
            const int base = (t - k) * 8 + k;
#pragma unroll
            for (int r = 0; r < 8; ++r) sh[x(base + r * Ns)] = a[r];
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < 8; ++r) a[r] = sh[x(t + 64 * r)];
            wave_lds_sync();
        }
    }
}

// In-place transforms of nlines LDS lines of N points (buf[line * ls + t])
// by wavefront lane groups: G = N/4 lanes per line (64 for N = 512, 8 points
// per lane), 64/G lines per wavefront at a time, the wavefronts of the block
// taking lines in turn.  Needs nlines % (64/G) == 0.  Ends with a workgroup
// barrier, as fft_lds.  PTW: tw is the per-pass table (stage_ptw512 for
// N = 512, stage_group_twiddles otherwise).
template <typename T, bool INV, int N, int NT = 0, bool PTW = false>
__device__ __forceinline__ void fft_lds_groups(cplx<T> *buf, int nlines, int ls, const cplx<T> *tw)
{
    constexpr int G = N == 512 ? 64 : N / 4, LPW = 64 / G;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (NT > 0 ? NT : (int)blockDim.x) >> 6;
    const int t = lane % G, lg = lane / G;
    for (int l0 = wv * LPW; l0 < nlines; l0 += nw * LPW) {
        cplx<T> *line = buf + (size_t)(l0 + lg) * ls;
        if constexpr (N == 512) {
            cplx<T> a[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) a[r] = line[t + 64 * r];
            wave_lds_sync();
            wave_fft512<T, INV, PTW>(a, line, tw, t);
#pragma unroll
            for (int r = 0; r < 8; ++r) line[t + 64 * r] = a[r];
        } else {
            cplx<T> a[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) a[r] = line[t + G * r];
            wave_lds_sync();
            group_fft<T, INV, N, PTW>(a, line, tw, t);
#pragma unroll
            for (int r = 0; r < 4; ++r) line[t + G * r] = a[r];
        }
        wave_lds_sync();
    }
    __syncthreads();
}

// The batched LDS transforms of a tile: lane groups for the lengths they
// serve (16 ... 512, powers of two, nlines a multiple of the lines per
// wavefront), otherwise fft_lds over sub-batches of at most sub lines.
// Whether fft_lds_tile transforms nlines lines of n points by lane groups.
__device__ inline bool fft_tile_groups(int n, int nlines)
{
    return n == 512 || n == 256 || (n == 128 && nlines % 2 == 0) || (n == 64 && nlines % 4 == 0);
}

// The twiddles fft_lds_tile<..., PTW = true> reads, staged into LDS at dst
// from tw[t] = exp(-2πi t / n): the per-pass table of the lane-group
// transform when it serves (n, nlines), otherwise tw itself (n entries).
template <typename T> __device__ inline const cplx<T> *stage_tile_twiddles(cplx<T> *dst, const cplx<T> *tw, int n, int nlines)
{
    if (!fft_tile_groups(n, nlines)) return stage_twiddles<T>(dst, tw, n);
    switch (n) {
        case 512: return stage_ptw512<T>(dst, tw);
        case 256: return stage_group_twiddles<T, 256>(dst, tw);
        case 128: return stage_group_twiddles<T, 128>(dst, tw);
        default: return stage_group_twiddles<T, 64>(dst, tw);
    }
}

// PTW: tw comes from stage_tile_twiddles (per-pass tables for the lane-group
// lengths), else it is exp(-2πi t / n), t < n.
template <typename T, int EPT, bool INV, int NT = 0, bool PTW = false>
__device__ __forceinline__ void fft_lds_tile(cplx<T> *buf, int nlines, int ls, const FFTDesc &d,
                                             const cplx<T> *tw, int sub)
{
    const int n = d.n;
    if (n == 512) return fft_lds_groups<T, INV, 512, NT, PTW>(buf, nlines, ls, tw);
    if (n == 256) return fft_lds_groups<T, INV, 256, NT, PTW>(buf, nlines, ls, tw);
    if (n == 128 && nlines % 2 == 0) return fft_lds_groups<T, INV, 128, NT, PTW>(buf, nlines, ls, tw);
    if (n == 64 && nlines % 4 == 0) return fft_lds_groups<T, INV, 64, NT, PTW>(buf, nlines, ls, tw);
    for (int sb = 0; sb < nlines; sb += sub)
        fft_lds<T, EPT, INV, NT>(buf + (size_t)sb * ls, min(sub, nlines - sb), ls, d, tw);
}

// Line tiles in HBM: a block holds tb lines (line b valid for b < nb) of n
// points; point t of line b lives at base + b + stride·t and at
// buf[b ls + t] in LDS.  When tb divides NT and NT divides tb·n
// (power-of-two n and tile width), element e of thread x is always line
// b = x mod tb, point t = x / tb + e·NT/tb (tile_walk): the global byte
// offset is a per-lane part fixed for the thread plus a block-uniform step
// (an SGPR), and the LDS offset the thread's base plus a uniform step, so no
// element pays an index division or a 64-bit address product (half the VALU
// work of an FP32 512-point pencil pass went there).  Other tiles take the
// per-element division.  Invalid lines take the byte offset 0x80000000,
// beyond the buffer range: they read zero and their writes are dropped.
constexpr int TILE_OOB = (int)0x80000000u;

template <typename T, int NT>
__device__ inline bool tile_walk(int tb, FastDiv ftb, int n, int stride, int &b, int &t0)
{
    static_assert((NT & (NT - 1)) == 0, "tile_walk: NT must be a power of two");
    const bool fast = (tb & (tb - 1)) == 0 && tb <= NT && ((tb * n) & (NT - 1)) == 0 &&
                      ((size_t)stride * n + tb) * sizeof(cplx<T>) < (size_t)0x7fffffff;
    t0 = fdiv((int)threadIdx.x, ftb);
    b = (int)threadIdx.x - t0 * tb;
    return fast;
}

// buf[b ls + t] = base[b + stride t], t < n (U loads in flight per thread on
// the division path).
template <typename T, int EPT, int NT, int U>
__device__ inline void tile_load(cplx<T> *buf, int ls, const cplx<T> *base, int stride, int nb, int tb, FastDiv ftb,
                                 int n)
{
    const int sz = (int)sizeof(cplx<T>);
    int b, t0;
    if (tile_walk<T, NT>(tb, ftb, n, stride, b, t0)) {
        const int S = NT / tb, E = tb * n / NT;
        const bufres r = make_bufres(base, (uint32_t)(((size_t)stride * (n - 1) + tb) * sz));
        const int v0 = b < nb ? (b + stride * t0) * sz : TILE_OOB;
        cplx<T> pre[EPT];
#pragma unroll
        for (int e = 0; e < EPT; ++e)
            if (e < E) pre[e] = bload<cplx<T>>(r, v0, e * S * stride * sz);
        cplx<T> *row = buf + b * ls + t0;
#pragma unroll
        for (int e = 0; e < EPT; ++e)
            if (e < E) row[e * S] = pre[e];
        return;
    }
    batched_for<U, cplx<T>>(
        tb * n,
        [&](int idx) {
            const int t = fdiv(idx, ftb), bb = idx - t * tb;
            return bb < nb ? base[bb + (size_t)stride * t] : mk<T>(T(0), T(0));
        },
        [&](int idx, cplx<T> v) {
            const int t = fdiv(idx, ftb), bb = idx - t * tb;
            buf[bb * ls + t] = v;
        });
}

// base[b + stride t] = buf[b ls + t], t < n; the LDS reads all issued before
// the stores (no branch between them).
template <typename T, int EPT, int NT>
__device__ inline void tile_store(const cplx<T> *buf, int ls, cplx<T> *base, int stride, int nb, int tb, FastDiv ftb,
                                  int n)
{
    const int sz = (int)sizeof(cplx<T>);
    int b, t0;
    if (tile_walk<T, NT>(tb, ftb, n, stride, b, t0)) {
        const int S = NT / tb, E = tb * n / NT;
        const bufres r = make_bufres(base, (uint32_t)(((size_t)stride * (n - 1) + tb) * sz));
        const int v0 = b < nb ? (b + stride * t0) * sz : TILE_OOB;
        const cplx<T> *row = buf + b * ls + t0;
        cplx<T> v[EPT];
#pragma unroll
        for (int e = 0; e < EPT; ++e)
            if (e < E) v[e] = row[e * S];
#pragma unroll
        for (int e = 0; e < EPT; ++e)
            if (e < E) bstore(r, v0, e * S * stride * sz, v[e]);
        return;
    }
    for (int idx = threadIdx.x; idx < tb * n; idx += NT) {
        const int t = fdiv(idx, ftb), bb = idx - t * tb;
        if (bb < nb) base[bb + (size_t)stride * t] = buf[bb * ls + t];
    }
}

}  // namespace dns
