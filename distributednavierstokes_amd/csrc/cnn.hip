// Host side of the sub-filter-stress CNN: a 4-layer 3×3×3 periodic CNN
// (3 → 64 → 64 → 64 → 6 channels, erf-GELU) on M³ LES grids, trained with
// Adam on filtered-DNS pairs held on the GPU; BF16 weights and activations
// with FP32 accumulation (v_mfma_f32_32x32x8_bf16).  C ABI: dns_cnn_*.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "closure.hpp"
#include "cnn_kernels.hpp"
#include "host_util.hpp"

namespace dns {

struct CNNLayer {
    int kin, kinp, nout, np;   // input channels (padded to kinp), outputs (padded to np)
    int kpt, noutp;            // transposed copy: rows kpt (kin padded to 32), cols noutp (nout padded to 8)
    size_t woff, boff;         // offsets of weights / bias in the flat parameter vector
    unsigned short *W = nullptr, *WT = nullptr;
};

// Planar (c, Vs) → channels-last BF16 with scale (one sample into one batch slot).
template <typename T>
__global__ void k_planar_to_bf16(const T *x, int cin, long Vs, float scale, unsigned short *out, int ldc)
{
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < Vs * ldc; e += (long)gridDim.x * blockDim.x) {
        const long v = e / ldc;
        const int c = (int)(e - v * ldc);
        out[e] = c < cin ? f2bf(scale * (float)x[(size_t)c * Vs + v]) : (unsigned short)0;
    }
}

// Planar (c, Vs) FP32 → interleaved (Vs, c) FP32 with scale.
__global__ void k_planar_to_interleaved(const float *x, int c_, long Vs, float scale, float *out)
{
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < Vs * c_; e += (long)gridDim.x * blockDim.x) {
        const long v = e / c_;
        const int c = (int)(e - v * c_);
        out[e] = scale * x[(size_t)c * Vs + v];
    }
}

// Interleaved (Vs, c) FP32 → planar (c, Vs) of T with scale.
template <typename T> __global__ void k_interleaved_to_planar(const float *y, int c_, long Vs, double scale, T *out)
{
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < Vs * c_; e += (long)gridDim.x * blockDim.x) {
        const int c = (int)(e / Vs);
        const long v = e - (long)c * Vs;
        out[e] = (T)(scale * (double)y[v * c_ + c]);
    }
}

template <typename T> __global__ void k_cast_f32(const T *x, long n, float *out)
{
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        out[e] = (float)x[e];
}

__global__ void __launch_bounds__(256) k_sumsq(const float *x, long n, double *part)
{
    __shared__ double s[256];
    double acc = 0;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        acc += (double)x[e] * x[e];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// dst = scale · Σ part[0..nparts) (one block).
__global__ void __launch_bounds__(256) k_sum_final(const double *part, int nparts, double scale, double *dst)
{
    __shared__ double s[256];
    double acc = 0;
    for (int q = threadIdx.x; q < nparts; q += blockDim.x) acc += part[q];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *dst = scale * s[0];
}

struct CNN {
    int M = 0, B = 0;
    long V = 0, Vs = 0; // B M³ voxels per batch, M³ per sample
    hipStream_t st = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    CNNLayer L[4];
    size_t nparam = 0;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr; // FP32 masters, Adam moments, gradient
    int adam_t = 0;
    // activations (BF16 channels-last) and buffers
    unsigned short *x0 = nullptr, *z[3] = {}, *a[3] = {}, *dy = nullptr, *dz[2] = {};
    float *y = nullptr, *t = nullptr;
    float *wpart = nullptr, *bpart = nullptr;
    double *lpart = nullptr, *lval = nullptr, *pinned = nullptr;
    int nsplit = 64, bblocks = 1024;
    static constexpr int kLossBlocks = 512;
    // dataset of filtered-DNS pairs (FP32 planar per sample: ū (3, M³), τ (6, M³))
    std::vector<float *> ds_u, ds_t;
    double su = 1.0, stau = 1.0; // normalisation: RMS of ū and τ components
    LaunchTimer timer;

    ~CNN()
    {
        if (st) hipStreamSynchronize(st);
        timer.flush();
        for (auto &l : L) {
            if (l.W) hipFree(l.W);
            if (l.WT) hipFree(l.WT);
        }
        void *ptrs[] = {w, m, v, g, x0, z[0], z[1], z[2], a[0], a[1], a[2], dy, dz[0], dz[1], y, t, wpart, bpart, lpart, lval};
        for (void *p : ptrs)
            if (p) hipFree(p);
        for (float *p : ds_u) hipFree(p);
        for (float *p : ds_t) hipFree(p);
        if (pinned) hipHostFree(pinned);
        if (ev_in) hipEventDestroy(ev_in);
        if (ev_out) hipEventDestroy(ev_out);
        if (st) hipStreamDestroy(st);
    }

    int create(int M_, int B_, uint64_t seed)
    {
        M = M_;
        B = B_;
        Vs = (long)M * M * M;
        V = (long)B * Vs;
        DNS_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        DNS_HIP_CHECK(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
        DNS_HIP_CHECK(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
        timer.st = st;
        const char *env = getenv("DNS_CNN_MFMA");
        mfma_k8 = env && !strcmp(env, "k8");
        const char *envc = getenv("DNS_CNN_CONV");
        row_conv = !(envc && !strcmp(envc, "direct"));
        conv_pf = row_conv && !(envc && !strcmp(envc, "rows"));
        if (const char *ew = getenv("DNS_CNN_WLDS")) conv_wlds = atoi(ew) != 0;
        const int kin[4] = {3, 64, 64, 64}, nout[4] = {64, 64, 64, 6};
        size_t off = 0;
        for (int l = 0; l < 4; ++l) {
            CNNLayer &c = L[l];
            c.kin = kin[l];
            c.kinp = (kin[l] + 15) / 16 * 16;
            c.nout = nout[l];
            c.np = (nout[l] + 31) / 32 * 32;
            c.kpt = (kin[l] + 31) / 32 * 32;
            c.noutp = (nout[l] + 15) / 16 * 16;
            c.woff = off;
            off += (size_t)27 * c.nout * c.kin;
            c.boff = off;
            off += c.nout;
            DNS_TRY(dalloc(&c.W, (size_t)27 * c.np * c.kinp));
            DNS_TRY(dalloc(&c.WT, (size_t)27 * c.kpt * c.noutp));
        }
        nparam = off;
        // uniform initialisation with variance 1/fan_in from a seeded generator (host, once)
        std::vector<float> h(nparam, 0.f);
        std::mt19937_64 gen(seed);
        for (int l = 0; l < 4; ++l) {
            const CNNLayer &c = L[l];
            const double fan_in = 27.0 * c.kin;
            const double bound = std::sqrt(3.0 / fan_in);
            std::uniform_real_distribution<double> U(-bound, bound);
            for (size_t t2 = 0; t2 < (size_t)27 * c.nout * c.kin; ++t2) h[c.woff + t2] = (float)U(gen);
        }
        DNS_TRY(dupload(&w, h));
        std::fill(h.begin(), h.end(), 0.f);
        DNS_TRY(dupload(&m, h));
        DNS_TRY(dupload(&v, h));
        DNS_TRY(dupload(&g, h));
        DNS_TRY(pack());
        DNS_TRY(dalloc(&x0, (size_t)V * 16));
        for (int q = 0; q < 3; ++q) {
            DNS_TRY(dalloc(&z[q], (size_t)V * 64));
            DNS_TRY(dalloc(&a[q], (size_t)V * 64));
        }
        DNS_TRY(dalloc(&dy, (size_t)V * 16));
        DNS_TRY(dalloc(&dz[0], (size_t)V * 64));
        DNS_TRY(dalloc(&dz[1], (size_t)V * 64));
        DNS_TRY(dalloc(&y, (size_t)V * 6));
        DNS_TRY(dalloc(&t, (size_t)V * 6));
        nsplit = (int)std::max<long>(1, std::min<long>(64, V / 4096));
        // the row-staged weight-gradient kernel splits the B·M² x-rows over at most 256 blocks
        row_wgrad = !mfma_k8 && M % 16 == 0 && M <= 256;
        const char *envw = getenv("DNS_CNN_WGRAD");
        rows_wgrad = row_wgrad && M <= 64 && envw && !strcmp(envw, "rows");
        planes_wgrad = row_wgrad && M <= 64 && !(envw && (!strcmp(envw, "rows") || !strcmp(envw, "row16")));
        if (const char *es = getenv("DNS_CNN_WSPLIT")) wgrad_splits = std::max(1, std::min(256, atoi(es)));
        const int nsplit16 = row_wgrad ? (int)std::min<long>(256, (long)B * M * M) : 0;
        DNS_TRY(dalloc(&wpart, (size_t)std::max(nsplit, nsplit16) * 27 * 64 * 64));
        DNS_TRY(dalloc(&bpart, (size_t)64 * bblocks));
        DNS_TRY(dalloc(&lpart, kLossBlocks));
        DNS_TRY(dalloc(&lval, 1));
        DNS_HIP_CHECK(hipHostMalloc((void **)&pinned, sizeof(double) * 1024));
        return DNS_OK;
    }

    int pack()
    {
        for (int l = 0; l < 4; ++l) {
            CNNLayer &c = L[l];
            hipLaunchKernelGGL(k_pack_weights, dim3(64), dim3(256), 0, st, w + c.woff, c.nout, c.kin, c.W, c.np, c.kinp,
                               c.WT, c.kpt, c.noutp);
        }
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    ConvGeom geom(long Vn) const { return ConvGeom{M, (int)(Vn / Vs), Vn}; }

    // 3x3x3 convolution launch: 4 waves of (64 voxels × 32·NB channels) on
    // the K = 16 MFMA; DNS_CNN_MFMA=k8 selects the 32x32x8 kernel instead.
    bool mfma_k8 = false;
    bool row_wgrad = false; // k_conv_wgrad16 (M % 16 == 0) instead of k_conv_wgrad
    bool rows_wgrad = false;   // k_conv_wgrad_rows (M % 16 == 0, M <= 64); DNS_CNN_WGRAD=rows
    bool planes_wgrad = false; // k_conv_wgrad_planes (M % 16 == 0, M <= 64), the default
    int wgrad_splits = 128;    // plane splits of k_conv_wgrad_planes (DNS_CNN_WSPLIT, sweep)
    bool row_conv = false;  // k_conv_rows (M = 32 or 64) instead of k_conv16; DNS_CNN_CONV=direct disables
    bool conv_pf = false;   // k_conv_rows_pf at M = 64 (kin 16 or 64); DNS_CNN_CONV=rows selects k_conv_rows
    bool conv_wlds = true;  // its weights shared through LDS; DNS_CNN_WLDS=0 streams them per wavefront
    template <int EPI> void conv(const ConvArgs &ca, int np, long Vn, hipStream_t s)
    {
        if (mfma_k8) {
            if (np % 64 == 0) {
                dim3 grd((unsigned)((Vn + 63) / 64), (unsigned)(np / 64));
                hipLaunchKernelGGL((k_conv<2, 2, EPI>), grd, dim3(256), 0, s, ca, geom(Vn));
            } else {
                dim3 grd((unsigned)((Vn + 127) / 128), (unsigned)((np + 31) / 32));
                hipLaunchKernelGGL((k_conv<4, 1, EPI>), grd, dim3(256), 0, s, ca, geom(Vn));
            }
            return;
        }
        if (row_conv && (M == 32 || M == 64) && ca.kin % 16 == 0 && (np == 32 || np == 64)) {
            const dim3 grd((unsigned)(M / 4), (unsigned)M, (unsigned)(Vn / Vs));
            const size_t lds = sizeof(unsigned short) * 6 * (size_t)M * (ca.kin + 8);
            if (conv_pf && M == 64 && (ca.kin == 16 || ca.kin == 64) && conv_wlds) {
                const size_t ldsw = lds + sizeof(unsigned short) * 2 * (size_t)np * (ca.kin + 8);
                if (ca.kin == 64) {
                    if (np == 64) hipLaunchKernelGGL((k_conv_rows_pf<2, 2, 64, EPI, true>), grd, dim3(256), ldsw, s, ca, geom(Vn));
                    else hipLaunchKernelGGL((k_conv_rows_pf<2, 1, 64, EPI, true>), grd, dim3(256), ldsw, s, ca, geom(Vn));
                } else {
                    if (np == 64) hipLaunchKernelGGL((k_conv_rows_pf<2, 2, 16, EPI, true>), grd, dim3(256), ldsw, s, ca, geom(Vn));
                    else hipLaunchKernelGGL((k_conv_rows_pf<2, 1, 16, EPI, true>), grd, dim3(256), ldsw, s, ca, geom(Vn));
                }
            } else if (conv_pf && M == 64 && (ca.kin == 16 || ca.kin == 64)) {
                if (ca.kin == 64) {
                    if (np == 64) hipLaunchKernelGGL((k_conv_rows_pf<2, 2, 64, EPI>), grd, dim3(256), lds, s, ca, geom(Vn));
                    else hipLaunchKernelGGL((k_conv_rows_pf<2, 1, 64, EPI>), grd, dim3(256), lds, s, ca, geom(Vn));
                } else {
                    if (np == 64) hipLaunchKernelGGL((k_conv_rows_pf<2, 2, 16, EPI>), grd, dim3(256), lds, s, ca, geom(Vn));
                    else hipLaunchKernelGGL((k_conv_rows_pf<2, 1, 16, EPI>), grd, dim3(256), lds, s, ca, geom(Vn));
                }
            } else if (M == 64) {
                if (np == 64) hipLaunchKernelGGL((k_conv_rows<2, 2, EPI>), grd, dim3(256), lds, s, ca, geom(Vn));
                else hipLaunchKernelGGL((k_conv_rows<2, 1, EPI>), grd, dim3(256), lds, s, ca, geom(Vn));
            } else {
                if (np == 64) hipLaunchKernelGGL((k_conv_rows<1, 2, EPI>), grd, dim3(256), lds, s, ca, geom(Vn));
                else hipLaunchKernelGGL((k_conv_rows<1, 1, EPI>), grd, dim3(256), lds, s, ca, geom(Vn));
            }
            return;
        }
        if (np % 64 == 0) {
            dim3 grd((unsigned)((Vn + 255) / 256), (unsigned)(np / 64));
            hipLaunchKernelGGL((k_conv16<2, 2, EPI>), grd, dim3(256), 0, s, ca, geom(Vn));
        } else {
            dim3 grd((unsigned)((Vn + 255) / 256), (unsigned)(np / 32));
            hipLaunchKernelGGL((k_conv16<2, 1, EPI>), grd, dim3(256), 0, s, ca, geom(Vn));
        }
    }

    // forward of the first nb batch slots of x0; the FP32 output lands in y (nb M³ × 6)
    int forward(hipStream_t s, int nb)
    {
        const long Vn = (long)nb * Vs;
        for (int l = 0; l < 4; ++l) {
            const CNNLayer &c = L[l];
            ConvArgs ca{};
            ca.in = l == 0 ? x0 : a[l - 1];
            ca.W = c.W;
            ca.bias = w + c.boff;
            ca.kin = c.kinp;
            ca.np = c.np;
            ca.nout = c.nout;
            ca.sign = 1;
            if (l < 3) {
                ca.out = a[l];
                ca.zbuf = z[l];
                ca.ldo = 64;
                conv<0>(ca, c.np, Vn, s);
            } else {
                ca.yout = y;
                ca.ldo = 6;
                conv<1>(ca, c.np, Vn, s);
            }
        }
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    int wgrad(int l, const unsigned short *dout, int ldd)
    {
        const CNNLayer &c = L[l];
        WGradArgs wa;
        wa.dout = dout;
        wa.in = l == 0 ? x0 : a[l - 1];
        wa.part = wpart;
        wa.ldd = ldd;
        wa.nout = c.nout;
        wa.kin = c.kinp;
        wa.np = c.np;
        wa.kp = c.kpt;
        int ns = nsplit;
        bool bias_done = false; // the plane kernel also forms the bias partials
        if (planes_wgrad && (c.np == 32 || c.np == 64) && (c.kpt == 32 || c.kpt == 64)) {
            bias_done = true;
            WGradPlaneArgs wp;
            wp.dout = dout;
            wp.in = wa.in;
            wp.part = wpart;
            wp.bpart = bpart;
            wp.nout = c.nout;
            wp.ldd = ldd;
            wp.kin = c.kinp;
            const long nplanes = (long)B * M;
            wp.planes_per_block = (int)((nplanes + wgrad_splits - 1) / wgrad_splits);
            ns = (int)((nplanes + wp.planes_per_block - 1) / wp.planes_per_block);
            wp.lds_d = wgrad_row_stride(M);
            wp.lds_i = wgrad_row_stride(M + 16);
            const int NC = (c.np / 32) * (c.kpt / 32);
            const size_t stage = sizeof(unsigned short) * ((size_t)c.np * wp.lds_d + (size_t)3 * c.kpt * wp.lds_i);
            const size_t lds = std::max(stage, NC < 4 ? sizeof(float) * 1024 * 9 * (size_t)NC : (size_t)0);
            const dim3 grd(ns, 3);
            if (c.np == 64 && c.kpt == 64)
                hipLaunchKernelGGL((k_conv_wgrad_planes<2, 2>), grd, dim3(256), lds, st, wp, geom(V));
            else if (c.np == 64)
                hipLaunchKernelGGL((k_conv_wgrad_planes<2, 1>), grd, dim3(256), lds, st, wp, geom(V));
            else if (c.kpt == 64)
                hipLaunchKernelGGL((k_conv_wgrad_planes<1, 2>), grd, dim3(256), lds, st, wp, geom(V));
            else
                hipLaunchKernelGGL((k_conv_wgrad_planes<1, 1>), grd, dim3(256), lds, st, wp, geom(V));
        } else if (rows_wgrad && (c.np == 32 || c.np == 64) && (c.kpt == 32 || c.kpt == 64)) {
            WGradRowArgs wr;
            wr.dout = dout;
            wr.in = wa.in;
            wr.part = wpart;
            wr.ldd = ldd;
            wr.kin = c.kinp;
            const long nrows = (long)B * M * M;
            wr.rows_per_block = (int)((nrows + 255) / 256);
            ns = (int)((nrows + wr.rows_per_block - 1) / wr.rows_per_block);
            wr.lds_d = wgrad_row_stride(M);
            wr.lds_i = wgrad_row_stride(M + 16);
            const size_t stage = sizeof(unsigned short) * ((size_t)c.np * wr.lds_d + (size_t)c.kpt * wr.lds_i);
            const size_t lds = std::max(stage, sizeof(float) * 1024 * (size_t)(c.np / 32) * (c.kpt / 32) * 3);
            const dim3 grd(ns, 9);
            if (c.np == 64 && c.kpt == 64)
                hipLaunchKernelGGL((k_conv_wgrad_rows<2, 2>), grd, dim3(256), lds, st, wr, geom(V));
            else if (c.np == 64)
                hipLaunchKernelGGL((k_conv_wgrad_rows<2, 1>), grd, dim3(256), lds, st, wr, geom(V));
            else if (c.kpt == 64)
                hipLaunchKernelGGL((k_conv_wgrad_rows<1, 2>), grd, dim3(256), lds, st, wr, geom(V));
            else
                hipLaunchKernelGGL((k_conv_wgrad_rows<1, 1>), grd, dim3(256), lds, st, wr, geom(V));
        } else if (row_wgrad) {
            WGrad16Args w16;
            w16.dout = dout;
            w16.in = wa.in;
            w16.part = wpart;
            w16.ldd = ldd;
            w16.nout = c.nout;
            w16.kin = c.kinp;
            w16.np = c.np;
            w16.kp = c.kpt;
            const long nrows = (long)B * M * M;
            w16.rows_per_block = (int)((nrows + 255) / 256);
            ns = (int)((nrows + w16.rows_per_block - 1) / w16.rows_per_block);
            const size_t lds = sizeof(unsigned short) * 4 * 64 * (size_t)(M + 8);
            hipLaunchKernelGGL(k_conv_wgrad16, dim3(ns, 9), dim3(256), lds, st, w16, geom(V));
        } else {
            wa.chunk = (V + nsplit - 1) / nsplit;
            wa.chunk = (wa.chunk + 63) / 64 * 64;
            hipLaunchKernelGGL(k_conv_wgrad, dim3(nsplit, 27), dim3(256), 0, st, wa, geom(V));
        }
        hipLaunchKernelGGL(k_wgrad_reduce, dim3(256), dim3(256), 0, st, wpart, ns, c.np, c.kpt, c.nout, c.kin,
                           g + c.woff);
        if (ldd <= 0 || ldd % 8 != 0 || 2048 % ldd != 0 || c.nout > ldd) {
            set_error("bias gradient: channel stride %d is not a multiple of 8 dividing 2048", ldd);
            return DNS_EINVAL;
        }
        if (bias_done) {
            hipLaunchKernelGGL(k_bias_reduce, dim3(c.nout), dim3(256), 0, st, bpart, ns, c.nout, g + c.boff);
        } else {
            hipLaunchKernelGGL(k_bias_grad, dim3(bblocks), dim3(256), 0, st, dout, ldd, c.nout, V, bpart);
            hipLaunchKernelGGL(k_bias_reduce, dim3(c.nout), dim3(256), 0, st, bpart, bblocks, c.nout, g + c.boff);
        }
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    // MSE of y against t over the first nb slots into *ldst (device), and dy
    int mse(int nb, double *ldst)
    {
        const long Vn = (long)nb * Vs;
        hipLaunchKernelGGL(k_mse, dim3(kLossBlocks), dim3(256), 0, st, y, t, 6, Vn, dy, 16, lpart);
        hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, st, lpart, kLossBlocks, 1.0 / ((double)Vn * 6), ldst);
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    // loss of the full batch (y against t) into *ldst (device) and the gradient into g
    int backward(double *ldst)
    {
        timer.begin(1);
        DNS_TRY(mse(B, ldst));
        DNS_TRY(wgrad(3, dy, 16));
        // dz3 = (W4ᵀ * dy) ⊙ gelu'(z3), then layers 3, 2
        const unsigned short *dout = dy;
        for (int l = 3; l >= 1; --l) {
            const CNNLayer &c = L[l];
            unsigned short *dst = dz[l & 1];
            ConvArgs ca{};
            ca.in = dout;
            ca.W = c.WT;
            ca.kin = c.noutp;
            ca.np = c.kpt;
            ca.nout = c.kin;
            ca.ldo = 64;
            ca.zprev = z[l - 1];
            ca.out = dst;
            ca.sign = -1;
            conv<2>(ca, c.kpt, V, st);
            DNS_TRY(wgrad(l - 1, dst, 64));
            dout = dst;
        }
        timer.end(1);
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    int read_scalar(const double *dsrc, double *out)
    {
        DNS_HIP_CHECK(hipMemcpyAsync(pinned, dsrc, sizeof(double), hipMemcpyDeviceToHost, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        *out = pinned[0];
        return DNS_OK;
    }

    int adam(float lr, float b1, float b2, float eps)
    {
        ++adam_t;
        AdamArgs aa;
        aa.w = w;
        aa.m = m;
        aa.v = v;
        aa.g = g;
        aa.count = (long)nparam;
        aa.lr = lr;
        aa.b1 = b1;
        aa.b2 = b2;
        aa.eps = eps;
        aa.c1 = (float)(1.0 / (1.0 - std::pow((double)b1, adam_t)));
        aa.c2 = (float)(1.0 / (1.0 - std::pow((double)b2, adam_t)));
        timer.begin(2);
        hipLaunchKernelGGL(k_adam, dim3(256), dim3(256), 0, st, aa);
        DNS_TRY(pack());
        timer.end(2);
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    // ------------------------------------------------------------ dataset
    template <typename T> int append(const T *ub, const T *tau, hipStream_t s)
    {
        DevTemp<float> ub_, tb_;
        DNS_TRY(ub_.alloc((size_t)3 * Vs));
        DNS_TRY(tb_.alloc((size_t)6 * Vs));
        float *du_ = ub_.p, *dt_ = tb_.p;
        ds_u.push_back(du_);
        ds_t.push_back(dt_);
        ub_.p = tb_.p = nullptr;  // now owned by the dataset
        hipLaunchKernelGGL((k_cast_f32<T>), dim3(1024), dim3(256), 0, s, ub, 3 * Vs, du_);
        hipLaunchKernelGGL((k_cast_f32<T>), dim3(1024), dim3(256), 0, s, tau, 6 * Vs, dt_);
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    // RMS of the ū and τ components over the dataset
    int fit_scales()
    {
        if (ds_u.empty()) {
            set_error("empty dataset");
            return DNS_EINVAL;
        }
        double acc[2] = {0, 0};
        for (int which = 0; which < 2; ++which) {
            const auto &set = which ? ds_t : ds_u;
            const long cnt = (which ? 6 : 3) * Vs;
            for (size_t q = 0; q < set.size(); ++q) {
                hipLaunchKernelGGL(k_sumsq, dim3(kLossBlocks), dim3(256), 0, st, set[q], cnt, lpart);
                hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, st, lpart, kLossBlocks, 1.0, lval);
                double s_ = 0;
                DNS_TRY(read_scalar(lval, &s_));
                acc[which] += s_;
            }
            acc[which] /= (double)set.size() * cnt;
        }
        if (!(acc[0] > 0) || !(acc[1] > 0)) {
            set_error("dataset has a zero field");
            return DNS_EINVAL;
        }
        su = std::sqrt(acc[0]);
        stau = std::sqrt(acc[1]);
        return DNS_OK;
    }

    // stages dataset samples idx[0..nb) into batch slots 0..nb (normalised)
    void stage(const int *idx, int nb)
    {
        for (int b = 0; b < nb; ++b) {
            hipLaunchKernelGGL((k_planar_to_bf16<float>), dim3(512), dim3(256), 0, st, ds_u[idx[b]], 3, Vs,
                               (float)(1.0 / su), x0 + (size_t)b * Vs * 16, 16);
            hipLaunchKernelGGL(k_planar_to_interleaved, dim3(512), dim3(256), 0, st, ds_t[idx[b]], 6, Vs,
                               (float)(1.0 / stau), t + (size_t)b * Vs * 6);
        }
    }

    // nsteps Adam steps on batches drawn from epoch-wise shuffles of the
    // dataset's first ntrain samples (all if ntrain <= 0; mt19937_64(seed));
    // the normalised loss of each step before its update goes to losses
    // (host, may be null).
    int train(int nsteps, float lr, float b1, float b2, float eps, uint64_t seed, int ntrain, double *losses)
    {
        const int nds = ntrain > 0 ? std::min(ntrain, (int)ds_u.size()) : (int)ds_u.size();
        if (nds == 0) {
            set_error("empty dataset");
            return DNS_EINVAL;
        }
        DevTemp<double> dlbuf;
        DNS_TRY(dlbuf.alloc((size_t)std::max(1, nsteps)));
        double *dl = dlbuf.p;
        std::mt19937_64 gen(seed);
        std::vector<int> perm(nds);
        size_t pos = nds;
        std::vector<int> idx(B);
        for (int it = 0; it < nsteps; ++it) {
            for (int b = 0; b < B; ++b) {
                if (pos == (size_t)nds) {
                    for (int q = 0; q < nds; ++q) perm[q] = q;
                    std::shuffle(perm.begin(), perm.end(), gen);
                    pos = 0;
                }
                idx[b] = perm[pos++];
            }
            stage(idx.data(), B);
            timer.begin(0);
            DNS_TRY(forward(st, B));
            timer.end(0);
            DNS_TRY(backward(dl + it));
            DNS_TRY(adam(lr, b1, b2, eps));
        }
        int rc = DNS_OK;
        if (losses && nsteps > 0) {
            if (hipMemcpyAsync(losses, dl, sizeof(double) * nsteps, hipMemcpyDeviceToHost, st) != hipSuccess) rc = DNS_EHIP;
        }
        if (hipStreamSynchronize(st) != hipSuccess) rc = DNS_EHIP;
        if (rc != DNS_OK) set_error("HIP error in training loop");
        return rc;
    }

    // normalised MSE of the network over dataset samples [first, first + count)
    int eval(int first, int count, double *out)
    {
        if (first < 0 || count < 1 || first + count > (int)ds_u.size()) {
            set_error("sample range outside the dataset");
            return DNS_EINVAL;
        }
        double acc = 0;
        std::vector<int> idx(B);
        for (int q0 = 0; q0 < count; q0 += B) {
            const int nb = std::min(B, count - q0);
            for (int b = 0; b < nb; ++b) idx[b] = first + q0 + b;
            stage(idx.data(), nb);
            DNS_TRY(forward(st, nb));
            DNS_TRY(mse(nb, lval));
            double s_ = 0;
            DNS_TRY(read_scalar(lval, &s_));
            acc += s_ * nb;
        }
        *out = acc / count;
        return DNS_OK;
    }

    // predicted τ (6, M³ planar, physical units) of dataset sample q (host)
    int predict(int q, float *tau)
    {
        if (q < 0 || q >= (int)ds_u.size()) {
            set_error("sample index outside the dataset");
            return DNS_EINVAL;
        }
        stage(&q, 1);
        DNS_TRY(forward(st, 1));
        float *tmp = reinterpret_cast<float *>(dz[0]); // free between training steps
        hipLaunchKernelGGL((k_interleaved_to_planar<float>), dim3(512), dim3(256), 0, st, y, 6, Vs, stau, tmp);
        DNS_HIP_CHECK(hipMemcpyAsync(tau, tmp, sizeof(float) * 6 * Vs, hipMemcpyDeviceToHost, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        return DNS_OK;
    }

    // τ of a physical velocity on stream s (slot 0 of the batch buffers)
    template <typename T> int stress(const T *vel, T *tau, hipStream_t s)
    {
        DNS_HIP_CHECK(hipEventRecord(ev_in, st));
        DNS_HIP_CHECK(hipStreamWaitEvent(s, ev_in, 0));
        hipLaunchKernelGGL((k_planar_to_bf16<T>), dim3(512), dim3(256), 0, s, vel, 3, Vs, (float)(1.0 / su), x0, 16);
        DNS_TRY(forward(s, 1));
        hipLaunchKernelGGL((k_interleaved_to_planar<T>), dim3(512), dim3(256), 0, s, y, 6, Vs, stau, tau);
        DNS_HIP_CHECK(hipGetLastError());
        DNS_HIP_CHECK(hipEventRecord(ev_out, s));
        DNS_HIP_CHECK(hipStreamWaitEvent(st, ev_out, 0));
        return DNS_OK;
    }
};

}  // namespace dns

struct dns_cnn {
    std::unique_ptr<dns::CNN> impl;
};

namespace dns {

template <typename T> int cnn_stress(dns_cnn *h, const T *vel, T *tau, int n, hipStream_t st)
{
    if (!h || !h->impl) {
        set_error("null dns_cnn handle");
        return DNS_EINVAL;
    }
    if (n != h->impl->M) {
        set_error("LES grid and network grid differ");
        return DNS_EINVAL;
    }
    return h->impl->stress<T>(vel, tau, st);
}

template <typename T> int cnn_append(dns_cnn *h, const T *ubar, const T *tau, int n, hipStream_t st)
{
    if (!h || !h->impl) {
        set_error("null dns_cnn handle");
        return DNS_EINVAL;
    }
    if (n != h->impl->M) {
        set_error("LES grid and network grid differ");
        return DNS_EINVAL;
    }
    CNN &c = *h->impl;
    DNS_HIP_CHECK(hipEventRecord(c.ev_in, c.st));
    DNS_HIP_CHECK(hipStreamWaitEvent(st, c.ev_in, 0));
    DNS_TRY(c.append<T>(ubar, tau, st));
    DNS_HIP_CHECK(hipStreamSynchronize(st));
    return DNS_OK;
}

template int cnn_stress<double>(dns_cnn *, const double *, double *, int, hipStream_t);
template int cnn_stress<float>(dns_cnn *, const float *, float *, int, hipStream_t);
template int cnn_append<double>(dns_cnn *, const double *, const double *, int, hipStream_t);
template int cnn_append<float>(dns_cnn *, const float *, const float *, int, hipStream_t);

}  // namespace dns

using namespace dns;

extern "C" {

int dns_cnn_create(int32_t M, int32_t batch, uint64_t seed, dns_cnn **out)
{
    if (!out) return DNS_EINVAL;
    *out = nullptr;
    if (M < 2 || batch < 1 || (long long)batch * M * M * M > (1LL << 27)) {
        set_error("invalid CNN grid or batch");
        return DNS_EINVAL;
    }
    std::unique_ptr<dns_cnn> h(new dns_cnn);
    h->impl.reset(new CNN());
    DNS_TRY(h->impl->create(M, batch, seed));
    *out = h.release();
    return DNS_OK;
}

int dns_cnn_destroy(dns_cnn *h)
{
    delete h;
    return DNS_OK;
}

int dns_cnn_num_params(dns_cnn *h, int64_t *n)
{
    if (!h || !n) return DNS_EINVAL;
    *n = (int64_t)h->impl->nparam;
    return DNS_OK;
}

// Flat FP32 parameter vector: per layer W[o][n][k] (27·nout·kin) then bias[nout].
int dns_cnn_set_params(dns_cnn *h, const float *host)
{
    if (!h || !host) return DNS_EINVAL;
    CNN &c = *h->impl;
    DNS_HIP_CHECK(hipMemcpyAsync(c.w, host, sizeof(float) * c.nparam, hipMemcpyHostToDevice, c.st));
    DNS_TRY(c.pack());
    DNS_HIP_CHECK(hipStreamSynchronize(c.st));
    return DNS_OK;
}

int dns_cnn_get_params(dns_cnn *h, float *host)
{
    if (!h || !host) return DNS_EINVAL;
    CNN &c = *h->impl;
    DNS_HIP_CHECK(hipMemcpyAsync(host, c.w, sizeof(float) * c.nparam, hipMemcpyDeviceToHost, c.st));
    DNS_HIP_CHECK(hipStreamSynchronize(c.st));
    return DNS_OK;
}

int dns_cnn_get_grads(dns_cnn *h, float *host)
{
    if (!h || !host) return DNS_EINVAL;
    CNN &c = *h->impl;
    DNS_HIP_CHECK(hipMemcpyAsync(host, c.g, sizeof(float) * c.nparam, hipMemcpyDeviceToHost, c.st));
    DNS_HIP_CHECK(hipStreamSynchronize(c.st));
    return DNS_OK;
}

// Forward on a host batch x (V × 3, interleaved, FP32, already normalised);
// y (V × 6) FP32 out.
int dns_cnn_forward(dns_cnn *h, const float *x, float *yout)
{
    if (!h || !x || !yout) return DNS_EINVAL;
    CNN &c = *h->impl;
    DevTemp<float> tbuf;
    DNS_TRY(tbuf.alloc((size_t)c.V * 3));
    float *tmp = tbuf.p;
    DNS_HIP_CHECK(hipMemcpyAsync(tmp, x, sizeof(float) * c.V * 3, hipMemcpyHostToDevice, c.st));
    hipLaunchKernelGGL(k_to_bf16, dim3(1024), dim3(256), 0, c.st, tmp, 3, 1.0f, c.V, c.x0, 16);
    DNS_TRY(c.forward(c.st, c.B));
    DNS_HIP_CHECK(hipMemcpyAsync(yout, c.y, sizeof(float) * c.V * 6, hipMemcpyDeviceToHost, c.st));
    DNS_HIP_CHECK(hipStreamSynchronize(c.st));
    return DNS_OK;
}

// Loss and gradient for a host batch (x: V × 3, t: V × 6, interleaved FP32).
int dns_cnn_loss_grad(dns_cnn *h, const float *x, const float *tgt, double *loss)
{
    if (!h || !x || !tgt) return DNS_EINVAL;
    CNN &c = *h->impl;
    DevTemp<float> tbuf;
    DNS_TRY(tbuf.alloc((size_t)c.V * 3));
    float *tmp = tbuf.p;
    DNS_HIP_CHECK(hipMemcpyAsync(tmp, x, sizeof(float) * c.V * 3, hipMemcpyHostToDevice, c.st));
    DNS_HIP_CHECK(hipMemcpyAsync(c.t, tgt, sizeof(float) * c.V * 6, hipMemcpyHostToDevice, c.st));
    hipLaunchKernelGGL(k_to_bf16, dim3(1024), dim3(256), 0, c.st, tmp, 3, 1.0f, c.V, c.x0, 16);
    DNS_TRY(c.forward(c.st, c.B));
    DNS_TRY(c.backward(c.lval));
    double l = 0;
    DNS_TRY(c.read_scalar(c.lval, &l));
    if (loss) *loss = l;
    return DNS_OK;
}

int dns_cnn_add_sample(dns_cnn *h, const float *ubar, const float *tau)
{
    if (!h || !ubar || !tau) return DNS_EINVAL;
    CNN &c = *h->impl;
    DevTemp<float> tbuf;
    DNS_TRY(tbuf.alloc((size_t)c.Vs * 9));
    float *tmp = tbuf.p;
    DNS_HIP_CHECK(hipMemcpyAsync(tmp, ubar, sizeof(float) * c.Vs * 3, hipMemcpyHostToDevice, c.st));
    DNS_HIP_CHECK(hipMemcpyAsync(tmp + 3 * c.Vs, tau, sizeof(float) * c.Vs * 6, hipMemcpyHostToDevice, c.st));
    DNS_TRY(c.append<float>(tmp, tmp + 3 * c.Vs, c.st));
    DNS_HIP_CHECK(hipStreamSynchronize(c.st));
    return DNS_OK;
}

int dns_cnn_dataset_size(dns_cnn *h, int32_t *n)
{
    if (!h || !n) return DNS_EINVAL;
    *n = (int32_t)h->impl->ds_u.size();
    return DNS_OK;
}

int dns_cnn_fit_scales(dns_cnn *h, double *su, double *stau)
{
    if (!h) return DNS_EINVAL;
    DNS_TRY(h->impl->fit_scales());
    if (su) *su = h->impl->su;
    if (stau) *stau = h->impl->stau;
    return DNS_OK;
}

int dns_cnn_set_scales(dns_cnn *h, double su, double stau)
{
    if (!h || !(su > 0) || !(stau > 0)) return DNS_EINVAL;
    h->impl->su = su;
    h->impl->stau = stau;
    return DNS_OK;
}

int dns_cnn_get_scales(dns_cnn *h, double *su, double *stau)
{
    if (!h || !su || !stau) return DNS_EINVAL;
    *su = h->impl->su;
    *stau = h->impl->stau;
    return DNS_OK;
}

int dns_cnn_train(dns_cnn *h, int32_t nsteps, float lr, float beta1, float beta2, float eps, uint64_t seed,
                  int32_t ntrain, double *losses)
{
    if (!h || nsteps < 0) return DNS_EINVAL;
    return h->impl->train(nsteps, lr, beta1, beta2, eps, seed, ntrain, losses);
}

int dns_cnn_eval(dns_cnn *h, int32_t first, int32_t count, double *loss)
{
    if (!h || !loss) return DNS_EINVAL;
    return h->impl->eval(first, count, loss);
}

int dns_cnn_predict(dns_cnn *h, int32_t sample, float *tau)
{
    if (!h || !tau) return DNS_EINVAL;
    return h->impl->predict(sample, tau);
}

int dns_cnn_timing(dns_cnn *h, int32_t enable)
{
    if (!h) return DNS_EINVAL;
    h->impl->timer.flush();
    for (int f = 0; f < 8; ++f) {
        h->impl->timer.ms[f] = 0;
        h->impl->timer.cnt[f] = 0;
    }
    h->impl->timer.on = enable != 0;
    return DNS_OK;
}

int dns_cnn_timing_read(dns_cnn *h, int32_t family, double *total_ms, int64_t *launches)
{
    if (!h || family < 0 || family >= 8 || !total_ms || !launches) return DNS_EINVAL;
    h->impl->timer.flush();
    *total_ms = h->impl->timer.ms[family];
    *launches = h->impl->timer.cnt[family];
    return DNS_OK;
}

int dns_cnn_adam_step(dns_cnn *h, float lr, float beta1, float beta2, float eps)
{
    if (!h) return DNS_EINVAL;
    return h->impl->adam(lr, beta1, beta2, eps);
}

int dns_cnn_sync(dns_cnn *h)
{
    if (!h) return DNS_EINVAL;
    DNS_HIP_CHECK(hipStreamSynchronize(h->impl->st));
    return DNS_OK;
}

}  // extern "C"
