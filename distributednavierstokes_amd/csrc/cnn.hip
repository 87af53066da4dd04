// Host side of the sub-filter-stress CNN: a 4-layer 3×3×3 periodic CNN
// (3 → 64 → 64 → 64 → 6 channels, erf-GELU) on M³ LES grids, trained with
// Adam on filtered-DNS pairs held on the GPU; BF16 weights and activations
// with FP32 accumulation (v_mfma_f32_32x32x8_bf16).  C ABI: dns_cnn_*.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "cnn_kernels.hpp"
#include "host_util.hpp"

namespace dns {

struct CNNLayer {
    int kin, kinp, nout, np;   // input channels (padded to kinp), outputs (padded to np)
    int kpt, noutp;            // transposed copy: rows kpt (kin padded to 32), cols noutp (nout padded to 8)
    size_t woff, boff;         // offsets of weights / bias in the flat parameter vector
    unsigned short *W = nullptr, *WT = nullptr;
};

struct CNN {
    int M = 0, B = 0;
    long V = 0;
    hipStream_t st = nullptr;
    CNNLayer L[4];
    size_t nparam = 0;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr; // FP32 masters, Adam moments, gradient
    int adam_t = 0;
    // activations (BF16 channels-last) and buffers
    unsigned short *x0 = nullptr, *z[3] = {}, *a[3] = {}, *dy = nullptr, *dz[2] = {};
    float *y = nullptr, *t = nullptr;
    float *wpart = nullptr, *bpart = nullptr;
    double *lpart = nullptr, *pinned = nullptr;
    int nsplit = 64, bblocks = 256;
    // dataset of filtered-DNS pairs (FP32 planar per sample: ū (3, M³), τ (6, M³))
    std::vector<float *> ds_u, ds_t;
    double su = 1.0, st_ = 1.0; // normalisation: RMS of ū and τ over the dataset
    LaunchTimer timer;

    ~CNN()
    {
        if (st) hipStreamSynchronize(st);
        for (auto &l : L) {
            if (l.W) hipFree(l.W);
            if (l.WT) hipFree(l.WT);
        }
        void *ptrs[] = {w, m, v, g, x0, z[0], z[1], z[2], a[0], a[1], a[2], dy, dz[0], dz[1], y, t, wpart, bpart, lpart};
        for (void *p : ptrs)
            if (p) hipFree(p);
        for (float *p : ds_u) hipFree(p);
        for (float *p : ds_t) hipFree(p);
        if (pinned) hipHostFree(pinned);
        if (st) hipStreamDestroy(st);
    }

    int create(int M_, int B_, uint64_t seed)
    {
        M = M_;
        B = B_;
        V = (long)B * M * M * M;
        DNS_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        timer.st = st;
        const int kin[4] = {3, 64, 64, 64}, nout[4] = {64, 64, 64, 6};
        size_t off = 0;
        for (int l = 0; l < 4; ++l) {
            CNNLayer &c = L[l];
            c.kin = kin[l];
            c.kinp = (kin[l] + 7) / 8 * 8;
            c.nout = nout[l];
            c.np = (nout[l] + 31) / 32 * 32;
            c.kpt = (kin[l] + 31) / 32 * 32;
            c.noutp = (nout[l] + 7) / 8 * 8;
            c.woff = off;
            off += (size_t)27 * c.nout * c.kin;
            c.boff = off;
            off += c.nout;
            DNS_TRY(dalloc(&c.W, (size_t)27 * c.np * c.kinp));
            DNS_TRY(dalloc(&c.WT, (size_t)27 * c.kpt * c.noutp));
        }
        nparam = off;
        // He-uniform initialisation from a seeded generator (host, once)
        std::vector<float> h(nparam, 0.f);
        std::mt19937_64 gen(seed);
        for (int l = 0; l < 4; ++l) {
            const CNNLayer &c = L[l];
            const double fan_in = 27.0 * c.kin;
            const double bound = std::sqrt(6.0 / fan_in) / std::sqrt(2.0);
            std::uniform_real_distribution<double> U(-bound, bound);
            for (size_t t2 = 0; t2 < (size_t)27 * c.nout * c.kin; ++t2) h[c.woff + t2] = (float)U(gen);
        }
        DNS_TRY(dupload(&w, h));
        std::fill(h.begin(), h.end(), 0.f);
        DNS_TRY(dupload(&m, h));
        DNS_TRY(dupload(&v, h));
        DNS_TRY(dupload(&g, h));
        DNS_TRY(pack());
        DNS_TRY(dalloc(&x0, (size_t)V * 8));
        for (int q = 0; q < 3; ++q) {
            DNS_TRY(dalloc(&z[q], (size_t)V * 64));
            DNS_TRY(dalloc(&a[q], (size_t)V * 64));
        }
        DNS_TRY(dalloc(&dy, (size_t)V * 8));
        DNS_TRY(dalloc(&dz[0], (size_t)V * 64));
        DNS_TRY(dalloc(&dz[1], (size_t)V * 64));
        DNS_TRY(dalloc(&y, (size_t)V * 6));
        DNS_TRY(dalloc(&t, (size_t)V * 6));
        nsplit = (int)std::max<long>(1, std::min<long>(64, V / 4096));
        DNS_TRY(dalloc(&wpart, (size_t)nsplit * 27 * 64 * 64));
        DNS_TRY(dalloc(&bpart, (size_t)64 * bblocks));
        DNS_TRY(dalloc(&lpart, 1024));
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

    ConvGeom geom() const { return ConvGeom{M, B, V}; }

    template <int WM, int WN, int EPI> void conv(const ConvArgs &ca, int np)
    {
        dim3 grd((unsigned)((V + 32 * WM - 1) / (32 * WM)), (unsigned)((np + 32 * WN - 1) / (32 * WN)));
        hipLaunchKernelGGL((k_conv<WM, WN, EPI>), grd, dim3(256), 0, st, ca, geom());
    }

    // forward of the batch held in x0; the FP32 output lands in y (V × 6)
    int forward()
    {
        timer.begin(0);
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
                conv<2, 2, 0>(ca, c.np);
            } else {
                ca.yout = y;
                ca.ldo = 6;
                conv<4, 1, 1>(ca, c.np);
            }
        }
        timer.end(0);
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
        wa.chunk = (V + nsplit - 1) / nsplit;
        wa.chunk = (wa.chunk + 63) / 64 * 64;
        hipLaunchKernelGGL(k_conv_wgrad, dim3(nsplit, 27), dim3(256), 0, st, wa, geom());
        hipLaunchKernelGGL(k_wgrad_reduce, dim3(256), dim3(256), 0, st, wpart, nsplit, c.np, c.kpt, c.nout, c.kin,
                           g + c.woff);
        hipLaunchKernelGGL(k_bias_grad, dim3(bblocks, c.nout), dim3(256), 0, st, dout, ldd, c.nout, V, bpart);
        hipLaunchKernelGGL(k_bias_reduce, dim3(1), dim3(64), 0, st, bpart, bblocks, c.nout, g + c.boff);
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    // loss of y against t and the full gradient into g
    int backward(double *loss)
    {
        timer.begin(1);
        const int nb = 512;
        hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, st, y, t, 6, V, dy, 8, lpart);
        DNS_TRY(wgrad(3, dy, 8));
        // dz3 = (W4ᵀ * dy) ⊙ gelu'(z3), then layers 3, 2
        const unsigned short *dout = dy;
        int ldd = 8;
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
            conv<2, 2, 2>(ca, c.kpt);
            DNS_TRY(wgrad(l - 1, dst, 64));
            dout = dst;
            ldd = 64;
        }
        (void)ldd;
        timer.end(1);
        DNS_HIP_CHECK(hipGetLastError());
        if (loss) {
            DNS_HIP_CHECK(hipMemcpyAsync(pinned, lpart, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
            DNS_HIP_CHECK(hipStreamSynchronize(st));
            double s = 0;
            for (int q = 0; q < nb; ++q) s += pinned[q];
            *loss = s / ((double)V * 6);
        }
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
};

// Planar (c, V) → channels-last BF16 with scale (one sample into batch slot b).
__global__ void k_planar_to_bf16(const float *x, int cin, long Vs, float scale, unsigned short *out, int ldc)
{
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < Vs * ldc; e += (long)gridDim.x * blockDim.x) {
        const long v = e / ldc;
        const int c = (int)(e - v * ldc);
        out[e] = c < cin ? f2bf(scale * x[(size_t)c * Vs + v]) : (unsigned short)0;
    }
}

// Planar (c, V) FP32 → interleaved (V, c) FP32 with scale.
__global__ void k_planar_to_interleaved(const float *x, int c_, long Vs, float scale, float *out)
{
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < Vs * c_; e += (long)gridDim.x * blockDim.x) {
        const long v = e / c_;
        const int c = (int)(e - v * c_);
        out[e] = scale * x[(size_t)c * Vs + v];
    }
}

template <typename T> __global__ void k_cast_f32(const T *x, long n, float *out)
{
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
        out[e] = (float)x[e];
}

__global__ void k_sumsq(const float *x, long n, double *part)
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

}  // namespace dns

struct dns_cnn {
    std::unique_ptr<dns::CNN> impl;
};

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
    float *tmp = nullptr;
    DNS_TRY(dalloc(&tmp, (size_t)c.V * 3));
    DNS_HIP_CHECK(hipMemcpyAsync(tmp, x, sizeof(float) * c.V * 3, hipMemcpyHostToDevice, c.st));
    hipLaunchKernelGGL(k_to_bf16, dim3(1024), dim3(256), 0, c.st, tmp, 3, 1.0f, c.V, c.x0, 8);
    DNS_TRY(c.forward());
    DNS_HIP_CHECK(hipMemcpyAsync(yout, c.y, sizeof(float) * c.V * 6, hipMemcpyDeviceToHost, c.st));
    DNS_HIP_CHECK(hipStreamSynchronize(c.st));
    hipFree(tmp);
    return DNS_OK;
}

// Loss and gradient for a host batch (x: V × 3, t: V × 6, interleaved FP32).
int dns_cnn_loss_grad(dns_cnn *h, const float *x, const float *tgt, double *loss)
{
    if (!h || !x || !tgt) return DNS_EINVAL;
    CNN &c = *h->impl;
    float *tmp = nullptr;
    DNS_TRY(dalloc(&tmp, (size_t)c.V * 3));
    DNS_HIP_CHECK(hipMemcpyAsync(tmp, x, sizeof(float) * c.V * 3, hipMemcpyHostToDevice, c.st));
    DNS_HIP_CHECK(hipMemcpyAsync(c.t, tgt, sizeof(float) * c.V * 6, hipMemcpyHostToDevice, c.st));
    hipLaunchKernelGGL(k_to_bf16, dim3(1024), dim3(256), 0, c.st, tmp, 3, 1.0f, c.V, c.x0, 8);
    DNS_TRY(c.forward());
    DNS_TRY(c.backward(loss));
    DNS_HIP_CHECK(hipStreamSynchronize(c.st));
    hipFree(tmp);
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
