// Host side of the pseudo-spectral path: spectral_setup, the IF-RK3 step
// orchestration, forcing/IC/diagnostic entry points and their C ABI.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "closure.hpp"
#include "host_util.hpp"
#include "spectral_kernels.hpp"

namespace dns {

// Wray RK3 (src/spectral/step.jl:16-20)
static const double W_A[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double W_B[3] = {1.0 / 4.0, 0.0, 0.0};
static const double W_DC[3] = {8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0};

constexpr int SP_XEPT = 8, SP_SEPT = 8, SP_MAXO = 6, SP_NTHR = 256, SP_REDB = 1024;

enum { SFAM_Y = 0, SFAM_Z = 1, SFAM_X = 2, SFAM_STAGE = 3, SFAM_FORCE = 4 };

struct SpecBase {
    virtual ~SpecBase() {}
    virtual int set_state(const double *h, int which) = 0;
    virtual int get_state(double *h, int which) = 0;
    virtual int set_phys(const double *h) = 0;
    virtual int get_phys(double *h) = 0;
    virtual int spec_to_phys() = 0;
    virtual int phys_to_spec() = 0;
    virtual int project() = 0;
    virtual int rhs() = 0;
    virtual int step(double dt, double *vmax) = 0;
    virtual int vsummax(double *out) = 0;
    virtual int shells(int s0, int ns, double *E) = 0;
    virtual int rescale(int s0, int ns, const double *f) = 0;
    virtual int noise(uint64_t seed) = 0;
    virtual int energy(double *e, double *diss) = 0;
    virtual int maxdiv(double *out) = 0;
    virtual int sync() = 0;
    virtual int products() = 0;
    virtual int get_products(double *h) = 0;
    // n, kcut, l, visc and precision (DNS_FP64 / DNS_MIXED) of the handle
    virtual void params(dns_spec_params *out) const = 0;
    LaunchTimer timer;
};

template <typename T> struct SpecImpl : SpecBase {
    // elements per thread of the y / z pencil FFT tiles (FP64: 8, fewer
    // registers; FP32: 16), measured at n = 512
    static constexpr int SP_EPT = sizeof(T) == 8 ? 8 : 16;
    SpecGeom g{};
    double visc = 0, l = 0;
    hipStream_t st = nullptr;
    size_t ns = 0; // K m m (one component)
    cplx<T> *uh = nullptr, *ustart = nullptr, *du = nullptr, *P = nullptr, *Q = nullptr;
    T *phys = nullptr;
    cplx<T> *twN = nullptr, *twx = nullptr, *twn = nullptr;
    FFTDesc fdn{}, fdh{};
    int tb = 1, tbf = 1, tbs = 1, lpb = 1;
    int xr_lpi = 0, xr_blocks = 0; // persistent RHS x pass: lines per block, blocks (0: k_sp_xmid)
    bool xr_wave = false;          // RHS x pass with one wavefront per line (k_sp_xrhs256, n = 512)
    size_t ldsxr = 0;
    size_t lds = 0, ldss = 0, ldsx = 0;
    double *red = nullptr, *redout = nullptr, *factors = nullptr, *pinned = nullptr;
    size_t redcap = 0;

    ~SpecImpl() override
    {
        if (st) hipStreamSynchronize(st);
        timer.flush();
        void *ptrs[] = {uh, ustart, du, P, Q, phys, twN, twx, twn, red, redout, factors, aux, tphys};
        for (void *p_ : ptrs)
            if (p_) hipFree(p_);
        if (pinned) hipHostFree(pinned);
        if (st) hipStreamDestroy(st);
    }

    int create(const dns_spec_params *S)
    {
        g.n = S->n;
        g.kcut = S->kcut;
        g.K = S->kcut + 1;
        g.m = 2 * S->kcut + 1;
        g.kfac = 6.283185307179586 / S->l; // T(2π / l), src/spectral/setup.jl:101
        visc = S->visc;
        l = S->l;
        ns = (size_t)g.K * g.m * g.m;
        if (!make_fft_desc(g.n, fdn) || !make_fft_desc(g.n / 2, fdh)) {
            set_error("unsupported grid: n must factor into 2, 3 and 5");
            return DNS_EUNSUPPORTED;
        }
        if (g.n > SP_EPT * SP_NTHR || g.n / 2 > SP_XEPT * SP_NTHR) {
            set_error("unsupported grid: n too large");
            return DNS_EUNSUPPORTED;
        }
        DNS_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        timer.st = st;
        DNS_TRY(dalloc(&uh, 3 * ns));
        DNS_TRY(dalloc(&ustart, 3 * ns));
        DNS_TRY(dalloc(&P, (size_t)g.K * g.n * g.n * 6));
        DNS_TRY(dalloc(&Q, (size_t)g.K * g.n * g.m * 6));
        DNS_TRY(dupload(&twn, twiddles<T>(g.n, g.n)));
        DNS_TRY(dupload(&twN, twiddles<T>(std::max(1, g.n / 2), std::max(1, g.n / 2))));
        DNS_TRY(dupload(&twx, twiddles<T>(g.n, g.n / 2 + 1)));

        int dev = 0, maxlds = 65536;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        if (maxlds <= 0) maxlds = 65536;
        const size_t sz = sizeof(cplx<T>);
        // y/z pencil tiles: FFT sub-batches of tbf lines; a block may load
        // tb >= tbf lines (wider rows measured no faster, so tb = tbf)
        tbf = std::max(1, std::min(64, SP_EPT * SP_NTHR / g.n));
        while (tbf > 1 && ((size_t)tbf * sp_line_stride<T>(g.n) + g.n) * sz > (size_t)maxlds) --tbf;
        tbf = std::min(tbf, g.K);
        tb = tbf;
        lds = ((size_t)tb * sp_line_stride<T>(g.n) + g.n) * sz; // tile + LDS twiddle table
        tbs = std::max(1, std::min({64, SP_SEPT * SP_NTHR / g.n, SP_MAXO * SP_NTHR / g.m}));
        while (tbs > 1 && ((size_t)tbs * sp_line_stride<T>(g.n) + g.n) * sz > (size_t)maxlds) --tbs;
        tbs = std::min(tbs, g.K);
        ldss = ((size_t)tbs * sp_line_stride<T>(g.n) + g.n) * sz;
        // per line: three real velocity lines and three half-length complex work lines
        const size_t perline = (size_t)6 * g.n * sizeof(T);
        {
            // lines per block of the fused x pass: LDS budget DNS_SP_XLDS bytes
            // (default 32 KB: one line per block in FP64, two in FP32, measured
            // fastest at n = 512)
            const char *e = getenv("DNS_SP_XLDS");
            const size_t budget = e ? (size_t)atol(e) : 32768;
            lpb = (int)std::max<size_t>(1, std::min<size_t>(16, budget / perline));
        }
        while (lpb > 1 && 3 * lpb * (g.n / 2) > SP_XEPT * SP_NTHR) --lpb;
        ldsx = perline * lpb + (size_t)(sizeof(T) == 4 ? g.n + 1 : g.n / 2) * sz; // + the LDS twiddle tables
        if (lds > (size_t)maxlds || ldss > (size_t)maxlds || ldsx + 2048 > (size_t)maxlds || tbs * g.m > SP_MAXO * SP_NTHR ||
            tbs * g.n > SP_SEPT * SP_NTHR) {
            set_error("spectral tiles do not fit the LDS limit");
            return DNS_EUNSUPPORTED;
        }
        // persistent RHS x pass: FP32 by default (measured 7.15 -> 6.7 ms per HIT 512
        // step; in FP64 one line per k_sp_xmid block stays faster, 13.2 vs 14.5 ms).
        // DNS_SP_XRHS=0/1 forces it off/on, DNS_SP_XLPI sets lines per block.
        {
            const char *e = getenv("DNS_SP_XRHS");
            const char *el = getenv("DNS_SP_XLPI");
            xr_lpi = el ? std::max(1, atoi(el)) : 16;
            if (sizeof(T) == 8 && !(e && atoi(e) == 1)) xr_lpi = 0;
            ldsxr = (size_t)3 * g.n * sizeof(T) + ((size_t)3 * (g.n / 2) + 3 * g.K + g.n + 1) * sz;
            if ((e && atoi(e) == 0) || 3 * g.K > XR_RP * SP_NTHR || ldsxr + 2048 > (size_t)maxlds) xr_lpi = 0;
            xr_blocks = xr_lpi ? (int)ceil_div((long)g.n * g.n, xr_lpi) : 0;
            // n = 512 (256-point half-length lines): one wavefront per line;
            // DNS_SP_XWAVE=0 keeps the block-per-line kernels
            const char *ew = getenv("DNS_SP_XWAVE");
            xr_wave = g.n == 512 && !(ew && atoi(ew) == 0);
        }
        redcap = std::max<size_t>((size_t)SP_REDB * 4, 2 * ((size_t)g.n * g.n / lpb + 16));
        DNS_TRY(dalloc(&red, redcap));
        DNS_TRY(dalloc(&redout, 64));
        DNS_TRY(dalloc(&factors, 4096));
        DNS_HIP_CHECK(hipHostMalloc((void **)&pinned, sizeof(double) * 4096));
        return DNS_OK;
    }

    void params(dns_spec_params *out) const override
    {
        out->n = g.n;
        out->kcut = g.kcut;
        out->l = l;
        out->visc = visc;
        out->precision = sizeof(T) == 8 ? DNS_FP64 : DNS_MIXED;
    }

    // --------------------------------------------------------- transfers
    cplx<T> *state_ptr(int which) { return which == 1 ? ustart : which == 2 ? du : uh; }

    int set_state(const double *h, int which) override
    {
        cplx<T> *d = state_ptr(which);
        if (!d) {
            set_error("state buffer not allocated");
            return DNS_EINVAL;
        }
        std::vector<cplx<T>> s(3 * ns);
        for (size_t t = 0; t < 3 * ns; ++t) s[t] = mk<T>((T)h[2 * t], (T)h[2 * t + 1]);
        DNS_HIP_CHECK(hipMemcpyAsync(d, s.data(), sizeof(cplx<T>) * 3 * ns, hipMemcpyHostToDevice, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        return DNS_OK;
    }

    int get_state(double *h, int which) override
    {
        cplx<T> *d = state_ptr(which);
        if (!d) {
            set_error("state buffer not allocated (call dns_spec_rhs first)");
            return DNS_EINVAL;
        }
        std::vector<cplx<T>> s(3 * ns);
        DNS_HIP_CHECK(hipMemcpyAsync(s.data(), d, sizeof(cplx<T>) * 3 * ns, hipMemcpyDeviceToHost, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        for (size_t t = 0; t < 3 * ns; ++t) {
            h[2 * t] = (double)s[t].x;
            h[2 * t + 1] = (double)s[t].y;
        }
        return DNS_OK;
    }

    int ensure_phys()
    {
        if (!phys) DNS_TRY(dalloc(&phys, (size_t)3 * g.n * g.n * g.n));
        return DNS_OK;
    }

    int set_phys(const double *h) override
    {
        DNS_TRY(ensure_phys());
        size_t N3 = (size_t)3 * g.n * g.n * g.n;
        std::vector<T> s(N3);
        for (size_t t = 0; t < N3; ++t) s[t] = (T)h[t];
        DNS_HIP_CHECK(hipMemcpyAsync(phys, s.data(), sizeof(T) * N3, hipMemcpyHostToDevice, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        return DNS_OK;
    }

    int get_phys(double *h) override
    {
        DNS_TRY(ensure_phys());
        size_t N3 = (size_t)3 * g.n * g.n * g.n;
        std::vector<T> s(N3);
        DNS_HIP_CHECK(hipMemcpyAsync(s.data(), phys, sizeof(T) * N3, hipMemcpyDeviceToHost, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        for (size_t t = 0; t < N3; ++t) h[t] = (double)s[t];
        return DNS_OK;
    }

    // ----------------------------------------------------------- passes
    int backward_yz(const cplx<T> *src = nullptr)
    {
        timer.begin(SFAM_Y);
        hipLaunchKernelGGL((k_sp_yback<T, SP_EPT, SP_NTHR>), dim3(ceil_div(g.K, tb), g.m, 3), dim3(SP_NTHR), lds, st,
                           src ? src : uh, Q, twn, tb, tbf, make_fastdiv(tb), g, fdn);
        timer.end(SFAM_Y);
        timer.begin(SFAM_Z);
        hipLaunchKernelGGL((k_sp_zback<T, SP_EPT, SP_NTHR>), dim3(ceil_div(g.K, tb), g.n, 3), dim3(SP_NTHR), lds, st, Q, P,
                           twn, tb, tbf, make_fastdiv(tb), g, fdn);
        timer.end(SFAM_Z);
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    // vpart: per-block velocity bounds (stage 1 and vsummax read them; later
    // stages write to the scratch half of the reduction buffer)
    template <int MODE> int xpass(double *vpart, T *physbuf = nullptr)
    {
        XMidArgs<T> a;
        a.P = P;
        a.phys = physbuf ? physbuf : phys;
        a.vpart = vpart;
        a.tw = twN;
        a.twx = twx;
        a.lpb = lpb;
        a.fN = make_fastdiv(g.n / 2);
        a.fn = make_fastdiv(g.n);
        a.fK = make_fastdiv(g.K);
        a.flpb = make_fastdiv(lpb);
        timer.begin(SFAM_X);
        if (MODE == XMID_RHS && xr_wave)
            hipLaunchKernelGGL((k_sp_xrhs256<T>), dim3(ceil_div((long)g.n * g.n, 4)), dim3(SP_NTHR),
                               (size_t)(TW4S_LEN + 257 + 4 * 256) * sizeof(cplx<T>), st, a, g);
        else if (MODE == XMID_RHS && xr_lpi)
            hipLaunchKernelGGL((k_sp_xrhs<T, SP_XEPT>), dim3(xr_blocks), dim3(SP_NTHR), ldsxr, st, a, g, fdh, xr_lpi);
        else
            hipLaunchKernelGGL((k_sp_xmid<T, SP_XEPT, MODE>), dim3(ceil_div((long)g.n * g.n, lpb)), dim3(SP_NTHR),
                               ldsx, st, a, g, fdh);
        timer.end(SFAM_X);
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    int zfwd(int ncomp)
    {
        timer.begin(SFAM_Z);
        hipLaunchKernelGGL((k_sp_zfwd<T, SP_EPT, SP_NTHR>), dim3(ceil_div(g.K, tb), g.n, ncomp), dim3(SP_NTHR), lds, st, P, Q,
                           twn, tb, tbf, make_fastdiv(tb), g, fdn);
        timer.end(SFAM_Z);
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    template <int MODE> int ystage(int stage, double dt, const cplx<T> *extra = nullptr)
    {
        StageArgs<T> s;
        s.Q = Q;
        s.uh = uh;
        s.ustart = ustart;
        s.du = du;
        s.extra = extra;
        s.tw = twn;
        s.tb = tbs;
        s.ftb = make_fastdiv(tbs);
        s.first = stage == 0;
        s.a = (T)W_A[stage];
        s.b = (T)W_B[stage];
        s.nudcdt = (T)(visc * W_DC[stage] * dt);
        s.dt = (T)dt;
        timer.begin(SFAM_STAGE);
        hipLaunchKernelGGL((k_sp_ystage<T, SP_SEPT, SP_MAXO, MODE, SP_NTHR>), dim3(ceil_div(g.K, tbs), g.m), dim3(SP_NTHR),
                           ldss, st, s, g, fdn);
        timer.end(SFAM_STAGE);
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    int spec_to_phys() override
    {
        DNS_TRY(ensure_phys());
        DNS_TRY(backward_yz());
        return xpass<XMID_C2R>(red);
    }

    // physical field of three spectral components at src into phys
    int spec_to_phys_from(const cplx<T> *src)
    {
        DNS_TRY(ensure_phys());
        DNS_TRY(backward_yz(src));
        return xpass<XMID_C2R>(red);
    }

    // σ̂_p = F[v_i v_j]/n³ (xx, yy, zz, xy, yz, zx) of the state's velocity,
    // written compactly as (K, m, m, 6) into P (free after the z pass).
    int products() override
    {
        DNS_TRY(backward_yz());
        DNS_TRY(xpass<XMID_RHS>(red + redcap / 2));
        DNS_TRY(zfwd(6));
        StageArgs<T> sa;
        sa.Q = Q;
        sa.uh = uh;
        sa.ustart = ustart;
        sa.du = P;
        sa.extra = nullptr;
        sa.tw = twn;
        sa.tb = tbs;
        sa.ftb = make_fastdiv(tbs);
        sa.first = 1;
        sa.a = sa.b = sa.nudcdt = sa.dt = T(0);
        hipLaunchKernelGGL((k_sp_ystage<T, SP_SEPT, SP_MAXO, 2, SP_NTHR>), dim3(ceil_div(g.K, tbs), g.m), dim3(SP_NTHR), ldss,
                           st, sa, g, fdn);
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    int get_products(double *h) override
    {
        const size_t n6 = 6 * ns;
        std::vector<cplx<T>> b(n6);
        DNS_HIP_CHECK(hipMemcpyAsync(b.data(), P, sizeof(cplx<T>) * n6, hipMemcpyDeviceToHost, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        for (size_t t = 0; t < n6; ++t) {
            h[2 * t] = (double)b[t].x;
            h[2 * t + 1] = (double)b[t].y;
        }
        return DNS_OK;
    }

    cplx<T> *aux = nullptr; // τ of the SFS recipe (K, m, m, 6), LES handles only
    int ensure_aux()
    {
        if (!aux) DNS_TRY(dalloc(&aux, 6 * ns));
        return DNS_OK;
    }

    int phys_to_spec() override
    {
        DNS_TRY(ensure_phys());
        DNS_TRY(xpass<XMID_R2C>(red));
        DNS_TRY(zfwd(3));
        hipLaunchKernelGGL((k_sp_yfwd<T, SP_EPT, SP_NTHR>), dim3(ceil_div(g.K, tb), g.m, 3), dim3(SP_NTHR), lds, st, Q, uh, twn,
                           tb, tbf, make_fastdiv(tb), g, fdn);
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    // Forward transform of three physical components at src (n³ planar each)
    // into the compact spectra at dst (K, m, m, 3).
    int phys_to_spec_into(T *src, cplx<T> *dst)
    {
        DNS_TRY(xpass<XMID_R2C>(red, src));
        DNS_TRY(zfwd(3));
        hipLaunchKernelGGL((k_sp_yfwd<T, SP_EPT, SP_NTHR>), dim3(ceil_div(g.K, tb), g.m, 3), dim3(SP_NTHR), lds, st, Q, dst, twn,
                           tb, tbf, make_fastdiv(tb), g, fdn);
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    // LES step with a learned sub-filter stress: each stage evaluates the
    // network on the physical velocity, transforms τ to (K, m, m, 6) and adds
    // it to σ̂ inside the fused stage (du_i = −ik_j(σ̂_ij + τ̂_ij)).
    T *tphys = nullptr; // physical τ (n³ × 6)
    int closure_step(dns_cnn *net, double dt, double *vmax)
    {
        DNS_TRY(ensure_phys());
        DNS_TRY(ensure_aux());
        if (!du) DNS_TRY(dalloc(&du, 3 * ns));
        const size_t n3 = (size_t)g.n * g.n * g.n;
        if (!tphys) DNS_TRY(dalloc(&tphys, 6 * n3));
        for (int s = 0; s < 3; ++s) {
            DNS_TRY(backward_yz());
            DNS_TRY(xpass<XMID_C2R>(red + redcap / 2));
            DNS_TRY(cnn_stress<T>(net, phys, tphys, g.n, st));
            DNS_TRY(phys_to_spec_into(tphys, aux));
            DNS_TRY(phys_to_spec_into(tphys + 3 * n3, aux + 3 * ns));
            DNS_TRY(backward_yz());
            DNS_TRY(xpass<XMID_RHS>(s == 0 ? red : red + redcap / 2));
            DNS_TRY(zfwd(6));
            DNS_TRY(ystage<0>(s, dt, aux));
        }
        if (vmax) DNS_TRY(vmax_final(rhs_parts(), vmax));
        return DNS_OK;
    }

    // Appends the physical (ū, τ) of the last SFS sample to the network's dataset.
    int append_sample(dns_cnn *net)
    {
        if (!aux) {
            set_error("no SFS sample on this handle");
            return DNS_EINVAL;
        }
        DNS_TRY(ensure_phys());
        const size_t n3 = (size_t)g.n * g.n * g.n;
        if (!tphys) DNS_TRY(dalloc(&tphys, 6 * n3));
        DNS_TRY(spec_to_phys_from(aux));
        DNS_HIP_CHECK(hipMemcpyAsync(tphys, phys, sizeof(T) * 3 * n3, hipMemcpyDeviceToDevice, st));
        DNS_TRY(spec_to_phys_from(aux + 3 * ns));
        DNS_HIP_CHECK(hipMemcpyAsync(tphys + 3 * n3, phys, sizeof(T) * 3 * n3, hipMemcpyDeviceToDevice, st));
        DNS_TRY(spec_to_phys_from(uh));
        return cnn_append<T>(net, phys, tphys, g.n, st);
    }

    int project() override
    {
        int nb = std::min<long>(SP_REDB, ceil_div((long)ns, SP_NTHR));
        hipLaunchKernelGGL((k_sp_project<T>), dim3(nb), dim3(SP_NTHR), 0, st, uh, g);
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    int rhs() override
    {
        if (!du) DNS_TRY(dalloc(&du, 3 * ns));
        DNS_TRY(backward_yz());
        DNS_TRY(xpass<XMID_RHS>(red));
        DNS_TRY(zfwd(6));
        return ystage<1>(0, 0.0);
    }

    // vmax partials written by the RHS x pass
    int rhs_parts() const
    {
        if (xr_wave) return (int)ceil_div((long)g.n * g.n, 4);
        return xr_lpi ? xr_blocks : (int)ceil_div((long)g.n * g.n, lpb);
    }

    int vmax_final(int nparts, double *out)
    {
        // two levels: up to 256 blocks into the scratch half of red, then one
        const int G = std::min(256, std::max(1, nparts / 1024));
        hipLaunchKernelGGL(k_max_mid, dim3(G), dim3(SP_NTHR), 0, st, red, nparts, 1, red + redcap / 2);
        hipLaunchKernelGGL(k_max_final, dim3(1), dim3(SP_NTHR), 0, st, red + redcap / 2, G, 1, redout);
        DNS_HIP_CHECK(hipGetLastError());
        DNS_HIP_CHECK(hipMemcpyAsync(pinned, redout, sizeof(double), hipMemcpyDeviceToHost, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        *out = pinned[0];
        return DNS_OK;
    }

    // spectral_step! (src/spectral/step.jl:68-92): three (rhs, IF stage,
    // projection) stages; vmax of the first stage's physical velocity.
    int step(double dt, double *vmax) override
    {
        for (int s = 0; s < 3; ++s) {
            DNS_TRY(backward_yz());
            DNS_TRY(xpass<XMID_RHS>(s == 0 ? red : red + redcap / 2));
            DNS_TRY(zfwd(6));
            DNS_TRY(ystage<0>(s, dt));
        }
        if (vmax) DNS_TRY(vmax_final(rhs_parts(), vmax));
        return DNS_OK;
    }

    int vsummax(double *out) override
    {
        DNS_TRY(backward_yz());
        DNS_TRY(xpass<XMID_VMAX>(red));
        return vmax_final(ceil_div((long)g.n * g.n, lpb), out);
    }

    long cube_count(int r) const
    {
        if (r >= g.kcut) return (long)ns;
        return (long)(r + 1) * (2 * r + 1) * (2 * r + 1);
    }

    int shells(int s0, int nsh, double *E) override
    {
        if (nsh <= 0 || nsh > 4096 || s0 < 0) {
            set_error("invalid shell range");
            return DNS_EINVAL;
        }
        int r = s0 + nsh - 1;
        long count = cube_count(r);
        // per-wavefront binning (nsh <= 2048) scales to more blocks than the
        // chunk-owner kernel
        int nb = (int)std::max<long>(1, std::min<long>(nsh <= 2048 ? 2048 : 256, (count + SP_NTHR - 1) / SP_NTHR));
        size_t need = (size_t)nb * nsh;
        if (need > redcap) {
            DNS_HIP_CHECK(hipStreamSynchronize(st));
            DNS_HIP_CHECK(hipFree(red));
            red = nullptr;
            redcap = 0;
            DNS_TRY(dalloc(&red, need));
            redcap = need;
        }
        timer.begin(SFAM_FORCE);
        if (nsh <= 2048)
            hipLaunchKernelGGL((k_sp_shells_w<T>), dim3(nb), dim3(SP_NTHR), sizeof(double) * 4 * nsh, st, uh, s0, nsh,
                               r, count, g, red);
        else
            hipLaunchKernelGGL((k_sp_shells<T>), dim3(nb), dim3(SP_NTHR),
                               sizeof(double) * nsh + (sizeof(double) + sizeof(int)) * SP_NTHR, st, uh, s0, nsh, r,
                               count, g, red);
        hipLaunchKernelGGL(k_bins_final, dim3(ceil_div(nsh, SP_NTHR)), dim3(SP_NTHR), 0, st, red, nb, nsh, factors);
        timer.end(SFAM_FORCE);
        DNS_HIP_CHECK(hipGetLastError());
        DNS_HIP_CHECK(hipMemcpyAsync(pinned, factors, sizeof(double) * nsh, hipMemcpyDeviceToHost, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        for (int q = 0; q < nsh; ++q) E[q] = pinned[q];
        return DNS_OK;
    }

    int rescale(int s0, int nsh, const double *f) override
    {
        if (nsh <= 0 || nsh > 4096 || s0 < 0) {
            set_error("invalid shell range");
            return DNS_EINVAL;
        }
        for (int q = 0; q < nsh; ++q) pinned[q] = f[q];
        DNS_HIP_CHECK(hipMemcpyAsync(factors, pinned, sizeof(double) * nsh, hipMemcpyHostToDevice, st));
        int r = s0 + nsh - 1;
        long count = cube_count(r);
        int nb = (int)std::max<long>(1, std::min<long>(SP_REDB, (count + SP_NTHR - 1) / SP_NTHR));
        timer.begin(SFAM_FORCE);
        hipLaunchKernelGGL((k_sp_rescale<T>), dim3(nb), dim3(SP_NTHR), 0, st, uh, s0, nsh, r, count, factors, g);
        timer.end(SFAM_FORCE);
        DNS_HIP_CHECK(hipGetLastError());
        DNS_HIP_CHECK(hipStreamSynchronize(st)); // the pinned staging buffer is reused
        return DNS_OK;
    }

    int noise(uint64_t seed) override
    {
        DNS_TRY(ensure_phys());
        size_t N3 = (size_t)3 * g.n * g.n * g.n;
        int nb = (int)std::min<size_t>(4096, (N3 + SP_NTHR - 1) / SP_NTHR);
        hipLaunchKernelGGL((k_sp_noise<T>), dim3(nb), dim3(SP_NTHR), 0, st, phys, seed, g.n);
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    int energy(double *e, double *diss) override
    {
        int nb = (int)std::min<long>(SP_REDB, ceil_div((long)ns, SP_NTHR));
        hipLaunchKernelGGL((k_sp_energy<T>), dim3(nb), dim3(SP_NTHR), 0, st, uh, g, reinterpret_cast<dd *>(red));
        hipLaunchKernelGGL(k_dd_final, dim3(2), dim3(SP_NTHR), 0, st, reinterpret_cast<dd *>(red), nb, 2, redout);
        DNS_HIP_CHECK(hipGetLastError());
        DNS_HIP_CHECK(hipMemcpyAsync(pinned, redout, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        if (e) *e = pinned[0];
        if (diss) *diss = visc * pinned[1];
        return DNS_OK;
    }

    int maxdiv(double *out) override
    {
        int nb = (int)std::min<long>(SP_REDB, ceil_div((long)ns, SP_NTHR));
        hipLaunchKernelGGL((k_sp_maxdiv<T>), dim3(nb), dim3(SP_NTHR), 0, st, uh, g, red);
        return vmax_final(nb, out);
    }

    int sync() override
    {
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        return DNS_OK;
    }
};

// sfs_sample! for one Gaussian bank cell (src/spectral/les.jl:313-382):
// LES state ← ū, LES aux ← τ̂ (trace-free), both on the LES handle's compact
// grid (n = M, kcut = M ÷ 3).  Clobbers the DNS pipeline buffers.
template <typename T>
int sfs_sample(SpecImpl<T> *D, SpecImpl<T> *L, double delta_fac, int protectshells)
{
    DNS_TRY(D->products());
    DNS_HIP_CHECK(hipStreamSynchronize(D->st));
    DNS_TRY(L->ensure_aux());
    const double M = L->g.n;
    const double delta = delta_fac * L->l / M;
    const double kb = L->g.kfac * (protectshells + 1);
    const int nb = (int)std::min<size_t>(SP_REDB, (L->ns + SP_NTHR - 1) / SP_NTHR);
    hipLaunchKernelGGL((k_les_gather<T>), dim3(nb), dim3(SP_NTHR), 0, L->st, D->uh, D->P, D->g, L->uh, L->aux, L->g,
                       kb * kb, delta);
    DNS_HIP_CHECK(hipGetLastError());
    DNS_TRY(L->products());
    hipLaunchKernelGGL((k_les_tau<T>), dim3(nb), dim3(SP_NTHR), 0, L->st, L->aux, L->P, L->g);
    DNS_HIP_CHECK(hipGetLastError());
    DNS_HIP_CHECK(hipStreamSynchronize(L->st));
    return DNS_OK;
}

template <typename T> int sfs_get(SpecImpl<T> *L, double *ubar, double *tau)
{
    if (ubar) DNS_TRY(L->get_state(ubar, 0));
    if (tau) {
        if (!L->aux) {
            set_error("no SFS sample on this handle");
            return DNS_EINVAL;
        }
        const size_t n6 = 6 * L->ns;
        std::vector<cplx<T>> b(n6);
        DNS_HIP_CHECK(hipMemcpyAsync(b.data(), L->aux, sizeof(cplx<T>) * n6, hipMemcpyDeviceToHost, L->st));
        DNS_HIP_CHECK(hipStreamSynchronize(L->st));
        for (size_t t = 0; t < n6; ++t) {
            tau[2 * t] = (double)b[t].x;
            tau[2 * t + 1] = (double)b[t].y;
        }
    }
    return DNS_OK;
}

// Physical ū (M³×3) and τ (M³×6) of the last sample, host copies.
template <typename T> int sfs_physical(SpecImpl<T> *L, double *ubar, double *tau)
{
    if (!L->aux) {
        set_error("no SFS sample on this handle");
        return DNS_EINVAL;
    }
    const size_t n3 = (size_t)L->g.n * L->g.n * L->g.n;
    std::vector<T> b(3 * n3);
    auto fetch = [&](double *dst) -> int {
        DNS_HIP_CHECK(hipMemcpyAsync(b.data(), L->phys, sizeof(T) * 3 * n3, hipMemcpyDeviceToHost, L->st));
        DNS_HIP_CHECK(hipStreamSynchronize(L->st));
        for (size_t t = 0; t < 3 * n3; ++t) dst[t] = (double)b[t];
        return DNS_OK;
    };
    if (ubar) {
        DNS_TRY(L->spec_to_phys_from(L->uh));
        DNS_TRY(fetch(ubar));
    }
    if (tau) {
        DNS_TRY(L->spec_to_phys_from(L->aux));
        DNS_TRY(fetch(tau));
        DNS_TRY(L->spec_to_phys_from(L->aux + 3 * L->ns));
        DNS_TRY(fetch(tau + 3 * n3));
    }
    return DNS_OK;
}

// ---------------------------------------------------------- snapshots
// spectral_save / spectral_load! (src/spectral/io.jl:29-104) on one GPU:
// <prefix>.bin holds the (kcut+1, m, m, 3) complex state column-major in
// global index order, component after component, as Complex{T} with T the
// handle's storage type (Float64, or Float32 in the mixed form); the
// <prefix>.toml sidecar carries format, version, eltype, n, kcut, m, l, visc,
// time, step, order, data and an optional [meta] table, keys sorted.  Both
// files are written under .tmp names and renamed when complete (sidecar
// first, then data, as the reference does).

// shortest decimal that reads back as v, with a '.' or exponent (a TOML float)
static std::string toml_float(double v)
{
    char buf[40];
    if (std::isnan(v)) return "nan";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    for (int p = 1; p <= 17; ++p) {
        snprintf(buf, sizeof buf, "%.*g", p, v);
        if (strtod(buf, nullptr) == v) break;
    }
    std::string out(buf);
    if (out.find_first_of(".en") == std::string::npos) out += ".0";
    return out;
}

static std::string toml_str(const std::string &v)
{
    std::string out = "\"";
    for (char c : v) {
        if (c == '"' || c == '\\') out += '\\';
        out += c;
    }
    return out + "\"";
}

static const char *eltype_name(int precision) { return precision == DNS_MIXED ? "Float32" : "Float64"; }

static int spec_save(SpecBase &impl, const char *prefix, double time, int64_t step, const char *meta)
{
    dns_spec_params P;
    impl.params(&P);
    const size_t K = (size_t)P.kcut + 1, m = (size_t)2 * P.kcut + 1, ns = K * m * m;
    std::vector<double> h(6 * ns);
    DNS_TRY(impl.get_state(h.data(), 0));
    const std::string data = std::string(prefix) + ".bin", side = std::string(prefix) + ".toml";
    const std::string datatmp = data + ".tmp", sidetmp = side + ".tmp";
    FILE *f = fopen(datatmp.c_str(), "wb");
    if (!f) {
        set_error("spectral_save: cannot open %s for writing", datatmp.c_str());
        return DNS_EINVAL;
    }
    size_t wrote = 0;
    if (P.precision == DNS_MIXED) {
        std::vector<float> hf(h.begin(), h.end());
        wrote = fwrite(hf.data(), sizeof(float), hf.size(), f);
    } else {
        wrote = fwrite(h.data(), sizeof(double), h.size(), f);
    }
    if (fclose(f) != 0 || wrote != h.size()) {
        remove(datatmp.c_str());
        set_error("spectral_save: short write to %s", datatmp.c_str());
        return DNS_EINVAL;
    }
    const size_t slash = data.find_last_of('/');
    const std::string base = slash == std::string::npos ? data : data.substr(slash + 1);
    std::string t;
    t += "data = " + toml_str(base) + "\n";
    t += "eltype = " + toml_str(eltype_name(P.precision)) + "\n";
    t += "format = " + toml_str("DistributedNavierStokes spectral state") + "\n";
    t += "kcut = " + std::to_string(P.kcut) + "\n";
    t += "l = " + toml_float(P.l) + "\n";
    t += "m = " + std::to_string(m) + "\n";
    t += "n = " + std::to_string(P.n) + "\n";
    t += "order = " +
         toml_str("column-major (kx, ky, kz, component) complex; kx = 0:kcut, ky/kz wrapped 0:kcut then -kcut:-1") +
         "\n";
    t += "step = " + std::to_string((long long)step) + "\n";
    t += "time = " + toml_float(time) + "\n";
    t += "version = 1\n";
    t += "visc = " + toml_float(P.visc) + "\n";
    if (meta && *meta) {
        t += "\n[meta]\n";
        t += meta;
        if (t.back() != '\n') t += "\n";
    }
    f = fopen(sidetmp.c_str(), "wb");
    if (!f) {
        remove(datatmp.c_str());
        set_error("spectral_save: cannot open %s for writing", sidetmp.c_str());
        return DNS_EINVAL;
    }
    wrote = fwrite(t.data(), 1, t.size(), f);
    if (fclose(f) != 0 || wrote != t.size()) {
        remove(datatmp.c_str());
        remove(sidetmp.c_str());
        set_error("spectral_save: short write to %s", sidetmp.c_str());
        return DNS_EINVAL;
    }
    if (rename(sidetmp.c_str(), side.c_str()) != 0 || rename(datatmp.c_str(), data.c_str()) != 0) {
        set_error("spectral_save: cannot rename the snapshot of %s into place", prefix);
        return DNS_EINVAL;
    }
    return DNS_OK;
}

// value text of a top-level `key = value` line of the sidecar ("" if absent)
static std::string toml_lookup(const std::string &text, const char *key)
{
    size_t pos = 0;
    const size_t klen = strlen(key);
    while (pos < text.size()) {
        size_t eol = text.find('\n', pos);
        if (eol == std::string::npos) eol = text.size();
        std::string line = text.substr(pos, eol - pos);
        pos = eol + 1;
        size_t a = line.find_first_not_of(" \t");
        if (a == std::string::npos) continue;
        if (line[a] == '[') break; // tables follow the top-level keys
        if (line.compare(a, klen, key) != 0) continue;
        size_t b = line.find_first_not_of(" \t", a + klen);
        if (b == std::string::npos || line[b] != '=') continue;
        b = line.find_first_not_of(" \t", b + 1);
        if (b == std::string::npos) return "";
        std::string v = line.substr(b);
        while (!v.empty() && (v.back() == ' ' || v.back() == '\r' || v.back() == '\t')) v.pop_back();
        if (v.size() >= 2 && v.front() == '"' && v.back() == '"') v = v.substr(1, v.size() - 2);
        return v;
    }
    return "";
}

static int spec_load(SpecBase &impl, const char *prefix, double *time, int64_t *step)
{
    dns_spec_params P;
    impl.params(&P);
    const std::string side = std::string(prefix) + ".toml", data = std::string(prefix) + ".bin";
    FILE *f = fopen(side.c_str(), "rb");
    if (!f) {
        set_error("spectral_load!: cannot open %s", side.c_str());
        return DNS_EINVAL;
    }
    std::string text;
    char buf[4096];
    size_t got;
    while ((got = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    fclose(f);
    const std::string eltype = toml_lookup(text, "eltype"), kc = toml_lookup(text, "kcut"),
                      ls = toml_lookup(text, "l");
    const long kcut = kc.empty() ? -1 : strtol(kc.c_str(), nullptr, 10);
    const double lv = ls.empty() ? NAN : strtod(ls.c_str(), nullptr);
    if (eltype != eltype_name(P.precision) || kcut != P.kcut || !(lv == P.l)) {
        set_error("snapshot %s (eltype %s, kcut %ld, l %s) does not match setup (eltype %s, kcut %d, l %.17g)",
                  prefix, eltype.c_str(), kcut, ls.c_str(), eltype_name(P.precision), P.kcut, P.l);
        return DNS_EINVAL;
    }
    const size_t K = (size_t)P.kcut + 1, m = (size_t)2 * P.kcut + 1, ns = K * m * m;
    std::vector<double> h(6 * ns);
    f = fopen(data.c_str(), "rb");
    if (!f) {
        set_error("spectral_load!: cannot open %s", data.c_str());
        return DNS_EINVAL;
    }
    size_t nread = 0;
    if (P.precision == DNS_MIXED) {
        std::vector<float> hf(h.size());
        nread = fread(hf.data(), sizeof(float), hf.size(), f);
        std::copy(hf.begin(), hf.end(), h.begin());
    } else {
        nread = fread(h.data(), sizeof(double), h.size(), f);
    }
    const bool extra = fgetc(f) != EOF;
    fclose(f);
    if (nread != h.size() || extra) {
        set_error("spectral_load!: %s does not hold %zu complex values", data.c_str(), 3 * ns);
        return DNS_EINVAL;
    }
    DNS_TRY(impl.set_state(h.data(), 0));
    if (time) {
        const std::string tv = toml_lookup(text, "time");
        *time = tv.empty() ? 0.0 : strtod(tv.c_str(), nullptr);
    }
    if (step) {
        const std::string sv = toml_lookup(text, "step");
        *step = sv.empty() ? 0 : strtoll(sv.c_str(), nullptr, 10);
    }
    return DNS_OK;
}

}  // namespace dns

struct dns_spec {
    std::unique_ptr<dns::SpecBase> impl;
};

using namespace dns;

#define SP_CHECK(h)                           \
    do {                                      \
        if (!(h) || !(h)->impl) {             \
            set_error("null dns_spec handle"); \
            return DNS_EINVAL;                \
        }                                     \
    } while (0)

extern "C" {

int dns_spec_create(const dns_spec_params *S, dns_spec **out)
{
    if (!out || !S) {
        set_error("null params or output handle");
        return DNS_EINVAL;
    }
    *out = nullptr;
    if (S->n % 2 != 0) {
        set_error("n must be even");
        return DNS_EINVAL;
    }
    if (S->kcut < 1) {
        set_error("kcut must be at least 1");
        return DNS_EINVAL;
    }
    if (3 * S->kcut > S->n) {
        set_error("aliasing: 3kcut ≤ n is required (2/3 rule)");
        return DNS_EINVAL;
    }
    if (S->precision != DNS_FP64 && S->precision != DNS_MIXED) {
        set_error("precision must be DNS_FP64 or DNS_MIXED");
        return DNS_EINVAL;
    }
    if ((long long)S->n * S->n * S->n * 3 >= (1LL << 40)) {
        set_error("grid too large");
        return DNS_EUNSUPPORTED;
    }
    std::unique_ptr<dns_spec> h(new dns_spec);
    int rc;
    if (S->precision == DNS_FP64) {
        auto *impl = new SpecImpl<double>();
        h->impl.reset(impl);
        rc = impl->create(S);
    } else {
        auto *impl = new SpecImpl<float>();
        h->impl.reset(impl);
        rc = impl->create(S);
    }
    if (rc != DNS_OK) return rc;
    *out = h.release();
    return DNS_OK;
}

int dns_spec_destroy(dns_spec *h)
{
    delete h;
    return DNS_OK;
}
int dns_spec_set_state(dns_spec *h, int32_t which, const double *host)
{
    SP_CHECK(h);
    if (!host || which < 0 || which > 1) return DNS_EINVAL;
    return h->impl->set_state(host, which);
}
int dns_spec_get_state(dns_spec *h, int32_t which, double *host)
{
    SP_CHECK(h);
    if (!host || which < 0 || which > 2) return DNS_EINVAL;
    return h->impl->get_state(host, which);
}
int dns_spec_set_physical(dns_spec *h, const double *host)
{
    SP_CHECK(h);
    if (!host) return DNS_EINVAL;
    return h->impl->set_phys(host);
}
int dns_spec_get_physical(dns_spec *h, double *host)
{
    SP_CHECK(h);
    if (!host) return DNS_EINVAL;
    return h->impl->get_phys(host);
}
int dns_spec_spec_to_phys(dns_spec *h)
{
    SP_CHECK(h);
    return h->impl->spec_to_phys();
}
int dns_spec_phys_to_spec(dns_spec *h)
{
    SP_CHECK(h);
    return h->impl->phys_to_spec();
}
int dns_spec_project(dns_spec *h)
{
    SP_CHECK(h);
    return h->impl->project();
}
int dns_spec_rhs(dns_spec *h)
{
    SP_CHECK(h);
    return h->impl->rhs();
}
int dns_spec_step(dns_spec *h, double dt, double *vmax)
{
    SP_CHECK(h);
    return h->impl->step(dt, vmax);
}
int dns_spec_vsummax(dns_spec *h, double *out)
{
    SP_CHECK(h);
    if (!out) return DNS_EINVAL;
    return h->impl->vsummax(out);
}
int dns_spec_shell_energies(dns_spec *h, int32_t s0, int32_t nshells, double *E)
{
    SP_CHECK(h);
    if (!E) return DNS_EINVAL;
    return h->impl->shells(s0, nshells, E);
}
int dns_spec_rescale_shells(dns_spec *h, int32_t s0, int32_t nshells, const double *factors)
{
    SP_CHECK(h);
    if (!factors) return DNS_EINVAL;
    return h->impl->rescale(s0, nshells, factors);
}
int dns_spec_noise(dns_spec *h, uint64_t seed)
{
    SP_CHECK(h);
    return h->impl->noise(seed);
}
int dns_spec_energy(dns_spec *h, double *out)
{
    SP_CHECK(h);
    if (!out) return DNS_EINVAL;
    return h->impl->energy(out, nullptr);
}
int dns_spec_dissipation(dns_spec *h, double *out)
{
    SP_CHECK(h);
    if (!out) return DNS_EINVAL;
    return h->impl->energy(nullptr, out);
}
int dns_spec_maxdiv(dns_spec *h, double *out)
{
    SP_CHECK(h);
    if (!out) return DNS_EINVAL;
    return h->impl->maxdiv(out);
}
int dns_spec_products(dns_spec *h)
{
    SP_CHECK(h);
    return h->impl->products();
}
int dns_spec_get_products(dns_spec *h, double *host)
{
    SP_CHECK(h);
    if (!host) return DNS_EINVAL;
    return h->impl->get_products(host);
}

#define SFS_DISPATCH(call64, call32)                                                  \
    do {                                                                              \
        if (auto *d64 = dynamic_cast<SpecImpl<double> *>(dns->impl.get())) {          \
            auto *l64 = dynamic_cast<SpecImpl<double> *>(les->impl.get());            \
            if (!l64) break;                                                          \
            return call64;                                                            \
        }                                                                             \
        if (auto *d32 = dynamic_cast<SpecImpl<float> *>(dns->impl.get())) {           \
            auto *l32 = dynamic_cast<SpecImpl<float> *>(les->impl.get());             \
            if (!l32) break;                                                          \
            return call32;                                                            \
        }                                                                             \
    } while (0)

int dns_sfs_sample(dns_spec *dns, dns_spec *les, double delta_fac, int32_t protectshells)
{
    SP_CHECK(dns);
    SP_CHECK(les);
    SpecBase *D = dns->impl.get(), *L = les->impl.get();
    (void)D;
    (void)L;
    SFS_DISPATCH(([&]() -> int {
                     if (l64->g.n % 2 || l64->g.kcut != l64->g.n / 3 || l64->g.n / 2 > d64->g.kcut ||
                         l64->l != d64->l) {
                         set_error("LES handle must have n = M (even), kcut = M ÷ 3, M/2 <= DNS kcut, same l");
                         return DNS_EINVAL;
                     }
                     return sfs_sample(d64, l64, delta_fac, protectshells);
                 }()),
                 ([&]() -> int {
                     if (l32->g.n % 2 || l32->g.kcut != l32->g.n / 3 || l32->g.n / 2 > d32->g.kcut ||
                         l32->l != d32->l) {
                         set_error("LES handle must have n = M (even), kcut = M ÷ 3, M/2 <= DNS kcut, same l");
                         return DNS_EINVAL;
                     }
                     return sfs_sample(d32, l32, delta_fac, protectshells);
                 }()));
    set_error("DNS and LES handles must have the same precision");
    return DNS_EINVAL;
}

int dns_sfs_get(dns_spec *les, double *ubar_hat, double *tau_hat)
{
    SP_CHECK(les);
    if (auto *l64 = dynamic_cast<SpecImpl<double> *>(les->impl.get())) return sfs_get(l64, ubar_hat, tau_hat);
    if (auto *l32 = dynamic_cast<SpecImpl<float> *>(les->impl.get())) return sfs_get(l32, ubar_hat, tau_hat);
    return DNS_EINVAL;
}

int dns_sfs_physical(dns_spec *les, double *ubar, double *tau)
{
    SP_CHECK(les);
    if (auto *l64 = dynamic_cast<SpecImpl<double> *>(les->impl.get())) return sfs_physical(l64, ubar, tau);
    if (auto *l32 = dynamic_cast<SpecImpl<float> *>(les->impl.get())) return sfs_physical(l32, ubar, tau);
    return DNS_EINVAL;
}

int dns_les_closure_step(dns_spec *les, dns_cnn *net, double dt, double *vmax)
{
    SP_CHECK(les);
    if (!net) {
        set_error("null dns_cnn handle");
        return DNS_EINVAL;
    }
    if (auto *l64 = dynamic_cast<SpecImpl<double> *>(les->impl.get())) return l64->closure_step(net, dt, vmax);
    if (auto *l32 = dynamic_cast<SpecImpl<float> *>(les->impl.get())) return l32->closure_step(net, dt, vmax);
    return DNS_EINVAL;
}

int dns_sfs_append(dns_spec *les, dns_cnn *net)
{
    SP_CHECK(les);
    if (!net) {
        set_error("null dns_cnn handle");
        return DNS_EINVAL;
    }
    if (auto *l64 = dynamic_cast<SpecImpl<double> *>(les->impl.get())) return l64->append_sample(net);
    if (auto *l32 = dynamic_cast<SpecImpl<float> *>(les->impl.get())) return l32->append_sample(net);
    return DNS_EINVAL;
}

int dns_spec_save(dns_spec *h, const char *prefix, double time, int64_t step, const char *meta_toml)
{
    SP_CHECK(h);
    if (!prefix || !*prefix) {
        set_error("null or empty prefix");
        return DNS_EINVAL;
    }
    return spec_save(*h->impl, prefix, time, step, meta_toml);
}
int dns_spec_load(dns_spec *h, const char *prefix, double *time, int64_t *step)
{
    SP_CHECK(h);
    if (!prefix || !*prefix) {
        set_error("null or empty prefix");
        return DNS_EINVAL;
    }
    return spec_load(*h->impl, prefix, time, step);
}
int dns_spec_sync(dns_spec *h)
{
    SP_CHECK(h);
    return h->impl->sync();
}
int dns_spec_timing(dns_spec *h, int32_t enable)
{
    SP_CHECK(h);
    h->impl->timer.flush();
    for (int f = 0; f < 8; ++f) {
        h->impl->timer.ms[f] = 0;
        h->impl->timer.cnt[f] = 0;
    }
    h->impl->timer.on = enable != 0;
    return DNS_OK;
}
int dns_spec_timing_read(dns_spec *h, int32_t family, double *total_ms, int64_t *launches)
{
    SP_CHECK(h);
    if (family < 0 || family >= 8 || !total_ms || !launches) return DNS_EINVAL;
    h->impl->timer.flush();
    *total_ms = h->impl->timer.ms[family];
    *launches = h->impl->timer.cnt[family];
    return DNS_OK;
}

}  // extern "C"
