// Host side of the finite-volume path: problem set-up (grid metrics,
// Poisson and CN coefficients, FFT plans), the RK3 stage orchestration and
// the C ABI declared in include/dns_rocm.h.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "fv_kernels.hpp"
#include "plain_kernels.hpp"
#include "host_util.hpp"

namespace dns {

static thread_local char g_errbuf[1024] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_errbuf, sizeof(g_errbuf), fmt, ap);
    va_end(ap);
}

const char *last_error() { return g_errbuf; }

// Modified wavenumber of the periodic 3-point Laplacian,
// modλ(k, n, Δ) = 2(cospi(2(k−1)/n) − 1)/Δ² (src/poisson.jl:12), k0 = k − 1.
double modlam(int k0, int n, double d)
{
    const long double pi = 3.141592653589793238462643383279502884L;
    int kk = std::min(k0, n - k0); // cos is even: keeps λ(k) == λ(n − k) exactly
    double c = kk == 0 ? 1.0 : (double)cosl(pi * (2.0L * (long double)kk / (long double)n));
    return 2.0 * (c - 1.0) / (d * d);
}

enum { FAM_MOM = 0, FAM_X = 1, FAM_Z = 2, FAM_Y = 3, FAM_PGRAD = 4, FAM_CN = 5 };

struct FVBase {
    virtual ~FVBase() {}
    virtual int set_field(int f, const double *h, int layout) = 0;
    virtual int get_field(int f, double *h, int layout) = 0;
    virtual int momentum(double A, double dt) = 0;
    virtual int project(int psi) = 0;
    virtual int step(double dt) = 0;
    virtual int cfl(double cfl, double *dt) = 0;
    virtual int advance_cfl(double cfl, int nsteps, double tend, double *t, double *dts, int *ndone) = 0;
    virtual int advance_cfl_noalloc(double cfl, int nsteps, double tend, double *t, double *dts, int *ndone) = 0;
    virtual int maxdiv(double *out) = 0;
    virtual int kinetic(double *out) = 0;
    virtual int chstats(double *acc) = 0;
    virtual int spectrum(double *E, int nbins) = 0;
    virtual int sync() = 0;
    virtual int describe(char *buf, int len) = 0;
    LaunchTimer timer;
};

static const double RK_A[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
static const double RK_B[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
static const double RK_ALPHA[3] = {1.0 / 3.0, 5.0 / 12.0, 1.0 / 4.0};

constexpr int X_EPT = 8, NTHR = 256;
constexpr int THOMAS_RMAX = 24, CN_RMAX = 32; // rows per thread segment of the tridiagonal kernels
// (a half-size variant of each kernel serves segments of at most RMAX / 2 rows)
constexpr int CH_NCHUNK = 16;
constexpr int RED_BLOCKS = 1024;

// Scratch of the allocation-free CFL loop (dns_fv_advance_cfl_noalloc), one
// file-scope array per GPU (160 KiB, held until the process ends).  Layout in
// doubles: [0] Δt, [1] t, [2] step counter (an int in its first 4 bytes),
// [NOALLOC_PART, +3·512) per-block maxima of the CFL reduction,
// [NOALLOC_RED, +3) reduced maxima, [NOALLOC_LOG, +NOALLOC_MAX_STEPS) the
// Δt log.  Host code never names the array in a launch or copy: it uses the
// address hipGetSymbolAddress gave for the current device (noalloc_scratch).
constexpr int NOALLOC_MAX_STEPS = 16384;
constexpr size_t NOALLOC_PART = 256, NOALLOC_RED = 2048, NOALLOC_LOG = 4096;
constexpr size_t NOALLOC_DOUBLES = NOALLOC_LOG + NOALLOC_MAX_STEPS;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

// Device address of `scratch` on device dev (the current device), resolved
// once per device and kept.
static int noalloc_scratch(int dev, double **out)
{
    static std::mutex mu;
    static double *addr[64] = {};
    if (dev < 0 || dev >= 64) {
        set_error("device ordinal %d out of range", dev);
        return DNS_EUNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        DNS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        addr[dev] = static_cast<double *>(p);
    }
    *out = addr[dev];
    return DNS_OK;
}

// Start of an allocation-free CFL loop: t and a zero step counter.  One thread.
__global__ void k_noalloc_start(double *s, double t)
{
    s[1] = t;
    *reinterpret_cast<int *>(s + 2) = 0;
}

template <typename T> struct FVImpl : FVBase {
    // elements per thread of the z / periodic-y FFT tiles: FP32 transforms a
    // whole 16-line tile in one LDS batch (measured 8 % faster at 512³ than
    // two batches of 8); FP64 keeps 8 (fewer registers)
    static constexpr int ZY_EPT = sizeof(T) == 8 ? 8 : 16;
    FVGeom g{};
    int imexy = 0;
    double visc = 0, f[3] = {0, 0, 0}, dx = 0, dz = 0, dymin = 0;
    std::vector<double> dyh, dyuh;
    hipStream_t st = nullptr;
    T *u[3] = {}, *v[3] = {}, *q[3] = {}, *p = nullptr, *delta = nullptr;
    cplx<T> *A = nullptr;
    double2 *Dth = nullptr;
    int thomas_tl = 0, thomas_nthr = 256; // lines and threads per block of the LDS Thomas kernel (0: sequential)
    int thomas_seg = 0;                   // rows per segment of the equal-segment Thomas kernel (0: not used)
    size_t ldst = 0;
    int cn_tl = 0, cn_nthr = 256; // x-lines and threads per block of the LDS CN kernel (0: sequential)
    int cn_seg = 0;               // rows per segment of the equal-segment CN kernel (0: not used)
    int cn_fused = 0;             // the three components in one launch (k_cn_seg3): 1 pressure rows re-read, 2 staged in LDS
    int cn_nl = 1;                // x-lines per thread of the equal-segment CN kernel
    size_t ldscn = 0;
    double *Dcn = nullptr;
    T *dy_t = nullptr, *rdy_t = nullptr, *rdyu_t = nullptr;
    double *dy_d = nullptr, *dyu_d = nullptr, *rdy_d = nullptr, *rdyu_d = nullptr;
    double *lam1 = nullptr, *lam2 = nullptr, *lam3 = nullptr, *ay = nullptr, *cy = nullptr, *iptab = nullptr;
    double *ydc[6] = {}, *cncoef = nullptr;
    FFTDesc fdx{}, fdz{}, fdy{};
    cplx<T> *twN = nullptr, *twx = nullptr, *twz = nullptr, *twy = nullptr;
    int lpb = 1, tbz = 1, tby = 1, tbzf = 1, tbyf = 1; // tile lines and FFT sub-batch lines
    int nthrz = 256, nthry = 256;                      // threads of the z / periodic-y FFT blocks
    size_t ldsx = 0, ldsz = 0, ldsy = 0;
    // red/redout: reduction partials of fixed size, captured by the CFL graphs
    // and therefore never reallocated; spectra use their own grow-only pair
    double *red = nullptr, *redout = nullptr, *upart = nullptr;
    size_t redcap = 0;
    double *spart = nullptr, *sout = nullptr;
    size_t spartcap = 0, soutcap = 0;
    double *pinned = nullptr;
    int N = 0;
    int nb_pgrad = 0;
    bool pgrad_vec = false; // k_pgrad_vec (16-byte accesses) instead of k_pgrad
    bool xwave = false;     // x passes by wavefront lane groups (nx/2 in 16..256, a power of two)
    bool umax_valid = false; // upart holds max |u_c| partials of the current u
    int momentum_kind = 5; // 5 LDS z-march (k_momentum_lz), 3 slab-ordered (k_momentum_slab), 1 direct (k_momentum)
    // reference-structured ("plain") step: extra arrays, allocated on first use
    bool plain = false;
    T *pr = nullptr, *pwork = nullptr, *pghost = nullptr, *pyb[3] = {};
    cplx<T> *pB = nullptr, *pC = nullptr;

    // HIP graphs of one fused step at a fixed Δt: the explicit form swaps the
    // u / u′ buffers once per step, so there is one graph per buffer parity
    // (index 0: u in the buffers allocated as u).  A step replays a graph
    // when the same Δt was used by the previous step and no per-launch timing
    // is on; DNS_FV_GRAPH=0 disables replay.
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    double gdt[2] = {0.0, 0.0};
    double last_dt = -1.0;
    T *ubuf0 = nullptr;
    bool graphs = true;
    // CFL-adaptive explicit steps with Δt computed on the device (advance_cfl):
    // graphs per buffer parity of [CFL reduction, Δt, step], keyed by (cfl, tend)
    hipGraphExec_t gcfl[2] = {nullptr, nullptr};
    double gcfl_key[2][2] = {{0, 0}, {0, 0}};
    const double *mom_dtp = nullptr; // set while capturing: momentum reads Δt here
    double *ddt = nullptr, *dtime = nullptr, *dlog = nullptr;
    int *dstep = nullptr;
    int nlog = 0;
    double *noalloc_s = nullptr; // this device's address of `scratch`

    ~FVImpl() override
    {
        if (st) hipStreamSynchronize(st);
        for (auto &ge : gexec)
            if (ge) hipGraphExecDestroy(ge);
        for (auto &ge : gcfl)
            if (ge) hipGraphExecDestroy(ge);
        for (void *q_ : {(void *)ddt, (void *)dtime, (void *)dlog, (void *)dstep})
            if (q_) hipFree(q_);
        timer.flush();
        void *ptrs[] = {u[0], u[1], u[2], v[0], v[1], v[2], q[0], q[1], q[2], p, delta, A, Dth, Dcn, dy_t, rdy_t,
                        rdyu_t, dy_d, dyu_d, rdy_d, rdyu_d, lam1, lam2, lam3, ay, cy, iptab, ydc[0], ydc[1], ydc[2],
                        ydc[3], ydc[4], ydc[5], cncoef, twN, twx, twz, twy, red, redout, upart, spart, sout,
                        pr, pwork, pghost, pyb[0], pyb[1], pyb[2], pB, pC};
        for (void *q_ : ptrs)
            if (q_) hipFree(q_);
        if (pinned) hipHostFree(pinned);
        if (st) hipStreamDestroy(st);
    }

    int create(const dns_fv_params *P)
    {
        const int nx = P->n[0], ny = P->n[1], nz = P->n[2];
        g.nx = nx;
        g.ny = ny;
        g.nz = nz;
        g.nxc = nx / 2 + 1;
        {
            const int per = 128 / (int)sizeof(cplx<T>); // complex values per 128-byte line
            g.ldc = (g.nxc + per - 1) / per * per;
        }
        g.wall = P->ywall ? 1 : 0;
        imexy = P->ydiffusion ? 1 : 0;
        {
            // DNS_MOMENTUM selects the stencil kernel: lz (default), slab, direct
            const char *e = getenv("DNS_MOMENTUM");
            momentum_kind = !e                       ? 5
                            : !strcmp(e, "direct") ? 1
                            : !strcmp(e, "slab")   ? 3
                                                   : 5;
            // the LDS z-march loads whole rows as 16-byte vectors
            if (momentum_kind == 5 && nx % (16 / (int)sizeof(T)) != 0) momentum_kind = 3;
        }
        visc = P->visc;
        for (int c = 0; c < 3; ++c) f[c] = P->bodyforce[c];
        N = nx * ny * nz;
        dx = (P->lims[1] - P->lims[0]) / nx;
        dz = (P->lims[5] - P->lims[4]) / nz;

        // griddata (src/setup.jl:148-193): extended faces t = 0..ny+2 hold
        // faces −1..ny+1; ghost faces mirror at walls or wrap periodically.
        const double *yf = P->yfaces;
        std::vector<double> yfe(ny + 3);
        for (int j = 0; j <= ny; ++j) yfe[j + 1] = yf[j];
        if (g.wall) {
            yfe[0] = 2 * yf[0] - yf[1];
            yfe[ny + 2] = 2 * yf[ny] - yf[ny - 1];
        } else {
            yfe[0] = yf[0] - (yf[ny] - yf[ny - 1]);
            yfe[ny + 2] = yf[ny] + (yf[1] - yf[0]);
        }
        dyh.assign(ny + 2, 0.0);
        dyuh.assign(ny + 2, 0.0);
        std::vector<double> yce(ny + 2);
        for (int c = 0; c < ny + 2; ++c) {
            dyh[c] = yfe[c + 1] - yfe[c];
            yce[c] = (yfe[c] + yfe[c + 1]) / 2;
        }
        for (int c = 0; c < ny + 1; ++c) dyuh[c] = yce[c + 1] - yce[c];
        dyuh[ny + 1] = dyh[ny + 1];
        dymin = *std::min_element(dyh.begin() + 1, dyh.begin() + ny + 1);
        if (!g.wall) {
            double mx = *std::max_element(dyh.begin() + 1, dyh.begin() + ny + 1);
            if (!(mx - dymin < 1e-12 * mx)) {
                set_error("periodic y requires a uniform grid (stretch = identity)");
                return DNS_EINVAL;
            }
        }

        DNS_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        timer.st = st;
        {
            int dev = 0;
            DNS_HIP_CHECK(hipGetDevice(&dev));
            DNS_TRY(noalloc_scratch(dev, &noalloc_s));
        }
        for (int c = 0; c < 3; ++c) {
            DNS_TRY(dalloc(&u[c], N));
            DNS_TRY(dalloc(&q[c], N));
            if (!imexy) DNS_TRY(dalloc(&v[c], N));
        }
        ubuf0 = u[0];
        if (const char *eg = getenv("DNS_FV_GRAPH")) graphs = atoi(eg) != 0;
        DNS_TRY(dalloc(&p, N));
        if (imexy) DNS_TRY(dalloc(&delta, N));
        DNS_TRY(dalloc(&A, (size_t)g.ldc * ny * nz));

        std::vector<T> dyt(ny + 2), rdyt(ny + 2), rdyut(ny + 2);
        std::vector<double> rdyh(ny + 2), rdyuh(ny + 2);
        for (int c = 0; c < ny + 2; ++c) {
            rdyh[c] = 1.0 / dyh[c];
            rdyuh[c] = 1.0 / dyuh[c];
            dyt[c] = (T)dyh[c];
            rdyt[c] = (T)rdyh[c];
            rdyut[c] = (T)rdyuh[c];
        }
        DNS_TRY(dupload(&dy_t, dyt));
        DNS_TRY(dupload(&rdy_t, rdyt));
        DNS_TRY(dupload(&rdyu_t, rdyut));
        DNS_TRY(dupload(&dy_d, dyh));
        DNS_TRY(dupload(&dyu_d, dyuh));
        DNS_TRY(dupload(&rdy_d, rdyh));
        DNS_TRY(dupload(&rdyu_d, rdyuh));

        // plan_poisson (src/poisson.jl:167-231)
        std::vector<double> l1(g.nxc), l3(nz);
        for (int k = 0; k < g.nxc; ++k) l1[k] = modlam(k, nx, dx);
        for (int k = 0; k < nz; ++k) l3[k] = modlam(k, nz, dz);
        DNS_TRY(dupload(&lam1, l1));
        DNS_TRY(dupload(&lam3, l3));
        if (!make_fft_desc(nx / 2, fdx) || !make_fft_desc(nz, fdz)) {
            set_error("unsupported grid: nx/2 and nz must factor into 2, 3 and 5 (nx even)");
            return DNS_EUNSUPPORTED;
        }
        if (nx % 2 != 0) {
            set_error("unsupported grid: nx must be even");
            return DNS_EUNSUPPORTED;
        }
        if (g.wall) {
            std::vector<double> a(ny), b(ny), c(ny);
            for (int j = 1; j <= ny; ++j) {
                double aa = j == 1 ? 0.0 : 1 / (dyh[j] * dyuh[j - 1]);
                double cc = j == ny ? 0.0 : 1 / (dyh[j] * dyuh[j]);
                a[j - 1] = aa;
                c[j - 1] = cc;
                b[j - 1] = -(aa + cc);
            }
            DNS_TRY(dupload(&ay, a));
            DNS_TRY(dupload(&cy, c));
            // inverse pivots of every line (kx, |kz|) of D_y + λ
            const int nzs = nz / 2 + 1;
            std::vector<double> ip((size_t)g.nxc * ny * nzs);
            for (int kz = 0; kz < nzs; ++kz)
                for (int kx = 0; kx < g.nxc; ++kx) {
                    double lam = l1[kx] + l3[kz];
                    bool mean = lam == 0.0;
                    double beta = mean ? 1.0 : b[0] + lam;
                    double cp = (mean ? 0.0 : c[0]) / beta;
                    ip[kx + (size_t)g.nxc * ((size_t)ny * kz)] = 1.0 / beta;
                    for (int j = 1; j < ny; ++j) {
                        beta = (b[j] + lam) - a[j] * cp;
                        cp = c[j] / beta;
                        ip[kx + (size_t)g.nxc * (j + (size_t)ny * kz)] = 1.0 / beta;
                    }
                }
            DNS_TRY(dupload(&iptab, ip));
        } else {
            std::vector<double> l2(ny);
            for (int k = 0; k < ny; ++k) l2[k] = modlam(k, ny, dyh[1]);
            DNS_TRY(dupload(&lam2, l2));
            if (!make_fft_desc(ny, fdy)) {
                set_error("unsupported grid: periodic ny must factor into 2, 3 and 5");
                return DNS_EUNSUPPORTED;
            }
        }
        if (imexy) {
            // ydiffusion_coefficients (src/ydiffusion.jl:31-51)
            std::vector<double> at(ny), bt(ny), ct(ny), an(ny, 0.0), bn(ny, 0.0), cn(ny, 0.0);
            for (int j = 1; j <= ny; ++j) {
                double aa = 1 / (dyh[j] * dyuh[j - 1]);
                double cc = 1 / (dyh[j] * dyuh[j]);
                at[j - 1] = j == 1 ? 0.0 : aa;
                ct[j - 1] = j == ny ? 0.0 : cc;
                bt[j - 1] = j == 1 ? -(cc + 2 * aa) : j == ny ? -(aa + 2 * cc) : -(aa + cc);
            }
            for (int j = 1; j <= ny - 1; ++j) {
                double aa = 1 / (dyuh[j] * dyh[j]);
                double cc = 1 / (dyuh[j] * dyh[j + 1]);
                an[j - 1] = j == 1 ? 0.0 : aa;
                cn[j - 1] = j == ny - 1 ? 0.0 : cc;
                bn[j - 1] = -(aa + cc);
            }
            std::vector<double> *sets[6] = {&at, &bt, &ct, &an, &bn, &cn};
            for (int s = 0; s < 6; ++s) DNS_TRY(dupload(&ydc[s], *sets[s]));
            DNS_TRY(dalloc(&cncoef, 12 * (size_t)ny)); // (stage, set): c', 1/β
        }

        // FFT tiles
        const int Nh = nx / 2;
        if (Nh > X_EPT * NTHR || nz > ZY_EPT * NTHR || (!g.wall && ny > ZY_EPT * NTHR)) {
            set_error("unsupported grid: FFT line longer than %d points", ZY_EPT * NTHR);
            return DNS_EUNSUPPORTED;
        }
        int dev = 0, maxlds = 65536;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        if (maxlds <= 0) maxlds = 65536;
        lpb = std::max(1, std::min(64, X_EPT * NTHR / std::max(1, Nh)));
        if (sizeof(T) == 8) lpb = std::min(lpb, 4); // FP64: four lines per block measured fastest (C2, C3)
        {
            // the wavefront-group x passes serve nx/2 = 16, 32, ..., 256;
            // DNS_FV_XWAVE=0 keeps the block kernels
            const char *ew = getenv("DNS_FV_XWAVE");
            xwave = (Nh == 16 || Nh == 32 || Nh == 64 || Nh == 128 || Nh == 256) && !(ew && atoi(ew) == 0);
        }
        if (const char *el = getenv("DNS_X_LPB")) // sweeps: fewer lines per x-pass block
            lpb = std::max(1, std::min(lpb, atoi(el)));
        ldsx = (size_t)(lpb + 1) * Nh * sizeof(cplx<T>); // lines + LDS twiddle table
        // z / periodic-y FFT tiles: a block loads tb lines and transforms them
        // in sub-batches of tbf lines.  In FP32, tiles narrower than 128-byte
        // rows move to 512 threads and at least 16 lines (measured faster); in
        // FP64 the 256-thread one-sub-batch tile measured best.
        auto pick_fft = [&](int n, int &tb, int &tbf, int &nthr) {
            const int rowlines = sizeof(T) == 4 ? (int)(128 / sizeof(cplx<T>)) : 1;
            nthr = (ZY_EPT * NTHR / n) < rowlines ? 2 * NTHR : NTHR;
            tbf = std::max(1, std::min(64, ZY_EPT * nthr / n));
            while (tbf > 1 && ((size_t)tbf * (n + 1) + n) * sizeof(cplx<T>) > (size_t)maxlds) --tbf;
            tbf = std::min(tbf, g.nxc);
            tb = std::min(g.nxc, std::max(tbf, rowlines));
            while (tb > tbf && ((size_t)tb * (n + 1) + n) * sizeof(cplx<T>) > (size_t)maxlds) --tb;
        };
        pick_fft(nz, tbz, tbzf, nthrz);
        ldsz = ((size_t)tbz * (nz + 1) + nz) * sizeof(cplx<T>); // tile + LDS twiddle table
        if (!g.wall) {
            pick_fft(ny, tby, tbyf, nthry);
            ldsy = ((size_t)tby * (ny + 1) + ny) * sizeof(cplx<T>);
        }
        if (g.wall) {
            // segmented tridiagonal kernels: a block of 256 or 512 threads owns TL
            // lines (128-byte rows where the segment length allows), each thread
            // one segment of at most RMAX rows in registers; DNS_THOMAS / DNS_CN =
            // sequential select the one-thread-per-line kernels with an HBM d'
            // scratch.
            auto pick_tile = [&](int elem_bytes, int rmax, auto need, int &tl, int &nthr, size_t &lds) {
                // 256 threads first (measured faster than 512), then the widest rows up to
                // 128 bytes whose segments fit rmax rows
                for (int nt : {256, 512})
                    for (int t = std::max(1, 128 / elem_bytes); t >= 1; t /= 2) {
                        const int rows = (ny + nt / t - 1) / (nt / t);
                        if (rows <= rmax && need(t, nt) <= (size_t)maxlds) {
                            tl = t;
                            nthr = nt;
                            lds = need(t, nt);
                            return;
                        }
                    }
                tl = 0;
            };
            const char *env = getenv("DNS_THOMAS");
            pick_tile((int)sizeof(cplx<T>), THOMAS_RMAX,
                      [&](int tl, int nt) {
                          (void)tl;
                          return (size_t)nt * 40 + (size_t)2 * ny * sizeof(double);
                      },
                      thomas_tl, thomas_nthr, ldst);
            if (env && !strcmp(env, "sequential")) thomas_tl = 0;
            if (const char *et = getenv("DNS_THOMAS_TL")) { // lines per block (sweep)
                const int tl = atoi(et);
                if (thomas_tl && tl >= 1 && 256 % tl == 0 && (ny + 256 / tl - 1) / (256 / tl) <= THOMAS_RMAX)
                    thomas_tl = tl;
            }
            if (!thomas_tl) DNS_TRY(dalloc(&Dth, (size_t)g.ldc * ny * nz));
            // equal segments of R rows (ny = S·R) use k_thomas_seg; DNS_THOMAS=tile keeps k_thomas_tile
            thomas_seg = 0;
            if (thomas_tl && thomas_nthr == 256 && !(env && !strcmp(env, "tile")) &&
                (size_t)g.ldc * ny * sizeof(cplx<T>) < ((size_t)1 << 31)) {
                const int S = thomas_nthr / thomas_tl, R = ny / S;
                if (ny % S == 0 && (R == 4 || R == 6 || R == 8 || R == 12 || R == 16 || R == 24)) thomas_seg = R;
            }
            if (imexy && (size_t)6 * ny * sizeof(double) > (size_t)maxlds) {
                set_error("unsupported grid: implicit y-diffusion coefficients need %zu bytes of LDS",
                          (size_t)6 * ny * sizeof(double));
                return DNS_EUNSUPPORTED;
            }
            if (imexy) {
                const char *envc = getenv("DNS_CN");
                pick_tile((int)sizeof(T), CN_RMAX,
                          [&](int tl, int nt) {
                              (void)tl;
                              return (size_t)3 * nt * sizeof(double) + (size_t)6 * ny * sizeof(double);
                          },
                          cn_tl, cn_nthr, ldscn);
                if (envc && !strcmp(envc, "sequential")) cn_tl = 0;
                if (const char *et = getenv("DNS_CN_TILE")) { // "TL:threads" override for tile sweeps
                    int tl = 0, nt = 0;
                    if (sscanf(et, "%d:%d", &tl, &nt) == 2 && tl > 0 && (nt == 256 || nt == 512) && nt % tl == 0 &&
                        (ny + nt / tl - 1) / (nt / tl) <= CN_RMAX) {
                        cn_tl = tl;
                        cn_nthr = nt;
                        ldscn = (size_t)3 * nt * sizeof(double) + (size_t)6 * ny * sizeof(double);
                    }
                }
                if (!cn_tl) DNS_TRY(dalloc(&Dcn, (size_t)N));
                // equal segments of R rows (ny = S·R) use k_cn_seg; DNS_CN=tile keeps k_cn_tile
                cn_seg = 0;
                // (its buffer offsets are 32-bit byte offsets within one z plane)
                if (cn_tl && cn_nthr == 256 && !(envc && !strcmp(envc, "tile")) &&
                    (size_t)nx * ny * sizeof(T) < ((size_t)1 << 31)) {
                    const int S = cn_nthr / cn_tl, R = ny / S;
                    if (ny % S == 0 && (R == 8 || R == 16 || R == 24 || R == 32) &&
                        ldscn + (size_t)ny * sizeof(double) <= (size_t)maxlds)
                        cn_seg = R;
                }
                // DNS_CN_NL=2: two x-lines per thread (pair accesses, block rows twice
                // as wide, same segments); one line measured as fast in FP64 and faster
                // in FP32 (the pair version spills there)
                cn_nl = 1;
                if (cn_seg && nx % 2 == 0) {
                    const char *en = getenv("DNS_CN_NL");
                    const size_t need = (size_t)6 * cn_nthr * sizeof(double) + (size_t)7 * ny * sizeof(double);
                    if (en && atoi(en) == 2 && need <= (size_t)maxlds) {
                        cn_nl = 2;
                        cn_tl *= 2;
                    }
                }
                // The three components in one launch (k_cn_seg3) with the pressure
                // rows staged once in LDS (2) or re-read through L2 (1), or one
                // launch per component (0).  Measured on the 512x384x512 channel:
                // FP64 2 is 2.4 % faster in steps/s than 0 (17 % fewer bytes at a
                // lower bandwidth: two blocks per CU) and 1 no faster; mixed 2 is
                // 4-5 % slower.  Default: 2 in FP64 where its LDS fits, else 0;
                // DNS_CN_FUSED=0|1|2 overrides (1 falls back from 2 when the LDS
                // does not fit)
                {
                    const char *ef = getenv("DNS_CN_FUSED");
                    cn_fused = ef ? atoi(ef) : (sizeof(T) == 8 ? 2 : 0);
                    if (!cn_seg || cn_nl != 1 || cn_fused < 0 || cn_fused > 2) cn_fused = 0;
                    if (cn_fused == 2 &&
                        (size_t)(3 + cn_seg + 1) * cn_nthr * sizeof(double) + (size_t)7 * ny * sizeof(double) > (size_t)maxlds)
                        cn_fused = ef ? 1 : 0;
                }
                if (cn_tl && !cn_seg && (ny + cn_nthr / cn_tl - 1) / (cn_nthr / cn_tl) > CN_RMAX) {
                    set_error("CN tile %d:%d leaves more than %d rows per thread", cn_tl, cn_nthr, CN_RMAX);
                    return DNS_EUNSUPPORTED;
                }
            }
        }
        if (ldsz > (size_t)maxlds || ldsy > (size_t)maxlds) {
            set_error("FFT tile does not fit the %d-byte LDS limit", maxlds);
            return DNS_EUNSUPPORTED;
        }
        DNS_TRY(dupload(&twN, twiddles<T>(std::max(1, Nh), std::max(1, Nh))));
        DNS_TRY(dupload(&twx, twiddles<T>(nx, Nh + 1)));
        DNS_TRY(dupload(&twz, twiddles<T>(nz, nz)));
        if (!g.wall) DNS_TRY(dupload(&twy, twiddles<T>(ny, ny)));

        redcap = std::max<size_t>({(size_t)RED_BLOCKS * 4, (size_t)16 * ny * CH_NCHUNK + 16});
        DNS_TRY(dalloc(&red, redcap));
        // vector pressure-gradient kernel unless nx is not a multiple of its
        // vector width (DNS_PGRAD=scalar keeps the one-cell-per-thread kernel)
        {
            const char *ep = getenv("DNS_PGRAD");
            pgrad_vec = nx % (16 / (int)sizeof(T)) == 0 && !(ep && !strcmp(ep, "scalar"));
        }
        nb_pgrad = ceil_div(pgrad_vec ? nx / (16 / (int)sizeof(T)) : nx, 64) * ceil_div(ny, 4) * nz;
        DNS_TRY(dalloc(&upart, (size_t)3 * nb_pgrad));
        DNS_TRY(dalloc(&redout, (size_t)8 * ny + 64));
        DNS_HIP_CHECK(hipHostMalloc((void **)&pinned, sizeof(double) * ((size_t)8 * ny + 64)));
        return DNS_OK;
    }

    // ------------------------------------------------------------ fields
    T *field_ptr(int fid)
    {
        switch (fid) {
            case DNS_FIELD_UX: return u[0];
            case DNS_FIELD_UY: return u[1];
            case DNS_FIELD_UZ: return u[2];
            case DNS_FIELD_P: return p;
            case DNS_FIELD_QX: return q[0];
            case DNS_FIELD_QY: return q[1];
            case DNS_FIELD_QZ: return q[2];
            case DNS_FIELD_DELTA: return delta ? delta : p;
        }
        return nullptr;
    }

    int set_field(int fid, const double *h, int layout) override
    {
        T *d = field_ptr(fid);
        if (!d || !h) {
            set_error("invalid field id %d or null pointer", fid);
            return DNS_EINVAL;
        }
        const int nx = g.nx, ny = g.ny, nz = g.nz;
        std::vector<T> s((size_t)N);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    size_t src = layout == DNS_LAYOUT_GHOSTED
                                     ? (size_t)(i + 1) + (size_t)(nx + 2) * ((j + 1) + (size_t)(ny + 2) * (k + 1))
                                     : (size_t)i + (size_t)nx * (j + (size_t)ny * k);
                    s[i + (size_t)nx * (j + (size_t)ny * k)] = (T)h[src];
                }
        // the top wall face of u.y is a boundary value (zero), src/halo.jl:31-33
        if (fid == DNS_FIELD_UY && g.wall)
            for (int k = 0; k < nz; ++k)
                for (int i = 0; i < nx; ++i) s[i + (size_t)nx * ((ny - 1) + (size_t)ny * k)] = T(0);
        DNS_HIP_CHECK(hipMemcpyAsync(d, s.data(), sizeof(T) * N, hipMemcpyHostToDevice, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        if (fid <= DNS_FIELD_UZ) umax_valid = false;
        return DNS_OK;
    }

    int get_field(int fid, double *h, int layout) override
    {
        T *d = field_ptr(fid);
        if (!d || !h) {
            set_error("invalid field id %d or null pointer", fid);
            return DNS_EINVAL;
        }
        const int nx = g.nx, ny = g.ny, nz = g.nz;
        std::vector<T> s((size_t)N);
        DNS_HIP_CHECK(hipMemcpyAsync(s.data(), d, sizeof(T) * N, hipMemcpyDeviceToHost, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        if (layout != DNS_LAYOUT_GHOSTED) {
            for (size_t t = 0; t < (size_t)N; ++t) h[t] = (double)s[t];
            return DNS_OK;
        }
        const int GX = nx + 2, GY = ny + 2, GZ = nz + 2;
        auto G = [&](int i, int j, int k) { return (size_t)i + (size_t)GX * (j + (size_t)GY * k); };
        std::fill(h, h + (size_t)GX * GY * GZ, 0.0);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) h[G(i + 1, j + 1, k + 1)] = (double)s[i + (size_t)nx * (j + (size_t)ny * k)];
        if (fid >= DNS_FIELD_QX && fid <= DNS_FIELD_QZ) return DNS_OK; // registers carry no halos
        int stag = fid == DNS_FIELD_UX ? 0 : fid == DNS_FIELD_UY ? 1 : fid == DNS_FIELD_UZ ? 2 : 3;
        // exchange_halo! order: x, then y, then z; slabs span the ghosts
        for (int k = 0; k < GZ; ++k)
            for (int j = 0; j < GY; ++j) {
                h[G(0, j, k)] = h[G(nx, j, k)];
                h[G(nx + 1, j, k)] = h[G(1, j, k)];
            }
        for (int k = 0; k < GZ; ++k)
            for (int i = 0; i < GX; ++i) {
                if (!g.wall) {
                    h[G(i, 0, k)] = h[G(i, ny, k)];
                    h[G(i, ny + 1, k)] = h[G(i, 1, k)];
                } else if (stag == 1) {
                    h[G(i, 0, k)] = 0.0;
                    h[G(i, ny, k)] = 0.0;
                    h[G(i, ny + 1, k)] = -h[G(i, ny - 1, k)];
                } else {
                    double sg = stag == 3 ? 1.0 : -1.0;
                    h[G(i, 0, k)] = sg * h[G(i, 1, k)];
                    h[G(i, ny + 1, k)] = sg * h[G(i, ny, k)];
                }
            }
        for (int j = 0; j < GY; ++j)
            for (int i = 0; i < GX; ++i) {
                h[G(i, j, 0)] = h[G(i, j, nz)];
                h[G(i, j, nz + 1)] = h[G(i, j, 1)];
            }
        return DNS_OK;
    }

    // ---------------------------------------------------------- kernels
    int launch_momentum(T **uu, T **vv, double Ac, double dt, double B, bool write_v)
    {
        MomentumArgs<T> a;
        a.ux = uu[0];
        a.uy = uu[1];
        a.uz = uu[2];
        a.qx = q[0];
        a.qy = q[1];
        a.qz = q[2];
        a.vx = vv ? vv[0] : nullptr;
        a.vy = vv ? vv[1] : nullptr;
        a.vz = vv ? vv[2] : nullptr;
        a.dy = dy_t;
        a.rdy = rdy_t;
        a.rdyu = rdyu_t;
        a.A = (T)Ac;
        a.dt = (T)dt;
        a.dtp = mom_dtp;
        a.nu = (T)visc;
        a.fx = (T)f[0];
        a.fy = (T)f[1];
        a.fz = (T)f[2];
        a.epsy = (T)(imexy ? 0.0 : 1.0);
        a.B = (T)B;
        a.rdx = (T)(1.0 / dx);
        a.rdx2 = (T)(1.0 / (dx * dx));
        a.rdz = (T)(1.0 / dz);
        a.rdz2 = (T)(1.0 / (dz * dz));
        timer.begin(FAM_MOM);
        const bool rq = Ac != 0.0;
        if (momentum_kind == 5) {
            // z chunks: the longest power-of-two march up to 32 planes that still
            // gives one full round of resident blocks (3 per CU in FP64, 6 in FP32)
            // over 256 CUs (measured: longer marches win until the grid stops
            // filling the chip; a partial last round costs a whole round)
            const int nbx = ceil_div(g.nx, 64), nby = ceil_div(g.ny, 8);
            const int zc = momentum_zc();
            dim3 grd(nbx, nby, ceil_div(g.nz, zc));
            if (write_v) {
                if (rq) hipLaunchKernelGGL((k_momentum_lz<T, true, true>), grd, dim3(256), 0, st, a, g, zc);
                else hipLaunchKernelGGL((k_momentum_lz<T, false, true>), grd, dim3(256), 0, st, a, g, zc);
            } else {
                if (rq) hipLaunchKernelGGL((k_momentum_lz<T, true, false>), grd, dim3(256), 0, st, a, g, zc);
                else hipLaunchKernelGGL((k_momentum_lz<T, false, false>), grd, dim3(256), 0, st, a, g, zc);
            }
            timer.end(FAM_MOM);
            DNS_HIP_CHECK(hipGetLastError());
            return DNS_OK;
        }
        if (momentum_kind == 3) {
            // y-slabs whose three-plane u window (3 planes × 3 components) stays within ~2.5 MB of L2
            const int nbx = ceil_div(g.nx, 64), nby = ceil_div(g.ny, 4);
            const size_t rowbytes = (size_t)g.nx * 4 * 9 * sizeof(T); // one block row, 3 planes × 3 components
            const int nslab = ceil_div(nby, std::max(1, std::min(nby, (int)((size_t)(2560 * 1024) / rowbytes))));
            const int nbys = ceil_div(nby, nslab);
            dim3 blk(64, 4, 1), grd((unsigned)nbx * nbys * g.nz * nslab);
            if (write_v) {
                if (rq) hipLaunchKernelGGL((k_momentum_slab<T, true, true>), grd, blk, 0, st, a, g, nbx, nbys);
                else hipLaunchKernelGGL((k_momentum_slab<T, false, true>), grd, blk, 0, st, a, g, nbx, nbys);
            } else {
                if (rq) hipLaunchKernelGGL((k_momentum_slab<T, true, false>), grd, blk, 0, st, a, g, nbx, nbys);
                else hipLaunchKernelGGL((k_momentum_slab<T, false, false>), grd, blk, 0, st, a, g, nbx, nbys);
            }
            timer.end(FAM_MOM);
            DNS_HIP_CHECK(hipGetLastError());
            return DNS_OK;
        }
        dim3 blk(64, 4, 1), grd(ceil_div(g.nx, 64), ceil_div(g.ny, 4), g.nz);
        if (write_v) {
            if (rq) hipLaunchKernelGGL((k_momentum<T, true, true>), grd, blk, 0, st, a, g);
            else hipLaunchKernelGGL((k_momentum<T, false, true>), grd, blk, 0, st, a, g);
        } else {
            if (rq) hipLaunchKernelGGL((k_momentum<T, true, false>), grd, blk, 0, st, a, g);
            else hipLaunchKernelGGL((k_momentum<T, false, false>), grd, blk, 0, st, a, g);
        }
        timer.end(FAM_MOM);
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    // poisson! of the divergence of uu, writing the pressure into psi
    // (src/poisson.jl:147-161), without transposes: x pass, z pass, y solve,
    // in-place z FFT of the (nxc, ny, nz) pencil A (k_fft_z, 256- or
    // 512-thread tiles)
    template <bool INV> void launch_fft_z(cplx<T> *A_)
    {
        const dim3 grd(ceil_div(g.nxc, tbz), g.ny);
        if (nthrz == 2 * NTHR)
            hipLaunchKernelGGL((k_fft_z<T, ZY_EPT, INV, 2 * NTHR>), grd, dim3(nthrz), ldsz, st, A_, twz, tbz, tbzf,
                               make_fastdiv(tbz), g, fdz);
        else
            hipLaunchKernelGGL((k_fft_z<T, ZY_EPT, INV, NTHR>), grd, dim3(nthrz), ldsz, st, A_, twz, tbz, tbzf,
                               make_fastdiv(tbz), g, fdz);
    }

    // z pass, x pass.
    int launch_poisson(T **uu, T *psi, int mode /*0 = divergence of uu*/)
    {
        XPassArgs<T> xa;
        xa.ux = uu[0];
        xa.uy = uu[1];
        xa.uz = uu[2];
        xa.A = A;
        xa.rdy = rdy_t;
        xa.rdx = (T)(1.0 / dx);
        xa.rdz = (T)(1.0 / dz);
        xa.tw = twN;
        xa.twx = twx;
        xa.lpb = lpb;
        xa.fN = make_fastdiv(g.nx / 2);
        xa.fnxc = make_fastdiv(g.nxc);
        xa.fny = make_fastdiv(g.ny);
        const int nlines = g.ny * g.nz;
        timer.begin(FAM_X);
        if (xwave) {
            const int N = g.nx / 2, lpbw = 256 / (N / 4);
            const dim3 grd(ceil_div(nlines, lpbw));
            const size_t lds = (size_t)(N + N + 1 + lpbw * N) * sizeof(cplx<T>);
            switch (N) {
                case 16: hipLaunchKernelGGL((k_rfft_xw<T, 16>), grd, dim3(NTHR), lds, st, xa, g); break;
                case 32: hipLaunchKernelGGL((k_rfft_xw<T, 32>), grd, dim3(NTHR), lds, st, xa, g); break;
                case 64: hipLaunchKernelGGL((k_rfft_xw<T, 64>), grd, dim3(NTHR), lds, st, xa, g); break;
                case 128: hipLaunchKernelGGL((k_rfft_xw<T, 128>), grd, dim3(NTHR), lds, st, xa, g); break;
                default: hipLaunchKernelGGL((k_rfft_xw<T, 256>), grd, dim3(NTHR), lds, st, xa, g); break;
            }
        } else {
            hipLaunchKernelGGL((k_rfft_x<T, X_EPT, 0>), dim3(ceil_div(nlines, lpb)), dim3(NTHR), ldsx, st, xa, g, fdx);
        }
        timer.end(FAM_X);
        DNS_HIP_CHECK(hipGetLastError());
        timer.begin(FAM_Z);
        launch_fft_z<false>(A);
        timer.end(FAM_Z);
        DNS_HIP_CHECK(hipGetLastError());
        timer.begin(FAM_Y);
        if (g.wall) {
            double scale = 1.0 / ((double)g.nx * g.nz);
            if (thomas_tl && thomas_seg) {
                const dim3 grd(ceil_div(g.nxc, thomas_tl), g.nz), blk(thomas_nthr);
#define DNS_THOMAS_SEG(R)                                                                                         \
    case R:                                                                                                       \
        hipLaunchKernelGGL((k_thomas_seg<T, R>), grd, blk, ldst, st, A, iptab, ay, cy, lam1, lam3, scale, thomas_tl, \
                           g);                                                                                    \
        break;
                switch (thomas_seg) {
                    DNS_THOMAS_SEG(4)
                    DNS_THOMAS_SEG(6)
                    DNS_THOMAS_SEG(8)
                    DNS_THOMAS_SEG(12)
                    DNS_THOMAS_SEG(16)
                    DNS_THOMAS_SEG(24)
                    default: break;
                }
#undef DNS_THOMAS_SEG
            } else if (thomas_tl) {
                const dim3 grd(ceil_div(g.nxc, thomas_tl), g.nz), blk(thomas_nthr);
                const int rows = ceil_div(g.ny, thomas_nthr / thomas_tl);
                if (rows <= THOMAS_RMAX / 2)
                    hipLaunchKernelGGL((k_thomas_tile<T, THOMAS_RMAX / 2>), grd, blk, ldst, st, A, iptab, ay, cy, lam1,
                                       lam3, scale, thomas_tl, g);
                else
                    hipLaunchKernelGGL((k_thomas_tile<T, THOMAS_RMAX>), grd, blk, ldst, st, A, iptab, ay, cy, lam1, lam3,
                                       scale, thomas_tl, g);
            } else {
                int lines = g.ldc * g.nz;
                hipLaunchKernelGGL((k_thomas_y<T>), dim3(ceil_div(lines, NTHR)), dim3(NTHR), 0, st, A, Dth, iptab, ay,
                                   cy, lam1, lam3, scale, g);
            }
        } else {
            double scale = 1.0 / ((double)g.nx * g.ny * g.nz);
            hipLaunchKernelGGL((k_fourier_y<T, ZY_EPT, true>), dim3(ceil_div(g.nxc, tby), g.nz), dim3(nthry), ldsy,
                               st, A, twy, lam1, lam2, lam3, scale, tby, tbyf, make_fastdiv(tby), g, fdy);
        }
        timer.end(FAM_Y);
        DNS_HIP_CHECK(hipGetLastError());
        timer.begin(FAM_Z);
        launch_fft_z<true>(A);
        timer.end(FAM_Z);
        DNS_HIP_CHECK(hipGetLastError());
        XInvArgs<T> xi{};
        xi.A = A;
        xi.out = psi;
        xi.tw = twN;
        xi.twx = twx;
        xi.lpb = lpb;
        xi.fN = make_fastdiv(g.nx / 2);
        xi.skip_dt = mom_dtp;
        timer.begin(FAM_X);
        if (xwave) {
            const int N = g.nx / 2, lpbw = 256 / (N / 4);
            const dim3 grd(ceil_div(nlines, lpbw));
            const size_t lds = (size_t)(N + N + 1 + lpbw * N) * sizeof(cplx<T>);
            switch (N) {
                case 16: hipLaunchKernelGGL((k_irfft_xw<T, 16>), grd, dim3(NTHR), lds, st, xi, g); break;
                case 32: hipLaunchKernelGGL((k_irfft_xw<T, 32>), grd, dim3(NTHR), lds, st, xi, g); break;
                case 64: hipLaunchKernelGGL((k_irfft_xw<T, 64>), grd, dim3(NTHR), lds, st, xi, g); break;
                case 128: hipLaunchKernelGGL((k_irfft_xw<T, 128>), grd, dim3(NTHR), lds, st, xi, g); break;
                default: hipLaunchKernelGGL((k_irfft_xw<T, 256>), grd, dim3(NTHR), lds, st, xi, g); break;
            }
        } else {
            hipLaunchKernelGGL((k_irfft_x<T, X_EPT>), dim3(ceil_div(nlines, lpb)), dim3(NTHR), ldsx, st, xi, g, fdx);
        }
        timer.end(FAM_X);
        DNS_HIP_CHECK(hipGetLastError());
        (void)mode;
        return DNS_OK;
    }

    int launch_pgrad(T **uu, T *psi, bool pupdate, double inv_adt)
    {
        PGradArgs<T> a{};
        a.ux = uu[0];
        a.uy = uu[1];
        a.uz = uu[2];
        a.psi = psi;
        a.p = p;
        a.rdyu = rdyu_t;
        a.rdx = (T)(1.0 / dx);
        a.rdz = (T)(1.0 / dz);
        a.s = T(1);
        a.inv_adt = (T)inv_adt;
        a.skip_dt = mom_dtp;
        timer.begin(FAM_PGRAD);
        if (pgrad_vec) {
            const dim3 blk(64, 4, 1), grd(ceil_div(g.nx / (16 / (int)sizeof(T)), 64), ceil_div(g.ny, 4), g.nz);
            if (pupdate) hipLaunchKernelGGL((k_pgrad_vec<T, true, true>), grd, blk, 0, st, a, g, upart);
            else hipLaunchKernelGGL((k_pgrad_vec<T, false, true>), grd, blk, 0, st, a, g, upart);
        } else {
            const dim3 blk(64, 4, 1), grd(ceil_div(g.nx, 64), ceil_div(g.ny, 4), g.nz);
            if (pupdate) hipLaunchKernelGGL((k_pgrad<T, true, true>), grd, blk, 0, st, a, g, upart);
            else hipLaunchKernelGGL((k_pgrad<T, false, true>), grd, blk, 0, st, a, g, upart);
        }
        timer.end(FAM_PGRAD);
        DNS_HIP_CHECK(hipGetLastError());
        umax_valid = true;
        return DNS_OK;
    }

    // equal-segment CN kernel for component c with R rows per thread
    template <int R, int NL> void cn_seg_launch_nl(int c, dim3 grd, dim3 blk, const CNArgs<T> &a)
    {
        const size_t lds = (size_t)3 * NL * cn_nthr * sizeof(double) + (size_t)7 * g.ny * sizeof(double);
        if (c == 0) hipLaunchKernelGGL((k_cn_seg<T, 0, R, NL>), grd, blk, lds, st, a, g, cn_tl);
        else if (c == 1) hipLaunchKernelGGL((k_cn_seg<T, 1, R, NL>), grd, blk, lds, st, a, g, cn_tl);
        else hipLaunchKernelGGL((k_cn_seg<T, 2, R, NL>), grd, blk, lds, st, a, g, cn_tl);
    }
    template <int R> void cn_seg_launch(int c, dim3 grd, dim3 blk, const CNArgs<T> &a)
    {
        if (cn_nl == 2) cn_seg_launch_nl<R, 2>(c, grd, blk, a);
        else cn_seg_launch_nl<R, 1>(c, grd, blk, a);
    }

    // the CN pivots of the three stages of a step at Δt (k_cn_coeffs3), read
    // by launch_cn(stage)
    void launch_cn_coeffs(double dt)
    {
        hipLaunchKernelGGL(k_cn_coeffs3, dim3(1), dim3(NTHR), (size_t)6 * g.ny * sizeof(double), st, ydc[0], ydc[1],
                           ydc[2], ydc[3], ydc[4], ydc[5], RK_ALPHA[0] * dt * visc / 2, RK_ALPHA[1] * dt * visc / 2,
                           RK_ALPHA[2] * dt * visc / 2, g.ny, cncoef);
    }

    int launch_cn(int stage, double dt)
    {
        const double B = RK_B[stage], alpha = RK_ALPHA[stage];
        const double theta = alpha * dt * visc / 2;
        const double *coef = cncoef + (size_t)stage * 4 * g.ny; // this stage's (set 0, set 1) pivots
        if (cn_fused) {
            CN3Args<T> a;
            for (int c = 0; c < 3; ++c) {
                a.u[c] = u[c];
                a.q[c] = q[c];
            }
            a.p = p;
            for (int set = 0; set < 2; ++set) {
                a.ya[set] = ydc[3 * set + 0];
                a.yb[set] = ydc[3 * set + 1];
                a.yc[set] = ydc[3 * set + 2];
                a.cp[set] = coef + set * 2 * g.ny;
                a.ip[set] = a.cp[set] + g.ny;
            }
            a.rdyu = rdyu_d;
            a.B = B;
            a.adt = alpha * dt;
            a.theta = theta;
            a.rdx = 1.0 / dx;
            a.rdz = 1.0 / dz;
            const dim3 grd(ceil_div(g.nx, cn_tl), g.nz), blk(cn_nthr);
            const size_t lds = (size_t)(3 + (cn_fused == 2 ? cn_seg + 1 : 0)) * cn_nthr * sizeof(double) +
                               (size_t)7 * g.ny * sizeof(double);
            timer.begin(FAM_CN);
            if (cn_fused == 2) {
                switch (cn_seg) {
                    case 8: hipLaunchKernelGGL((k_cn_seg3<T, 8, true>), grd, blk, lds, st, a, g, cn_tl); break;
                    case 16: hipLaunchKernelGGL((k_cn_seg3<T, 16, true>), grd, blk, lds, st, a, g, cn_tl); break;
                    case 24: hipLaunchKernelGGL((k_cn_seg3<T, 24, true>), grd, blk, lds, st, a, g, cn_tl); break;
                    default: hipLaunchKernelGGL((k_cn_seg3<T, 32, true>), grd, blk, lds, st, a, g, cn_tl); break;
                }
            } else {
                switch (cn_seg) {
                    case 8: hipLaunchKernelGGL((k_cn_seg3<T, 8, false>), grd, blk, lds, st, a, g, cn_tl); break;
                    case 16: hipLaunchKernelGGL((k_cn_seg3<T, 16, false>), grd, blk, lds, st, a, g, cn_tl); break;
                    case 24: hipLaunchKernelGGL((k_cn_seg3<T, 24, false>), grd, blk, lds, st, a, g, cn_tl); break;
                    default: hipLaunchKernelGGL((k_cn_seg3<T, 32, false>), grd, blk, lds, st, a, g, cn_tl); break;
                }
            }
            timer.end(FAM_CN);
            DNS_HIP_CHECK(hipGetLastError());
            return DNS_OK;
        }
        for (int c = 0; c < 3; ++c) {
            CNArgs<T> a;
            a.u = u[c];
            a.q = q[c];
            a.p = p;
            a.D = Dcn;
            const int set = c == 1 ? 1 : 0;
            a.a = ydc[3 * set + 0];
            a.b = ydc[3 * set + 1];
            a.c = ydc[3 * set + 2];
            a.cp = coef + set * 2 * g.ny;
            a.ip = a.cp + g.ny;
            a.rdyu = rdyu_d;
            a.B = B;
            a.adt = alpha * dt;
            a.theta = theta;
            a.rdx = 1.0 / dx;
            a.rdz = 1.0 / dz;
            timer.begin(FAM_CN); // one component per timed launch
            if (cn_tl && cn_seg) {
                const dim3 grd(ceil_div(g.nx, cn_tl), g.nz), blk(cn_nthr);
                switch (cn_seg) {
                    case 8: cn_seg_launch<8>(c, grd, blk, a); break;
                    case 16: cn_seg_launch<16>(c, grd, blk, a); break;
                    case 24: cn_seg_launch<24>(c, grd, blk, a); break;
                    default: cn_seg_launch<32>(c, grd, blk, a); break;
                }
            } else if (cn_tl) {
                dim3 grd(ceil_div(g.nx, cn_tl), g.nz);
                const dim3 blk(cn_nthr);
                if (ceil_div(g.ny, cn_nthr / cn_tl) <= CN_RMAX / 2) {
                    if (c == 0) hipLaunchKernelGGL((k_cn_tile<T, 0, CN_RMAX / 2>), grd, blk, ldscn, st, a, g, cn_tl);
                    else if (c == 1) hipLaunchKernelGGL((k_cn_tile<T, 1, CN_RMAX / 2>), grd, blk, ldscn, st, a, g, cn_tl);
                    else hipLaunchKernelGGL((k_cn_tile<T, 2, CN_RMAX / 2>), grd, blk, ldscn, st, a, g, cn_tl);
                } else {
                    if (c == 0) hipLaunchKernelGGL((k_cn_tile<T, 0, CN_RMAX>), grd, blk, ldscn, st, a, g, cn_tl);
                    else if (c == 1) hipLaunchKernelGGL((k_cn_tile<T, 1, CN_RMAX>), grd, blk, ldscn, st, a, g, cn_tl);
                    else hipLaunchKernelGGL((k_cn_tile<T, 2, CN_RMAX>), grd, blk, ldscn, st, a, g, cn_tl);
                }
            } else {
                dim3 blk(64), grd(ceil_div(g.nx, 64), g.nz);
                if (c == 0) hipLaunchKernelGGL((k_cn<T, 0>), grd, blk, 0, st, a, g);
                else if (c == 1) hipLaunchKernelGGL((k_cn<T, 1>), grd, blk, 0, st, a, g);
                else hipLaunchKernelGGL((k_cn<T, 2>), grd, blk, 0, st, a, g);
            }
            timer.end(FAM_CN);
        }
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    int momentum(double Ac, double dt) override { return launch_momentum(u, nullptr, Ac, dt, 0.0, false); }

    int project(int psi_field) override
    {
        T *psi = psi_field == DNS_FIELD_DELTA ? (delta ? delta : p) : p;
        DNS_TRY(launch_poisson(u, psi, 0));
        DNS_TRY(launch_pgrad(u, psi, false, 0.0));
        return DNS_OK;
    }

    // ------------------------------------------------------------ plain step
    // The reference's single-rank kernel sequence (src/timestep.jl:41-63):
    // momentum!, separate RK broadcasts, exchange_halo! plane copies with a
    // synchronisation, divergence into r, poisson! with the copies between
    // the rfft/fft/solve passes, copybox!, per-component gradients and, with
    // CN, imex_update! with its staged copies and HBM Thomas scratch.
    int plain_alloc()
    {
        if (pr) return DNS_OK;
        const size_t pen = (size_t)g.ldc * g.ny * g.nz;
        DNS_TRY(dalloc(&pr, (size_t)N));
        DNS_TRY(dalloc(&pwork, (size_t)N));
        DNS_TRY(dalloc(&pghost, (size_t)6 * std::max(g.nx, g.ny) * std::max(g.ny, g.nz)));
        DNS_TRY(dalloc(&pB, pen));
        DNS_TRY(dalloc(&pC, pen));
        if (g.wall && !Dth) DNS_TRY(dalloc(&Dth, pen));
        if (imexy) {
            for (int c = 0; c < 3; ++c) DNS_TRY(dalloc(&pyb[c], (size_t)N));
            if (!Dcn) DNS_TRY(dalloc(&Dcn, (size_t)N));
        }
        return DNS_OK;
    }

    void plain_copy(T *dst, const T *src)
    {
        hipLaunchKernelGGL((k_plain_copy<T>), dim3(2048), dim3(NTHR), 0, st, dst, src, (long)N);
    }
    void plain_copy_pencil(cplx<T> *dst, const cplx<T> *src)
    {
        hipLaunchKernelGGL((k_plain_copy<cplx<T>>), dim3(2048), dim3(NTHR), 0, st, dst, src,
                           (long)g.ldc * g.ny * g.nz);
    }

    int plain_halo(T *const *fields, int nf)
    {
        for (int c = 0; c < nf; ++c)
            for (int face = 0; face < 6; ++face)
                hipLaunchKernelGGL((k_plain_halo_face<T>), dim3(64), dim3(NTHR), 0, st, fields[c], pghost, face, g);
        DNS_HIP_CHECK(hipGetLastError());
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        return DNS_OK;
    }

    int plain_project(T *psi)
    {
        const dim3 blk(64, 4, 1), grd(ceil_div(g.nx, 64), ceil_div(g.ny, 4), g.nz);
        timer.begin(FAM_X);
        hipLaunchKernelGGL((k_plain_div<T>), grd, blk, 0, st, u[0], u[1], u[2], rdy_t, (T)(1.0 / dx), (T)(1.0 / dz), pr,
                           g);
        XPassArgs<T> xa{};
        xa.ux = pr;
        xa.A = A;
        xa.tw = twN;
        xa.twx = twx;
        xa.lpb = lpb;
        xa.fN = make_fastdiv(g.nx / 2);
        xa.fnxc = make_fastdiv(g.nxc);
        xa.fny = make_fastdiv(g.ny);
        const int nlines = g.ny * g.nz;
        hipLaunchKernelGGL((k_rfft_x<T, X_EPT, 1>), dim3(ceil_div(nlines, lpb)), dim3(NTHR), ldsx, st, xa, g, fdx);
        timer.end(FAM_X);
        timer.begin(FAM_Z);
        plain_copy_pencil(pB, A);
        launch_fft_z<false>(pB);
        plain_copy_pencil(pC, pB);
        timer.end(FAM_Z);
        timer.begin(FAM_Y);
        if (g.wall) {
            const int lines = g.ldc * g.nz;
            hipLaunchKernelGGL((k_thomas_y<T>), dim3(ceil_div(lines, NTHR)), dim3(NTHR), 0, st, pC, Dth, iptab, ay, cy,
                               lam1, lam3, 1.0 / ((double)g.nx * g.nz), g);
        } else {
            hipLaunchKernelGGL((k_fourier_y<T, ZY_EPT, true>), dim3(ceil_div(g.nxc, tby), g.nz), dim3(nthry), ldsy,
                               st, pC, twy, lam1, lam2, lam3, 1.0 / ((double)g.nx * g.ny * g.nz), tby, tbyf, make_fastdiv(tby), g, fdy);
        }
        timer.end(FAM_Y);
        timer.begin(FAM_Z);
        plain_copy_pencil(pB, pC);
        launch_fft_z<true>(pB);
        plain_copy_pencil(A, pB);
        timer.end(FAM_Z);
        timer.begin(FAM_X);
        XInvArgs<T> xi{};
        xi.A = A;
        xi.out = pwork;
        xi.tw = twN;
        xi.twx = twx;
        xi.lpb = lpb;
        xi.fN = make_fastdiv(g.nx / 2);
        hipLaunchKernelGGL((k_irfft_x<T, X_EPT>), dim3(ceil_div(nlines, lpb)), dim3(NTHR), ldsx, st, xi, g, fdx);
        plain_copy(psi, pwork); // copybox! into the field interior
        timer.end(FAM_X);
        DNS_HIP_CHECK(hipGetLastError());
        DNS_TRY(plain_halo(&psi, 1));
        timer.begin(FAM_PGRAD);
        hipLaunchKernelGGL((k_plain_pgrad<T, 0>), grd, blk, 0, st, u[0], psi, rdyu_t, (T)(1.0 / dx), (T)(1.0 / dz), T(1), g);
        hipLaunchKernelGGL((k_plain_pgrad<T, 1>), grd, blk, 0, st, u[1], psi, rdyu_t, (T)(1.0 / dx), (T)(1.0 / dz), T(1), g);
        hipLaunchKernelGGL((k_plain_pgrad<T, 2>), grd, blk, 0, st, u[2], psi, rdyu_t, (T)(1.0 / dx), (T)(1.0 / dz), T(1), g);
        timer.end(FAM_PGRAD);
        DNS_HIP_CHECK(hipGetLastError());
        return plain_halo(u, 3);
    }

    int plain_imex(double B, double alpha, double dt)
    {
        const double theta = alpha * dt * visc / 2;
        const dim3 blk(64, 4, 1), grd(ceil_div(g.nx, 64), ceil_div(g.ny, 4), g.nz);
        timer.begin(FAM_CN);
        hipLaunchKernelGGL(k_cn_coeffs, dim3(1), dim3(NTHR), (size_t)6 * g.ny * sizeof(double), st, ydc[0], ydc[1], ydc[2],
                           ydc[3], ydc[4], ydc[5], theta, g.ny, cncoef);
        for (int c = 0; c < 3; ++c) {
            const int set = c == 1 ? 1 : 0;
            const double *ca = ydc[3 * set], *cb = ydc[3 * set + 1], *cc = ydc[3 * set + 2];
            const double *cp = cncoef + set * 2 * g.ny, *ip = cp + g.ny;
            plain_copy(pyb[0], u[c]);   // uⁿ staged
            plain_copy(pyb[1], pyb[0]); // y-local layout (a copy on one rank)
            if (c == 0)
                hipLaunchKernelGGL((k_plain_imexv<T, 0>), grd, blk, 0, st, u[c], q[c], p, rdyu_t, B, alpha * dt,
                                   1.0 / dx, 1.0 / dz, g);
            else if (c == 1)
                hipLaunchKernelGGL((k_plain_imexv<T, 1>), grd, blk, 0, st, u[c], q[c], p, rdyu_t, B, alpha * dt,
                                   1.0 / dx, 1.0 / dz, g);
            else
                hipLaunchKernelGGL((k_plain_imexv<T, 2>), grd, blk, 0, st, u[c], q[c], p, rdyu_t, B, alpha * dt,
                                   1.0 / dx, 1.0 / dz, g);
            plain_copy(pyb[0], u[c]);   // v staged
            plain_copy(pyb[2], pyb[0]); // y-local layout
            hipLaunchKernelGGL((k_plain_imexrhs<T>), grd, blk, 0, st, pyb[2], pyb[1], ca, cb, cc, theta, g);
            hipLaunchKernelGGL((k_plain_imexsolve<T>), dim3(ceil_div(g.nx, 64), g.nz), dim3(64), 0, st, pyb[0], pyb[2],
                               Dcn, ca, cp, ip, theta, g);
            plain_copy(pyb[1], pyb[0]); // back to the x-local layout
            plain_copy(u[c], pyb[1]);
        }
        timer.end(FAM_CN);
        DNS_HIP_CHECK(hipGetLastError());
        DNS_HIP_CHECK(hipStreamSynchronize(st)); // imex_update! synchronises (src/ydiffusion.jl:153)
        return DNS_OK;
    }

    int step_plain(double dt)
    {
        DNS_TRY(plain_alloc());
        for (int s = 0; s < 3; ++s) {
            DNS_TRY(momentum_plain(RK_A[s], dt));
            if (!imexy) {
                timer.begin(FAM_MOM);
                for (int c = 0; c < 3; ++c)
                    hipLaunchKernelGGL((k_plain_axpy<T>), dim3(2048), dim3(NTHR), 0, st, u[c], q[c], (T)RK_B[s], (long)N);
                timer.end(FAM_MOM);
                DNS_TRY(plain_halo(u, 3));
                DNS_TRY(plain_project(p));
            } else {
                DNS_TRY(plain_imex(RK_B[s], RK_ALPHA[s], dt));
                DNS_TRY(plain_halo(u, 3));
                DNS_TRY(plain_project(delta));
                hipLaunchKernelGGL((k_plain_axpy<T>), dim3(2048), dim3(NTHR), 0, st, p, delta,
                                   (T)(1.0 / (RK_ALPHA[s] * dt)), (long)N);
                DNS_TRY(plain_halo(&p, 1));
            }
        }
        umax_valid = false;
        DNS_HIP_CHECK(hipGetLastError());
        return DNS_OK;
    }

    // z planes per block of the LDS z-march momentum kernel
    int momentum_zc() const
    {
        const int nbx = ceil_div(g.nx, 64), nby = ceil_div(g.ny, 8);
        const long round = 256L * (sizeof(T) == 8 ? 3 : 6);
        int zc = 32;
        while (zc > 2 && (long)nbx * nby * ceil_div(g.nz, zc) < round) zc /= 2;
        if (const char *ez = getenv("DNS_MOM_ZC")) zc = std::max(1, atoi(ez)); // sweeps
        return zc;
    }

    // the kernel instances and launch shapes this handle's step uses
    int describe(char *buf, int len) override
    {
        const char *mom = momentum_kind == 5 ? "lz" : momentum_kind == 3 ? "slab" : "direct";
        char xdesc[32];
        if (xwave) snprintf(xdesc, sizeof xdesc, "wave groups of %d", g.nx / 8);
        else snprintf(xdesc, sizeof xdesc, "lpb=%d", lpb);
        char ys[64], cs[96];
        if (!g.wall) snprintf(ys, sizeof ys, "fourier n=%d tile=%d/%d thr=%d", g.ny, tby, tbyf, nthry);
        else if (thomas_tl && thomas_seg) snprintf(ys, sizeof ys, "seg R=%d tl=%d thr=%d", thomas_seg, thomas_tl, thomas_nthr);
        else if (thomas_tl) snprintf(ys, sizeof ys, "tile tl=%d thr=%d", thomas_tl, thomas_nthr);
        else snprintf(ys, sizeof ys, "sequential");
        if (!imexy) snprintf(cs, sizeof cs, "none");
        else if (cn_fused) snprintf(cs, sizeof cs, "seg3 R=%d tl=%d thr=%d p=%s", cn_seg, cn_tl, cn_nthr, cn_fused == 2 ? "lds" : "l2");
        else if (cn_tl && cn_seg) snprintf(cs, sizeof cs, "seg R=%d tl=%d thr=%d nl=%d", cn_seg, cn_tl, cn_nthr, cn_nl);
        else if (cn_tl) snprintf(cs, sizeof cs, "tile tl=%d thr=%d", cn_tl, cn_nthr);
        else snprintf(cs, sizeof cs, "sequential");
        const int n = snprintf(buf, (size_t)len,
                               "prec=%s; momentum=%s zc=%d; xpass=half %d %s; zpass=n %d tile=%d/%d thr=%d; "
                               "ysolve=%s; cn=%s; pgrad=%s",
                               plain ? "plain" : sizeof(T) == 8 ? "fp64" : "mixed", mom, momentum_zc(), g.nx / 2, xdesc,
                               g.nz, tbz, tbzf, nthrz, ys, cs, pgrad_vec ? "vec" : "scalar");
        if (n < 0 || n >= len) {
            set_error("describe: buffer of %d bytes too small", len);
            return DNS_EINVAL;
        }
        return DNS_OK;
    }

    // momentum_kernel! as one launch writing q only (src/operators.jl:44-178)
    int momentum_plain(double Ac, double dt)
    {
        const int keep = momentum_kind;
        momentum_kind = 1;
        const int rc = launch_momentum(u, nullptr, Ac, dt, 0.0, false);
        momentum_kind = keep;
        return rc;
    }

    int step(double dt) override
    {
        if (!std::isfinite(dt)) {
            set_error("step!: Δt must be finite");
            return DNS_EINVAL;
        }
        if (imexy && !(dt > 0)) {
            // the lagged-pressure update p += δ/(αΔt) (src/timestep.jl:51) divides by Δt
            set_error("step!: Δt must be positive with implicit y-diffusion");
            return DNS_EINVAL;
        }
        if (plain) return step_plain(dt);
        const bool replay = graphs && !timer.on && dt == last_dt;
        last_dt = dt;
        if (!replay) return step_launch(dt);
        const int key = u[0] == ubuf0 ? 0 : 1;
        if (gexec[key] && gdt[key] == dt) {
            DNS_HIP_CHECK(hipGraphLaunch(gexec[key], st));
            // the host side of a step: the explicit form's three u / u′ swaps
            // (one net swap) and the CFL partials left by the last pgrad
            if (!imexy)
                for (int c = 0; c < 3; ++c) std::swap(u[c], v[c]);
            umax_valid = true;
            return DNS_OK;
        }
        if (gexec[key]) {
            DNS_HIP_CHECK(hipGraphExecDestroy(gexec[key]));
            gexec[key] = nullptr;
        }
        // capture the launches of one step (its host-side state changes happen
        // now), then run the graph once for this step
        hipGraph_t graph = nullptr;
        DNS_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = step_launch(dt);
        const hipError_t ec = hipStreamEndCapture(st, &graph);
        if (rc != DNS_OK) {
            if (graph) hipGraphDestroy(graph);
            return rc;
        }
        DNS_HIP_CHECK(ec);
        const hipError_t ei = hipGraphInstantiate(&gexec[key], graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        DNS_HIP_CHECK(ei);
        gdt[key] = dt;
        DNS_HIP_CHECK(hipGraphLaunch(gexec[key], st));
        return DNS_OK;
    }

    int step_launch(double dt)
    {
        if (imexy) launch_cn_coeffs(dt);
        for (int s = 0; s < 3; ++s) {
            if (!imexy) {
                DNS_TRY(launch_momentum(u, v, RK_A[s], dt, RK_B[s], true));
                DNS_TRY(launch_poisson(v, p, 0));
                DNS_TRY(launch_pgrad(v, p, false, 0.0));
                for (int c = 0; c < 3; ++c) std::swap(u[c], v[c]);
            } else {
                DNS_TRY(launch_momentum(u, nullptr, RK_A[s], dt, 0.0, false));
                DNS_TRY(launch_cn(s, dt));
                DNS_TRY(launch_poisson(u, delta, 0));
                DNS_TRY(launch_pgrad(u, delta, true, 1.0 / (RK_ALPHA[s] * dt)));
            }
        }
        return DNS_OK;
    }

    int read_back(size_t count)
    {
        DNS_HIP_CHECK(hipMemcpyAsync(pinned, redout, sizeof(double) * count, hipMemcpyDeviceToHost, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        return DNS_OK;
    }

    // Up to nsteps CFL-adaptive steps of solve! (src/timestep.jl:92-111): Δt =
    // min(cfl_timestep, tend − t) each step, stopping once t >= tend −
    // 1e-10(|tend| + 1) as the reference loop does.  The explicit form
    // computes Δt on the device from the pressure-gradient partials and
    // replays one captured graph per buffer parity, so no step waits for the
    // host (steps after the end time leave u unchanged); other forms (or
    // per-launch timing) run the host loop.  t is advanced; dts (optional)
    // receives the Δt of every step taken, *ndone their count.
    int advance_cfl(double cflv, int nsteps, double tend, double *t_io, double *dts, int *ndone) override
    {
        *ndone = 0;
        if (nsteps <= 0) return DNS_OK;
        double t = *t_io;
        int done = 0;
        const double tstop = std::isinf(tend) ? tend : tend - 1e-10 * (std::fabs(tend) + 1); // +inf: no end time
        auto host_step = [&]() -> int {
            double d = 0;
            DNS_TRY(cfl(cflv, &d));
            d = std::min(d, tend - t);
            DNS_TRY(step(d));
            t += d;
            if (dts) dts[done] = d;
            ++done;
            return DNS_OK;
        };
        if (plain || imexy || timer.on || !graphs) {
            while (done < nsteps && t < tstop) DNS_TRY(host_step());
            *t_io = t;
            *ndone = done;
            return DNS_OK;
        }
        if (!umax_valid && t < tstop) DNS_TRY(host_step()); // the device Δt needs a pressure pass's partials
        if (done == nsteps || !(t < tstop)) {
            *t_io = t;
            *ndone = done;
            return DNS_OK;
        }
        if (!ddt) {
            DNS_TRY(dalloc(&ddt, 1));
            DNS_TRY(dalloc(&dtime, 1));
            DNS_TRY(dalloc(&dstep, 1));
        }
        if (nlog < nsteps) {
            if (dlog) DNS_HIP_CHECK(hipFree(dlog));
            dlog = nullptr;
            DNS_TRY(dalloc(&dlog, (size_t)nsteps));
            nlog = nsteps;
            for (auto &ge : gcfl)
                if (ge) {
                    DNS_HIP_CHECK(hipGraphExecDestroy(ge));
                    ge = nullptr;
                }
        }
        const int step0 = done;
        DNS_HIP_CHECK(hipMemcpyAsync(dtime, &t, sizeof(double), hipMemcpyHostToDevice, st));
        DNS_HIP_CHECK(hipMemcpyAsync(dstep, &step0, sizeof(int), hipMemcpyHostToDevice, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st)); // the host sources are locals
        const double eps = 2.220446049250313e-16;
        const double invdy2 = 1 / (dymin * dymin);
        const double diff = 1 / (2 * visc * (1 / (dx * dx) + invdy2 + 1 / (dz * dz)) + eps);
        const int G = std::min(512, std::max(1, nb_pgrad / 1024));
        for (; done < nsteps; ++done) {
            const int key = u[0] == ubuf0 ? 0 : 1;
            if (!gcfl[key] || gcfl_key[key][0] != cflv || gcfl_key[key][1] != tend) {
                if (gcfl[key]) {
                    DNS_HIP_CHECK(hipGraphExecDestroy(gcfl[key]));
                    gcfl[key] = nullptr;
                }
                hipGraph_t graph = nullptr;
                DNS_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                hipLaunchKernelGGL(k_max_mid, dim3(G), dim3(NTHR), 0, st, upart, nb_pgrad, 3, red);
                hipLaunchKernelGGL(k_max_final, dim3(1), dim3(NTHR), 0, st, red, G, 3, redout);
                hipLaunchKernelGGL(k_cfl_dt, dim3(1), dim3(1), 0, st, redout, dx, dymin, dz, diff, cflv, tend, dtime,
                                   ddt, dlog, dstep);
                mom_dtp = ddt;
                const int rc = step_launch(0.0);
                mom_dtp = nullptr;
                const hipError_t ec = hipStreamEndCapture(st, &graph);
                if (rc != DNS_OK) {
                    if (graph) hipGraphDestroy(graph);
                    return rc;
                }
                DNS_HIP_CHECK(ec);
                const hipError_t ei = hipGraphInstantiate(&gcfl[key], graph, nullptr, nullptr, 0);
                hipGraphDestroy(graph);
                DNS_HIP_CHECK(ei);
                gcfl_key[key][0] = cflv;
                gcfl_key[key][1] = tend;
                // the capture already advanced the host state (buffer swaps)
                DNS_HIP_CHECK(hipGraphLaunch(gcfl[key], st));
            } else {
                DNS_HIP_CHECK(hipGraphLaunch(gcfl[key], st));
                for (int c = 0; c < 3; ++c) std::swap(u[c], v[c]);
            }
            umax_valid = true;
        }
        last_dt = -1.0; // the fixed-Δt graphs stay valid, but the next step launches directly
        DNS_HIP_CHECK(hipMemcpyAsync(pinned, dtime, sizeof(double), hipMemcpyDeviceToHost, st));
        DNS_HIP_CHECK(hipMemcpyAsync(pinned + 1, dstep, sizeof(int), hipMemcpyDeviceToHost, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        int taken = 0;
        std::memcpy(&taken, pinned + 1, sizeof(int));
        if (dts && taken > step0) {
            DNS_HIP_CHECK(hipMemcpyAsync(dts + step0, dlog + step0, sizeof(double) * (taken - step0),
                                         hipMemcpyDeviceToHost, st));
            DNS_HIP_CHECK(hipStreamSynchronize(st));
        }
        *t_io = pinned[0];
        *ndone = taken;
        return DNS_OK;
    }

    // advance_cfl without any allocation: the device Δt, time, step counter,
    // Δt log and CFL partials live in the file-scope scratch, the steps are
    // direct launches (no graph is captured or instantiated) and the host
    // copies go to the pinned buffer of the handle or to the caller's arrays.
    // Same kernels and arithmetic as advance_cfl, hence the same results.
    int advance_cfl_noalloc(double cflv, int nsteps, double tend, double *t_io, double *dts, int *ndone) override
    {
        *ndone = 0;
        if (nsteps < 0 || nsteps > NOALLOC_MAX_STEPS) {
            set_error("advance_cfl_noalloc: nsteps must be in 0..%d (the size of its Δt log)", NOALLOC_MAX_STEPS);
            return DNS_EINVAL;
        }
        if (plain) {
            set_error("advance_cfl_noalloc: the plain form allocates its reference-structured arrays on first use");
            return DNS_EUNSUPPORTED;
        }
        if (timer.on) {
            set_error("advance_cfl_noalloc: per-launch timing creates events; disable it first");
            return DNS_EUNSUPPORTED;
        }
        if (nsteps == 0) return DNS_OK;
        double t = *t_io;
        int done = 0;
        const double tstop = std::isinf(tend) ? tend : tend - 1e-10 * (std::fabs(tend) + 1);
        auto host_step = [&]() -> int {
            double d = 0;
            DNS_TRY(cfl(cflv, &d));
            d = std::min(d, tend - t);
            if (imexy && !(d > 0)) {
                set_error("step!: Δt must be positive with implicit y-diffusion");
                return DNS_EINVAL;
            }
            DNS_TRY(step_launch(d));
            last_dt = -1.0;
            t += d;
            if (dts) dts[done] = d;
            ++done;
            return DNS_OK;
        };
        if (imexy) {
            while (done < nsteps && t < tstop) DNS_TRY(host_step());
            *t_io = t;
            *ndone = done;
            return DNS_OK;
        }
        if (!umax_valid && t < tstop) DNS_TRY(host_step()); // the device Δt needs a pressure pass's partials
        if (done == nsteps || !(t < tstop)) {
            *t_io = t;
            *ndone = done;
            return DNS_OK;
        }
        double *sdt = noalloc_s, *stime = noalloc_s + 1, *spart = noalloc_s + NOALLOC_PART,
               *sred = noalloc_s + NOALLOC_RED, *slog = noalloc_s + NOALLOC_LOG;
        int *sstep = reinterpret_cast<int *>(noalloc_s + 2);
        const int step0 = done;
        hipLaunchKernelGGL(k_noalloc_start, dim3(1), dim3(1), 0, st, noalloc_s, t);
        const double eps = 2.220446049250313e-16;
        const double diff = 1 / (2 * visc * (1 / (dx * dx) + 1 / (dymin * dymin) + 1 / (dz * dz)) + eps);
        const int G = std::min(512, std::max(1, nb_pgrad / 1024));
        for (; done < nsteps; ++done) {
            hipLaunchKernelGGL(k_max_mid, dim3(G), dim3(NTHR), 0, st, upart, nb_pgrad, 3, spart);
            hipLaunchKernelGGL(k_max_final, dim3(1), dim3(NTHR), 0, st, spart, G, 3, sred);
            hipLaunchKernelGGL(k_cfl_dt, dim3(1), dim3(1), 0, st, sred, dx, dymin, dz, diff, cflv, tend, stime, sdt,
                               slog, sstep);
            mom_dtp = sdt;
            const int rc = step_launch(0.0);
            mom_dtp = nullptr;
            DNS_TRY(rc);
            umax_valid = true;
        }
        last_dt = -1.0;
        DNS_HIP_CHECK(hipGetLastError());
        DNS_HIP_CHECK(hipMemcpyAsync(pinned, stime, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        int taken = 0;
        std::memcpy(&taken, pinned + 1, sizeof(int));
        if (dts && taken > 0) {
            DNS_HIP_CHECK(hipMemcpyAsync(dts + step0, slog, sizeof(double) * taken, hipMemcpyDeviceToHost, st));
            DNS_HIP_CHECK(hipStreamSynchronize(st));
        }
        *t_io = pinned[0];
        *ndone = step0 + taken;
        return DNS_OK;
    }

    int cfl(double cflv, double *dt) override
    {
        if (umax_valid) {
            // partial maxima left by the pressure-gradient pass that produced u,
            // reduced in two levels (up to 512 blocks, then one)
            const int G = std::min(512, std::max(1, nb_pgrad / 1024));
            hipLaunchKernelGGL(k_max_mid, dim3(G), dim3(NTHR), 0, st, upart, nb_pgrad, 3, red);
            hipLaunchKernelGGL(k_max_final, dim3(1), dim3(NTHR), 0, st, red, G, 3, redout);
        } else {
            const int nb = std::min(RED_BLOCKS, ceil_div(N, NTHR));
            hipLaunchKernelGGL((k_maxabs3<T>), dim3(nb), dim3(NTHR), 0, st, u[0], u[1], u[2], N, red);
            hipLaunchKernelGGL(k_max_final, dim3(1), dim3(NTHR), 0, st, red, nb, 3, redout);
        }
        DNS_HIP_CHECK(hipGetLastError());
        DNS_TRY(read_back(3));
        const double eps = 2.220446049250313e-16;
        double conv = std::min({dx / (pinned[0] + eps), dymin / (pinned[1] + eps), dz / (pinned[2] + eps)});
        double invdy2 = imexy ? 0.0 : 1 / (dymin * dymin);
        double diff = 1 / (2 * visc * (1 / (dx * dx) + invdy2 + 1 / (dz * dz)) + eps);
        *dt = std::min(cflv * conv, cflv * diff);
        return DNS_OK;
    }

    int maxdiv(double *out) override
    {
        const int nb = std::min(RED_BLOCKS, ceil_div(N, NTHR));
        hipLaunchKernelGGL((k_maxdiv<T>), dim3(nb), dim3(NTHR), 0, st, u[0], u[1], u[2], rdy_d, 1.0 / dx, 1.0 / dz, g,
                           red);
        hipLaunchKernelGGL(k_max_final, dim3(1), dim3(NTHR), 0, st, red, nb, 1, redout);
        DNS_HIP_CHECK(hipGetLastError());
        DNS_TRY(read_back(1));
        *out = pinned[0];
        return DNS_OK;
    }

    int kinetic(double *out) override
    {
        const int nb = std::min(RED_BLOCKS, ceil_div(N, NTHR));
        hipLaunchKernelGGL((k_kinetic<T>), dim3(nb), dim3(NTHR), 0, st, u[0], u[1], u[2], dy_d, dyu_d, dx * dz, g,
                           reinterpret_cast<dd *>(red));
        hipLaunchKernelGGL(k_dd_final, dim3(1), dim3(NTHR), 0, st, reinterpret_cast<dd *>(red), nb, 1, redout);
        DNS_HIP_CHECK(hipGetLastError());
        DNS_TRY(read_back(1));
        *out = pinned[0];
        return DNS_OK;
    }

    int chstats(double *acc) override
    {
        const int ny = g.ny, nf = g.wall ? ny - 1 : ny;
        hipLaunchKernelGGL((k_chstats<T>), dim3(ny, CH_NCHUNK), dim3(NTHR), 0, st, u[0], u[1], u[2], g, nf, CH_NCHUNK,
                           reinterpret_cast<dd *>(red));
        hipLaunchKernelGGL(k_dd_final, dim3(std::min(8 * ny, 1024)), dim3(NTHR), 0, st, reinterpret_cast<dd *>(red),
                           CH_NCHUNK, 8 * ny, redout);
        DNS_HIP_CHECK(hipGetLastError());
        DNS_TRY(read_back((size_t)8 * ny));
        for (int s = 0; s < 4; ++s)
            for (int j = 0; j < ny; ++j) acc[s * ny + j] += pinned[s * ny + j];
        for (int s = 0; s < 4; ++s)
            for (int j = 0; j < nf; ++j) acc[4 * ny + s * nf + j] += pinned[(4 + s) * ny + j];
        return DNS_OK;
    }

    int spectrum(double *E, int nbins) override
    {
        if (g.wall) {
            set_error("spectra require a fully periodic box");
            return DNS_EINVAL;
        }
        if (nbins <= 0 || nbins > 16384) {
            set_error("nbins out of range");
            return DNS_EINVAL;
        }
        const int NC = g.nxc * g.ny * g.nz;
        // per-wavefront binning (nbins <= 2048) scales to more blocks than the
        // chunk-owner kernel
        const int nb = std::min(nbins <= 2048 ? 2048 : 256, ceil_div(NC, NTHR));
        const size_t need = (size_t)nb * nbins;
        // grow-only buffers of their own: `red` is captured by the CFL graphs
        if (need > spartcap || (size_t)nbins > soutcap) {
            DNS_HIP_CHECK(hipStreamSynchronize(st));
            if (need > spartcap) {
                if (spart) DNS_HIP_CHECK(hipFree(spart));
                spart = nullptr;
                spartcap = 0;
                DNS_TRY(dalloc(&spart, need));
                spartcap = need;
            }
            if ((size_t)nbins > soutcap) {
                if (sout) DNS_HIP_CHECK(hipFree(sout));
                sout = nullptr;
                soutcap = 0;
                DNS_TRY(dalloc(&sout, (size_t)nbins));
                soutcap = nbins;
            }
        }
        std::vector<double> hostE(nbins);
        const double N2 = (double)N * (double)N;
        for (int c = 0; c < 3; ++c) {
            XPassArgs<T> xa;
            xa.ux = u[c];
            xa.uy = u[c];
            xa.uz = u[c];
            xa.A = A;
            xa.rdy = rdy_t;
            xa.rdx = T(0);
            xa.rdz = T(0);
            xa.tw = twN;
            xa.twx = twx;
            xa.lpb = lpb;
        xa.fN = make_fastdiv(g.nx / 2);
        xa.fnxc = make_fastdiv(g.nxc);
        xa.fny = make_fastdiv(g.ny);
            hipLaunchKernelGGL((k_rfft_x<T, X_EPT, 1>), dim3(ceil_div(g.ny * g.nz, lpb)), dim3(NTHR), ldsx, st, xa, g,
                               fdx);
            launch_fft_z<false>(A);
            hipLaunchKernelGGL((k_fourier_y<T, ZY_EPT, false>), dim3(ceil_div(g.nxc, tby), g.nz), dim3(nthry), ldsy,
                               st, A, twy, lam1, lam2, lam3, 1.0, tby, tbyf, make_fastdiv(tby), g, fdy);
            if (nbins <= 2048)
                hipLaunchKernelGGL((k_binshells_w<T>), dim3(nb), dim3(NTHR), sizeof(double) * 4 * nbins, st, A, g,
                                   nbins, 1.0 / N2, spart);
            else
                hipLaunchKernelGGL((k_binshells<T>), dim3(nb), dim3(NTHR),
                                   sizeof(double) * nbins + (sizeof(double) + sizeof(int)) * NTHR, st, A, g, nbins,
                                   1.0 / N2, spart);
            hipLaunchKernelGGL(k_bins_final, dim3(ceil_div(nbins, NTHR)), dim3(NTHR), 0, st, spart, nb, nbins, sout);
            DNS_HIP_CHECK(hipGetLastError());
            DNS_HIP_CHECK(hipMemcpyAsync(hostE.data(), sout, sizeof(double) * nbins, hipMemcpyDeviceToHost, st));
            DNS_HIP_CHECK(hipStreamSynchronize(st));
            for (int b = 0; b < nbins; ++b) E[b] += hostE[b];
        }
        return DNS_OK;
    }

    int sync() override
    {
        DNS_HIP_CHECK(hipStreamSynchronize(st));
        return DNS_OK;
    }
};

}  // namespace dns

struct dns_fv {
    std::unique_ptr<dns::FVBase> impl;
};

using namespace dns;

extern "C" {

const char *dns_last_error(void) { return dns::last_error(); }

const char *dns_version(void) { return "libdnsrocm 0.1 (gfx950)"; }

int dns_fv_create(const dns_fv_params *P, dns_fv **out)
{
    if (!out) {
        set_error("null output handle");
        return DNS_EINVAL;
    }
    *out = nullptr;
    if (!P || !P->yfaces) {
        set_error("null params or yfaces");
        return DNS_EINVAL;
    }
    if (P->n[0] < 2 || P->n[1] < 2 || P->n[2] < 2) {
        set_error("local block smaller than ghost width; use fewer ranks");
        return DNS_EINVAL;
    }
    // 32-bit element indices; the complex pencil's x extent is padded to whole
    // 128-byte lines (at most 16 elements)
    if ((long long)P->n[0] * P->n[1] * P->n[2] >= (1LL << 31) ||
        (long long)((P->n[0] / 2 + 1 + 15) / 16 * 16) * P->n[1] * P->n[2] >= (1LL << 31)) {
        set_error("grid too large for 32-bit indexing");
        return DNS_EUNSUPPORTED;
    }
    if (P->ydiffusion && !P->ywall) {
        set_error("implicit y-diffusion requires bc[2] == :wall");
        return DNS_EINVAL;
    }
    if (P->precision != DNS_FP64 && P->precision != DNS_MIXED && P->precision != DNS_PLAIN_FP64) {
        set_error("precision must be DNS_FP64, DNS_MIXED or DNS_PLAIN_FP64");
        return DNS_EINVAL;
    }
    std::unique_ptr<dns_fv> h(new dns_fv);
    int rc;
    if (P->precision == DNS_FP64 || P->precision == DNS_PLAIN_FP64) {
        auto *impl = new FVImpl<double>();
        h->impl.reset(impl);
        impl->plain = P->precision == DNS_PLAIN_FP64;
        rc = impl->create(P);
    } else {
        auto *impl = new FVImpl<float>();
        h->impl.reset(impl);
        rc = impl->create(P);
    }
    if (rc != DNS_OK) return rc;
    *out = h.release();
    return DNS_OK;
}

int dns_fv_destroy(dns_fv *h)
{
    delete h;
    return DNS_OK;
}

#define FV_CHECK(h)                         \
    do {                                    \
        if (!(h) || !(h)->impl) {           \
            set_error("null dns_fv handle"); \
            return DNS_EINVAL;              \
        }                                   \
    } while (0)

int dns_fv_set_field(dns_fv *h, int32_t field, const double *host, int32_t layout)
{
    FV_CHECK(h);
    return h->impl->set_field(field, host, layout);
}
int dns_fv_get_field(dns_fv *h, int32_t field, double *host, int32_t layout)
{
    FV_CHECK(h);
    return h->impl->get_field(field, host, layout);
}
int dns_fv_momentum(dns_fv *h, double A, double dt)
{
    FV_CHECK(h);
    return h->impl->momentum(A, dt);
}
int dns_fv_project(dns_fv *h, int32_t psi)
{
    FV_CHECK(h);
    if (psi != DNS_FIELD_P && psi != DNS_FIELD_DELTA) {
        set_error("psi must be DNS_FIELD_P or DNS_FIELD_DELTA");
        return DNS_EINVAL;
    }
    return h->impl->project(psi);
}
int dns_fv_step(dns_fv *h, double dt)
{
    FV_CHECK(h);
    return h->impl->step(dt);
}
int dns_fv_cfl_timestep(dns_fv *h, double cfl, double *dt)
{
    FV_CHECK(h);
    if (!dt) return DNS_EINVAL;
    return h->impl->cfl(cfl, dt);
}
int dns_fv_advance_cfl(dns_fv *h, double cfl, int32_t nsteps, double tend, double *t, double *dts, int32_t *ndone)
{
    FV_CHECK(h);
    if (!t || nsteps < 0) return DNS_EINVAL;
    int taken = 0;
    const int rc = h->impl->advance_cfl(cfl, nsteps, tend, t, dts, &taken);
    if (ndone) *ndone = taken;
    return rc;
}
int dns_fv_advance_cfl_noalloc(dns_fv *h, double cfl, int32_t nsteps, double tend, double *t, double *dts,
                               int32_t *ndone)
{
    FV_CHECK(h);
    if (!t) {
        set_error("null t");
        return DNS_EINVAL;
    }
    int taken = 0;
    const int rc = h->impl->advance_cfl_noalloc(cfl, nsteps, tend, t, dts, &taken);
    if (ndone) *ndone = taken;
    return rc;
}
int dns_fv_maxdiv(dns_fv *h, double *out)
{
    FV_CHECK(h);
    if (!out) return DNS_EINVAL;
    return h->impl->maxdiv(out);
}
int dns_fv_kinetic_energy(dns_fv *h, double *out)
{
    FV_CHECK(h);
    if (!out) return DNS_EINVAL;
    return h->impl->kinetic(out);
}
int dns_fv_channelstats_sample(dns_fv *h, double *acc)
{
    FV_CHECK(h);
    if (!acc) return DNS_EINVAL;
    return h->impl->chstats(acc);
}
int dns_fv_spectrum_sample(dns_fv *h, double *E, int32_t nbins)
{
    FV_CHECK(h);
    if (!E) return DNS_EINVAL;
    return h->impl->spectrum(E, nbins);
}
int dns_fv_sync(dns_fv *h)
{
    FV_CHECK(h);
    return h->impl->sync();
}
int dns_fv_describe(dns_fv *h, char *buf, int32_t len)
{
    FV_CHECK(h);
    if (!buf || len <= 0) return DNS_EINVAL;
    return h->impl->describe(buf, len);
}
int dns_fv_timing(dns_fv *h, int32_t enable)
{
    FV_CHECK(h);
    h->impl->timer.flush();
    for (int f = 0; f < 8; ++f) {
        h->impl->timer.ms[f] = 0;
        h->impl->timer.cnt[f] = 0;
    }
    h->impl->timer.on = enable != 0;
    return DNS_OK;
}
int dns_fv_timing_read(dns_fv *h, int32_t family, double *total_ms, int64_t *launches)
{
    FV_CHECK(h);
    if (family < 0 || family >= 8 || !total_ms || !launches) return DNS_EINVAL;
    h->impl->timer.flush();
    *total_ms = h->impl->timer.ms[family];
    *launches = h->impl->timer.cnt[family];
    return DNS_OK;
}

}  // extern "C"
