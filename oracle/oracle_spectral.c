/*
 * FP64 CPU oracle for the pseudo-spectral HIT path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load this file's library; the HIP product never links or calls it.
 *
 * Restates (reference file:line):
 *   compactfreq           src/spectral/operators.jl:7
 *   pad_y / truncate_y    src/spectral/fft.jl:181-193
 *   spec_to_phys!         src/spectral/fft.jl:200-214
 *   phys_to_spec!         src/spectral/fft.jl:221-232
 *   project_kernel!       src/spectral/operators.jl:10-29
 *   spectral_rhs!         src/spectral/operators.jl:35-79
 *   wray3, ifstage        src/spectral/step.jl:16-45
 *   vsummax               src/spectral/step.jl:54-60
 *   spectral_step!        src/spectral/step.jl:68-92
 *   shell_energies        src/spectral/forcing.jl:12-24
 *   rescale_shell!        src/spectral/forcing.jl:27-35
 *   splitmix64, noise     src/spectral/ic.jl:31-48
 *   spectral_energy etc.  src/spectral/operators.jl:87-110
 *
 * Layouts (column-major, first index fastest):
 *   state uh  (K, m, m, 3) complex, K = kcut+1, m = 2kcut+1, compact
 *             frequency order j <= kcut ? j : j - m (0-based);
 *   physical  (n, n, n, 3) real, cell-centred transform grid.
 * The transforms follow the reference's stage order (x, then z, then y
 * forward; the reverse backward), dropping killed modes between stages;
 * FFTs come from oracle_fft.c instead of FFTW.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_fft.h"

uint64_t ospec_splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

typedef struct ospec {
    int n, kcut, K, m;
    double l, visc, kfac;
    double complex *ax; /* (K, n, n) x-stage buffer, one component */
    double complex *ay; /* (K, n, m) */
    double *v, *rbuf;   /* (n, n, n, 3) */
    double complex *sig; /* (K, m, m, 3) */
} ospec;

static long SIDX(const ospec *h, int i, int j, int k, int c)
{
    return (long)i + (long)h->K * ((long)j + (long)h->m * ((long)k + (long)h->m * c));
}

int ospec_compactfreq(int j0, int m, int kcut) { return j0 <= kcut ? j0 : j0 - m; }

ospec *ospec_create(int n, int kcut, double l, double visc)
{
    ospec *h = calloc(1, sizeof(ospec));
    h->n = n;
    h->kcut = kcut;
    h->K = kcut + 1;
    h->m = 2 * kcut + 1;
    h->l = l;
    h->visc = visc;
    h->kfac = 6.283185307179586 / l;
    long n3 = (long)n * n * n;
    h->ax = calloc((size_t)h->K * n * n, sizeof(double complex));
    h->ay = calloc((size_t)h->K * n * h->m, sizeof(double complex));
    h->v = calloc((size_t)n3 * 3, sizeof(double));
    h->rbuf = calloc((size_t)n3 * 3, sizeof(double));
    h->sig = calloc((size_t)h->K * h->m * h->m * 3, sizeof(double complex));
    return h;
}

void ospec_destroy(ospec *h)
{
    if (!h) return;
    free(h->ax);
    free(h->ay);
    free(h->v);
    free(h->rbuf);
    free(h->sig);
    free(h);
}

double *ospec_vbuf(ospec *h) { return h->v; }
double *ospec_rbuf(ospec *h) { return h->rbuf; }

/* spec_to_phys! for components c0..c0+nc-1 of uh into v (physical). */
void ospec_spec_to_phys(ospec *h, const double complex *uh, double *v)
{
    const int n = h->n, K = h->K, m = h->m, kc = h->kcut;
    for (int c = 0; c < 3; ++c) {
        /* pad y + bfft y: ay(i, y, kz) */
#pragma omp parallel for
        for (int k = 0; k < m; ++k)
            for (int i = 0; i < K; ++i) {
                double complex *line = h->ay + i + (long)K * n * k;
                for (int jf = 0; jf < n; ++jf) {
                    double complex val = 0;
                    if (!(kc < jf && jf < n - kc)) {
                        int jc = jf <= kc ? jf : jf - (n - m);
                        val = uh[SIDX(h, i, jc, k, c)];
                    }
                    line[(long)K * jf] = val;
                }
                offt_strided(line, n, K, +1);
            }
        /* expand kz to fine positions (band zeroed) + bfft z: ax(i, y, z) */
#pragma omp parallel for
        for (int y = 0; y < n; ++y)
            for (int i = 0; i < K; ++i) {
                double complex *line = h->ax + i + (long)K * y;
                long zs = (long)K * n;
                for (int kf = 0; kf < n; ++kf) {
                    double complex val = 0;
                    if (!(kc < kf && kf < n - kc)) {
                        int kcmp = kf <= kc ? kf : kf - (n - m);
                        val = h->ay[i + (long)K * (y + (long)n * kcmp)];
                    }
                    line[zs * kf] = val;
                }
                offt_strided(line, n, zs, +1);
            }
        /* rows kx > kcut are zero; brfft x */
#pragma omp parallel for
        for (int z = 0; z < n; ++z)
            for (int y = 0; y < n; ++y) {
                double complex *X = malloc(sizeof(double complex) * (n / 2 + 1));
                for (int i = 0; i <= n / 2; ++i) X[i] = i < K ? h->ax[i + (long)K * (y + (long)n * z)] : 0;
                offt_c2r(X, v + (long)n * (y + (long)n * (z + (long)n * c)), n);
                free(X);
            }
    }
}

/* phys_to_spec! of 3 physical components w into uh (× 1/n³). */
void ospec_phys_to_spec(ospec *h, const double *w, double complex *uh)
{
    const int n = h->n, K = h->K, m = h->m, kc = h->kcut;
    const double invn3 = 1.0 / ((double)n * n * n);
    for (int c = 0; c < 3; ++c) {
#pragma omp parallel for
        for (int z = 0; z < n; ++z)
            for (int y = 0; y < n; ++y) {
                double complex *X = malloc(sizeof(double complex) * (n / 2 + 1));
                offt_r2c(w + (long)n * (y + (long)n * (z + (long)n * c)), X, n);
                for (int i = 0; i < K; ++i) h->ax[i + (long)K * (y + (long)n * z)] = X[i];
                free(X);
            }
#pragma omp parallel for
        for (int y = 0; y < n; ++y)
            for (int i = 0; i < K; ++i) {
                long zs = (long)K * n;
                double complex *line = h->ax + i + (long)K * y;
                offt_strided(line, n, zs, -1);
                for (int kcmp = 0; kcmp < m; ++kcmp) {
                    int kf = kcmp <= kc ? kcmp : kcmp + (n - m);
                    h->ay[i + (long)K * (y + (long)n * kcmp)] = line[zs * kf];
                }
            }
#pragma omp parallel for
        for (int k = 0; k < m; ++k)
            for (int i = 0; i < K; ++i) {
                double complex *line = h->ay + i + (long)K * n * k;
                offt_strided(line, n, K, -1);
                for (int jc = 0; jc < m; ++jc) {
                    int jf = jc <= kc ? jc : jc + (n - m);
                    uh[SIDX(h, i, jc, k, c)] = line[(long)K * jf] * invn3;
                }
            }
    }
}

static void kvec(const ospec *h, int i, int j, int k, double *k1, double *k2, double *k3)
{
    *k1 = h->kfac * i;
    *k2 = h->kfac * ospec_compactfreq(j, h->m, h->kcut);
    *k3 = h->kfac * ospec_compactfreq(k, h->m, h->kcut);
}

void ospec_project(const ospec *h, double complex *uh)
{
#pragma omp parallel for
    for (int k = 0; k < h->m; ++k)
        for (int j = 0; j < h->m; ++j)
            for (int i = 0; i < h->K; ++i) {
                double k1, k2, k3;
                kvec(h, i, j, k, &k1, &k2, &k3);
                double kk = k1 * k1 + k2 * k2 + k3 * k3;
                double complex ux = uh[SIDX(h, i, j, k, 0)], uy = uh[SIDX(h, i, j, k, 1)], uz = uh[SIDX(h, i, j, k, 2)];
                double complex p = (k1 * ux + k2 * uy + k3 * uz) / kk;
                if (kk == 0) p = 0;
                uh[SIDX(h, i, j, k, 0)] = ux - k1 * p;
                uh[SIDX(h, i, j, k, 1)] = uy - k2 * p;
                uh[SIDX(h, i, j, k, 2)] = uz - k3 * p;
            }
}

double ospec_vsummax(const ospec *h, const double *v)
{
    long n3 = (long)h->n * h->n * h->n;
    double m = 0;
    for (long t = 0; t < n3; ++t) {
        double s = fabs(v[t]) + fabs(v[t + n3]) + fabs(v[t + 2 * n3]);
        if (s > m) m = s;
    }
    return m;
}

/* spectral_rhs!(du, uh): returns vsummax of the physical velocity. */
double ospec_rhs(ospec *h, double complex *du, const double complex *uh)
{
    long n3 = (long)h->n * h->n * h->n;
    ospec_spec_to_phys(h, uh, h->v);
    double vmax = ospec_vsummax(h, h->v);
    for (long t = 0; t < 3 * n3; ++t) h->rbuf[t] = h->v[t] * h->v[t];
    ospec_phys_to_spec(h, h->rbuf, h->sig);
#pragma omp parallel for
    for (int k = 0; k < h->m; ++k)
        for (int j = 0; j < h->m; ++j)
            for (int i = 0; i < h->K; ++i) {
                double k1, k2, k3;
                kvec(h, i, j, k, &k1, &k2, &k3);
                du[SIDX(h, i, j, k, 0)] = -I * k1 * h->sig[SIDX(h, i, j, k, 0)];
                du[SIDX(h, i, j, k, 1)] = -I * k2 * h->sig[SIDX(h, i, j, k, 1)];
                du[SIDX(h, i, j, k, 2)] = -I * k3 * h->sig[SIDX(h, i, j, k, 2)];
            }
    for (long t = 0; t < n3; ++t) {
        double a = h->v[t], b = h->v[t + n3], c = h->v[t + 2 * n3];
        h->rbuf[t] = a * b;          /* xy */
        h->rbuf[t + n3] = b * c;     /* yz */
        h->rbuf[t + 2 * n3] = c * a; /* zx */
    }
    ospec_phys_to_spec(h, h->rbuf, h->sig);
#pragma omp parallel for
    for (int k = 0; k < h->m; ++k)
        for (int j = 0; j < h->m; ++j)
            for (int i = 0; i < h->K; ++i) {
                double k1, k2, k3;
                kvec(h, i, j, k, &k1, &k2, &k3);
                double complex sxy = h->sig[SIDX(h, i, j, k, 0)], syz = h->sig[SIDX(h, i, j, k, 1)],
                               szx = h->sig[SIDX(h, i, j, k, 2)];
                du[SIDX(h, i, j, k, 0)] -= I * (k2 * sxy + k3 * szx);
                du[SIDX(h, i, j, k, 1)] -= I * (k1 * sxy + k3 * syz);
                du[SIDX(h, i, j, k, 2)] -= I * (k1 * szx + k2 * syz);
            }
    return vmax;
}

static const double W_A[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double W_B[3] = {1.0 / 4.0, 0.0, 0.0};
static const double W_DC[3] = {8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0};

/* spectral_step!(uh, s, Δt, cache): returns vsummax of stage 1. */
double ospec_step(ospec *h, double complex *uh, double dt, double complex *ustart, double complex *du)
{
    long ns = (long)h->K * h->m * h->m * 3;
    memcpy(ustart, uh, sizeof(double complex) * ns);
    double vmax = 0;
    for (int s = 0; s < 3; ++s) {
        double v = ospec_rhs(h, du, uh);
        if (s == 0) vmax = v;
        double nudc = h->visc * W_DC[s] * dt;
#pragma omp parallel for
        for (int k = 0; k < h->m; ++k)
            for (int j = 0; j < h->m; ++j)
                for (int i = 0; i < h->K; ++i) {
                    double k1, k2, k3;
                    kvec(h, i, j, k, &k1, &k2, &k3);
                    double kk = k1 * k1 + k2 * k2 + k3 * k3;
                    double phi = exp(-nudc * kk);
                    for (int c = 0; c < 3; ++c) {
                        double complex st = ustart[SIDX(h, i, j, k, c)], d = du[SIDX(h, i, j, k, c)];
                        uh[SIDX(h, i, j, k, c)] = phi * (st + W_A[s] * dt * d);
                        ustart[SIDX(h, i, j, k, c)] = phi * (st + W_B[s] * dt * d);
                    }
                }
        ospec_project(h, uh);
    }
    return vmax;
}

static int ishell(const ospec *h, int i, int j, int k)
{
    long a = i, b = ospec_compactfreq(j, h->m, h->kcut), c = ospec_compactfreq(k, h->m, h->kcut);
    return (int)floor(sqrt((double)(a * a + b * b + c * c)));
}

/* shell_energies for shells s0..s0+ns-1 (contiguous), E[s - s0]. */
void ospec_shell_energies(const ospec *h, const double complex *uh, int s0, int ns, double *E)
{
    for (int q = 0; q < ns; ++q) E[q] = 0;
    for (int k = 0; k < h->m; ++k)
        for (int j = 0; j < h->m; ++j)
            for (int i = 0; i < h->K; ++i) {
                int s = ishell(h, i, j, k) - s0;
                if (s < 0 || s >= ns) continue;
                double w = i == 0 ? 1.0 : 2.0;
                double e = 0;
                for (int c = 0; c < 3; ++c) {
                    double complex a = uh[SIDX(h, i, j, k, c)];
                    e += creal(a) * creal(a) + cimag(a) * cimag(a);
                }
                E[s] += w * e / 2;
            }
}

void ospec_rescale_shell(const ospec *h, double complex *uh, int sh, double factor)
{
    for (int k = 0; k < h->m; ++k)
        for (int j = 0; j < h->m; ++j)
            for (int i = 0; i < h->K; ++i)
                if (ishell(h, i, j, k) == sh)
                    for (int c = 0; c < 3; ++c) uh[SIDX(h, i, j, k, c)] *= factor;
}

/* noise_kernel!: v[i, j, k, c] for the whole grid (single rank, o = 0). */
void ospec_noise(const ospec *h, uint64_t seed, double *v)
{
    const uint64_t n = (uint64_t)h->n;
    const long nn = h->n;
#pragma omp parallel for
    for (long k = 0; k < nn; ++k)
        for (long c = 0; c < 3; ++c)
            for (long j = 0; j < nn; ++j)
                for (long i = 0; i < nn; ++i) {
                    uint64_t key = (((uint64_t)c * n + (uint64_t)k) * n + (uint64_t)j) * n + (uint64_t)i;
                    uint64_t r1 = ospec_splitmix64(seed + 2 * key), r2 = ospec_splitmix64(seed + 2 * key + 1);
                    double u1 = ((double)(r1 >> 11) + 0.5) * 0x1p-53;
                    double u2 = ((double)(r2 >> 11) + 0.5) * 0x1p-53;
                    /* cospi(2 u2) via the exactly reduced argument */
                    long double a = 2.0L * (long double)u2;
                    double cs = (double)cosl(3.141592653589793238462643383279502884L * a);
                    v[i + nn * (j + nn * (k + nn * c))] = sqrt(-2 * log(u1)) * cs;
                }
}

double ospec_energy(const ospec *h, const double complex *uh)
{
    double e = 0;
    long ns = (long)h->K * h->m * h->m * 3;
    for (long t = 0; t < ns; ++t) {
        int i = (int)(t % h->K);
        double complex a = uh[t];
        double a2 = creal(a) * creal(a) + cimag(a) * cimag(a);
        e += (i == 0 ? 1.0 : 2.0) * a2;
    }
    return e / 2;
}

double ospec_dissipation(const ospec *h, const double complex *uh)
{
    double s = 0;
    for (int k = 0; k < h->m; ++k)
        for (int j = 0; j < h->m; ++j)
            for (int i = 0; i < h->K; ++i) {
                double k1, k2, k3;
                kvec(h, i, j, k, &k1, &k2, &k3);
                double a2 = 0;
                for (int c = 0; c < 3; ++c) {
                    double complex a = uh[SIDX(h, i, j, k, c)];
                    a2 += creal(a) * creal(a) + cimag(a) * cimag(a);
                }
                s += (i == 0 ? 1.0 : 2.0) * a2 * (k1 * k1 + k2 * k2 + k3 * k3);
            }
    return h->visc * s;
}

double ospec_maxdiv(const ospec *h, const double complex *uh)
{
    double mx = 0;
    for (int k = 0; k < h->m; ++k)
        for (int j = 0; j < h->m; ++j)
            for (int i = 0; i < h->K; ++i) {
                double k1, k2, k3;
                kvec(h, i, j, k, &k1, &k2, &k3);
                double complex d = k1 * uh[SIDX(h, i, j, k, 0)] + k2 * uh[SIDX(h, i, j, k, 1)] + k3 * uh[SIDX(h, i, j, k, 2)];
                if (cabs(d) > mx) mx = cabs(d);
            }
    return mx;
}

/* σ̂ = F[v_i v_j]/n³ at the retained modes for the six products
 * (xx, yy, zz, xy, yz, zx) of the physical velocity of uh -- the two product
 * batches of spectral_rhs! (src/spectral/operators.jl:66-75) and of
 * sfs_sample! (src/spectral/les.jl:323-337).  sig6: (K, m, m, 6). */
void ospec_products(ospec *h, const double complex *uh, double complex *sig6)
{
    long n3 = (long)h->n * h->n * h->n;
    long ns3 = (long)h->K * h->m * h->m * 3;
    ospec_spec_to_phys(h, uh, h->v);
    for (long t = 0; t < 3 * n3; ++t) h->rbuf[t] = h->v[t] * h->v[t];
    ospec_phys_to_spec(h, h->rbuf, sig6);
    for (long t = 0; t < n3; ++t) {
        double a = h->v[t], b = h->v[t + n3], c = h->v[t + 2 * n3];
        h->rbuf[t] = a * b;
        h->rbuf[t + n3] = b * c;
        h->rbuf[t + 2 * n3] = c * a;
    }
    ospec_phys_to_spec(h, h->rbuf, sig6 + ns3);
}
