/*
 * FP64 CPU oracle for the finite-volume path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load this file's library; the HIP product in distributednavierstokes_amd/
 * never links or calls it.
 *
 * A plain restatement of the reference's CPU path, in the reference's own
 * data layout: ghosted column-major (x fastest) arrays of extent
 * (nx+2, ny+2, nz+2), ghost width w = 1.  Every index below is the
 * reference's 1-based ghosted index minus one, and the y metric arrays
 * dy[j], dyu[j] are the reference's Δy[J], Δyu[J] at the same J = j + 1.
 * The structure is reference-literal (one loop nest per reference kernel,
 * separate halo fills, the Thomas c' array in memory) so that this file also
 * serves as the CPU baseline.  FFTs use oracle_fft.c instead of FFTW.
 *
 * Functions follow (reference file:line):
 *   griddata            src/setup.jl:148-193
 *   activeranges        src/setup.jl:132-138
 *   exchange_halo       src/halo.jl:26-45, 81-133 (single-rank branches 93-104)
 *   divergence          src/operators.jl:10-42
 *   momentum            src/operators.jl:44-178
 *   pgrad               src/operators.jl:180-220
 *   plan_poisson        src/poisson.jl:12, 56-120
 *   thomas / modedivide src/poisson.jl:14-49, 122-140
 *   poisson             src/poisson.jl:147-161
 *   project             src/timestep.jl:25-31
 *   step                src/timestep.jl:10-14, 41-63
 *   cfl_timestep        src/timestep.jl:69-80
 *   ydiffusion_coeffs   src/ydiffusion.jl:31-51
 *   imex_update         src/ydiffusion.jl:71-154
 *   maxdiv, kinetic_en. src/fields.jl:38-60
 *   channelstats        src/stats.jl:39-58
 *   spectrum sample     src/stats.jl:120-157
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_fft.h"

typedef struct ofv {
    int nx, ny, nz, wall, imexy;
    double dx, dz, visc, f[3], epsy;
    double *dy, *dyu; /* ny+2 entries: cells 0..ny+1 (0 and ny+1 are ghosts) */
    double dymin;
    double *yfg; /* global faces 0..ny */
    double *ycg; /* global cell centres 1..ny */
    double *lam1, *lam2, *lam3;
    double *ay, *by, *cy;             /* Poisson y rows (walls) */
    double *at, *bt, *ct, *an, *bn, *cn; /* CN y-diffusion rows */
    double *r;                        /* nx*ny*nz */
    double complex *ahat;             /* (nx/2+1)*ny*nz */
    double *work;                     /* (nx/2+1)*ny*nz, Thomas c' */
    double *ybuf, *ybuf2, *ybuf3;     /* nx*ny*nz, CN lines */
} ofv;

#define GX (h->nx + 2)
#define GY (h->ny + 2)
#define GZ (h->nz + 2)
#define G(i, j, k) ((long)(i) + (long)GX * ((long)(j) + (long)GY * (long)(k)))
#define R(i, j, k) ((long)(i) + (long)h->nx * ((long)(j) + (long)h->ny * (long)(k)))
#define NXC (h->nx / 2 + 1)
#define C(i, j, k) ((long)(i) + (long)NXC * ((long)(j) + (long)h->ny * (long)(k)))

static double modlam(int k0, int n, double d)
{
    /* modλ(k, n, Δ) = 2(cospi(2(k-1)/n) - 1)/Δ², k0 = k - 1 */
    const long double pi = 3.141592653589793238462643383279502884L;
    double c = (double)cosl(pi * (2.0L * (long double)k0 / (long double)n));
    if (k0 == 0) c = 1.0;
    return 2.0 * (c - 1.0) / (d * d);
}

/* OpenMP threads of every later oracle call (the CPU baseline times one
 * thread and all host cores); returns the count in effect. */
#ifdef _OPENMP
#include <omp.h>
#endif
int oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

void ofv_destroy(ofv *h)
{
    if (!h) return;
    free(h->dy); free(h->dyu); free(h->yfg); free(h->ycg);
    free(h->lam1); free(h->lam2); free(h->lam3);
    free(h->ay); free(h->by); free(h->cy);
    free(h->at); free(h->bt); free(h->ct); free(h->an); free(h->bn); free(h->cn);
    free(h->r); free(h->ahat); free(h->work);
    free(h->ybuf); free(h->ybuf2); free(h->ybuf3);
    free(h);
}

/* n[3]; lims[6] = (x0, x1, y0, y1, z0, z1); yf[ny+1] = y faces 0..ny
 * (yface(j) = y0 + Ly*stretch(j/ny), evaluated by the caller). */
int ofv_create(const int *n, const double *lims, int wall, const double *yf, double visc,
               const double *force, int imexy, ofv **out)
{
    *out = NULL;
    if (imexy && !wall) return -2; /* implicit y-diffusion requires bc[2] == :wall */
    ofv *h = calloc(1, sizeof(ofv));
    int nx = n[0], ny = n[1], nz = n[2];
    h->nx = nx; h->ny = ny; h->nz = nz; h->wall = wall; h->imexy = imexy;
    h->dx = (lims[1] - lims[0]) / nx;
    h->dz = (lims[5] - lims[4]) / nz;
    h->visc = visc;
    for (int c = 0; c < 3; ++c) h->f[c] = force[c];
    h->epsy = imexy ? 0.0 : 1.0;

    /* griddata: extended faces t = 0..ny+2 hold faces -1..ny+1 */
    double *yfe = malloc(sizeof(double) * (ny + 3));
    for (int j = 0; j <= ny; ++j) yfe[j + 1] = yf[j];
    if (!wall) {
        yfe[0] = yf[0] - (yf[ny] - yf[ny - 1]);
        yfe[ny + 2] = yf[ny] + (yf[1] - yf[0]);
    } else {
        yfe[0] = 2 * yf[0] - yf[1];
        yfe[ny + 2] = 2 * yf[ny] - yf[ny - 1];
    }
    h->dy = malloc(sizeof(double) * (ny + 2));
    h->dyu = malloc(sizeof(double) * (ny + 2));
    double *yce = malloc(sizeof(double) * (ny + 2));
    for (int c = 0; c < ny + 2; ++c) {
        h->dy[c] = yfe[c + 1] - yfe[c];
        yce[c] = (yfe[c] + yfe[c + 1]) / 2;
    }
    for (int c = 0; c < ny + 1; ++c) h->dyu[c] = yce[c + 1] - yce[c];
    h->dyu[ny + 1] = h->dy[ny + 1];
    h->dymin = h->dy[1];
    for (int c = 1; c <= ny; ++c)
        if (h->dy[c] < h->dymin) h->dymin = h->dy[c];
    h->yfg = malloc(sizeof(double) * (ny + 1));
    for (int j = 0; j <= ny; ++j) h->yfg[j] = yfe[j + 1];
    h->ycg = malloc(sizeof(double) * ny);
    for (int j = 1; j <= ny; ++j) h->ycg[j - 1] = yce[j];
    free(yfe);
    free(yce);

    /* plan_poisson */
    h->lam1 = malloc(sizeof(double) * (nx / 2 + 1));
    for (int k = 0; k <= nx / 2; ++k) h->lam1[k] = modlam(k, nx, h->dx);
    h->lam3 = malloc(sizeof(double) * nz);
    for (int k = 0; k < nz; ++k) h->lam3[k] = modlam(k, nz, h->dz);
    if (wall) {
        h->ay = calloc(ny, sizeof(double));
        h->by = calloc(ny, sizeof(double));
        h->cy = calloc(ny, sizeof(double));
        for (int j = 1; j <= ny; ++j) {
            double a = j == 1 ? 0.0 : 1 / (h->dy[j] * h->dyu[j - 1]);
            double c = j == ny ? 0.0 : 1 / (h->dy[j] * h->dyu[j]);
            h->ay[j - 1] = a; h->cy[j - 1] = c; h->by[j - 1] = -(a + c);
        }
    } else {
        double mx = h->dy[1], mn = h->dy[1];
        for (int c = 1; c <= ny; ++c) {
            if (h->dy[c] > mx) mx = h->dy[c];
            if (h->dy[c] < mn) mn = h->dy[c];
        }
        if (!(mx - mn < 1e-12 * mx)) { ofv_destroy(h); return -3; } /* periodic y needs uniform grid */
        h->lam2 = malloc(sizeof(double) * ny);
        for (int k = 0; k < ny; ++k) h->lam2[k] = modlam(k, ny, h->dy[1]);
    }
    long N = (long)nx * ny * nz, NC = (long)(nx / 2 + 1) * ny * nz;
    h->r = calloc(N, sizeof(double));
    h->ahat = calloc(NC, sizeof(double complex));
    h->work = calloc(NC, sizeof(double));

    if (imexy) { /* ydiffusion_coefficients */
        h->at = calloc(ny, sizeof(double)); h->bt = calloc(ny, sizeof(double)); h->ct = calloc(ny, sizeof(double));
        h->an = calloc(ny, sizeof(double)); h->bn = calloc(ny, sizeof(double)); h->cn = calloc(ny, sizeof(double));
        for (int j = 1; j <= ny; ++j) {
            double a = 1 / (h->dy[j] * h->dyu[j - 1]);
            double c = 1 / (h->dy[j] * h->dyu[j]);
            h->at[j - 1] = j == 1 ? 0.0 : a;
            h->ct[j - 1] = j == ny ? 0.0 : c;
            h->bt[j - 1] = j == 1 ? -(c + 2 * a) : j == ny ? -(a + 2 * c) : -(a + c);
        }
        for (int j = 1; j <= ny - 1; ++j) {
            double a = 1 / (h->dyu[j] * h->dy[j]);
            double c = 1 / (h->dyu[j] * h->dy[j + 1]);
            h->an[j - 1] = j == 1 ? 0.0 : a;
            h->cn[j - 1] = j == ny - 1 ? 0.0 : c;
            h->bn[j - 1] = -(a + c);
        }
        h->ybuf = calloc(N, sizeof(double));
        h->ybuf2 = calloc(N, sizeof(double));
        h->ybuf3 = calloc(N, sizeof(double));
    }
    *out = h;
    return 0;
}

/* Accessors for tests: metrics and Poisson coefficient rows. */
void ofv_griddata(const ofv *h, double *dy, double *dyu, double *dymin, double *yfg, double *ycg)
{
    memcpy(dy, h->dy, sizeof(double) * (h->ny + 2));
    memcpy(dyu, h->dyu, sizeof(double) * (h->ny + 2));
    *dymin = h->dymin;
    memcpy(yfg, h->yfg, sizeof(double) * (h->ny + 1));
    memcpy(ycg, h->ycg, sizeof(double) * h->ny);
}

double *ofv_rbuf(ofv *h) { return h->r; }

/* Last active ghosted y index of component c (0 = x, 1 = y, 2 = z, 3 = p). */
static int jhi(const ofv *h, int c) { return (c == 1 && h->wall) ? h->ny - 1 : h->ny; }

/* ---------------------------------------------------------------- halos */

/* stag: 0 = x, 1 = y, 2 = z, 3 = p */
void ofv_exchange_halo(const ofv *h, double *phi, int stag)
{
    int nx = h->nx, ny = h->ny, nz = h->nz;
    /* d = 1: x periodic, slabs span full (y, z) extents */
    for (int k = 0; k < GZ; ++k)
        for (int j = 0; j < GY; ++j) {
            phi[G(0, j, k)] = phi[G(nx, j, k)];
            phi[G(nx + 1, j, k)] = phi[G(1, j, k)];
        }
    /* d = 2: y */
    for (int k = 0; k < GZ; ++k)
        for (int i = 0; i < GX; ++i) {
            if (!h->wall) {
                phi[G(i, 0, k)] = phi[G(i, ny, k)];
                phi[G(i, ny + 1, k)] = phi[G(i, 1, k)];
            } else if (stag == 1) {
                phi[G(i, 0, k)] = 0.0;       /* lo wall face */
                phi[G(i, ny, k)] = 0.0;      /* hi wall face */
                phi[G(i, ny + 1, k)] = -phi[G(i, ny - 1, k)];
            } else {
                double s = stag == 3 ? 1.0 : -1.0;
                phi[G(i, 0, k)] = s * phi[G(i, 1, k)];
                phi[G(i, ny + 1, k)] = s * phi[G(i, ny, k)];
            }
        }
    /* d = 3: z periodic */
    for (int j = 0; j < GY; ++j)
        for (int i = 0; i < GX; ++i) {
            phi[G(i, j, 0)] = phi[G(i, j, nz)];
            phi[G(i, j, nz + 1)] = phi[G(i, j, 1)];
        }
}

void ofv_exchange_halo_u(const ofv *h, double *ux, double *uy, double *uz)
{
    ofv_exchange_halo(h, ux, 0);
    ofv_exchange_halo(h, uy, 1);
    ofv_exchange_halo(h, uz, 2);
}

/* ------------------------------------------------------------ operators */

void ofv_divergence(const ofv *h, double *r, const double *ux, const double *uy, const double *uz)
{
#pragma omp parallel for
    for (int k = 1; k <= h->nz; ++k)
        for (int j = 1; j <= h->ny; ++j)
            for (int i = 1; i <= h->nx; ++i)
                r[R(i - 1, j - 1, k - 1)] = (ux[G(i, j, k)] - ux[G(i - 1, j, k)]) / h->dx +
                                            (uy[G(i, j, k)] - uy[G(i, j - 1, k)]) / h->dy[j] +
                                            (uz[G(i, j, k)] - uz[G(i, j, k - 1)]) / h->dz;
}

void ofv_momentum(const ofv *h, double *qx, double *qy, double *qz, const double *ux,
                  const double *uy, const double *uz, double A, double dt)
{
    const double dx = h->dx, dz = h->dz, nu = h->visc, ey = h->epsy;
    const double *Dy = h->dy, *Dyu = h->dyu;
    const int jymax = jhi(h, 1);
#pragma omp parallel for
    for (int k = 1; k <= h->nz; ++k)
        for (int j = 1; j <= h->ny; ++j)
            for (int i = 1; i <= h->nx; ++i) {
                /* x momentum */
                double uce = (ux[G(i, j, k)] + ux[G(i + 1, j, k)]) / 2;
                double ucw = (ux[G(i - 1, j, k)] + ux[G(i, j, k)]) / 2;
                double vn = (uy[G(i, j, k)] + uy[G(i + 1, j, k)]) / 2;
                double un = (ux[G(i, j, k)] + ux[G(i, j + 1, k)]) / 2;
                double vs = (uy[G(i, j - 1, k)] + uy[G(i + 1, j - 1, k)]) / 2;
                double us = (ux[G(i, j - 1, k)] + ux[G(i, j, k)]) / 2;
                double wt = (uz[G(i, j, k)] + uz[G(i + 1, j, k)]) / 2;
                double ut = (ux[G(i, j, k)] + ux[G(i, j, k + 1)]) / 2;
                double wb = (uz[G(i, j, k - 1)] + uz[G(i + 1, j, k - 1)]) / 2;
                double ub = (ux[G(i, j, k - 1)] + ux[G(i, j, k)]) / 2;
                double conv = (uce * uce - ucw * ucw) / dx + (vn * un - vs * us) / Dy[j] +
                              (wt * ut - wb * ub) / dz;
                double diff =
                    (ux[G(i + 1, j, k)] - 2 * ux[G(i, j, k)] + ux[G(i - 1, j, k)]) / (dx * dx) +
                    ey * ((ux[G(i, j + 1, k)] - ux[G(i, j, k)]) / Dyu[j] -
                          (ux[G(i, j, k)] - ux[G(i, j - 1, k)]) / Dyu[j - 1]) / Dy[j] +
                    (ux[G(i, j, k + 1)] - 2 * ux[G(i, j, k)] + ux[G(i, j, k - 1)]) / (dz * dz);
                qx[G(i, j, k)] = A * qx[G(i, j, k)] + dt * (-conv + nu * diff + h->f[0]);

                /* y momentum */
                double ue = (ux[G(i, j, k)] * Dy[j] + ux[G(i, j + 1, k)] * Dy[j + 1]) / (2 * Dyu[j]);
                double ve = (uy[G(i, j, k)] + uy[G(i + 1, j, k)]) / 2;
                double uw = (ux[G(i - 1, j, k)] * Dy[j] + ux[G(i - 1, j + 1, k)] * Dy[j + 1]) / (2 * Dyu[j]);
                double vw = (uy[G(i - 1, j, k)] + uy[G(i, j, k)]) / 2;
                double vcn = (uy[G(i, j, k)] + uy[G(i, j + 1, k)]) / 2;
                double vcs = (uy[G(i, j - 1, k)] + uy[G(i, j, k)]) / 2;
                double wtv = (uz[G(i, j, k)] * Dy[j] + uz[G(i, j + 1, k)] * Dy[j + 1]) / (2 * Dyu[j]);
                double vt = (uy[G(i, j, k)] + uy[G(i, j, k + 1)]) / 2;
                double wbv = (uz[G(i, j, k - 1)] * Dy[j] + uz[G(i, j + 1, k - 1)] * Dy[j + 1]) / (2 * Dyu[j]);
                double vb = (uy[G(i, j, k - 1)] + uy[G(i, j, k)]) / 2;
                conv = (ue * ve - uw * vw) / dx + (vcn * vcn - vcs * vcs) / Dyu[j] + (wtv * vt - wbv * vb) / dz;
                diff = (uy[G(i + 1, j, k)] - 2 * uy[G(i, j, k)] + uy[G(i - 1, j, k)]) / (dx * dx) +
                       ey * ((uy[G(i, j + 1, k)] - uy[G(i, j, k)]) / Dy[j + 1] -
                             (uy[G(i, j, k)] - uy[G(i, j - 1, k)]) / Dy[j]) / Dyu[j] +
                       (uy[G(i, j, k + 1)] - 2 * uy[G(i, j, k)] + uy[G(i, j, k - 1)]) / (dz * dz);
                if (j <= jymax) qy[G(i, j, k)] = A * qy[G(i, j, k)] + dt * (-conv + nu * diff + h->f[1]);

                /* z momentum */
                ue = (ux[G(i, j, k)] + ux[G(i, j, k + 1)]) / 2;
                double we = (uz[G(i, j, k)] + uz[G(i + 1, j, k)]) / 2;
                uw = (ux[G(i - 1, j, k)] + ux[G(i - 1, j, k + 1)]) / 2;
                double ww = (uz[G(i - 1, j, k)] + uz[G(i, j, k)]) / 2;
                double vnw = (uy[G(i, j, k)] + uy[G(i, j, k + 1)]) / 2;
                double wn = (uz[G(i, j, k)] + uz[G(i, j + 1, k)]) / 2;
                double vsw = (uy[G(i, j - 1, k)] + uy[G(i, j - 1, k + 1)]) / 2;
                double ws = (uz[G(i, j - 1, k)] + uz[G(i, j, k)]) / 2;
                double wct = (uz[G(i, j, k)] + uz[G(i, j, k + 1)]) / 2;
                double wcb = (uz[G(i, j, k - 1)] + uz[G(i, j, k)]) / 2;
                conv = (ue * we - uw * ww) / dx + (vnw * wn - vsw * ws) / Dy[j] + (wct * wct - wcb * wcb) / dz;
                diff = (uz[G(i + 1, j, k)] - 2 * uz[G(i, j, k)] + uz[G(i - 1, j, k)]) / (dx * dx) +
                       ey * ((uz[G(i, j + 1, k)] - uz[G(i, j, k)]) / Dyu[j] -
                             (uz[G(i, j, k)] - uz[G(i, j - 1, k)]) / Dyu[j - 1]) / Dy[j] +
                       (uz[G(i, j, k + 1)] - 2 * uz[G(i, j, k)] + uz[G(i, j, k - 1)]) / (dz * dz);
                qz[G(i, j, k)] = A * qz[G(i, j, k)] + dt * (-conv + nu * diff + h->f[2]);
            }
}

/* φ ← φ - s ∂p/∂c on the active range of component c (0, 1, 2). */
void ofv_pgrad(const ofv *h, double *phi, const double *p, int c, double s)
{
    int jmax = jhi(h, c);
#pragma omp parallel for
    for (int k = 1; k <= h->nz; ++k)
        for (int j = 1; j <= jmax; ++j)
            for (int i = 1; i <= h->nx; ++i) {
                if (c == 0) phi[G(i, j, k)] -= s * (p[G(i + 1, j, k)] - p[G(i, j, k)]) / h->dx;
                else if (c == 1) phi[G(i, j, k)] -= s * (p[G(i, j + 1, k)] - p[G(i, j, k)]) / h->dyu[j];
                else phi[G(i, j, k)] -= s * (p[G(i, j, k + 1)] - p[G(i, j, k)]) / h->dz;
            }
}

/* -------------------------------------------------------------- Poisson */

static void solve_y(ofv *h)
{
    int nxc = NXC, ny = h->ny, nz = h->nz;
    if (h->wall) {
#pragma omp parallel for
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nxc; ++i) {
                double complex *x = h->ahat;
                double *cp = h->work;
                double lam = h->lam1[i] + h->lam3[k];
                double beta = lam == 0 ? 1.0 : h->by[0] + lam;
                cp[C(i, 0, k)] = (lam == 0 ? 0.0 : h->cy[0]) / beta;
                x[C(i, 0, k)] = (lam == 0 ? 0.0 : x[C(i, 0, k)]) / beta;
                for (int j = 1; j < ny; ++j) {
                    beta = (h->by[j] + lam) - h->ay[j] * cp[C(i, j - 1, k)];
                    cp[C(i, j, k)] = h->cy[j] / beta;
                    x[C(i, j, k)] = (x[C(i, j, k)] - h->ay[j] * x[C(i, j - 1, k)]) / beta;
                }
                for (int j = ny - 2; j >= 0; --j) x[C(i, j, k)] -= cp[C(i, j, k)] * x[C(i, j + 1, k)];
            }
    } else {
#pragma omp parallel for
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nxc; ++i) offt_strided(h->ahat + C(i, 0, k), ny, nxc, -1);
#pragma omp parallel for
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nxc; ++i) {
                    double lam = h->lam1[i] + h->lam2[j] + h->lam3[k];
                    h->ahat[C(i, j, k)] = lam == 0 ? 0.0 : h->ahat[C(i, j, k)] / lam;
                }
#pragma omp parallel for
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nxc; ++i) {
                offt_strided(h->ahat + C(i, 0, k), ny, nxc, +1);
                for (int j = 0; j < ny; ++j) h->ahat[C(i, j, k)] /= ny;
            }
    }
}

/* Forward rfft(x) -> fft(z) of a real (nx, ny, nz) array into h->ahat. */
static void forward_xz(ofv *h, const double *r)
{
    int nx = h->nx, ny = h->ny, nz = h->nz, nxc = NXC;
#pragma omp parallel for
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) offt_r2c(r + R(0, j, k), h->ahat + C(0, j, k), nx);
#pragma omp parallel for
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nxc; ++i) offt_strided(h->ahat + C(i, j, 0), nz, (long)nxc * ny, -1);
}

/* poisson!(p): the divergence in h->r; writes p's interior (not its halos). */
void ofv_poisson(ofv *h, double *p)
{
    int nx = h->nx, ny = h->ny, nz = h->nz, nxc = NXC;
    forward_xz(h, h->r);
    solve_y(h);
#pragma omp parallel for
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nxc; ++i) {
            offt_strided(h->ahat + C(i, j, 0), nz, (long)nxc * ny, +1);
            for (int k = 0; k < nz; ++k) h->ahat[C(i, j, k)] /= nz;
        }
#pragma omp parallel for
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) {
            offt_c2r(h->ahat + C(0, j, k), h->r + R(0, j, k), nx);
            for (int i = 0; i < nx; ++i) h->r[R(i, j, k)] /= nx;
        }
#pragma omp parallel for
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) p[G(i, j, k)] = h->r[R(i - 1, j - 1, k - 1)];
}

void ofv_pressuregrad_update(const ofv *h, double *ux, double *uy, double *uz, const double *p)
{
    ofv_pgrad(h, ux, p, 0, 1.0);
    ofv_pgrad(h, uy, p, 1, 1.0);
    ofv_pgrad(h, uz, p, 2, 1.0);
}

void ofv_project(ofv *h, double *ux, double *uy, double *uz, double *psi)
{
    ofv_divergence(h, h->r, ux, uy, uz);
    ofv_poisson(h, psi);
    ofv_exchange_halo(h, psi, 3);
    ofv_pressuregrad_update(h, ux, uy, uz, psi);
    ofv_exchange_halo_u(h, ux, uy, uz);
}

/* ------------------------------------------------- CN y-diffusion (imex) */

void ofv_imex_update(ofv *h, double *u[3], double *const q[3], const double *p, double B, double alpha,
                     double dt)
{
    int nx = h->nx, ny = h->ny, nz = h->nz;
    double theta = alpha * dt * h->visc / 2;
    for (int c = 0; c < 3; ++c) {
        const double *a = c == 1 ? h->an : h->at, *b = c == 1 ? h->bn : h->bt, *cc = c == 1 ? h->cn : h->ct;
        double *phi = u[c];
        /* uⁿ lines */
#pragma omp parallel for
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) h->ybuf[R(i - 1, j - 1, k - 1)] = phi[G(i, j, k)];
        /* v = uⁿ + B q - αΔt ∂p on the active range */
        int jmax = jhi(h, c);
#pragma omp parallel for
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= jmax; ++j)
                for (int i = 1; i <= nx; ++i) phi[G(i, j, k)] += B * q[c][G(i, j, k)];
        ofv_pgrad(h, phi, p, c, alpha * dt);
#pragma omp parallel for
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) h->ybuf2[R(i - 1, j - 1, k - 1)] = phi[G(i, j, k)];
        /* rhs = v + θ Dy uⁿ, then the Thomas solve with c' in ybuf */
#pragma omp parallel for
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i) {
                double *rhs = h->ybuf3, *ut = h->ybuf, *vt = h->ybuf2;
                rhs[R(i, 0, k)] = vt[R(i, 0, k)] + theta * (b[0] * ut[R(i, 0, k)] + cc[0] * ut[R(i, 1, k)]);
                for (int j = 1; j < ny - 1; ++j)
                    rhs[R(i, j, k)] = vt[R(i, j, k)] + theta * (a[j] * ut[R(i, j - 1, k)] + b[j] * ut[R(i, j, k)] +
                                                               cc[j] * ut[R(i, j + 1, k)]);
                rhs[R(i, ny - 1, k)] = vt[R(i, ny - 1, k)] +
                                       theta * (a[ny - 1] * ut[R(i, ny - 2, k)] + b[ny - 1] * ut[R(i, ny - 1, k)]);
                double *x = rhs, *cp = h->ybuf; /* uⁿ no longer needed for this line */
                double beta = 1 - theta * b[0];
                cp[R(i, 0, k)] = -theta * cc[0] / beta;
                x[R(i, 0, k)] = x[R(i, 0, k)] / beta;
                for (int j = 1; j < ny; ++j) {
                    beta = (1 - theta * b[j]) + theta * a[j] * cp[R(i, j - 1, k)];
                    cp[R(i, j, k)] = -theta * cc[j] / beta;
                    x[R(i, j, k)] = (x[R(i, j, k)] + theta * a[j] * x[R(i, j - 1, k)]) / beta;
                }
                for (int j = ny - 2; j >= 0; --j) x[R(i, j, k)] -= cp[R(i, j, k)] * x[R(i, j + 1, k)];
            }
#pragma omp parallel for
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) phi[G(i, j, k)] = h->ybuf3[R(i - 1, j - 1, k - 1)];
    }
}

/* --------------------------------------------------------------- step */

static const double RK_A[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
static const double RK_B[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
static const double RK_ALPHA[3] = {1.0 / 3.0, 5.0 / 12.0, 1.0 / 4.0};

void ofv_step(ofv *h, double *ux, double *uy, double *uz, double *p, double *qx, double *qy, double *qz,
              double *delta, double dt)
{
    double *u[3] = {ux, uy, uz};
    double *q[3] = {qx, qy, qz};
    for (int s = 0; s < 3; ++s) {
        ofv_momentum(h, qx, qy, qz, ux, uy, uz, RK_A[s], dt);
        if (h->imexy) {
            ofv_imex_update(h, u, q, p, RK_B[s], RK_ALPHA[s], dt);
            ofv_exchange_halo_u(h, ux, uy, uz);
            ofv_project(h, ux, uy, uz, delta);
            double ad = RK_ALPHA[s] * dt;
#pragma omp parallel for
            for (int k = 1; k <= h->nz; ++k)
                for (int j = 1; j <= h->ny; ++j)
                    for (int i = 1; i <= h->nx; ++i) p[G(i, j, k)] += delta[G(i, j, k)] / ad;
            ofv_exchange_halo(h, p, 3);
        } else {
            for (int c = 0; c < 3; ++c) {
                int jmax = jhi(h, c);
                double *phi = u[c], *qc = q[c];
#pragma omp parallel for
                for (int k = 1; k <= h->nz; ++k)
                    for (int j = 1; j <= jmax; ++j)
                        for (int i = 1; i <= h->nx; ++i) phi[G(i, j, k)] += RK_B[s] * qc[G(i, j, k)];
            }
            ofv_exchange_halo_u(h, ux, uy, uz);
            ofv_project(h, ux, uy, uz, p);
        }
    }
}

static double maxabs_interior(const ofv *h, const double *phi)
{
    double m = 0;
#pragma omp parallel for reduction(max : m)
    for (int k = 1; k <= h->nz; ++k)
        for (int j = 1; j <= h->ny; ++j)
            for (int i = 1; i <= h->nx; ++i) {
                double a = fabs(phi[G(i, j, k)]);
                if (a > m) m = a;
            }
    return m;
}

double ofv_cfl_timestep(const ofv *h, const double *ux, const double *uy, const double *uz, double cfl)
{
    const double eps = 2.220446049250313e-16;
    double cx = h->dx / (maxabs_interior(h, ux) + eps);
    double cy = h->dymin / (maxabs_interior(h, uy) + eps);
    double cz = h->dz / (maxabs_interior(h, uz) + eps);
    double conv = fmin(cx, fmin(cy, cz));
    double invdy2 = h->imexy ? 0.0 : 1 / (h->dymin * h->dymin);
    double diff = 1 / (2 * h->visc * (1 / (h->dx * h->dx) + invdy2 + 1 / (h->dz * h->dz)) + eps);
    return fmin(cfl * conv, cfl * diff);
}

double ofv_maxdiv(ofv *h, const double *ux, const double *uy, const double *uz)
{
    ofv_divergence(h, h->r, ux, uy, uz);
    long N = (long)h->nx * h->ny * h->nz;
    double m = 0;
    for (long t = 0; t < N; ++t)
        if (fabs(h->r[t]) > m) m = fabs(h->r[t]);
    return m;
}

double ofv_kinetic_energy(const ofv *h, const double *ux, const double *uy, const double *uz)
{
    const double *u[3] = {ux, uy, uz};
    double e = 0;
    for (int c = 0; c < 3; ++c) {
        const double *d2 = c == 1 ? h->dyu : h->dy;
        int jmax = jhi(h, c);
        for (int k = 1; k <= h->nz; ++k)
            for (int j = 1; j <= jmax; ++j)
                for (int i = 1; i <= h->nx; ++i) {
                    double v = u[c][G(i, j, k)];
                    e += v * v * h->dx * d2[j] * h->dz / 2;
                }
    }
    return e;
}

/* ---------------------------------------------------------- statistics */

/* acc: Uc, Wc, uuc, wwc (ny each), then Vf, vvf, Uf, uvf (nf each),
 * nf = ny-1 (walls) or ny; the plane sums of one sample are added. */
void ofv_channelstats_sample(const ofv *h, const double *ux, const double *uy, const double *uz, double *acc)
{
    int nx = h->nx, ny = h->ny, nz = h->nz, nf = jhi(h, 1);
    double *Uc = acc, *Wc = acc + ny, *uuc = acc + 2 * ny, *wwc = acc + 3 * ny;
    double *Vf = acc + 4 * ny, *vvf = Vf + nf, *Uf = vvf + nf, *uvf = Uf + nf;
    for (int j = 1; j <= ny; ++j) {
        double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                double a = ux[G(i, j, k)], b = uz[G(i, j, k)];
                s1 += a; s2 += b; s3 += a * a; s4 += b * b;
            }
        Uc[j - 1] += s1; Wc[j - 1] += s2; uuc[j - 1] += s3; wwc[j - 1] += s4;
    }
    for (int j = 1; j <= nf; ++j) {
        double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                double v = uy[G(i, j, k)];
                double uf = (ux[G(i, j, k)] + ux[G(i, j + 1, k)]) / 2;
                double vx = (uy[G(i, j, k)] + uy[G(i + 1, j, k)]) / 2;
                s1 += v; s2 += v * v; s3 += uf; s4 += uf * vx;
            }
        Vf[j - 1] += s1; vvf[j - 1] += s2; Uf[j - 1] += s3; uvf[j - 1] += s4;
    }
}

/* Adds one sample of ½|û|²/N² per integer shell κ = round(|k|) for the three
 * components (fully periodic boxes only). */
void ofv_spectrum_sample(ofv *h, const double *ux, const double *uy, const double *uz, double *E, int nbins)
{
    int nx = h->nx, ny = h->ny, nz = h->nz, nxc = NXC;
    const double *u[3] = {ux, uy, uz};
    double N2 = (double)nx * ny * nz;
    N2 = N2 * N2;
    for (int c = 0; c < 3; ++c) {
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) h->r[R(i - 1, j - 1, k - 1)] = u[c][G(i, j, k)];
        forward_xz(h, h->r);
#pragma omp parallel for
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nxc; ++i) offt_strided(h->ahat + C(i, 0, k), ny, nxc, -1);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nxc; ++i) {
                    long kx = i;
                    long ky = j > ny / 2 ? j - ny : j;
                    long kz = k > nz / 2 ? k - nz : k;
                    int kap = (int)lround(sqrt((double)(kx * kx + ky * ky + kz * kz)));
                    double wgt = (0 < kx && 2 * kx < nx) ? 2.0 : 1.0;
                    double complex a = h->ahat[C(i, j, k)];
                    if (kap < nbins) E[kap] += wgt * (creal(a) * creal(a) + cimag(a) * cimag(a)) / 2 / N2;
                }
    }
}
