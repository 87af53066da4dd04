// Kernels of the reference-structured ("plain") FP64 step: one kernel per
// reference kernel, separate broadcasts, materialised transposes/copies and
// an HBM d' array in the tridiagonal solves, as the reference's single-rank
// GPU path runs them (src/timestep.jl:41-63, src/operators.jl:10-220,
// src/poisson.jl:147-161, src/ydiffusion.jl:71-154, src/halo.jl:81-133).
// It is the baseline the fused step is measured against, not a product path.
#pragma once

#include "fv_kernels.hpp"

namespace dns {

// y ← y + s x over n values (`@. v += B*q`, src/timestep.jl:54-58, and
// p += δ/(αΔt), src/timestep.jl:51)
template <typename T> __global__ void __launch_bounds__(256) k_plain_axpy(T *y, const T *x, T s, long n)
{
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        y[t] += s * x[t];
}

// dst ← src (the single-rank transposes of src/transpose.jl:96-101 and copybox!)
template <typename V> __global__ void __launch_bounds__(256) k_plain_copy(V *dst, const V *src, long n)
{
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        dst[t] = src[t];
}

// One face plane of a field copied into the side ghost buffer (the plane
// copies of exchange_halo!; the fused kernels apply the same ghost rules
// inline, so the plain step only pays their traffic and launches).
template <typename T>
__global__ void __launch_bounds__(256) k_plain_halo_face(const T *f, T *ghost, int face, FVGeom g)
{
    const int dim = face / 2, hi = face % 2;
    const int n0 = dim == 0 ? g.ny : g.nx, n1 = dim == 2 ? g.ny : g.nz;
    const int plane = hi ? (dim == 0 ? g.nx - 1 : dim == 1 ? g.ny - 1 : g.nz - 1) : 0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n0 * n1; t += gridDim.x * blockDim.x) {
        const int a = t % n0, b = t / n0;
        int i, j, k;
        if (dim == 0) { i = plane; j = a; k = b; }
        else if (dim == 1) { i = a; j = plane; k = b; }
        else { i = a; j = b; k = plane; }
        ghost[(size_t)face * (size_t)(g.nx > g.ny ? g.nx : g.ny) * (size_t)(g.ny > g.nz ? g.ny : g.nz) + t] =
            f[IX(i, j, k, g)];
    }
}

// divergence_kernel! (src/operators.jl:10-42) into the ghost-free r
template <typename T>
__global__ void __launch_bounds__(256) k_plain_div(const T *ux, const T *uy, const T *uz, const T *rdy, T rdx,
                                                   T rdz, T *r, FVGeom g)
{
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const int jj = j + 1;
    const T dxu = ux[IX(i, j, k, g)] - ux[IX(wrapm(i, g.nx), j, k, g)];
    const T dyv = ldf(uy, i, jj, k, g) - ldf(uy, i, jj - 1, k, g);
    const T dzw = uz[IX(i, j, k, g)] - uz[IX(i, j, wrapm(k, g.nz), g)];
    r[IX(i, j, k, g)] = dxu * rdx + dyv * rdy[jj] + dzw * rdz;
}

// pgrad{x,y,z}_kernel! (src/operators.jl:180-220): u_c −= s ∂ψ on active DOFs
template <typename T, int COMP>
__global__ void __launch_bounds__(256) k_plain_pgrad(T *u, const T *psi, const T *rdyu, T rdx, T rdz, T s, FVGeom g)
{
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const int c = IX(i, j, k, g);
    if (COMP == 0) u[c] -= s * (psi[IX(wrapp(i, g.nx), j, k, g)] - psi[c]) * rdx;
    else if (COMP == 1) {
        if (!(g.wall && j == g.ny - 1)) u[c] -= s * (ldc(psi, i, j + 2, k, g, T(1)) - psi[c]) * rdyu[j + 1];
    } else u[c] -= s * (psi[IX(i, j, wrapp(k, g.nz), g)] - psi[c]) * rdz;
}

// imex_update! register step (src/ydiffusion.jl:119-154): v = u + B q − αΔt ∂p
// on the active rows (written over the staged copy of u)
template <typename T, int COMP>
__global__ void __launch_bounds__(256) k_plain_imexv(T *v, const T *q, const T *p, const T *rdyu, double B,
                                                     double adt, double rdx, double rdz, FVGeom g)
{
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    if (COMP == 1 && g.wall && j == g.ny - 1) return;
    const int c = IX(i, j, k, g);
    double dpdc;
    if (COMP == 0) dpdc = ((double)p[IX(wrapp(i, g.nx), j, k, g)] - (double)p[c]) * rdx;
    else if (COMP == 1) dpdc = ((double)ldc(p, i, j + 2, k, g, T(1)) - (double)p[c]) * rdyu[j + 1];
    else dpdc = ((double)p[IX(i, j, wrapp(k, g.nz), g)] - (double)p[c]) * rdz;
    v[c] = T((double)v[c] + B * (double)q[c] - adt * dpdc);
}

// imexrhs_kernel! (src/ydiffusion.jl:71-91): rhs = v + θ Dy uⁿ per column
template <typename T>
__global__ void __launch_bounds__(256) k_plain_imexrhs(T *rhs, const T *un, const double *a, const double *b,
                                                       const double *c, double theta, FVGeom g)
{
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    double lap = b[j] * (double)un[IX(i, j, k, g)];
    if (j > 0) lap += a[j] * (double)un[IX(i, j - 1, k, g)];
    if (j < g.ny - 1) lap += c[j] * (double)un[IX(i, j + 1, k, g)];
    const int cc = IX(i, j, k, g);
    rhs[cc] = T((double)rhs[cc] + theta * lap);
}

// imexsolve_kernel! (src/ydiffusion.jl:94-109): Thomas per (i, k) column with
// the pivots of k_cn_coeffs, d' stored in the HBM array D
template <typename T>
__global__ void __launch_bounds__(64) k_plain_imexsolve(T *x, const T *rhs, double *D, const double *a,
                                                        const double *cp, const double *ip, double theta, FVGeom g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (i >= g.nx) return;
    double dp = 0.0;
    for (int j = 0; j < g.ny; ++j) {
        const double r = (double)rhs[IX(i, j, k, g)];
        dp = j == 0 ? r * ip[0] : (r + theta * a[j] * dp) * ip[j];
        D[IX(i, j, k, g)] = dp;
    }
    double xv = dp;
    x[IX(i, g.ny - 1, k, g)] = T(xv);
    for (int j = g.ny - 2; j >= 0; --j) {
        xv = D[IX(i, j, k, g)] - cp[j] * xv;
        x[IX(i, j, k, g)] = T(xv);
    }
}

}  // namespace dns
