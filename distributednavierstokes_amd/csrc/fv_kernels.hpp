// Device kernels of the staggered finite-volume time step (gfx950).
//
// Storage: every field is an interior-only column-major array (x fastest)
// of extent (nx, ny, nz); there are no ghost layers in HBM.  The ghost
// values the reference keeps in its (nx+2, ny+2, nz+2) arrays
// (src/halo.jl:19-45, 81-133) are produced inline by the load helpers below:
// periodic directions wrap, and at walls cell-centred velocities mirror
// antisymmetrically, the pressure mirrors symmetrically and the wall-normal
// velocity is zero on the wall faces with an antisymmetric ghost face.
// y metric arrays keep the reference's ghosted indexing: dy[jj], dyu[jj] for
// jj = 0..ny+1 equal Δy[J], Δyu[J] at J = jj + 1 (src/setup.jl:148-193).
// u.y row j stores face j+1 (the top face of cell j+1); with walls the last
// row is the top wall face and always holds zero.
#pragma once

#include <type_traits>

#include "common.hpp"

// Loads kept in flight per thread by the HBM→LDS staging loops (capped by the
// tile's elements per thread).
#ifndef FV_LOAD_MAX
#define FV_LOAD_MAX 16
#endif
#define FV_LOAD_U(e) ((e) < FV_LOAD_MAX ? (e) : FV_LOAD_MAX)
#include "fft_lds.hpp"
#include "reduce.hpp"

namespace dns {

// nxc = nx/2 + 1 rfft coefficients per x-line; the complex pencil A is
// stored with leading dimension ldc >= nxc rounded up to whole 128-byte
// lines, so every tile of the z and y passes starts on a cache-line boundary.
struct FVGeom {
    int nx, ny, nz, nxc, ldc;
    int wall;
};

__device__ inline int IX(int i, int j, int k, const FVGeom &g) { return i + g.nx * (j + g.ny * k); }
__device__ inline int CX(int i, int j, int k, const FVGeom &g) { return i + g.ldc * (j + g.ny * k); }

__device__ inline int wrapm(int i, int n) { return i == 0 ? n - 1 : i - 1; }
__device__ inline int wrapp(int i, int n) { return i == n - 1 ? 0 : i + 1; }

// Cell-centred field at ghosted row jj (0..ny+1); s = -1 for tangential
// velocities (no-slip mirror), +1 for the pressure (Neumann mirror).
template <typename T>
__device__ inline T ldc(const T *__restrict__ f, int i, int jj, int k, const FVGeom &g, T s)
{
    if (jj < 1) {
        if (g.wall) return s * f[IX(i, 0, k, g)];
        jj = g.ny;
    } else if (jj > g.ny) {
        if (g.wall) return s * f[IX(i, g.ny - 1, k, g)];
        jj = 1;
    }
    return f[IX(i, jj - 1, k, g)];
}

// Wall-normal face field at ghosted face index jj (0..ny+1).
template <typename T> __device__ inline T ldf(const T *__restrict__ f, int i, int jj, int k, const FVGeom &g)
{
    if (jj < 1) {
        if (g.wall) return T(0);
        jj = g.ny;
    } else if (jj > g.ny) {
        if (g.wall) return -f[IX(i, g.ny - 2, k, g)];
        jj = 1;
    }
    return f[IX(i, jj - 1, k, g)];
}

// ------------------------------------------------------------------------
// Fused momentum right-hand side and Williamson register update
// (src/operators.jl:44-178 and src/timestep.jl:54-58):
//   q ← A q + Δt (−C(u) + ν D(u) + f)        all three components
//   v ← u + B q                               (explicit mode, WRITE_V)
// One thread per cell; block (64, 4), grid (nx/64, ny/4, nz).
// ------------------------------------------------------------------------
template <typename T> struct MomentumArgs {
    const T *ux, *uy, *uz;
    T *qx, *qy, *qz;
    T *vx, *vy, *vz;
    const T *dy, *rdy, *rdyu; // ghosted y metrics and reciprocals
    T A, dt, nu, fx, fy, fz, epsy, B;
    T rdx, rdx2, rdz, rdz2;
    const double *dtp; // when set, Δt is read here (computed on the device) instead of dt
};

// The three momentum components at one cell from its 27 stencil values
// (uXabc: component X at offsets a, b, c in x, y, z; p = +1, m = -1).
// NT: the q and u′ streams (read or written once per launch) bypass L2
// residency with non-temporal accesses, leaving L2 to the u planes that
// neighbouring z planes re-read.
template <typename T> __device__ inline T ld_stream(const T *p, bool nt) { return nt ? __builtin_nontemporal_load(p) : *p; }
template <typename T> __device__ inline void st_stream(T v, T *p, bool nt)
{
    if (nt) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Device-computed Δt: read once at kernel entry, before any store, so the
// load is a uniform scalar load and Δt stays in scalar registers.
template <typename T> __device__ inline void resolve_dt(MomentumArgs<T> &a)
{
    if (a.dtp) a.dt = T(*a.dtp);
}

// Metrics of row j (ghosted index jj = j + 1): Δy at jj and jj+1, their
// reciprocals, 1/Δyu at jj and jj−1.
template <typename T> struct RowMetrics {
    T Dyj, Dyp, rDyj, rDyp, rDyuj, rDyum;
};
template <typename T> __device__ inline RowMetrics<T> row_metrics(const MomentumArgs<T> &a, int jj)
{
    return {a.dy[jj], a.dy[jj + 1], a.rdy[jj], a.rdy[jj + 1], a.rdyu[jj], a.rdyu[jj - 1]};
}

// The update itself with the row metrics and the old q values (qo*, read
// only when READ_Q) supplied by the caller.
template <typename T, bool READ_Q, bool WRITE_V, bool NT>
__device__ inline void momentum_core(const MomentumArgs<T> &a, const FVGeom &g, int c, int j, const RowMetrics<T> &m,
                                     T qox, T qoy, T qoz, T ux000, T uxp00, T uxm00, T ux0p0, T ux0m0, T ux00p,
                                     T ux00m, T uxmp0, T uxm0p, T uy000, T uyp00, T uym00, T uy0p0, T uy0m0, T uy00p,
                                     T uy00m, T uypm0, T uy0mp, T uz000, T uzp00, T uzm00, T uz0p0, T uz0m0, T uz00p,
                                     T uz00m, T uzp0m, T uz0pm)
{
    const T Dyj = m.Dyj, Dyp = m.Dyp;
    const T rDyj = m.rDyj, rDyp = m.rDyp, rDyuj = m.rDyuj, rDyum = m.rDyum;
    const T h = T(0.5);
    const T dt = a.dt;

    // x momentum at the right x-face of the cell
    {
        T uce = (ux000 + uxp00) * h, ucw = (uxm00 + ux000) * h;
        T vn = (uy000 + uyp00) * h, un = (ux000 + ux0p0) * h;
        T vs = (uy0m0 + uypm0) * h, us = (ux0m0 + ux000) * h;
        T wt = (uz000 + uzp00) * h, ut = (ux000 + ux00p) * h;
        T wb = (uz00m + uzp0m) * h, ub = (ux00m + ux000) * h;
        T conv = (uce * uce - ucw * ucw) * a.rdx + (vn * un - vs * us) * rDyj + (wt * ut - wb * ub) * a.rdz;
        T diff = (uxp00 - T(2) * ux000 + uxm00) * a.rdx2 +
                 a.epsy * ((ux0p0 - ux000) * rDyuj - (ux000 - ux0m0) * rDyum) * rDyj +
                 (ux00p - T(2) * ux000 + ux00m) * a.rdz2;
        T q = dt * (-conv + a.nu * diff + a.fx);
        if (READ_Q) q += a.A * qox;
        st_stream(q, a.qx + c, NT);
        if (WRITE_V) st_stream(ux000 + a.B * q, a.vx + c, NT);
    }
    // y momentum at the top y-face; the staggered volume straddles cells
    // jj and jj+1, hence the Δy-weighted transporting velocities
    {
        const bool active = !(g.wall && j == g.ny - 1);
        if (active) {
            T hw = h * rDyuj;
            T ue = (ux000 * Dyj + ux0p0 * Dyp) * hw, ve = (uy000 + uyp00) * h;
            T uw = (uxm00 * Dyj + uxmp0 * Dyp) * hw, vw = (uym00 + uy000) * h;
            T vcn = (uy000 + uy0p0) * h, vcs = (uy0m0 + uy000) * h;
            T wtv = (uz000 * Dyj + uz0p0 * Dyp) * hw, vt = (uy000 + uy00p) * h;
            T wbv = (uz00m * Dyj + uz0pm * Dyp) * hw, vb = (uy00m + uy000) * h;
            T conv = (ue * ve - uw * vw) * a.rdx + (vcn * vcn - vcs * vcs) * rDyuj + (wtv * vt - wbv * vb) * a.rdz;
            T diff = (uyp00 - T(2) * uy000 + uym00) * a.rdx2 +
                     a.epsy * ((uy0p0 - uy000) * rDyp - (uy000 - uy0m0) * rDyj) * rDyuj +
                     (uy00p - T(2) * uy000 + uy00m) * a.rdz2;
            T q = dt * (-conv + a.nu * diff + a.fy);
            if (READ_Q) q += a.A * qoy;
            st_stream(q, a.qy + c, NT);
            if (WRITE_V) st_stream(uy000 + a.B * q, a.vy + c, NT);
        } else if (WRITE_V) {
            st_stream(uy000, a.vy + c, NT);
        }
    }
    // z momentum at the back z-face
    {
        T ue = (ux000 + ux00p) * h, we = (uz000 + uzp00) * h;
        T uw = (uxm00 + uxm0p) * h, ww = (uzm00 + uz000) * h;
        T vnw = (uy000 + uy00p) * h, wn = (uz000 + uz0p0) * h;
        T vsw = (uy0m0 + uy0mp) * h, ws = (uz0m0 + uz000) * h;
        T wct = (uz000 + uz00p) * h, wcb = (uz00m + uz000) * h;
        T conv = (ue * we - uw * ww) * a.rdx + (vnw * wn - vsw * ws) * rDyj + (wct * wct - wcb * wcb) * a.rdz;
        T diff = (uzp00 - T(2) * uz000 + uzm00) * a.rdx2 +
                 a.epsy * ((uz0p0 - uz000) * rDyuj - (uz000 - uz0m0) * rDyum) * rDyj +
                 (uz00p - T(2) * uz000 + uz00m) * a.rdz2;
        T q = dt * (-conv + a.nu * diff + a.fz);
        if (READ_Q) q += a.A * qoz;
        st_stream(q, a.qz + c, NT);
        if (WRITE_V) st_stream(uz000 + a.B * q, a.vz + c, NT);
    }
}

template <typename T, bool READ_Q, bool WRITE_V, bool NT = false>
__device__ inline void momentum_cell(const MomentumArgs<T> &a, const FVGeom &g, int c, int j, int jj, T ux000,
                                     T uxp00, T uxm00, T ux0p0, T ux0m0, T ux00p, T ux00m, T uxmp0, T uxm0p, T uy000,
                                     T uyp00, T uym00, T uy0p0, T uy0m0, T uy00p, T uy00m, T uypm0, T uy0mp, T uz000,
                                     T uzp00, T uzm00, T uz0p0, T uz0m0, T uz00p, T uz00m, T uzp0m, T uz0pm)
{
    const T qox = READ_Q ? ld_stream(a.qx + c, NT) : T(0);
    const T qoy = READ_Q && !(g.wall && j == g.ny - 1) ? ld_stream(a.qy + c, NT) : T(0);
    const T qoz = READ_Q ? ld_stream(a.qz + c, NT) : T(0);
    momentum_core<T, READ_Q, WRITE_V, NT>(a, g, c, j, row_metrics(a, jj), qox, qoy, qoz, ux000, uxp00, uxm00, ux0p0,
                                          ux0m0, ux00p, ux00m, uxmp0, uxm0p, uy000, uyp00, uym00, uy0p0, uy0m0, uy00p,
                                          uy00m, uypm0, uy0mp, uz000, uzp00, uzm00, uz0p0, uz0m0, uz00p, uz00m, uzp0m,
                                          uz0pm);
}

template <typename T, bool READ_Q, bool WRITE_V>
__global__ void __launch_bounds__(256) k_momentum(MomentumArgs<T> a, FVGeom g)
{
    resolve_dt(a);
    int bx, by, bz;
    xcd_block3(bx, by, bz);
    const int i = bx * 64 + threadIdx.x;
    const int j = by * 4 + threadIdx.y;
    const int k = bz;
    if (i >= g.nx || j >= g.ny) return;
    const int ip = wrapp(i, g.nx), im = wrapm(i, g.nx);
    const int kp = wrapp(k, g.nz), km = wrapm(k, g.nz);
    const int jj = j + 1;
    const T m1 = T(-1);
    const T *ux = a.ux, *uy = a.uy, *uz = a.uz;

    const T ux000 = ux[IX(i, j, k, g)], uxp00 = ux[IX(ip, j, k, g)], uxm00 = ux[IX(im, j, k, g)];
    const T ux0p0 = ldc(ux, i, jj + 1, k, g, m1), ux0m0 = ldc(ux, i, jj - 1, k, g, m1);
    const T ux00p = ux[IX(i, j, kp, g)], ux00m = ux[IX(i, j, km, g)];
    const T uxmp0 = ldc(ux, im, jj + 1, k, g, m1), uxm0p = ux[IX(im, j, kp, g)];

    const T uy000 = ldf(uy, i, jj, k, g), uyp00 = ldf(uy, ip, jj, k, g), uym00 = ldf(uy, im, jj, k, g);
    const T uy0p0 = ldf(uy, i, jj + 1, k, g), uy0m0 = ldf(uy, i, jj - 1, k, g);
    const T uy00p = ldf(uy, i, jj, kp, g), uy00m = ldf(uy, i, jj, km, g);
    const T uypm0 = ldf(uy, ip, jj - 1, k, g), uy0mp = ldf(uy, i, jj - 1, kp, g);

    const T uz000 = uz[IX(i, j, k, g)], uzp00 = uz[IX(ip, j, k, g)], uzm00 = uz[IX(im, j, k, g)];
    const T uz0p0 = ldc(uz, i, jj + 1, k, g, m1), uz0m0 = ldc(uz, i, jj - 1, k, g, m1);
    const T uz00p = uz[IX(i, j, kp, g)], uz00m = uz[IX(i, j, km, g)];
    const T uzp0m = uz[IX(ip, j, km, g)], uz0pm = ldc(uz, i, jj + 1, km, g, m1);

    momentum_cell<T, READ_Q, WRITE_V>(a, g, IX(i, j, k, g), j, jj, ux000, uxp00, uxm00, ux0p0, ux0m0, ux00p, ux00m,
                                      uxmp0, uxm0p, uy000, uyp00, uym00, uy0p0, uy0m0, uy00p, uy00m, uypm0, uy0mp,
                                      uz000, uzp00, uzm00, uz0p0, uz0m0, uz00p, uz00m, uzp0m, uz0pm);
}

// Slab-ordered variant of k_momentum: blocks are ordered x-fastest within
// y-slabs of `nbys` block rows, then z, then slab, and each XCD takes a
// contiguous run of that order, so one XCD streams through consecutive z
// planes of one slab whose three-plane u window fits its L2; q and u′ are
// streamed non-temporally.  1D grid of nbx · nbys · nz · nslab blocks.
template <typename T, bool READ_Q, bool WRITE_V>
__global__ void __launch_bounds__(256) k_momentum_slab(MomentumArgs<T> a, FVGeom g, int nbx, int nbys)
{
    resolve_dt(a);
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    int t = lb;
    const int bx = t % nbx;
    t /= nbx;
    const int byl = t % nbys;
    t /= nbys;
    const int k = t % g.nz, slab = t / g.nz;
    const int i = bx * 64 + threadIdx.x;
    const int j = (slab * nbys + byl) * 4 + threadIdx.y;
    if (i >= g.nx || j >= g.ny) return;
    const int ip = wrapp(i, g.nx), im = wrapm(i, g.nx);
    const int kp = wrapp(k, g.nz), km = wrapm(k, g.nz);
    const int jj = j + 1;
    const T m1 = T(-1);
    const T *ux = a.ux, *uy = a.uy, *uz = a.uz;

    const T ux000 = ux[IX(i, j, k, g)], uxp00 = ux[IX(ip, j, k, g)], uxm00 = ux[IX(im, j, k, g)];
    const T ux0p0 = ldc(ux, i, jj + 1, k, g, m1), ux0m0 = ldc(ux, i, jj - 1, k, g, m1);
    const T ux00p = ux[IX(i, j, kp, g)], ux00m = ux[IX(i, j, km, g)];
    const T uxmp0 = ldc(ux, im, jj + 1, k, g, m1), uxm0p = ux[IX(im, j, kp, g)];

    const T uy000 = ldf(uy, i, jj, k, g), uyp00 = ldf(uy, ip, jj, k, g), uym00 = ldf(uy, im, jj, k, g);
    const T uy0p0 = ldf(uy, i, jj + 1, k, g), uy0m0 = ldf(uy, i, jj - 1, k, g);
    const T uy00p = ldf(uy, i, jj, kp, g), uy00m = ldf(uy, i, jj, km, g);
    const T uypm0 = ldf(uy, ip, jj - 1, k, g), uy0mp = ldf(uy, i, jj - 1, kp, g);

    const T uz000 = uz[IX(i, j, k, g)], uzp00 = uz[IX(ip, j, k, g)], uzm00 = uz[IX(im, j, k, g)];
    const T uz0p0 = ldc(uz, i, jj + 1, k, g, m1), uz0m0 = ldc(uz, i, jj - 1, k, g, m1);
    const T uz00p = uz[IX(i, j, kp, g)], uz00m = uz[IX(i, j, km, g)];
    const T uzp0m = uz[IX(ip, j, km, g)], uz0pm = ldc(uz, i, jj + 1, km, g, m1);

    momentum_cell<T, READ_Q, WRITE_V, true>(a, g, IX(i, j, k, g), j, jj, ux000, uxp00, uxm00, ux0p0, ux0m0, ux00p,
                                            ux00m, uxmp0, uxm0p, uy000, uyp00, uym00, uy0p0, uy0m0, uy00p, uy00m, uypm0,
                                            uy0mp, uz000, uzp00, uzm00, uz0p0, uz0m0, uz00p, uz00m, uzp0m, uz0pm);
}

// LDS z-march with staged plane loads: a block owns a 64 × 8 (x, y) tile
// and marches over `zc` planes, keeping planes k−1, k, k+1 of the three
// components (with a one-cell halo) in LDS.  The 26 neighbours come from
// LDS, so a cell costs ~3 vector-memory instructions for u instead of 27.
// Plane loads are 16-byte vector loads of whole tile rows (plus one scalar
// load per halo column); each thread's rows, ghost-row sources and signs
// are fixed for the whole march, so a plane costs a few adds per load.  The
// loads of plane k+3 are issued before plane k is computed, and plane k+2
// (issued one plane earlier) lands in LDS after it: two planes in flight per
// thread.  The old q values are read before the plane loads.  Rows outside the wall follow the ldc/ldf ghost rules.
// Grid (ceil(nx/64), ceil(ny/8), ceil(nz/zc)) of 256 threads, XCD-ordered;
// LDS 3 planes × 3 components × 10 rows × (64 + 2V) + metric rows.
template <typename T> struct LZ {
    static constexpr int TX = 64, TY = 8, SY = TY + 2;
    static constexpr int V = 16 / sizeof(T);          // elements per 16-byte load
    static constexpr int P = TX + 2 * V;              // LDS row pitch; interior at V.., halos at V−1 and V+w
    static constexpr int LR = TX / V;                 // vector lanes per tile row
    static constexpr int NROW = 3 * SY;               // tile rows per plane (three components)
    static constexpr int RP = 256 / LR;               // tile rows one pass of 256 threads covers
    static constexpr int NI = (NROW + RP - 1) / RP;   // vector items per thread per plane
    static constexpr int PL = NROW * P;               // one plane slot
};

template <typename T, bool READ_Q, bool WRITE_V>
__global__ void __launch_bounds__(256) k_momentum_lz(MomentumArgs<T> a, FVGeom g, int zc)
{
    resolve_dt(a);
    using L = LZ<T>;
    using VT = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
    __shared__ __align__(16) T sm[3 * L::PL];
    __shared__ T met[6][L::TY];
    int bx, by, bz;
    xcd_block3(bx, by, bz);
    const int x0 = bx * L::TX, y0 = by * L::TY, k0 = bz * zc;
    const int kend = min(k0 + zc, g.nz);
    const int w = min(L::TX, g.nx - x0); // tile width (a multiple of V)
    const int tid = threadIdx.x;
    const size_t plane = (size_t)g.nx * g.ny;
    const T *fld[3] = {a.ux, a.uy, a.uz};

    // this thread's vector items: tile row r = r0 + s·RP, lane v (fixed)
    const int v = tid % L::LR, r0 = tid / L::LR;
    const bool vlane = v * L::V < w;
    int voff[L::NI];   // element offset of the source row within a plane (−1: no item)
    T vsgn[L::NI];     // ghost sign (0 clears the row)
    int vdst[L::NI];   // LDS offset within a slot
    const T *vsrc[L::NI];
    auto ghost = [&](int f, int j, int &js, T &sg) {
        // source row and sign of tile row j of component f (ldc / ldf rules)
        sg = T(1);
        js = j;
        if (j < 0) {
            if (g.wall) {
                js = 0;
                sg = f == 1 ? T(0) : T(-1);
            } else js = g.ny - 1;
        } else if (j >= g.ny) {
            if (j > g.ny) {
                js = g.ny - 1;
                sg = T(0); // past the first ghost row: never read
            } else if (g.wall) {
                js = f == 1 ? g.ny - 2 : g.ny - 1;
                sg = T(-1);
            } else js = 0;
        }
    };
#pragma unroll
    for (int s2 = 0; s2 < L::NI; ++s2) {
        const int r = r0 + s2 * L::RP;
        voff[s2] = -1;
        vsgn[s2] = T(0);
        vdst[s2] = 0;
        vsrc[s2] = a.ux;
        if (r < L::NROW && vlane) {
            const int f = r / L::SY, ly = r - f * L::SY;
            int js;
            T sg;
            ghost(f, y0 - 1 + ly, js, sg);
            voff[s2] = js * g.nx + x0 + v * L::V;
            vsgn[s2] = sg;
            vdst[s2] = r * L::P + L::V + v * L::V;
            vsrc[s2] = fld[f];
        }
    }
    // halo columns: threads 0 .. 2·NROW−1, one (row, side) each
    const bool hlane = tid < 2 * L::NROW;
    int hoff = 0, hdst = 0;
    T hsgn = T(0);
    const T *hsrc = a.ux;
    if (hlane) {
        const int r = tid >> 1, side = tid & 1;
        const int f = r / L::SY, ly = r - f * L::SY;
        int js;
        ghost(f, y0 - 1 + ly, js, hsgn);
        const int i = side ? (x0 + w == g.nx ? 0 : x0 + w) : (x0 == 0 ? g.nx - 1 : x0 - 1);
        hoff = js * g.nx + i;
        hdst = r * L::P + (side ? L::V + w : L::V - 1);
        hsrc = fld[f];
    }
    // row metrics of the tile rows
    if (tid < L::TY) {
        const int j = min(y0 + tid, g.ny - 1), jj = j + 1;
        met[0][tid] = a.dy[jj];
        met[1][tid] = a.dy[jj + 1];
        met[2][tid] = a.rdy[jj];
        met[3][tid] = a.rdy[jj + 1];
        met[4][tid] = a.rdyu[jj];
        met[5][tid] = a.rdyu[jj - 1];
    }

    // two register sets alternate: while plane k is computed, planes k+2
    // (issued one plane earlier) and k+3 are in flight
    VT vA[L::NI], vB[L::NI];
    T hA = T(0), hB = T(0);
    auto issue = [&](int k, VT (&vr)[L::NI], T &hr) {
        const int kw = k < 0 ? k + g.nz : (k >= g.nz ? k - g.nz : k);
        const size_t pk = plane * kw;
#pragma unroll
        for (int s2 = 0; s2 < L::NI; ++s2)
            if (voff[s2] >= 0) vr[s2] = *reinterpret_cast<const VT *>(vsrc[s2] + pk + voff[s2]);
        if (hlane) hr = hsrc[pk + hoff];
    };
    auto land = [&](int slot, VT (&vr)[L::NI], T hr) {
        T *dst = sm + slot * L::PL;
#pragma unroll
        for (int s2 = 0; s2 < L::NI; ++s2) {
            if (voff[s2] >= 0) {
                VT x = vr[s2];
                T *e = reinterpret_cast<T *>(&x);
#pragma unroll
                for (int q = 0; q < L::V; ++q) e[q] = vsgn[s2] == T(0) ? T(0) : vsgn[s2] * e[q];
                *reinterpret_cast<VT *>(dst + vdst[s2]) = x;
            }
        }
        if (hlane) dst[hdst] = hsgn == T(0) ? T(0) : hsgn * hr;
    };

    issue(k0 - 1, vA, hA);
    land(0, vA, hA);
    issue(k0, vA, hA);
    land(1, vA, hA);
    issue(k0 + 1, vA, hA);
    land(2, vA, hA);
    if (k0 + 1 < kend) issue(k0 + 2, vA, hA);
    __syncthreads();
    const int tx = tid & 63, ty = __builtin_amdgcn_readfirstlane(tid >> 6);
    int sM = 0, s0 = 1, sP = 2;
    // one plane: old q values, the loads of plane k+3, the update of plane k,
    // then plane k+2 (in cur) lands in the slot plane k−1 leaves
    auto plane_step = [&](int k, VT (&cur)[L::NI], T &curh, VT (&nxt)[L::NI], T &nxth) {
        const int i = x0 + tx;
        T qo[2][3];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int j = y0 + ty + 4 * h;
            const size_t c = (size_t)i + (size_t)g.nx * j + plane * k;
            qo[h][0] = qo[h][1] = qo[h][2] = T(0);
            if (READ_Q && tx < w && j < g.ny) {
                qo[h][0] = ld_stream(a.qx + c, true);
                if (!(g.wall && j == g.ny - 1)) qo[h][1] = ld_stream(a.qy + c, true);
                qo[h][2] = ld_stream(a.qz + c, true);
            }
        }
        if (k + 2 < kend) issue(k + 3, nxt, nxth);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int ly = ty + 4 * h, j = y0 + ly;
            if (tx < w && j < g.ny) {
                const int cc = (ly + 1) * L::P + L::V + tx;
                const T *X0 = sm + s0 * L::PL, *Y0 = X0 + L::SY * L::P, *Z0 = Y0 + L::SY * L::P;
                const T *XP = sm + sP * L::PL, *YP = XP + L::SY * L::P, *ZP = YP + L::SY * L::P;
                const T *XM = sm + sM * L::PL, *YM = XM + L::SY * L::P, *ZM = YM + L::SY * L::P;
                const RowMetrics<T> m{met[0][ly], met[1][ly], met[2][ly], met[3][ly], met[4][ly], met[5][ly]};
                const int P = L::P;
                momentum_core<T, READ_Q, WRITE_V, true>(
                    a, g, IX(i, j, k, g), j, m, qo[h][0], qo[h][1], qo[h][2], X0[cc], X0[cc + 1], X0[cc - 1],
                    X0[cc + P], X0[cc - P], XP[cc], XM[cc], X0[cc - 1 + P], XP[cc - 1], Y0[cc], Y0[cc + 1],
                    Y0[cc - 1], Y0[cc + P], Y0[cc - P], YP[cc], YM[cc], Y0[cc + 1 - P], YP[cc - P], Z0[cc],
                    Z0[cc + 1], Z0[cc - 1], Z0[cc + P], Z0[cc - P], ZP[cc], ZM[cc], ZM[cc + 1], ZM[cc + P]);
            }
        }
        __syncthreads(); // every wave is done with slot sM
        if (k + 1 < kend) {
            land(sM, cur, curh);
            const int t = sM;
            sM = s0;
            s0 = sP;
            sP = t;
        }
        __syncthreads();
    };
    for (int k = k0; k < kend; k += 2) {
        plane_step(k, vA, hA, vB, hB);
        if (k + 1 < kend) plane_step(k + 1, vB, hB, vA, hA);
    }
}

// ------------------------------------------------------------------------
// x pass, forward: builds the real x-lines (the backward-difference
// divergence of src/operators.jl:10-26, MODE 0, or a plain field copy,
// MODE 1), transforms them with a half-length complex FFT in LDS and
// unpacks the nx/2+1 rfft coefficients into A(kx, j, k).
// Lines L = j + ny k; `lpb` lines per block; d.n = nx/2.
// twx[t] = exp(-2πi t / nx), t = 0..nx/2.
// ------------------------------------------------------------------------
template <typename T> struct XPassArgs {
    const T *ux, *uy, *uz; // MODE 0: velocity; MODE 1: ux is the field
    cplx<T> *A;
    const T *rdy;
    T rdx, rdz;
    const cplx<T> *tw;  // length nx/2
    const cplx<T> *twx; // length nx/2+1
    int lpb;
    FastDiv fN, fnxc, fny; // division by nx/2, nx/2+1 and ny
};

template <typename T, int EPT, int MODE>
__global__ void __launch_bounds__(256) k_rfft_x(XPassArgs<T> a, FVGeom g, FFTDesc d)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
    const int N = d.n, lpb = a.lpb;
    const int nlines = g.ny * g.nz;
    const int L0 = xcd_block(blockIdx.x, gridDim.x) * lpb;
    const cplx<T> *twl = stage_twiddles<T>(buf + (size_t)lpb * N, a.tw, N);
    batched_for<FV_LOAD_U(EPT), cplx<T>>(
        lpb * N,
        [&](int idx) {
            int b = fdiv(idx, a.fN), t = idx - b * N;
            int L = L0 + b;
            cplx<T> z = mk<T>(T(0), T(0));
            if (L < nlines) {
                int k = fdiv(L, a.fny), j = L - k * g.ny;
                int i0 = 2 * t, i1 = 2 * t + 1;
                if (MODE == 0) {
                    // cells i0, i1 = i0 + 1 as one pair load per row (i0 even)
                    const int km = wrapm(k, g.nz), jj = j + 1;
                    const T rdyj = a.rdy[jj];
                    const int c0 = IX(i0, j, k, g);
                    const cplx<T> X = *reinterpret_cast<const cplx<T> *>(a.ux + c0);
                    const T xm = a.ux[IX(wrapm(i0, g.nx), j, k, g)];
                    const cplx<T> Y0 = *reinterpret_cast<const cplx<T> *>(a.uy + c0);
                    // face below: j − 1, the wall (0) or the periodic image ny − 1
                    cplx<T> Ym = mk<T>(T(0), T(0));
                    if (j > 0) Ym = *reinterpret_cast<const cplx<T> *>(a.uy + c0 - g.nx);
                    else if (!g.wall) Ym = *reinterpret_cast<const cplx<T> *>(a.uy + IX(i0, g.ny - 1, k, g));
                    const cplx<T> Z0 = *reinterpret_cast<const cplx<T> *>(a.uz + c0);
                    const cplx<T> Zm = *reinterpret_cast<const cplx<T> *>(a.uz + IX(i0, j, km, g));
                    const T y0 = Y0.x - Ym.x, y1 = Y0.y - Ym.y;
                    const T z0 = Z0.x - Zm.x, z1 = Z0.y - Zm.y;
                    z.x = (X.x - xm) * a.rdx + y0 * rdyj + z0 * a.rdz;
                    z.y = (X.y - X.x) * a.rdx + y1 * rdyj + z1 * a.rdz;
                } else {
                    z.x = a.ux[IX(i0, j, k, g)];
                    z.y = a.ux[IX(i1, j, k, g)];
                }
            }
            return z;
        },
        [&](int idx, cplx<T> v) { buf[idx] = v; });
    __syncthreads();
    fft_lds<T, EPT, false>(buf, lpb, N, d, twl);
    const int nxc = N + 1;
    for (int idx = threadIdx.x; idx < lpb * nxc; idx += blockDim.x) {
        int b = fdiv(idx, a.fnxc), kx = idx - b * nxc;
        int L = L0 + b;
        if (L >= nlines) continue;
        cplx<T> Zk = buf[b * N + (kx == N ? 0 : kx)];
        cplx<T> Zc = cconj<T>(buf[b * N + (kx == 0 ? 0 : N - kx)]);
        cplx<T> s = cadd<T>(Zk, Zc), df = csub<T>(Zk, Zc);
        // X = (Zk + Zc)/2 − (i/2) w^kx (Zk − Zc)
        cplx<T> wd = cmul<T>(a.twx[kx], df);
        cplx<T> X = mk<T>(T(0.5) * (s.x + wd.y), T(0.5) * (s.y - wd.x));
        a.A[kx + (size_t)g.ldc * L] = X;
    }
}

// ------------------------------------------------------------------------
// x pass, backward (c2r): A(kx, j, k) → real lines out(i, j, k) as the
// unnormalised backward real transform (FFTW c2r semantics: imaginary parts
// of the kx = 0 and kx = nx/2 coefficients are ignored).
// ------------------------------------------------------------------------
template <typename T> struct XInvArgs {
    const cplx<T> *A;
    T *out;
    const cplx<T> *tw;  // length nx/2
    const cplx<T> *twx; // length nx/2+1
    int lpb;
    FastDiv fN; // division by nx/2
    // device Δt of the CFL loop (null otherwise): out is left as it is when
    // Δt is 0 (after the loop reached its end time)
    const double *skip_dt;
};

template <typename T, int EPT>
__global__ void __launch_bounds__(256) k_irfft_x(XInvArgs<T> a, FVGeom g, FFTDesc d)
{
    if (a.skip_dt && *a.skip_dt == 0.0) return; // uniform over the grid
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
    const int N = d.n, lpb = a.lpb;
    const int nlines = g.ny * g.nz;
    const int L0 = xcd_block(blockIdx.x, gridDim.x) * lpb;
    const cplx<T> *twl = stage_twiddles<T>(buf + (size_t)lpb * N, a.tw, N);
    batched_for<FV_LOAD_U(EPT), cplx<T>>(
        lpb * N,
        [&](int idx) {
            int b = fdiv(idx, a.fN), k = idx - b * N;
            int L = L0 + b;
            cplx<T> Z = mk<T>(T(0), T(0));
            if (L < nlines) {
                cplx<T> Xk = a.A[k + (size_t)g.ldc * L];
                cplx<T> Xn = a.A[(N - k) + (size_t)g.ldc * L];
                if (k == 0) {
                    Xk.y = T(0);
                    Xn.y = T(0);
                }
                cplx<T> Xc = cconj<T>(Xn);
                cplx<T> ev = cadd<T>(Xk, Xc);                     // 2 Ev
                cplx<T> od = cmulc<T>(csub<T>(Xk, Xc), a.twx[k]); // 2 Od = w^-k (Xk − Xc)
                // Z = Ev + i Od, and brfft = 2 IDFT(Z): fold the 2 in
                Z = mk<T>(ev.x - od.y, ev.y + od.x);
            }
            return Z;
        },
        [&](int idx, cplx<T> v) { buf[idx] = v; });
    __syncthreads();
    fft_lds<T, EPT, true>(buf, lpb, N, d, twl);
    for (int idx = threadIdx.x; idx < lpb * N; idx += blockDim.x) {
        int b = fdiv(idx, a.fN), t = idx - b * N;
        int L = L0 + b;
        if (L >= nlines) continue;
        cplx<T> z = buf[idx];
        reinterpret_cast<cplx<T> *>(a.out + (size_t)g.nx * L)[t] = z;
    }
}

// ------------------------------------------------------------------------
// x passes with the transforms held by wavefront lane groups (nx/2 = N in
// 16..256): G = N/4 lanes per line, 64/G lines per wavefront, 4 wavefronts
// per block; the transform runs in registers with the group's own LDS line
// for the exchanges (group_fft), so no workgroup barrier follows the twiddle
// staging.  Same arithmetic as k_rfft_x<MODE 0> / k_irfft_x with the radix
// order 4·…·4(·2) instead of 8·…
// LDS: the per-pass transform twiddles (stage_group_twiddles, at most N) and
// the N+1 real-transform twiddles, then N complex per line (256/G lines per
// block).
// ------------------------------------------------------------------------
template <typename T, int N>
__global__ void __launch_bounds__(256) k_rfft_xw(XPassArgs<T> a, FVGeom g)
{
    constexpr int G = N / 4, LPB = 256 / G;
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *tw = reinterpret_cast<cplx<T> *>(smem_raw);
    cplx<T> *twx = tw + N;
    const int t = threadIdx.x % G, lb = threadIdx.x / G;
    cplx<T> *sh = twx + (N + 1) + lb * N;
    stage_group_twiddles<T, N>(tw, a.tw);
    for (int q = threadIdx.x; q <= N; q += blockDim.x) twx[q] = a.twx[q];
    const int nlines = g.ny * g.nz;
    const int L = xcd_block(blockIdx.x, gridDim.x) * LPB + lb;
    const bool live = L < nlines;
    const int LL = live ? L : 0;
    const int k = fdiv(LL, a.fny), j = LL - k * g.ny;
    const int km = wrapm(k, g.nz);
    const T rdyj = a.rdy[j + 1];
    cplx<T> z[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        // divergence of cells i0 = 2(t + rG) and i0 + 1 (one pair load per row)
        const int i0 = 2 * (t + r * G);
        const int c0 = IX(i0, j, k, g);
        const cplx<T> X = *reinterpret_cast<const cplx<T> *>(a.ux + c0);
        const T xm = a.ux[IX(wrapm(i0, g.nx), j, k, g)];
        const cplx<T> Y0 = *reinterpret_cast<const cplx<T> *>(a.uy + c0);
        // face below: j − 1, the wall (0) or the periodic image ny − 1
        cplx<T> Ym = mk<T>(T(0), T(0));
        if (j > 0) Ym = *reinterpret_cast<const cplx<T> *>(a.uy + c0 - g.nx);
        else if (!g.wall) Ym = *reinterpret_cast<const cplx<T> *>(a.uy + IX(i0, g.ny - 1, k, g));
        const cplx<T> Z0 = *reinterpret_cast<const cplx<T> *>(a.uz + c0);
        const cplx<T> Zm = *reinterpret_cast<const cplx<T> *>(a.uz + IX(i0, j, km, g));
        const T y0 = Y0.x - Ym.x, y1 = Y0.y - Ym.y;
        const T z0 = Z0.x - Zm.x, z1 = Z0.y - Zm.y;
        z[r].x = (X.x - xm) * a.rdx + y0 * rdyj + z0 * a.rdz;
        z[r].y = (X.y - X.x) * a.rdx + y1 * rdyj + z1 * a.rdz;
    }
    __syncthreads(); // the twiddles are staged
    group_fft<T, false, N, true>(z, sh, tw, t);
#pragma unroll
    for (int r = 0; r < 4; ++r) sh[t + r * G] = z[r];
    wave_lds_sync();
    if (live) {
        cplx<T> *out = a.A + (size_t)g.ldc * L;
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            // kx = t + rG for r < 4, and kx = N (r = 4, lane t = 0)
            const int kx = r < 4 ? t + r * G : N;
            if (r == 4 && t != 0) continue;
            const cplx<T> Zk = r < 4 ? z[r] : z[0];
            const cplx<T> Zc = cconj<T>(sh[(N - kx) & (N - 1)]);
            const cplx<T> sm = cadd<T>(Zk, Zc), df = csub<T>(Zk, Zc);
            // X = (Zk + Zc)/2 − (i/2) w^kx (Zk − Zc)
            const cplx<T> wd = cmul<T>(twx[kx], df);
            out[kx] = mk<T>(T(0.5) * (sm.x + wd.y), T(0.5) * (sm.y - wd.x));
        }
    }
}

template <typename T, int N>
__global__ void __launch_bounds__(256) k_irfft_xw(XInvArgs<T> a, FVGeom g)
{
    if (a.skip_dt && *a.skip_dt == 0.0) return; // uniform over the grid
    constexpr int G = N / 4, LPB = 256 / G;
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *tw = reinterpret_cast<cplx<T> *>(smem_raw);
    cplx<T> *twx = tw + N;
    const int t = threadIdx.x % G, lb = threadIdx.x / G;
    cplx<T> *sh = twx + (N + 1) + lb * N;
    stage_group_twiddles<T, N>(tw, a.tw);
    for (int q = threadIdx.x; q <= N; q += blockDim.x) twx[q] = a.twx[q];
    const int nlines = g.ny * g.nz;
    const int L = xcd_block(blockIdx.x, gridDim.x) * LPB + lb;
    const bool live = L < nlines;
    const cplx<T> *row = a.A + (size_t)g.ldc * (live ? L : 0);
    cplx<T> xk[4], xn[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int k = t + r * G;
        xk[r] = row[k];
        xn[r] = row[N - k];
    }
    __syncthreads(); // the twiddles are staged
    cplx<T> z[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int k = t + r * G;
        cplx<T> Xk = xk[r], Xn = xn[r];
        if (k == 0) {
            Xk.y = T(0);
            Xn.y = T(0);
        }
        const cplx<T> Xc = cconj<T>(Xn);
        const cplx<T> ev = cadd<T>(Xk, Xc);                  // 2 Ev
        const cplx<T> od = cmulc<T>(csub<T>(Xk, Xc), twx[k]); // 2 Od = w^-k (Xk − Xc)
        z[r] = mk<T>(ev.x - od.y, ev.y + od.x);
    }
    group_fft<T, true, N, true>(z, sh, tw, t);
    if (live) {
        cplx<T> *out = reinterpret_cast<cplx<T> *>(a.out + (size_t)g.nx * L);
#pragma unroll
        for (int r = 0; r < 4; ++r) out[t + r * G] = z[r];
    }
}

// ------------------------------------------------------------------------
// z pass: in-place complex FFT along z of A(kx, j, kz) for a tile of `tb`
// consecutive kx at fixed j (coalesced tb·16-byte row segments).
// ------------------------------------------------------------------------
// NT: the block size (256 or 512).  The line tile walks of fft_lds.hpp
// (tile_load / tile_store).
template <typename T, int EPT, bool INV, int NT>
__global__ void __launch_bounds__(512) k_fft_z(cplx<T> *A, const cplx<T> *tw, int tb, int tbf, FastDiv ftb, FVGeom g,
                                                  FFTDesc d)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
    const int n = d.n, ls = n + 1;
    const cplx<T> *twl = stage_tile_twiddles<T>(buf + (size_t)tb * ls, tw, n, tb);
    int bxr, byr;
    xcd_block2(bxr, byr);
    const int kx0 = bxr * tb, j = byr, nb = min(tb, g.nxc - kx0), stride = g.ldc * g.ny;
    cplx<T> *line0 = A + CX(kx0, j, 0, g);
    tile_load<T, EPT, NT, FV_LOAD_U(EPT)>(buf, ls, line0, stride, nb, tb, ftb, n);
    __syncthreads();
    fft_lds_tile<T, EPT, INV, 0, true>(buf, tb, ls, d, twl, tbf);
    tile_store<T, EPT, NT>(buf, ls, line0, stride, nb, tb, ftb, n);
}

// ------------------------------------------------------------------------
// y solve, periodic y (src/poisson.jl:43-49, 136-138): fft y, divide by
// λ1+λ2+λ3 (mean mode → 0), ifft y, all in LDS for a tile of `tb`
// consecutive kx at fixed kz.  `scale` folds the 1/(nx ny nz)
// normalisation of the inverse transforms.  With DIVIDE = false only the
// forward y transform is applied (spectrum sampling).
// ------------------------------------------------------------------------
template <typename T, int EPT, bool DIVIDE>
__global__ void __launch_bounds__(512) k_fourier_y(cplx<T> *A, const cplx<T> *tw, const double *lam1,
                                                   const double *lam2, const double *lam3, double scale, int tb,
                                                   int tbf, FastDiv ftb, FVGeom g, FFTDesc d)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
    const int n = d.n, ls = n + 1;
    const cplx<T> *twl = stage_tile_twiddles<T>(buf + (size_t)tb * ls, tw, n, tb);
    int bxr, byr;
    xcd_block2(bxr, byr);
    const int kx0 = bxr * tb, kz = byr;
    batched_for<FV_LOAD_U(EPT), cplx<T>>(
        tb * n,
        [&](int idx) {
            int t = fdiv(idx, ftb), kx = kx0 + (idx - t * tb);
            return kx < g.nxc ? A[CX(kx, t, kz, g)] : mk<T>(T(0), T(0));
        },
        [&](int idx, cplx<T> v) {
            int t = fdiv(idx, ftb), b = idx - t * tb;
            buf[b * ls + t] = v;
        });
    __syncthreads();
    fft_lds_tile<T, EPT, false, 0, true>(buf, tb, ls, d, twl, tbf);
    if (DIVIDE) {
        for (int idx = threadIdx.x; idx < tb * n; idx += blockDim.x) {
            int t = fdiv(idx, ftb), b = idx - t * tb;
            int kx = kx0 + b;
            if (kx >= g.nxc) continue;
            double lam = lam1[kx] + lam2[t] + lam3[kz];
            cplx<T> v = buf[b * ls + t];
            double f = lam == 0.0 ? 0.0 : scale / lam;
            buf[b * ls + t] = mk<T>(T((double)v.x * f), T((double)v.y * f));
        }
        __syncthreads();
        fft_lds_tile<T, EPT, true, 0, true>(buf, tb, ls, d, twl, tbf);
    }
    for (int idx = threadIdx.x; idx < tb * n; idx += blockDim.x) {
        int t = fdiv(idx, ftb), b = idx - t * tb;
        int kx = kx0 + b;
        if (kx < g.nxc) A[CX(kx, t, kz, g)] = buf[b * ls + t];
    }
}

// ------------------------------------------------------------------------
// y solve with walls (src/poisson.jl:14-41): batched Thomas for
// (D_y + λ1 + λ3) x = r̂ in FP64, one thread per (kx, kz) line, lines
// flattened as L = kx + nxc kz so neighbouring lanes read neighbouring kx.
// The inverse pivots ip[kx, j, |kz|] = 1/β_j of every line are computed
// once at setup (the operator never changes); the mean mode (λ == 0) pins
// row 1 exactly as the reference does.  The forward sweep stores d' in the
// FP64 scratch D; `scale` folds 1/(nx nz).
// ------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_thomas_y(cplx<T> *__restrict__ A, double2 *__restrict__ D,
                                                  const double *__restrict__ ip, const double *__restrict__ ay,
                                                  const double *__restrict__ cy, const double *__restrict__ lam1,
                                                  const double *__restrict__ lam3, double scale, FVGeom g)
{
    constexpr int CH = 8; // rows loaded ahead of the dependent recurrence
    const int L = blockIdx.x * blockDim.x + threadIdx.x;
    const int kx = L % g.ldc, kz = L / g.ldc;
    if (kx >= g.nxc || kz >= g.nz) return;
    const int kzs = kz <= g.nz - kz ? kz : g.nz - kz;
    const bool mean = (lam1[kx] + lam3[kz]) == 0.0;
    const int ny = g.ny;
    const size_t ipbase = (size_t)kx + (size_t)g.nxc * (size_t)ny * kzs, ips = g.nxc;
    const size_t base = (size_t)kx + (size_t)g.ldc * (size_t)ny * kz, cs = g.ldc;
    // forward sweep: d'_j = (r_j − a_j d'_{j−1}) / β_j
    double2 dp = make_double2(0.0, 0.0);
    for (int j0 = 0; j0 < ny; j0 += CH) {
        cplx<T> xr[CH];
        double wr[CH], ar[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int j = j0 + q;
            if (j < ny) {
                xr[q] = A[base + j * cs];
                wr[q] = ip[ipbase + j * ips];
                ar[q] = ay[j];
            }
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int j = j0 + q;
            if (j < ny) {
                if (j == 0) {
                    dp = mean ? make_double2(0.0, 0.0)
                              : make_double2((double)xr[q].x * scale * wr[q], (double)xr[q].y * scale * wr[q]);
                } else {
                    dp = make_double2(((double)xr[q].x * scale - ar[q] * dp.x) * wr[q],
                                      ((double)xr[q].y * scale - ar[q] * dp.y) * wr[q]);
                }
                D[base + j * cs] = dp;
            }
        }
    }
    // back substitution: x_j = d'_j − c'_j x_{j+1}, c'_j = c_j / β_j
    double2 xn = dp;
    A[base + (size_t)(ny - 1) * cs] = mk<T>(T(xn.x), T(xn.y));
    for (int j1 = ny - 2; j1 >= 0; j1 -= CH) {
        double2 dr[CH];
        double cr[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int j = j1 - q;
            if (j >= 0) {
                dr[q] = D[base + j * cs];
                cr[q] = (j == 0 && mean) ? 0.0 : cy[j] * ip[ipbase + j * ips];
            }
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int j = j1 - q;
            if (j >= 0) {
                xn = make_double2(dr[q].x - cr[q] * xn.x, dr[q].y - cr[q] * xn.y);
                A[base + j * cs] = mk<T>(T(xn.x), T(xn.y));
            }
        }
    }
}

// Segmented variant of the same solve: a block owns TL consecutive kx lines
// of one kz plane, and each line's rows are split into S = blockDim/TL
// segments of at most RMAX rows held in registers by one thread (the data,
// then d', then x, all FP64, with the rows' pivots).  The forward sweep and
// the back substitution -- both first-order affine recurrences
// x_j = α_j x_{j∓1} + β_j -- run segment-parallel: every segment solves its
// rows from a zero inflow while accumulating Π α, one thread per line chains
// the S segment maps through LDS, then every segment adds (Π α) x_in.  HBM
// traffic: r read and x written once, the pivot table read once; LDS holds
// only the y coefficients and the segment maps, so several blocks share a
// CU.  Same pivots and mean-mode rule as k_thomas_y.
// grid (ceil(nxc / TL), nz); LDS 2 ny + 5 blockDim doubles.
template <typename T, int RMAX>
__global__ void __launch_bounds__(512) k_thomas_tile(cplx<T> *__restrict__ A, const double *__restrict__ ip,
                                                     const double *__restrict__ ay, const double *__restrict__ cy,
                                                     const double *__restrict__ lam1, const double *__restrict__ lam3,
                                                     double scale, int TL, FVGeom g)
{
    extern __shared__ unsigned char smem_raw[];
    const int ny = g.ny, S = blockDim.x / TL;
    double2 *Bs = reinterpret_cast<double2 *>(smem_raw); // [S][TL] segment offsets
    double2 *Xin = Bs + blockDim.x;                      // [S][TL] segment inflows
    double *As = reinterpret_cast<double *>(Xin + blockDim.x);
    double *sa = As + blockDim.x, *sc = sa + ny;
    int bxr, byr;
    xcd_block2(bxr, byr);
    const int kx0 = bxr * TL, kz = byr;
    const int kzs = kz <= g.nz - kz ? kz : g.nz - kz;
    const int l = threadIdx.x % TL, sg = threadIdx.x / TL, kx = kx0 + l;
    const bool valid = kx < g.nxc;
    const int R = (ny + S - 1) / S, j0 = min(ny, sg * R), j1 = min(ny, j0 + R);
    cplx<T> *Al = A + (size_t)g.ldc * (size_t)ny * kz + (valid ? kx : 0);
    const double *ipl = ip + (size_t)g.nxc * (size_t)ny * kzs + (valid ? kx : 0);
    double2 x[RMAX];
    double w[RMAX];
#pragma unroll
    for (int q = 0; q < RMAX; ++q) {
        const int j = j0 + q;
        x[q] = make_double2(0.0, 0.0);
        w[q] = 0.0;
        if (valid && j < j1) {
            const cplx<T> a = Al[(size_t)g.ldc * j];
            x[q] = make_double2((double)a.x * scale, (double)a.y * scale);
            w[q] = ipl[(size_t)g.nxc * j];
        }
    }
    for (int j = threadIdx.x; j < ny; j += blockDim.x) {
        sa[j] = ay[j];
        sc[j] = cy[j];
    }
    const bool mean = valid && (lam1[kx] + lam3[kz]) == 0.0;
    __syncthreads();
    // forward sweep d'_j = −a_j ip_j d'_{j−1} + ip_j r_j (row 0: no coupling; 0 for the mean mode)
    {
        double P = 1.0;
        double2 y = make_double2(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < RMAX; ++q) {
            const int j = j0 + q;
            if (j < j1) {
                const double al = j == 0 ? 0.0 : -sa[j] * w[q];
                const double be = (j == 0 && mean) ? 0.0 : w[q];
                y = make_double2(al * y.x + be * x[q].x, al * y.y + be * x[q].y);
                x[q] = y;
                P *= al;
            }
        }
        As[threadIdx.x] = P;
        Bs[threadIdx.x] = y;
    }
    __syncthreads();
    if (sg == 0) {
        double2 xi = make_double2(0.0, 0.0);
        for (int q = 0; q < S; ++q) {
            Xin[q * TL + l] = xi;
            const double a = As[q * TL + l];
            const double2 b = Bs[q * TL + l];
            xi = make_double2(a * xi.x + b.x, a * xi.y + b.y);
        }
    }
    __syncthreads();
    {
        const double2 xi = Xin[threadIdx.x];
        double P = 1.0;
#pragma unroll
        for (int q = 0; q < RMAX; ++q) {
            const int j = j0 + q;
            if (j < j1) {
                P *= j == 0 ? 0.0 : -sa[j] * w[q];
                x[q] = make_double2(x[q].x + P * xi.x, x[q].y + P * xi.y);
            }
        }
    }
    __syncthreads(); // the segment maps of the forward chain are read; reuse As/Bs
    // back substitution x_j = d'_j − c'_j x_{j+1}, c'_j = c_j ip_j (0 on the last row and row 0 of the mean mode)
    {
        double P = 1.0;
        double2 y = make_double2(0.0, 0.0);
#pragma unroll
        for (int q = RMAX - 1; q >= 0; --q) {
            const int j = j0 + q;
            if (j < j1) {
                const double al = (j == ny - 1 || (j == 0 && mean)) ? 0.0 : -sc[j] * w[q];
                y = make_double2(al * y.x + x[q].x, al * y.y + x[q].y);
                x[q] = y;
                P *= al;
            }
        }
        As[threadIdx.x] = P;
        Bs[threadIdx.x] = y;
    }
    __syncthreads();
    if (sg == 0) {
        double2 xi = make_double2(0.0, 0.0);
        for (int q = S - 1; q >= 0; --q) {
            Xin[q * TL + l] = xi;
            const double a = As[q * TL + l];
            const double2 b = Bs[q * TL + l];
            xi = make_double2(a * xi.x + b.x, a * xi.y + b.y);
        }
    }
    __syncthreads();
    const double2 xi = Xin[threadIdx.x];
    double P = 1.0;
#pragma unroll
    for (int q = RMAX - 1; q >= 0; --q) {
        const int j = j0 + q;
        if (j < j1) {
            P *= (j == ny - 1 || (j == 0 && mean)) ? 0.0 : -sc[j] * w[q];
            if (valid) Al[(size_t)g.ldc * j] = mk<T>(T(x[q].x + P * xi.x), T(x[q].y + P * xi.y));
        }
    }
}

// k_thomas_tile for equal segments (ny = S·R): no row needs a bounds test,
// and the rows of r and of the pivot table are read through buffer
// resources (one offset register per array for all R rows, a scalar offset
// per row), all issued before the first use.  Register budget: two blocks
// of 256 threads per CU.  Same recurrences, pivots and mean-mode rule.
// grid (ceil(nxc / TL), nz); LDS 2 ny + 5 blockDim doubles.
template <typename T, int R>
__global__ void __launch_bounds__(256, 2) k_thomas_seg(cplx<T> *__restrict__ A, const double *__restrict__ ip,
                                                       const double *__restrict__ ay, const double *__restrict__ cy,
                                                       const double *__restrict__ lam1, const double *__restrict__ lam3,
                                                       double scale, int TL, FVGeom g)
{
    extern __shared__ unsigned char smem_raw[];
    const int ny = g.ny, S = blockDim.x / TL;
    double2 *Bs = reinterpret_cast<double2 *>(smem_raw); // [S][TL] segment offsets
    double2 *Xin = Bs + blockDim.x;                      // [S][TL] segment inflows
    double *As = reinterpret_cast<double *>(Xin + blockDim.x);
    double *sa = As + blockDim.x, *sc = sa + ny;
    int bxr, byr;
    xcd_block2(bxr, byr);
    const int kx0 = bxr * TL, kz = byr;
    const int kzs = kz <= g.nz - kz ? kz : g.nz - kz;
    const int l = threadIdx.x % TL, sg = threadIdx.x / TL, kx = kx0 + l;
    const bool valid = kx < g.nxc;
    const int kxl = valid ? kx : 0;
    const int j0 = sg * R;
    // one kz plane of A and one |kz| plane of the pivot table per buffer range
    const bufres ra = make_bufres(A + (size_t)g.ldc * ny * kz, (uint32_t)((size_t)g.ldc * ny * sizeof(cplx<T>)));
    const bufres rw = make_bufres(ip + (size_t)g.nxc * ny * kzs, (uint32_t)((size_t)g.nxc * ny * sizeof(double)));
    const int arow = g.ldc * (int)sizeof(cplx<T>), wrow = g.nxc * (int)sizeof(double);
    const int va = (kxl + g.ldc * j0) * (int)sizeof(cplx<T>), vw = (kxl + g.nxc * j0) * (int)sizeof(double);
    double2 x[R];
    double w[R];
    {
        cplx<T> ar[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            ar[q] = bload<cplx<T>>(ra, va, arow * q);
            w[q] = bload<double>(rw, vw, wrow * q);
        }
#pragma unroll
        for (int q = 0; q < R; ++q) x[q] = make_double2((double)ar[q].x * scale, (double)ar[q].y * scale);
    }
    for (int j = threadIdx.x; j < ny; j += blockDim.x) {
        sa[j] = ay[j];
        sc[j] = cy[j];
    }
    const bool mean = (lam1[kxl] + lam3[kz]) == 0.0;
    __syncthreads();
    // forward sweep d'_j = −a_j ip_j d'_{j−1} + ip_j r_j (row 0: no coupling; 0 for the mean mode)
    {
        double P = 1.0;
        double2 y = make_double2(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int j = j0 + q;
            const double al = j == 0 ? 0.0 : -sa[j] * w[q];
            const double be = (j == 0 && mean) ? 0.0 : w[q];
            y = make_double2(al * y.x + be * x[q].x, al * y.y + be * x[q].y);
            x[q] = y;
            P *= al;
        }
        As[threadIdx.x] = P;
        Bs[threadIdx.x] = y;
    }
    __syncthreads();
    if (sg == 0) {
        double2 xi = make_double2(0.0, 0.0);
        for (int q = 0; q < S; ++q) {
            Xin[q * TL + l] = xi;
            const double a = As[q * TL + l];
            const double2 b = Bs[q * TL + l];
            xi = make_double2(a * xi.x + b.x, a * xi.y + b.y);
        }
    }
    __syncthreads();
    {
        const double2 xi = Xin[threadIdx.x];
        double P = 1.0;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int j = j0 + q;
            P *= j == 0 ? 0.0 : -sa[j] * w[q];
            x[q] = make_double2(x[q].x + P * xi.x, x[q].y + P * xi.y);
        }
    }
    __syncthreads(); // the segment maps of the forward chain are read; reuse As/Bs
    // back substitution x_j = d'_j − c'_j x_{j+1}, c'_j = c_j ip_j (0 on the last row and row 0 of the mean mode)
    {
        double P = 1.0;
        double2 y = make_double2(0.0, 0.0);
#pragma unroll
        for (int q = R - 1; q >= 0; --q) {
            const int j = j0 + q;
            const double al = (j == ny - 1 || (j == 0 && mean)) ? 0.0 : -sc[j] * w[q];
            y = make_double2(al * y.x + x[q].x, al * y.y + x[q].y);
            x[q] = y;
            P *= al;
        }
        As[threadIdx.x] = P;
        Bs[threadIdx.x] = y;
    }
    __syncthreads();
    if (sg == 0) {
        double2 xi = make_double2(0.0, 0.0);
        for (int q = S - 1; q >= 0; --q) {
            Xin[q * TL + l] = xi;
            const double a = As[q * TL + l];
            const double2 b = Bs[q * TL + l];
            xi = make_double2(a * xi.x + b.x, a * xi.y + b.y);
        }
    }
    __syncthreads();
    if (!valid) return;
    const double2 xi = Xin[threadIdx.x];
    double P = 1.0;
#pragma unroll
    for (int q = R - 1; q >= 0; --q) {
        const int j = j0 + q;
        P *= (j == ny - 1 || (j == 0 && mean)) ? 0.0 : -sc[j] * w[q];
        bstore(ra, va, arow * q, mk<T>(T(x[q].x + P * xi.x), T(x[q].y + P * xi.y)));
    }
}

// Δt of a CFL-adaptive step on the device (cfl_timestep, src/timestep.jl:69-80,
// the same expression as the host's): from the per-component velocity maxima,
// clamped to the time left before tend; advances the device time and logs Δt.
// Once t >= tend − 1e-10(|tend| + 1) (the end test of solve!,
// src/timestep.jl:102) it writes Δt = 0 and neither advances t nor counts a
// step: the step that follows leaves u unchanged (q = 0, u′ = u, and the
// pressure-gradient pass skips on Δt = 0).  One thread.
__global__ void k_cfl_dt(const double *umax, double dx, double dymin, double dz, double diff, double cfl,
                         double tend, double *t, double *dt, double *dt_log, int *step)
{
    const double eps = 2.220446049250313e-16;
    const double tnow = *t;
    const double tol = isinf(tend) ? 0.0 : 1e-10 * (fabs(tend) + 1); // tend = +inf: no end time
    if (!(tnow < tend - tol)) {
        *dt = 0.0;
        return;
    }
    const double conv = fmin(fmin(dx / (umax[0] + eps), dymin / (umax[1] + eps)), dz / (umax[2] + eps));
    double d = fmin(cfl * conv, cfl * diff);
    d = fmin(d, tend - tnow);
    *dt = d;
    *t = tnow + d;
    dt_log[*step] = d;
    *step += 1;
}

// ------------------------------------------------------------------------
// Pressure-gradient update on active DOFs (src/operators.jl:180-220):
//   u_c ← u_c − s ∂ψ/∂c,  plus (CN) p ← p + ψ/(αΔt)   (src/timestep.jl:51)
// ------------------------------------------------------------------------
template <typename T> struct PGradArgs {
    T *ux, *uy, *uz;
    const T *psi;
    T *p;
    const T *rdyu;
    T rdx, rdz, s, inv_adt;
    // device Δt of the CFL loop (null otherwise): the pass does nothing when
    // it is 0, i.e. after the loop reached its end time
    const double *skip_dt;
};

template <typename T, bool PUPDATE, bool MAXU>
__global__ void __launch_bounds__(256) k_pgrad(PGradArgs<T> a, FVGeom g, double *umax_part)
{
    if (a.skip_dt && *a.skip_dt == 0.0) return; // uniform over the grid
    int bx, by, bz;
    xcd_block3(bx, by, bz);
    const int i = bx * 64 + threadIdx.x;
    const int j = by * 4 + threadIdx.y;
    const int k = bz;
    double m0 = 0, m1 = 0, m2 = 0;
    if (i < g.nx && j < g.ny) {
        const int c = IX(i, j, k, g);
        const T pc = a.psi[c];
        const T pxp = a.psi[IX(wrapp(i, g.nx), j, k, g)];
        const T pzp = a.psi[IX(i, j, wrapp(k, g.nz), g)];
        const T ux = a.ux[c] - a.s * (pxp - pc) * a.rdx;
        a.ux[c] = ux;
        T uy = a.uy[c];
        if (!(g.wall && j == g.ny - 1)) {
            const T pyp = ldc(a.psi, i, j + 2, k, g, T(1));
            uy -= a.s * (pyp - pc) * a.rdyu[j + 1];
            a.uy[c] = uy;
        }
        const T uz = a.uz[c] - a.s * (pzp - pc) * a.rdz;
        a.uz[c] = uz;
        if (PUPDATE) a.p[c] += pc * a.inv_adt;
        m0 = fabs((double)ux);
        m1 = fabs((double)uy);
        m2 = fabs((double)uz);
    }
    if (MAXU) {
        // per-block max |u_c| of the updated field: the CFL bound of the
        // next step (src/timestep.jl:72-76) without another pass over u
        __shared__ double s[3][256];
        const int t = threadIdx.x + 64 * threadIdx.y;
        s[0][t] = m0;
        s[1][t] = m1;
        s[2][t] = m2;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (t < w)
                for (int q = 0; q < 3; ++q) s[q][t] = fmax(s[q][t], s[q][t + w]);
            __syncthreads();
        }
        if (t < 3) {
            const int lb = bx + gridDim.x * (by + gridDim.y * bz);
            umax_part[(size_t)lb * 3 + t] = s[t][0];
        }
    }
}

// k_pgrad with V = 16 / sizeof(T) consecutive x cells per thread, so that
// every stream moves as 16-byte vector accesses (a quarter or half of the
// memory instructions); the x+1 neighbour of the last cell is one scalar
// load.  Same arithmetic per cell.  Needs nx % V == 0.
// Block (64, 4), grid (ceil(nx / (64 V)), ceil(ny / 4), nz).
template <typename T, bool PUPDATE, bool MAXU>
__global__ void __launch_bounds__(256) k_pgrad_vec(PGradArgs<T> a, FVGeom g, double *umax_part)
{
    if (a.skip_dt && *a.skip_dt == 0.0) return; // uniform over the grid
    constexpr int V = 16 / sizeof(T);
    using VT = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
    int bx, by, bz;
    xcd_block3(bx, by, bz);
    const int i0 = (bx * 64 + threadIdx.x) * V;
    const int j = by * 4 + threadIdx.y;
    const int k = bz;
    double m0 = 0, m1 = 0, m2 = 0;
    if (i0 < g.nx && j < g.ny) {
        const int c = IX(i0, j, k, g);
        const bool active = !(g.wall && j == g.ny - 1);
        // y+1 row of ψ (ghost rule of ldc with s = +1; unused on the inactive wall row)
        const int jp = j + 1 < g.ny ? j + 1 : (g.wall ? g.ny - 1 : 0);
        VT pcv = *reinterpret_cast<const VT *>(a.psi + c);
        const T pnext = a.psi[IX(i0 + V == g.nx ? 0 : i0 + V, j, k, g)];
        VT pzv = *reinterpret_cast<const VT *>(a.psi + IX(i0, j, wrapp(k, g.nz), g));
        VT pyv = *reinterpret_cast<const VT *>(a.psi + IX(i0, jp, k, g));
        VT uxv = *reinterpret_cast<const VT *>(a.ux + c);
        VT uyv = *reinterpret_cast<const VT *>(a.uy + c);
        VT uzv = *reinterpret_cast<const VT *>(a.uz + c);
        VT ppv{};
        if (PUPDATE) ppv = *reinterpret_cast<const VT *>(a.p + c);
        const T *pc = reinterpret_cast<const T *>(&pcv), *pz = reinterpret_cast<const T *>(&pzv),
                *py = reinterpret_cast<const T *>(&pyv);
        T *ux = reinterpret_cast<T *>(&uxv), *uy = reinterpret_cast<T *>(&uyv), *uz = reinterpret_cast<T *>(&uzv),
          *pp = reinterpret_cast<T *>(&ppv);
        const T rdy = a.rdyu[j + 1];
#pragma unroll
        for (int q = 0; q < V; ++q) {
            const T pxp = q + 1 < V ? pc[q + 1] : pnext;
            ux[q] = ux[q] - a.s * (pxp - pc[q]) * a.rdx;
            if (active) uy[q] -= a.s * (py[q] - pc[q]) * rdy;
            uz[q] = uz[q] - a.s * (pz[q] - pc[q]) * a.rdz;
            if (PUPDATE) pp[q] += pc[q] * a.inv_adt;
            m0 = fmax(m0, fabs((double)ux[q]));
            m1 = fmax(m1, fabs((double)uy[q]));
            m2 = fmax(m2, fabs((double)uz[q]));
        }
        *reinterpret_cast<VT *>(a.ux + c) = uxv;
        if (active) *reinterpret_cast<VT *>(a.uy + c) = uyv;
        *reinterpret_cast<VT *>(a.uz + c) = uzv;
        if (PUPDATE) *reinterpret_cast<VT *>(a.p + c) = ppv;
    }
    if (MAXU) {
        __shared__ double s[3][256];
        const int t = threadIdx.x + 64 * threadIdx.y;
        s[0][t] = m0;
        s[1][t] = m1;
        s[2][t] = m2;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (t < w)
                for (int q = 0; q < 3; ++q) s[q][t] = fmax(s[q][t], s[q][t + w]);
            __syncthreads();
        }
        if (t < 3) {
            const int lb = bx + gridDim.x * (by + gridDim.y * bz);
            umax_part[(size_t)lb * 3 + t] = s[t][0];
        }
    }
}

// ------------------------------------------------------------------------
// Semi-implicit y-diffusion (src/ydiffusion.jl:71-154), one velocity
// component per launch, one thread per (i, k) column, all in FP64:
//   v   = uⁿ + B q − αΔt ∂p              (active rows only)
//   rhs = v + θ Dy uⁿ,    θ = ½ αΔtν
//   (I − θ Dy) u* = rhs                   Thomas, c'/β from k_cn_coeffs
// The matrix is the same for every column, so its pivots come from one
// tiny kernel per stage; d' goes to the FP64 scratch D.
// COMP: 0 = x, 1 = y, 2 = z.
// ------------------------------------------------------------------------
template <typename T> struct CNArgs {
    T *u;
    const T *q, *p;
    double *D;
    const double *a, *b, *c;   // Dy rows
    const double *cp, *ip;     // c'_j and 1/β_j
    const double *rdyu;        // 1/Δyu (ghosted index)
    double B, adt, theta, rdx, rdz;
};

template <typename T, int COMP>
__global__ void __launch_bounds__(256) k_cn(CNArgs<T> a, FVGeom g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= g.nx) return;
    const int ny = g.ny;
    const int ip1 = wrapp(i, g.nx), kp1 = wrapp(k, g.nz);
    const int jact = (COMP == 1 && g.wall) ? ny - 1 : ny; // active rows
    auto vrow = [&](int j, double un) -> double {
        if (j >= jact) return un;
        const int c = IX(i, j, k, g);
        double dpdc;
        if (COMP == 0) dpdc = ((double)a.p[IX(ip1, j, k, g)] - (double)a.p[c]) * a.rdx;
        else if (COMP == 1) dpdc = ((double)ldc(a.p, i, j + 2, k, g, T(1)) - (double)a.p[c]) * a.rdyu[j + 1];
        else dpdc = ((double)a.p[IX(i, j, kp1, g)] - (double)a.p[c]) * a.rdz;
        return un + a.B * (double)a.q[c] - a.adt * dpdc;
    };
    double um = 0.0, u0 = (double)a.u[IX(i, 0, k, g)];
    double up = ny > 1 ? (double)a.u[IX(i, 1, k, g)] : 0.0;
    double dp = 0.0;
    for (int j = 0; j < ny; ++j) {
        double lap = a.b[j] * u0;
        if (j > 0) lap += a.a[j] * um;
        if (j < ny - 1) lap += a.c[j] * up;
        double rhs = vrow(j, u0) + a.theta * lap;
        dp = (j == 0) ? rhs * a.ip[0] : (rhs + a.theta * a.a[j] * dp) * a.ip[j];
        a.D[IX(i, j, k, g)] = dp;
        um = u0;
        u0 = up;
        up = (j + 2 < ny) ? (double)a.u[IX(i, j + 2, k, g)] : 0.0;
    }
    double x = dp;
    a.u[IX(i, ny - 1, k, g)] = T(x);
    for (int j = ny - 2; j >= 0; --j) {
        x = a.D[IX(i, j, k, g)] - a.cp[j] * x;
        a.u[IX(i, j, k, g)] = T(x);
    }
}

// Segmented variant of k_cn: a block owns TL consecutive x-lines of one z
// plane; each line's rows are split into S = blockDim/TL segments of at most
// RMAX rows held in registers by one thread (rhs, then d', then x, FP64).
// The forward sweep (d'_j = θ a_j ip_j d'_{j−1} + ip_j rhs_j) and the back
// substitution (x_j = d'_j − c'_j x_{j+1}) run as segment-parallel affine
// recurrences (local solve from zero inflow with Π α, chained per line by
// one thread through LDS, then corrected); the row coefficients are staged
// in LDS.  Each thread loads CH rows of u, p, q ahead of use.  u is updated
// in place: a line's rows are read only by this block, and all reads
// precede the first barrier.
// grid (ceil(nx / TL), nz); LDS 6 ny + 3 blockDim doubles.
template <typename T, int COMP, int RMAX>
__global__ void __launch_bounds__(512) k_cn_tile(CNArgs<T> a, FVGeom g, int TL)
{
    extern __shared__ unsigned char smem_raw[];
    constexpr int CH = 8;
    const int ny = g.ny, S = blockDim.x / TL;
    double *As = reinterpret_cast<double *>(smem_raw), *Bs = As + blockDim.x, *Xin = Bs + blockDim.x;
    double *ca = Xin + blockDim.x, *cb = ca + ny, *cc = cb + ny, *cf = cc + ny, *ci = cf + ny, *cq = ci + ny;
    for (int j = threadIdx.x; j < ny; j += blockDim.x) {
        ca[j] = a.a[j];
        cb[j] = a.b[j];
        cc[j] = a.c[j];
        cf[j] = j == 0 ? 0.0 : a.theta * a.a[j] * a.ip[j]; // forward α
        ci[j] = a.ip[j];
        cq[j] = j == ny - 1 ? 0.0 : -a.cp[j]; // backward α
    }
    const int l = threadIdx.x % TL, sg = threadIdx.x / TL;
    int bxr, byr;
    xcd_block2(bxr, byr);
    const int i = bxr * TL + l, k = byr;
    const bool valid = i < g.nx;
    const int ii = valid ? i : 0; // loads of an idle lane stay in bounds; its results are discarded
    const int ip1 = wrapp(ii, g.nx), kp1 = wrapp(k, g.nz);
    const int jact = (COMP == 1 && g.wall) ? ny - 1 : ny;
    const int R = (ny + S - 1) / S, j0 = min(ny, sg * R), j1 = min(ny, j0 + R);
    auto uat = [&](int j) -> double { return (j >= 0 && j < ny) ? (double)a.u[IX(ii, j, k, g)] : 0.0; };
    // right-hand sides of the segment's rows: rhs_j = u_j + B q_j − αΔt ∂p + θ (Dy u)_j
    double x[RMAX];
    {
        double um = uat(j0 - 1);
#pragma unroll
        for (int c0 = 0; c0 < RMAX; c0 += CH) {
            double ur[CH + 1], vr[CH];
#pragma unroll
            for (int q = 0; q <= CH; ++q) ur[q] = (j0 + c0 + q <= j1) ? uat(j0 + c0 + q) : 0.0;
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                const int j = j0 + c0 + q;
                vr[q] = 0.0;
                if (j < j1 && j < jact) {
                    const int c = IX(ii, j, k, g);
                    double dpdc;
                    if (COMP == 0) dpdc = ((double)a.p[IX(ip1, j, k, g)] - (double)a.p[c]) * a.rdx;
                    else if (COMP == 1) dpdc = ((double)ldc(a.p, ii, j + 2, k, g, T(1)) - (double)a.p[c]) * a.rdyu[j + 1];
                    else dpdc = ((double)a.p[IX(ii, j, kp1, g)] - (double)a.p[c]) * a.rdz;
                    vr[q] = a.B * (double)a.q[c] - a.adt * dpdc;
                }
            }
            if (c0 == 0) __syncthreads(); // the coefficient rows are staged before first use
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                const int j = j0 + c0 + q;
                x[c0 + q] = 0.0;
                if (j < j1) {
                    const double u0 = ur[q], up = ur[q + 1];
                    double lap = cb[j] * u0;
                    if (j > 0) lap += ca[j] * um;
                    if (j < ny - 1) lap += cc[j] * up;
                    x[c0 + q] = u0 + vr[q] + a.theta * lap;
                    um = u0;
                }
            }
        }
    }
    // forward sweep, local to the segment
    {
        double P = 1.0, y = 0.0;
#pragma unroll
        for (int q = 0; q < RMAX; ++q) {
            const int j = j0 + q;
            if (j < j1) {
                y = cf[j] * y + ci[j] * x[q];
                x[q] = y;
                P *= cf[j];
            }
        }
        As[threadIdx.x] = P;
        Bs[threadIdx.x] = y;
    }
    __syncthreads();
    if (sg == 0) {
        double xi = 0.0;
        for (int q = 0; q < S; ++q) {
            Xin[q * TL + l] = xi;
            xi = As[q * TL + l] * xi + Bs[q * TL + l];
        }
    }
    __syncthreads();
    {
        const double xi = Xin[threadIdx.x];
        double P = 1.0;
#pragma unroll
        for (int q = 0; q < RMAX; ++q) {
            const int j = j0 + q;
            if (j < j1) {
                P *= cf[j];
                x[q] += P * xi;
            }
        }
    }
    __syncthreads(); // the forward chain's maps are consumed; reuse As/Bs
    // back substitution, local to the segment
    {
        double P = 1.0, y = 0.0;
#pragma unroll
        for (int q = RMAX - 1; q >= 0; --q) {
            const int j = j0 + q;
            if (j < j1) {
                y = cq[j] * y + x[q];
                x[q] = y;
                P *= cq[j];
            }
        }
        As[threadIdx.x] = P;
        Bs[threadIdx.x] = y;
    }
    __syncthreads();
    if (sg == 0) {
        double xi = 0.0;
        for (int q = S - 1; q >= 0; --q) {
            Xin[q * TL + l] = xi;
            xi = As[q * TL + l] * xi + Bs[q * TL + l];
        }
    }
    __syncthreads();
    const double xi = Xin[threadIdx.x];
    double P = 1.0;
#pragma unroll
    for (int q = RMAX - 1; q >= 0; --q) {
        const int j = j0 + q;
        if (j < j1) {
            P *= cq[j];
            if (valid) a.u[IX(i, j, k, g)] = T(x[q] + P * xi);
        }
    }
}

// k_cn_tile for the common case of equal segments: every line's ny rows
// split into S segments of exactly R rows (ny = S·R), so no row needs a
// bounds test.  Each thread owns NL neighbouring x-lines (NL = 2: pair
// accesses of 8 or 16 bytes, a block row of TL lines is TL·sizeof(T) bytes
// wide) over the R rows of its segment; S = blockDim·NL / TL.  All loads of
// a chunk of CH rows are issued unconditionally (the row past a line end
// reads 0 through the buffer range check, the row before it is clamped, and
// both couplings are cut by zero coefficients), so they are all in flight
// together; the segment's rows stay in registers through both sweeps as in
// k_cn_tile.  Loads go through buffer resources (one offset register per
// array for all rows) and the register budget is two blocks per CU (three
// spilled and measured slower).
// grid (ceil(nx / TL), nz); LDS 7 ny + 3 NL blockDim doubles.
template <typename T, int NL> struct CNVec;
template <typename T> struct CNVec<T, 1> {
    using type = T;
    __device__ static T at(const T &v, int) { return v; }
    __device__ static T make(const double *x) { return T(x[0]); }
};
template <> struct CNVec<float, 2> {
    using type = float2;
    __device__ static float at(const float2 &v, int e) { return e == 0 ? v.x : v.y; }
    __device__ static float2 make(const double *x) { return make_float2((float)x[0], (float)x[1]); }
};
template <> struct CNVec<double, 2> {
    using type = double2;
    __device__ static double at(const double2 &v, int e) { return e == 0 ? v.x : v.y; }
    __device__ static double2 make(const double *x) { return make_double2(x[0], x[1]); }
};

template <typename T, int COMP, int R, int NL>
__global__ void __launch_bounds__(256, 2) k_cn_seg(CNArgs<T> a, FVGeom g, int TL)
{
    using CV = CNVec<T, NL>;
    using VT = typename CV::type;
    extern __shared__ unsigned char smem_raw[];
    constexpr int CH = R % 8 == 0 ? 8 : R;
    const int ny = g.ny, TLT = TL / NL, S = blockDim.x / TLT; // threads per block row, segments
    double *As = reinterpret_cast<double *>(smem_raw), *Bs = As + NL * blockDim.x, *Xin = Bs + NL * blockDim.x;
    double *ca = Xin + NL * blockDim.x, *cb = ca + ny, *cc = cb + ny, *cf = cc + ny, *ci = cf + ny, *cq = ci + ny;
    double *cr = cq + ny; // 1/Δyu of the normal rows
    for (int j = threadIdx.x; j < ny; j += blockDim.x) {
        if (COMP == 1) cr[j] = a.rdyu[j + 1];
        ca[j] = j == 0 ? 0.0 : a.a[j]; // Dy rows with the couplings past the line ends cut
        cb[j] = a.b[j];
        cc[j] = j == ny - 1 ? 0.0 : a.c[j];
        cf[j] = j == 0 ? 0.0 : a.theta * a.a[j] * a.ip[j];
        ci[j] = a.ip[j];
        cq[j] = j == ny - 1 ? 0.0 : -a.cp[j];
    }
    const int l = threadIdx.x % TLT, sg = threadIdx.x / TLT;
    int bxr, byr;
    xcd_block2(bxr, byr);
    const int i = bxr * TL + NL * l, k = byr; // first of this thread's NL lines
    const bool valid = i < g.nx;
    const int ii = valid ? i : 0;
    const int j0 = sg * R;
    // one z plane per buffer range: rows past the top wall read 0
    const size_t plane = (size_t)g.nx * ny;
    const uint32_t pbytes = (uint32_t)(plane * sizeof(T));
    const int rowb = g.nx * (int)sizeof(T);
    const bufres ru = make_bufres(a.u + plane * k, pbytes), rq = make_bufres(a.q + plane * k, pbytes);
    const bufres rp = make_bufres(a.p + plane * k, pbytes);
    // pressure neighbour: x+1 (same plane), y+1 (next row) or z+1 (next plane)
    const bufres rpn = COMP == 2 ? make_bufres(a.p + plane * wrapp(k, g.nz), pbytes) : rp;
    const int vo = (ii + g.nx * j0) * (int)sizeof(T);
    // x+1 of the last of the NL lines (the others are in the pair itself)
    const int von = COMP == 0 ? (wrapp(ii + NL - 1, g.nx) + g.nx * j0) * (int)sizeof(T) : vo;
    const int jact = (COMP == 1 && g.wall) ? ny - 1 : ny;
    const double rdn = COMP == 0 ? a.rdx : a.rdz;
    double x[NL][R];
    double um[NL];
    {
        const VT v = bload<VT>(ru, sg > 0 ? vo - rowb : vo, 0); // coefficient ca[0] = 0 cuts row −1
#pragma unroll
        for (int e = 0; e < NL; ++e) um[e] = (double)CV::at(v, e);
    }
#pragma unroll
    for (int c0 = 0; c0 < R; c0 += CH) {
        VT ur[CH + 1], qr[CH], pr[CH], pn[CH];
        T pxl[CH]; // COMP 0: x+1 of the last line
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int so = rowb * (c0 + q);
            ur[q] = bload<VT>(ru, vo, so);
            qr[q] = bload<VT>(rq, vo, so);
            pr[q] = bload<VT>(rp, vo, so);
            if (COMP == 0) pxl[q] = bload<T>(rpn, von, so);
            else pn[q] = bload<VT>(rpn, vo, COMP == 1 ? so + rowb : so);
        }
        ur[CH] = bload<VT>(ru, vo, rowb * (c0 + CH));
        if (c0 == 0) __syncthreads(); // the coefficient rows are staged before first use
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int j = j0 + c0 + q;
            const double rd = COMP == 1 ? cr[j] : rdn;
#pragma unroll
            for (int e = 0; e < NL; ++e) {
                const double pc = (double)CV::at(pr[q], e);
                const double pnb = COMP == 0 ? (e + 1 < NL ? (double)CV::at(pr[q], e + 1) : (double)pxl[q])
                                             : (double)CV::at(pn[q], e);
                const double dpdc = (pnb - pc) * rd;
                const double v = j < jact ? a.B * (double)CV::at(qr[q], e) - a.adt * dpdc : 0.0;
                const double u0 = (double)CV::at(ur[q], e);
                double lap = cb[j] * u0;
                lap += ca[j] * um[e];
                lap += cc[j] * (double)CV::at(ur[q + 1], e);
                x[e][c0 + q] = u0 + v + a.theta * lap;
                um[e] = u0;
            }
        }
    }
    // forward sweep, local to the segment
#pragma unroll
    for (int e = 0; e < NL; ++e) {
        double P = 1.0, y = 0.0;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int j = j0 + q;
            y = cf[j] * y + ci[j] * x[e][q];
            x[e][q] = y;
            P *= cf[j];
        }
        As[NL * threadIdx.x + e] = P;
        Bs[NL * threadIdx.x + e] = y;
    }
    __syncthreads();
    if (sg == 0) {
        for (int e = 0; e < NL; ++e) {
            double xi = 0.0;
            for (int q = 0; q < S; ++q) {
                const int t = NL * (q * TLT + l) + e;
                Xin[t] = xi;
                xi = As[t] * xi + Bs[t];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < NL; ++e) {
        const double xi = Xin[NL * threadIdx.x + e];
        double P = 1.0;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            P *= cf[j0 + q];
            x[e][q] += P * xi;
        }
    }
    __syncthreads();
    // back substitution, local to the segment
#pragma unroll
    for (int e = 0; e < NL; ++e) {
        double P = 1.0, y = 0.0;
#pragma unroll
        for (int q = R - 1; q >= 0; --q) {
            const int j = j0 + q;
            y = cq[j] * y + x[e][q];
            x[e][q] = y;
            P *= cq[j];
        }
        As[NL * threadIdx.x + e] = P;
        Bs[NL * threadIdx.x + e] = y;
    }
    __syncthreads();
    if (sg == 0) {
        for (int e = 0; e < NL; ++e) {
            double xi = 0.0;
            for (int q = S - 1; q >= 0; --q) {
                const int t = NL * (q * TLT + l) + e;
                Xin[t] = xi;
                xi = As[t] * xi + Bs[t];
            }
        }
    }
    __syncthreads();
    if (!valid) return;
    double xi[NL], P[NL];
#pragma unroll
    for (int e = 0; e < NL; ++e) {
        xi[e] = Xin[NL * threadIdx.x + e];
        P[e] = 1.0;
    }
#pragma unroll
    for (int q = R - 1; q >= 0; --q) {
        double out[NL];
#pragma unroll
        for (int e = 0; e < NL; ++e) {
            P[e] *= cq[j0 + q];
            out[e] = x[e][q] + P[e] * xi[e];
        }
        bstore(ru, vo, rowb * q, CV::make(out));
    }
}

// The three velocity components of the CN y-diffusion in one launch
// (k_cn_seg with one x-line per thread, R rows per segment): the pressure
// rows of the segment, p(k) for rows j0 .. j0 + R, are loaded once and kept
// in LDS for all three components (the per-component launches read the
// plane three times; in registers they would spill); each component then
// loads its u and q rows, the neighbour pressure it needs (x + 1 for u, the
// next row — already staged — for v, the k + 1 plane for w), and runs the same segmented
// sweeps and chains as k_cn_seg, one after the other through the block's
// LDS.  The y component uses the normal-row coefficient set, x and z the
// tangential one: the LDS coefficient rows are restaged between components.
// Same arithmetic in the same order as k_cn_seg for each component.
// grid (ceil(nx / TL), nz); LDS 7 ny + (R + 4) blockDim doubles.
template <typename T> struct CN3Args {
    T *u[3];
    const T *q[3];
    const T *p;
    const double *ya[2], *yb[2], *yc[2]; // Dy rows: set 0 (x, z), set 1 (y)
    const double *cp[2], *ip[2];         // c'_j and 1/β_j of each set
    const double *rdyu;                  // 1/Δyu (ghosted index)
    double B, adt, theta, rdx, rdz;
};

template <typename T, int COMP, int R, bool PKL>
__device__ __forceinline__ void cn3_component(const CN3Args<T> &a, const FVGeom &g, const double *pk, int TL,
                                              double *As, double *Bs, double *Xin, const double *ca, const double *cb,
                                              const double *cc, const double *cf, const double *ci, const double *cq,
                                              const double *cr)
{
    constexpr int CH = R % 8 == 0 ? 8 : R;
    const int ny = g.ny, S = blockDim.x / TL;
    const int l = threadIdx.x % TL, sg = threadIdx.x / TL;
    int bxr, byr;
    xcd_block2(bxr, byr);
    const int i = bxr * TL + l, k = byr;
    const bool valid = i < g.nx;
    const int ii = valid ? i : 0;
    const int j0 = sg * R;
    const size_t plane = (size_t)g.nx * ny;
    const uint32_t pbytes = (uint32_t)(plane * sizeof(T));
    const int rowb = g.nx * (int)sizeof(T);
    const bufres ru = make_bufres(a.u[COMP] + plane * k, pbytes), rq = make_bufres(a.q[COMP] + plane * k, pbytes);
    // neighbour pressure: x + 1 (same plane), y + 1 (next row) or the k + 1 plane
    const bufres rp = make_bufres(a.p + plane * k, pbytes);
    const bufres rpn = make_bufres(a.p + plane * (COMP == 2 ? wrapp(k, g.nz) : k), pbytes);
    const int vo = (ii + g.nx * j0) * (int)sizeof(T);
    const int von = COMP == 0 ? (wrapp(ii, g.nx) + g.nx * j0) * (int)sizeof(T) : vo;
    const int jact = (COMP == 1 && g.wall) ? ny - 1 : ny;
    const double rdn = COMP == 0 ? a.rdx : a.rdz;
    double x[R];
    double um = (double)bload<T>(ru, sg > 0 ? vo - rowb : vo, 0); // coefficient ca[0] = 0 cuts row −1
#pragma unroll
    for (int c0 = 0; c0 < R; c0 += CH) {
        // the chunk's global offsets and LDS rows pass through an empty asm
        // that depends on the previous chunk's result: neither its loads nor
        // its coefficient reads are hoisted above that chunk (registers)
        int vc = vo, vn = von, jb = j0 + c0;
        asm volatile("" : "+v"(vc), "+v"(vn), "+v"(jb) : "v"(x[c0 > 0 ? c0 - 1 : 0]));
        T ur[CH + 1], qr[CH], pn[CH], pr[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int so = rowb * (c0 + q);
            ur[q] = bload<T>(ru, vc, so);
            qr[q] = bload<T>(rq, vc, so);
            if (!PKL) pr[q] = bload<T>(rp, vc, so);
            if (COMP != 1) pn[q] = bload<T>(rpn, vn, so);
            else if (!PKL) pn[q] = bload<T>(rp, vc, so + rowb);
        }
        ur[CH] = bload<T>(ru, vc, rowb * (c0 + CH));
        if (c0 == 0) __syncthreads(); // (as k_cn_seg: keeps the coefficient reads after the first loads)
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int j = jb + q;
            const double rd = COMP == 1 ? cr[j] : rdn;
            const double pc = PKL ? pk[c0 + q] : (double)pr[q];
            const double pnb = COMP == 1 && PKL ? pk[c0 + q + 1] : (double)pn[q];
            const double dpdc = (pnb - pc) * rd;
            const double v = j < jact ? a.B * (double)qr[q] - a.adt * dpdc : 0.0;
            const double u0 = (double)ur[q];
            double lap = cb[j] * u0;
            lap += ca[j] * um;
            lap += cc[j] * (double)ur[q + 1];
            x[c0 + q] = u0 + v + a.theta * lap;
            um = u0;
        }
    }
    // forward sweep, local to the segment (its coefficient rows are read
    // after the right-hand side is complete, as above)
    int js = j0;
    asm volatile("" : "+v"(js) : "v"(x[R - 1]));
    {
        double P = 1.0, y = 0.0;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int j = js + q;
            y = cf[j] * y + ci[j] * x[q];
            x[q] = y;
            P *= cf[j];
        }
        As[threadIdx.x] = P;
        Bs[threadIdx.x] = y;
    }
    __syncthreads();
    if (sg == 0) {
        double xi = 0.0;
        for (int q = 0; q < S; ++q) {
            const int t = q * TL + l;
            Xin[t] = xi;
            xi = As[t] * xi + Bs[t];
        }
    }
    __syncthreads();
    {
        const double xi = Xin[threadIdx.x];
        double P = 1.0;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            P *= cf[js + q];
            x[q] += P * xi;
        }
    }
    __syncthreads();
    // back substitution, local to the segment
    int jq = j0;
    asm volatile("" : "+v"(jq) : "v"(x[R - 1]));
    {
        double P = 1.0, y = 0.0;
#pragma unroll
        for (int q = R - 1; q >= 0; --q) {
            const int j = jq + q;
            y = cq[j] * y + x[q];
            x[q] = y;
            P *= cq[j];
        }
        As[threadIdx.x] = P;
        Bs[threadIdx.x] = y;
    }
    __syncthreads();
    if (sg == 0) {
        double xi = 0.0;
        for (int q = S - 1; q >= 0; --q) {
            const int t = q * TL + l;
            Xin[t] = xi;
            xi = As[t] * xi + Bs[t];
        }
    }
    __syncthreads();
    if (valid) {
        const double xi = Xin[threadIdx.x];
        double P = 1.0;
#pragma unroll
        for (int q = R - 1; q >= 0; --q) {
            P *= cq[jq + q];
            bstore(ru, vo, rowb * q, T(x[q] + P * xi));
        }
    }
}

template <typename T, int R, bool PKL>
__global__ void __launch_bounds__(256, 2) k_cn_seg3(CN3Args<T> a, FVGeom g, int TL)
{
    extern __shared__ unsigned char smem_raw[];
    const int ny = g.ny;
    double *As = reinterpret_cast<double *>(smem_raw), *Bs = As + blockDim.x, *Xin = Bs + blockDim.x;
    double *ca = Xin + blockDim.x, *cb = ca + ny, *cc = cb + ny, *cf = cc + ny, *ci = cf + ny, *cq = ci + ny;
    double *cr = cq + ny;            // 1/Δyu of the normal rows
    double *pks = cr + ny;           // the threads' pressure rows, R + 1 each
    auto stage = [&](int set) {
        for (int j = threadIdx.x; j < ny; j += blockDim.x) {
            ca[j] = j == 0 ? 0.0 : a.ya[set][j]; // Dy rows with the couplings past the line ends cut
            cb[j] = a.yb[set][j];
            cc[j] = j == ny - 1 ? 0.0 : a.yc[set][j];
            cf[j] = j == 0 ? 0.0 : a.theta * a.ya[set][j] * a.ip[set][j];
            ci[j] = a.ip[set][j];
            cq[j] = j == ny - 1 ? 0.0 : -a.cp[set][j];
        }
    };
    for (int j = threadIdx.x; j < ny; j += blockDim.x) cr[j] = a.rdyu[j + 1];
    stage(0);
    // this thread's pressure rows j0 .. j0 + R of plane k (past the plane: 0)
    const int l = threadIdx.x % TL, sg = threadIdx.x / TL;
    int bxr, byr;
    xcd_block2(bxr, byr);
    const int i = bxr * TL + l, k = byr;
    const int ii = i < g.nx ? i : 0;
    const size_t plane = (size_t)g.nx * ny;
    const bufres rp = make_bufres(a.p + plane * k, (uint32_t)(plane * sizeof(T)));
    const int vo = (ii + g.nx * sg * R) * (int)sizeof(T), rowb = g.nx * (int)sizeof(T);
    double *pk = pks + (size_t)threadIdx.x * (R + 1); // odd stride: the lanes' rows on distinct banks
    if (PKL) {
#pragma unroll
        for (int q = 0; q <= R; ++q) pk[q] = (double)bload<T>(rp, vo, rowb * q);
    }
    __syncthreads(); // the coefficient rows are staged
    cn3_component<T, 0, R, PKL>(a, g, pk, TL, As, Bs, Xin, ca, cb, cc, cf, ci, cq, cr);
    __syncthreads(); // the x component's last reads of the rows are done
    stage(1);
    __syncthreads();
    cn3_component<T, 1, R, PKL>(a, g, pk, TL, As, Bs, Xin, ca, cb, cc, cf, ci, cq, cr);
    __syncthreads();
    stage(0);
    __syncthreads();
    cn3_component<T, 2, R, PKL>(a, g, pk, TL, As, Bs, Xin, ca, cb, cc, cf, ci, cq, cr);
}

// Pivots of I − θ tridiag(a, b, c) for the tangential (set 0) and normal
// (set 1) rows: out[set][0][j] = c'_j, out[set][1][j] = 1/β_j.
__global__ void __launch_bounds__(256) k_cn_coeffs(const double *__restrict__ at, const double *__restrict__ bt,
                                                   const double *__restrict__ ct, const double *__restrict__ an,
                                                   const double *__restrict__ bn, const double *__restrict__ cn,
                                                   double theta, int ny, double *__restrict__ out)
{
    // the rows are staged in LDS so the two sequential recurrences (one per
    // set) wait on LDS, not on global memory; dynamic LDS 6·ny doubles
    extern __shared__ unsigned char smem_raw[];
    double *rows = reinterpret_cast<double *>(smem_raw); // [set][a, b, c][ny], reused for [set][c', 1/β][ny]
    for (int j = threadIdx.x; j < ny; j += blockDim.x) {
        rows[0 * ny + j] = at[j];
        rows[1 * ny + j] = bt[j];
        rows[2 * ny + j] = ct[j];
        rows[3 * ny + j] = an[j];
        rows[4 * ny + j] = bn[j];
        rows[5 * ny + j] = cn[j];
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int set = threadIdx.x;
        double *a = rows + 3 * set * ny, *b = a + ny, *c = b + ny;
        double beta = 1.0 - theta * b[0];
        double cp = -theta * c[0] / beta;
        c[0] = cp;     // c'_0 (row c is consumed at the same index)
        b[0] = 1.0 / beta; // 1/β_0
        for (int j = 1; j < ny; ++j) {
            beta = (1.0 - theta * b[j]) + theta * a[j] * cp;
            cp = -theta * c[j] / beta;
            c[j] = cp;
            b[j] = 1.0 / beta;
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < ny; j += blockDim.x) {
        for (int set = 0; set < 2; ++set) {
            out[set * 2 * ny + j] = rows[(3 * set + 2) * ny + j];      // c'_j
            out[set * 2 * ny + ny + j] = rows[(3 * set + 1) * ny + j]; // 1/β_j
        }
    }
}

// The pivots of k_cn_coeffs for the three RK stages of a step at once
// (theta_s = α_s Δt ν / 2): the six recurrences (stage, set) run on six
// threads side by side, the rows staged in LDS, the results written straight
// to out[(2 s + set) · 2 ny ...] (c'_j, then 1/β_j), the same arithmetic per
// recurrence as k_cn_coeffs.  One launch per step instead of three.
__global__ void __launch_bounds__(256) k_cn_coeffs3(const double *__restrict__ at, const double *__restrict__ bt,
                                                    const double *__restrict__ ct, const double *__restrict__ an,
                                                    const double *__restrict__ bn, const double *__restrict__ cn,
                                                    double theta0, double theta1, double theta2, int ny,
                                                    double *__restrict__ out)
{
    extern __shared__ unsigned char smem_raw[];
    double *rows = reinterpret_cast<double *>(smem_raw); // [set][a, b, c][ny]
    for (int j = threadIdx.x; j < ny; j += blockDim.x) {
        rows[0 * ny + j] = at[j];
        rows[1 * ny + j] = bt[j];
        rows[2 * ny + j] = ct[j];
        rows[3 * ny + j] = an[j];
        rows[4 * ny + j] = bn[j];
        rows[5 * ny + j] = cn[j];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int stage = threadIdx.x >> 1, set = threadIdx.x & 1;
        const double theta = stage == 0 ? theta0 : stage == 1 ? theta1 : theta2;
        const double *a = rows + 3 * set * ny, *b = a + ny, *c = b + ny;
        double *cpo = out + (size_t)(2 * stage + set) * 2 * ny, *ipo = cpo + ny;
        double beta = 1.0 - theta * b[0];
        double cp = -theta * c[0] / beta;
        cpo[0] = cp;
        ipo[0] = 1.0 / beta;
        for (int j = 1; j < ny; ++j) {
            beta = (1.0 - theta * b[j]) + theta * a[j] * cp;
            cp = -theta * c[j] / beta;
            cpo[j] = cp;
            ipo[j] = 1.0 / beta;
        }
    }
}

// ------------------------------------------------------------ reductions

// Per-block maxima of |u_c| over the interior (CFL, src/timestep.jl:69-80).
template <typename T>
__global__ void __launch_bounds__(256) k_maxabs3(const T *ux, const T *uy, const T *uz, int N, double *part)
{
    __shared__ double s[3][256];
    double m0 = 0, m1 = 0, m2 = 0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < N; t += gridDim.x * blockDim.x) {
        m0 = fmax(m0, fabs((double)ux[t]));
        m1 = fmax(m1, fabs((double)uy[t]));
        m2 = fmax(m2, fabs((double)uz[t]));
    }
    s[0][threadIdx.x] = m0;
    s[1][threadIdx.x] = m1;
    s[2][threadIdx.x] = m2;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int c = 0; c < 3; ++c) s[c][threadIdx.x] = fmax(s[c][threadIdx.x], s[c][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 3) part[blockIdx.x * 3 + threadIdx.x] = s[threadIdx.x][0];
}

// Per-block max |div u| computed in FP64 (src/fields.jl:38-42).
template <typename T>
__global__ void __launch_bounds__(256) k_maxdiv(const T *ux, const T *uy, const T *uz, const double *rdy, double rdx,
                                                double rdz, FVGeom g, double *part)
{
    __shared__ double s[256];
    const int N = g.nx * g.ny * g.nz;
    double m = 0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < N; t += gridDim.x * blockDim.x) {
        int i = t % g.nx, r = t / g.nx, j = r % g.ny, k = r / g.ny;
        int jj = j + 1;
        double d = ((double)ux[t] - (double)ux[IX(wrapm(i, g.nx), j, k, g)]) * rdx +
                   ((double)ldf(uy, i, jj, k, g) - (double)ldf(uy, i, jj - 1, k, g)) * rdy[jj] +
                   ((double)uz[t] - (double)uz[IX(i, j, wrapm(k, g.nz), g)]) * rdz;
        m = fmax(m, fabs(d));
    }
    s[threadIdx.x] = m;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Kinetic energy ½ Σ u² Δx Δ2[j] Δz over active DOFs (src/fields.jl:45-60),
// compensated: per-thread double-double, block tree, per-block partials.
template <typename T>
__global__ void __launch_bounds__(256) k_kinetic(const T *ux, const T *uy, const T *uz, const double *dy,
                                                 const double *dyu, double dxdz, FVGeom g, dd *part)
{
    __shared__ dd s[256];
    const int N = g.nx * g.ny * g.nz;
    dd acc{0.0, 0.0};
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < N; t += gridDim.x * blockDim.x) {
        int j = (t / g.nx) % g.ny, jj = j + 1;
        double a = (double)ux[t], b = (double)uy[t], c = (double)uz[t];
        double ey = (g.wall && j == g.ny - 1) ? 0.0 : b * b * dxdz * dyu[jj] / 2;
        acc = dd_add1(acc, a * a * dxdz * dy[jj] / 2);
        acc = dd_add1(acc, ey);
        acc = dd_add1(acc, c * c * dxdz * dy[jj] / 2);
    }
    dd r = block_dd_sum(acc, s);
    if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// Channel statistics plane sums (src/stats.jl:39-58): block (j, chunk) sums
// over a slice of the (x, z) plane; slots (cell row j): U, W, uu, ww;
// (face row j, j < nf): V, vv, Uf, uv.  part[(slot*ny + j)*nchunk + chunk].
template <typename T>
__global__ void __launch_bounds__(256) k_chstats(const T *ux, const T *uy, const T *uz, FVGeom g, int nf,
                                                 int nchunk, dd *part)
{
    __shared__ dd s[256];
    const int j = blockIdx.x, chunk = blockIdx.y;
    const int P = g.nx * g.nz;
    dd acc[8];
    for (int q = 0; q < 8; ++q) acc[q] = dd{0.0, 0.0};
    const bool face = j < nf;
    for (int t = chunk * blockDim.x + threadIdx.x; t < P; t += nchunk * blockDim.x) {
        int i = t % g.nx, k = t / g.nx;
        int c = IX(i, j, k, g);
        double a = (double)ux[c], w = (double)uz[c];
        acc[0] = dd_add1(acc[0], a);
        acc[1] = dd_add1(acc[1], w);
        acc[2] = dd_add1(acc[2], a * a);
        acc[3] = dd_add1(acc[3], w * w);
        if (face) {
            double v = (double)uy[c];
            double uf = (a + (double)ldc(ux, i, j + 2, k, g, T(-1))) / 2;
            double vx = (v + (double)uy[IX(wrapp(i, g.nx), j, k, g)]) / 2;
            acc[4] = dd_add1(acc[4], v);
            acc[5] = dd_add1(acc[5], v * v);
            acc[6] = dd_add1(acc[6], uf);
            acc[7] = dd_add1(acc[7], uf * vx);
        }
    }
    for (int q = 0; q < 8; ++q) {
        dd r = block_dd_sum(acc[q], s);
        if (threadIdx.x == 0) part[((size_t)q * g.ny + j) * nchunk + chunk] = r;
    }
}

// Shell binning of one transformed component (src/stats.jl:143-157):
// κ = round(|k|), weight 2 for 0 < kx < nx/2, adds w|Â|²/2/N² to per-block
// bins; part[block * nbins + κ].  Deterministic: the block stages 256
// (κ, e) pairs in LDS, then the thread that owns bin b adds that chunk's
// contributions to b in index order, so every run sums in the same order
// (the grid depends only on the sizes).  Dynamic LDS: nbins doubles +
// 256 doubles + 256 ints.
template <typename T>
__global__ void __launch_bounds__(256) k_binshells(const cplx<T> *A, FVGeom g, int nbins, double invN2, double *part)
{
    extern __shared__ unsigned char smem_raw[];
    double *bins = reinterpret_cast<double *>(smem_raw);
    double *val = bins + nbins;
    int *key = reinterpret_cast<int *>(val + 256);
    const int tid = threadIdx.x;
    for (int b = tid; b < nbins; b += 256) bins[b] = 0.0;
    const int N = g.nxc * g.ny * g.nz;
    for (int base = blockIdx.x * 256; base < N; base += gridDim.x * 256) {
        const int t = base + tid;
        int kap = -1;
        double e = 0.0;
        if (t < N) {
            const int kx = t % g.nxc, r = t / g.nxc, j = r % g.ny, k = r / g.ny;
            const int ky = j > g.ny / 2 ? j - g.ny : j;
            const int kz = k > g.nz / 2 ? k - g.nz : k;
            kap = (int)rint(sqrt((double)(kx * kx + ky * ky + kz * kz)));
            const double wgt = (kx > 0 && 2 * kx < g.nx) ? 2.0 : 1.0;
            const cplx<T> a = A[kx + (size_t)g.ldc * r];
            e = wgt * ((double)a.x * (double)a.x + (double)a.y * (double)a.y) / 2 * invN2;
            if (kap >= nbins) kap = -1;
        }
        __syncthreads(); // the previous chunk's readers are done with key/val
        key[tid] = kap;
        val[tid] = e;
        __syncthreads();
        const int cnt = min(256, N - base);
        for (int b = tid; b < nbins; b += 256) {
            double acc = bins[b];
            for (int q = 0; q < cnt; ++q)
                if (key[q] == b) acc += val[q];
            bins[b] = acc;
        }
    }
    __syncthreads();
    for (int b = tid; b < nbins; b += 256) part[(size_t)blockIdx.x * nbins + b] = bins[b];
}

// k_binshells with per-wavefront bins (4 nbins doubles of LDS, nbins ≤
// 2048): each wavefront bins its 64 elements per pass with wave_bin_add; the
// four arrays are summed in wave order.  Deterministic.
template <typename T>
__global__ void __launch_bounds__(256) k_binshells_w(const cplx<T> *A, FVGeom g, int nbins, double invN2, double *part)
{
    extern __shared__ unsigned char smem_raw[];
    double *bins = reinterpret_cast<double *>(smem_raw); // [4][nbins]
    const int tid = threadIdx.x, wave = tid >> 6;
    for (int b = tid; b < 4 * nbins; b += 256) bins[b] = 0.0;
    __syncthreads();
    const int N = g.nxc * g.ny * g.nz;
    for (int base = blockIdx.x * 256; base < N; base += gridDim.x * 256) {
        const int t = base + tid;
        int kap = -1;
        double e = 0.0;
        if (t < N) {
            const int kx = t % g.nxc, r = t / g.nxc, j = r % g.ny, k = r / g.ny;
            const int ky = j > g.ny / 2 ? j - g.ny : j;
            const int kz = k > g.nz / 2 ? k - g.nz : k;
            kap = (int)rint(sqrt((double)(kx * kx + ky * ky + kz * kz)));
            const double wgt = (kx > 0 && 2 * kx < g.nx) ? 2.0 : 1.0;
            const cplx<T> a = A[kx + (size_t)g.ldc * r];
            e = wgt * ((double)a.x * (double)a.x + (double)a.y * (double)a.y) / 2 * invN2;
            if (kap >= nbins) kap = -1;
        }
        wave_bin_add(bins + wave * nbins, kap, e);
    }
    __syncthreads();
    for (int b = tid; b < nbins; b += 256)
        part[(size_t)blockIdx.x * nbins + b] = ((bins[b] + bins[nbins + b]) + bins[2 * nbins + b]) + bins[3 * nbins + b];
}

}  // namespace dns
