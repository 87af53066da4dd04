/*
 * libdnsrocm -- C ABI of the MI355X (gfx950) single-device time-step hot
 * path of DistributedNavierStokes.
 *
 * Every entry point takes plain C types (handles, host pointers, sizes),
 * returns an int status (DNS_OK or a negative DNS_E* code) and never throws.
 * dns_last_error() returns a thread-local message for the last failure.
 * A handle owns its device memory and one HIP stream; entry points that
 * return host-visible values synchronise that stream.  One handle is used
 * by one host thread at a time.
 *
 * Field layouts exchanged with the host are column-major (x fastest):
 *   DNS_LAYOUT_GHOSTED  (nx+2, ny+2, nz+2), ghost width 1, exactly the
 *                       reference's `scalarfield`/`vectorfield` arrays
 *                       (src/fields.jl:4-9); on output the ghosts are
 *                       filled as exchange_halo! would (src/halo.jl:81-133);
 *                       on input ghosts are ignored.
 *   DNS_LAYOUT_INTERIOR (nx, ny, nz).
 */
#ifndef DNS_ROCM_H
#define DNS_ROCM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DNS_OK 0
#define DNS_EINVAL (-1)
#define DNS_EUNSUPPORTED (-2)
#define DNS_EHIP (-3)
#define DNS_ENOMEM (-4)

#define DNS_FP64 0  /* FP64 storage and arithmetic */
#define DNS_MIXED 1 /* FP32 storage, stencils and FFTs; FP64 Thomas solves and reductions */
/* FP64 with the reference's single-rank kernel sequence (one kernel per
 * reference kernel, separate RK broadcasts, materialised copies/transposes,
 * halo plane copies with synchronisation, HBM Thomas scratch): the baseline
 * the fused forms are measured against (FV handles only). */
#define DNS_PLAIN_FP64 2

#define DNS_LAYOUT_INTERIOR 0
#define DNS_LAYOUT_GHOSTED 1

/* Field identifiers of an FV handle. */
#define DNS_FIELD_UX 0
#define DNS_FIELD_UY 1
#define DNS_FIELD_UZ 2
#define DNS_FIELD_P 3     /* pressure (lagged pressure in semi-implicit mode) */
#define DNS_FIELD_QX 4    /* RK register q */
#define DNS_FIELD_QY 5
#define DNS_FIELD_QZ 6
#define DNS_FIELD_DELTA 7 /* projection increment δ (semi-implicit mode) */

/* Message of the last failed call on this thread ("" if none). */
const char *dns_last_error(void);

/* Library version string. */
const char *dns_version(void);

/* ------------------------------------------------------------------ FV
 * Staggered symmetry-preserving finite-volume box/channel solver. */

typedef struct dns_fv dns_fv;

/* Replaces the keyword arguments of `setup` (src/setup.jl:34-113).  The
 * `stretch` closure cannot cross the ABI: the caller evaluates the global
 * y faces yfaces[j] = lims[2][1] + Ly*stretch(j/ny), j = 0..ny
 * (src/setup.jl:155-156), and the library builds the ghost-extended metrics
 * of `griddata` (src/setup.jl:148-193) from them. */
typedef struct dns_fv_params {
    int32_t n[3];          /* cells (nx, ny, nz) */
    double lims[6];        /* (x0, x1, y0, y1, z0, z1) */
    int32_t ywall;         /* 0: bc[2] == :periodic, 1: bc[2] == :wall */
    const double *yfaces;  /* ny+1 host values */
    double visc;           /* ν */
    double bodyforce[3];   /* f */
    int32_t ydiffusion;    /* 0: :explicit, 1: :implicit (per-stage CN in y) */
    int32_t precision;     /* DNS_FP64, DNS_MIXED or DNS_PLAIN_FP64 */
} dns_fv_params;

/* setup(; ...) -- src/setup.jl:34-113 (with plan_poisson, src/poisson.jl:56-120,
 * and plan_ydiffusion, src/ydiffusion.jl:54-68).  Allocates all device
 * state (u, p, q, δ and the pipeline buffers) once; fields start at zero. */
int dns_fv_create(const dns_fv_params *params, dns_fv **out);
int dns_fv_destroy(dns_fv *h);

/* Host <-> device field copies (one component per call); ghosted output
 * fills the halos like exchange_halo! (src/halo.jl:81-133). */
int dns_fv_set_field(dns_fv *h, int32_t field, const double *host, int32_t layout);
int dns_fv_get_field(dns_fv *h, int32_t field, double *host, int32_t layout);

/* momentum!(q, u, A, Δt, s) -- src/operators.jl:147-178: q ← A q + Δt F(u). */
int dns_fv_momentum(dns_fv *h, double A, double dt);

/* project!(u, ψ, s) -- src/timestep.jl:25-31, with ψ = DNS_FIELD_P or
 * DNS_FIELD_DELTA. */
int dns_fv_project(dns_fv *h, int32_t psi_field);

/* step!(u, p, q, δ, s, Δt) -- src/timestep.jl:41-63: one Williamson RK3 step
 * (explicit or semi-implicit per the handle).  Asynchronous.  Δt must be
 * finite, and positive with implicit y-diffusion (the lagged-pressure update
 * p += δ/(αΔt), src/timestep.jl:51, divides by it): DNS_EINVAL otherwise. */
int dns_fv_step(dns_fv *h, double dt);

/* cfl_timestep(u, s; cfl) -- src/timestep.jl:69-80. */
int dns_fv_cfl_timestep(dns_fv *h, double cfl, double *dt);

/* The CFL branch of solve! (src/timestep.jl:92-111) for at most nsteps
 * steps: while t < tend − 1e-10(|tend| + 1) (src/timestep.jl:102),
 * Δt = min(cfl_timestep(u, s; cfl), tend − t) then step!, with *t advanced,
 * the Δt of every step taken stored in dts (nsteps entries, may be null) and
 * the number of steps taken in *ndone (may be null); tend = +inf sets no
 * end time (nsteps steps are taken).  In the explicit form
 * Δt is computed on the device and the steps replay captured HIP graphs
 * without waiting for the host (graphs replayed after the end time leave
 * the state unchanged); returns after the last step. */
int dns_fv_advance_cfl(dns_fv *h, double cfl, int32_t nsteps, double tend, double *t, double *dts, int32_t *ndone);

/* No counterpart in the reference: dns_fv_advance_cfl for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, semantics and
 * results; it allocates nothing: the device Δt, time, step counter, Δt log
 * and CFL reduction partials sit in a static 160 KiB device array of the
 * library (one per GPU, kept until the process ends), the steps are direct
 * launches (no graph), and nsteps is limited to 16384 (DNS_EINVAL beyond).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (several handles on concurrent host threads) race.  Not available for
 * DNS_PLAIN_FP64 handles or with kernel timing on (DNS_EUNSUPPORTED). */
int dns_fv_advance_cfl_noalloc(dns_fv *h, double cfl, int32_t nsteps, double tend, double *t, double *dts,
                               int32_t *ndone);

/* maxdiv(u, s) -- src/fields.jl:38-42 (divergence residual in FP64). */
int dns_fv_maxdiv(dns_fv *h, double *out);

/* kinetic_energy(u, s) -- src/fields.jl:45-60 (compensated FP64 sum). */
int dns_fv_kinetic_energy(dns_fv *h, double *out);

/* One sample of sample_channelstats! (src/stats.jl:39-58): adds the (x, z)
 * plane sums to the host accumulator acc laid out as
 * Uc, Wc, uuc, wwc (ny values each) then Vf, vvf, Uf, uvf (nf values each),
 * nf = ny-1 with walls, ny otherwise. */
int dns_fv_channelstats_sample(dns_fv *h, double *acc);

/* One sample of sample_spectrum! (src/stats.jl:120-157): adds ½|û|²/N² per
 * integer shell κ = round(|k|) to E[0..nbins-1] (fully periodic boxes). */
int dns_fv_spectrum_sample(dns_fv *h, double *E, int32_t nbins);

/* Waits for all work queued on the handle's stream. */
int dns_fv_sync(dns_fv *h);

/* No counterpart in the reference (diagnostics): writes a one-line
 * description of the kernel instances and launch shapes the handle's step
 * uses (momentum march, FFT line lengths and tiles, y-solve and CN kernels
 * with their segment lengths) into buf (len bytes, NUL-terminated). */
int dns_fv_describe(dns_fv *h, char *buf, int32_t len);

/* Kernel timing hook used by the benchmark: when enabled, the library
 * records HIP events around every launch of the named kernel family on the
 * handle's stream and accumulates their durations.
 * family: 0 momentum, 1 x-pass, 2 z-pass, 3 y-solve, 4 pgrad, 5 CN (one
 * velocity component per launch). */
int dns_fv_timing(dns_fv *h, int32_t enable);
int dns_fv_timing_read(dns_fv *h, int32_t family, double *total_ms, int64_t *launches);

/* ------------------------------------------------------------ spectral
 * Pseudo-spectral HIT solver on a cubic periodic box, 2/3-truncated state.
 * State layout (host and device): (K, m, m, 3) complex, K = kcut+1,
 * m = 2kcut+1, first index fastest, compact frequency order
 * (src/spectral/operators.jl:7); host buffers interleave (re, im) doubles.
 * Physical layout: (n, n, n, 3) real on the cell-centred transform grid. */

typedef struct dns_spec dns_spec;

/* Replaces the keyword arguments of `spectral_setup`
 * (src/spectral/setup.jl:33-48); kcut must be given (n ÷ 3 by default in the
 * reference). */
typedef struct dns_spec_params {
    int32_t n;
    int32_t kcut;
    double l;
    double visc;
    int32_t precision; /* DNS_FP64, or DNS_MIXED = FP32 FFTs/state with FP64 reductions */
} dns_spec_params;

/* spectral_setup(; ...) -- src/spectral/setup.jl:33-158. */
int dns_spec_create(const dns_spec_params *params, dns_spec **out);
int dns_spec_destroy(dns_spec *h);

/* which: 0 = uh (the state), 1 = ustart (the second Wray register),
 * 2 = du (the last dns_spec_rhs result; read only). */
int dns_spec_set_state(dns_spec *h, int32_t which, const double *host);
int dns_spec_get_state(dns_spec *h, int32_t which, double *host);

/* The physical buffer (s.fft.v / s.fft.rbuf of the reference). */
int dns_spec_set_physical(dns_spec *h, const double *host);
int dns_spec_get_physical(dns_spec *h, double *host);

/* spec_to_phys!(v, uh, s) -- src/spectral/fft.jl:200-214 (state → physical). */
int dns_spec_spec_to_phys(dns_spec *h);
/* phys_to_spec!(uh, w, s) -- src/spectral/fft.jl:221-232 (physical → state). */
int dns_spec_phys_to_spec(dns_spec *h);
/* spectral_project!(uh, s) -- src/spectral/operators.jl:25-29. */
int dns_spec_project(dns_spec *h);
/* spectral_rhs!(du, uh, s) -- src/spectral/operators.jl:60-79 (into du). */
int dns_spec_rhs(dns_spec *h);
/* spectral_step!(uh, s, Δt, cache) -- src/spectral/step.jl:68-92; writes
 * vsummax of the first stage to *vmax when vmax is not NULL (synchronises);
 * asynchronous otherwise. */
int dns_spec_step(dns_spec *h, double dt, double *vmax);
/* vsummax(spec_to_phys(uh)) -- src/spectral/step.jl:54-60, 140-145. */
int dns_spec_vsummax(dns_spec *h, double *out);
/* shell_energies(uh, s, s0:s0+nshells-1) -- src/spectral/forcing.jl:12-24. */
int dns_spec_shell_energies(dns_spec *h, int32_t s0, int32_t nshells, double *E);
/* rescale_shell! for shells s0..s0+nshells-1 by factors[] -- src/spectral/forcing.jl:27-35. */
int dns_spec_rescale_shells(dns_spec *h, int32_t s0, int32_t nshells, const double *factors);
/* noise_kernel! into the physical buffer -- src/spectral/ic.jl:38-48. */
int dns_spec_noise(dns_spec *h, uint64_t seed);
/* spectral_energy / spectral_dissipation / spectral_maxdiv --
 * src/spectral/operators.jl:87-110 (compensated FP64 sums). */
int dns_spec_energy(dns_spec *h, double *out);
int dns_spec_dissipation(dns_spec *h, double *out);
int dns_spec_maxdiv(dns_spec *h, double *out);
/* σ̂ = F[v_i v_j]/n³ of the state's velocity at the retained modes, six
 * products in the order xx, yy, zz, xy, yz, zx (the product batches of
 * spectral_rhs!, src/spectral/operators.jl:66-75, and of sfs_sample!,
 * src/spectral/les.jl:323-337), kept in the handle's product buffer until
 * the next pipeline call; dns_spec_get_products copies (K, m, m, 6). */
int dns_spec_products(dns_spec *h);
int dns_spec_get_products(dns_spec *h, double *host);
int dns_spec_sync(dns_spec *h);

/* spectral_save(prefix, uh, s; time, step, meta) -- src/spectral/io.jl:29-75:
 * writes the state to <prefix>.bin, the (kcut+1, m, m, 3) complex array
 * column-major in global index order, one component after the other, as
 * Complex{Float64} (Complex{Float32} for DNS_MIXED handles), and the TOML
 * sidecar <prefix>.toml (format, version, eltype, n, kcut, m, l, visc, time,
 * step, order, data; keys sorted).  meta_toml (may be NULL) holds TOML
 * `key = value` lines stored under a [meta] table.  Atomic: both files
 * appear under their final names only when complete. */
int dns_spec_save(dns_spec *h, const char *prefix, double time, int64_t step, const char *meta_toml);
/* spectral_load!(uh, prefix, s) -- src/spectral/io.jl:77-104: reads a
 * snapshot into the state; its eltype, kcut and l must match the handle's
 * (DNS_EINVAL with the reference's message otherwise), n need not.  time and
 * step (may be NULL) receive the sidecar values. */
int dns_spec_load(dns_spec *h, const char *prefix, double *time, int64_t *step);

/* -------------------------------------------- SFS data (learned closure)
 * sfs_sample! for one Gaussian bank cell (src/spectral/les.jl:313-382):
 * ū = T⅔(G(cut(û))) and τ = T⅔(G(cut(σ̂)) − F[ū_i ū_j]/M³), trace-free, with
 * G = exp(−Δ²k²/24), Δ = delta_fac·l/M, shells |k| <= protectshells left
 * unfiltered.  `les` is a spectral handle with n = M, kcut = M ÷ 3 (the modes
 * T⅔ keeps), the same l and precision, and M/2 <= the DNS kcut.  The results
 * stay on `les`: its state holds ū, its SFS buffer holds τ̂ (6 components). */
int dns_sfs_sample(dns_spec *dns, dns_spec *les, double delta_fac, int32_t protectshells);
/* Copies ū (K, m, m, 3) and τ̂ (K, m, m, 6) of the LES handle (either may be NULL). */
int dns_sfs_get(dns_spec *les, double *ubar_hat, double *tau_hat);
/* Physical ū (M, M, M, 3) and τ (M, M, M, 6) on the LES grid (host). */
int dns_sfs_physical(dns_spec *les, double *ubar, double *tau);

/* Kernel timing hook, as dns_fv_timing; family: 0 y pass, 1 z pass,
 * 2 fused x pass, 3 y pass + IF stage, 4 forcing. */
int dns_spec_timing(dns_spec *h, int32_t enable);
int dns_spec_timing_read(dns_spec *h, int32_t family, double *total_ms, int64_t *launches);

/* ------------------------------------------------ learned SFS closure CNN
 * No counterpart in the reference (it ends at the SFS training data,
 * src/spectral/les.jl:313-382): a 4-layer 3×3×3 periodic CNN, channels
 * 3 → 64 → 64 → 64 → 6 with erf-GELU, on M³ grids in batches of B, BF16
 * weights and activations with FP32 accumulation (gfx950 MFMA), Adam on FP32
 * master weights.  Parameter vector (FP32): per layer W[o][n][k] (27·nout·kin,
 * o ↔ offset (o%3−1, (o/3)%3−1, o/9−1) in (x, y, z)) then bias[nout].
 * Host batches are interleaved [b][z][y][x][c] FP32. */
typedef struct dns_cnn dns_cnn;
int dns_cnn_create(int32_t M, int32_t batch, uint64_t seed, dns_cnn **out);
int dns_cnn_destroy(dns_cnn *h);
int dns_cnn_num_params(dns_cnn *h, int64_t *n);
int dns_cnn_set_params(dns_cnn *h, const float *params);
int dns_cnn_get_params(dns_cnn *h, float *params);
/* Gradient of the last dns_cnn_loss_grad, same layout as the parameters. */
int dns_cnn_get_grads(dns_cnn *h, float *grads);
/* y (B·M³ × 6) of the batch x (B·M³ × 3). */
int dns_cnn_forward(dns_cnn *h, const float *x, float *y);
/* MSE loss of the network on (x, t) and its gradient (kept on the device). */
int dns_cnn_loss_grad(dns_cnn *h, const float *x, const float *t, double *loss);
/* One Adam update of the parameters with the current gradient. */
int dns_cnn_adam_step(dns_cnn *h, float lr, float beta1, float beta2, float eps);
int dns_cnn_sync(dns_cnn *h);
/* Device dataset of (ū, τ) pairs, planar FP32 host arrays (3·M³, 6·M³). */
int dns_cnn_add_sample(dns_cnn *h, const float *ubar, const float *tau);
int dns_cnn_dataset_size(dns_cnn *h, int32_t *n);
/* Normalisation: ū/su in, τ/stau out.  fit_scales sets the RMS of the
 * dataset's components; set/get handle saved values. */
int dns_cnn_fit_scales(dns_cnn *h, double *su, double *stau);
int dns_cnn_set_scales(dns_cnn *h, double su, double stau);
int dns_cnn_get_scales(dns_cnn *h, double *su, double *stau);
/* nsteps Adam steps on batches of epoch-wise shuffles (mt19937_64(seed)) of
 * the first ntrain dataset samples (all if ntrain <= 0); losses[nsteps]
 * (normalised MSE before each update) may be NULL. */
int dns_cnn_train(dns_cnn *h, int32_t nsteps, float lr, float beta1, float beta2, float eps, uint64_t seed,
                  int32_t ntrain, double *losses);
/* Normalised MSE over dataset samples [first, first + count). */
int dns_cnn_eval(dns_cnn *h, int32_t first, int32_t count, double *loss);
/* Predicted τ (6·M³ planar, physical units) of one dataset sample. */
int dns_cnn_predict(dns_cnn *h, int32_t sample, float *tau);
/* Kernel timing, families 0 forward, 1 backward, 2 Adam + weight packing. */
int dns_cnn_timing(dns_cnn *h, int32_t enable);
int dns_cnn_timing_read(dns_cnn *h, int32_t family, double *total_ms, int64_t *launches);

/* ------------------------------------------------ closure LES (no counterpart
 * in the reference; it ends at the SFS data, src/spectral/les.jl:313-382)
 * Appends the physical (ū, τ) of the last dns_sfs_sample on `les` to the
 * network's dataset, on the device. */
int dns_sfs_append(dns_spec *les, dns_cnn *net);
/* spectral_step! (src/spectral/step.jl:68-92) of an LES handle (n = the
 * network's M) with the learned stress: each stage adds τ̂ = F[net(ū)] to σ̂,
 * du_i = −ik_j(σ̂_ij + τ̂_ij).  vmax as dns_spec_step. */
int dns_les_closure_step(dns_spec *les, dns_cnn *net, double dt, double *vmax);

#ifdef __cplusplus
}
#endif

#endif /* DNS_ROCM_H */
