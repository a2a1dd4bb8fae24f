/*
 * spectral_rs.h -- C ABI of the MI355X (gfx950) implementation of the
 * spectral_rs per-timestep hot path: particle move + current/density deposit,
 * binomial current filter, spectral Maxwell solve through a 2D FFT, and the
 * relativistic Boris push.
 *
 * The reference (pcrumley/spectral_rs, Rust) has no FFI; its seam is the
 * method set that `run()` calls.  Each entry point below replaces one of those
 * methods (file:line relative to the reference root) and keeps its meaning:
 * same grids (padded (Ny+2)x(Nx+2) row-major, 1-cell ghost ring), same cell
 * convention (ghost-offset ix in 1..=Nx, iy in 1..=Ny, dx,dy in [-0.5,0.5)),
 * same normalisations.  Where the reference panics through assert!, the call
 * returns SP_EINVARIANT and sp_last_error() names the assert it mirrors.
 *
 * Threading: one host thread per context.  Calls are queued on the context's
 * HIP stream and return before the GPU finishes, except the getters/setters
 * and sp_sync(), which wait.  Invariant violations raised by a kernel are
 * reported by the next call that waits.  No device memory is allocated after
 * sp_create()/sp_species_*()/sp_checkpoint_load() (like the reference, which
 * allocates only in Flds::new and Prtl::new), except by the optional
 * features that say so (denoiser, autotuner, ensembles).
 *
 * Precision: SP_F32 is the reference's default build (Float = f32); SP_F64 is
 * `--features dprec` (Cargo.toml:27-31, src/lib.rs:15-19).  Real arrays
 * passed through `void*` hold float (SP_F32) or double (SP_F64) elements.
 */
#ifndef SPECTRAL_RS_H
#define SPECTRAL_RS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum sp_status {
  SP_OK = 0,
  SP_EINVAL = -1,      /* bad argument (config validation, src/lib.rs:75-113) */
  SP_EHIP = -2,        /* HIP runtime error */
  SP_EINVARIANT = -3,  /* a reference assert! would have fired */
  SP_ENOMEM = -4
};

enum sp_precision { SP_F32 = 0, SP_F64 = 1 };

/* Behaviour when a particle reaches ix <= delta.  The reference has no left
 * boundary and asserts (src/prtls.rs:163-170); SP_LEFT_REFLECT mirrors the
 * right wall (src/prtls.rs:172-188) at cell delta+1. */
enum sp_left_edge { SP_LEFT_ASSERT = 0, SP_LEFT_REFLECT = 1 };

/* Grid ids.  0..10 name the padded real grids of Flds (src/flds/mod.rs:15-37);
 * the same ids name their spectra for sp_get_spectrum/sp_set_spectrum/
 * sp_fft2d, and SP_BXW..SP_BZW name the persistent half-step spectra
 * b_*_wrk (src/flds/mod.rs:31-33). */
enum sp_grid {
  SP_EX = 0, SP_EY, SP_EZ, SP_BX, SP_BY, SP_BZ, SP_JX, SP_JY, SP_JZ,
  SP_RHO /* Flds.dsty */, SP_DENS, SP_BXW, SP_BYW, SP_BZW
};

/* Mirrors Params + Setup of config.toml (src/lib.rs:34-66) plus the cargo
 * features as run-time flags. */
typedef struct sp_params {
  int32_t size_x, size_y; /* even (src/lib.rs:76-87) */
  int32_t delta;          /* 2*delta < size_x (src/lib.rs:88-92) */
  int32_t dens;           /* particles per species per cell */
  int32_t n_pass;         /* binomial filter passes */
  double dt, c;           /* reference units: 1/omega_pe and c/omega_pe cells */
  double gamma_inj;       /* > 1 (src/lib.rs:93-95) */
  int32_t precision;      /* enum sp_precision */
  int32_t periodic_x;     /* cargo feature `periodic` (src/prtls.rs:136-143) */
  int32_t checked;        /* 1 = the asserts of the default build are checked */
  int32_t left_edge;      /* enum sp_left_edge */
  int32_t sort_interval;  /* steps between cell sorts of the particles, 0 = never */
  int32_t study;          /* SP_F32 precision-study variants (enum sp_study bits), 0 = none */
  int32_t reserved[2];
} sp_params;

/* Precision-study variants of the SP_F32 path (no reference counterpart),
 * for comparing it with an all-FP32 code: SP_STUDY_ABS_POS32 keeps each
 * position to the resolution of an absolute FP32 coordinate (after every
 * position update the offset becomes fl(fl(ix + dx) - ix), the rounding an
 * x stored as one float would get); SP_STUDY_TWIDDLE32 generates the FFT
 * twiddles in FP32 instead of FP64; SP_STUDY_SUM32 accumulates the
 * diagnostics' energy and momentum sums in FP32. */
enum sp_study { SP_STUDY_ABS_POS32 = 1, SP_STUDY_TWIDDLE32 = 2, SP_STUDY_SUM32 = 4 };

/* FP64 diagnostics; none has a reference counterpart.  Energies share one
 * unit: with J and rho divided by dens in the solve (src/flds/mod.rs:80-84,
 * :127), energy_e + energy_b + sum(energy_kin) is the conserved total. */
typedef struct sp_diag {
  double energy_e;       /* sum |E|^2 / 2 over the interior cells */
  double energy_b;       /* sum |B|^2 / 2 */
  double energy_kin[8];  /* species 0..7: m c^2 sum (gamma - 1) / dens */
  double momentum[3];    /* all species: sum m p / dens (p = gamma v) */
  /* Gauss law of the last field solve, relative RMS over the modes off k=0
   * and off the Nyquist lines: |k.E + i rho| / |rho| before the correction
   * (charge-continuity error of the deposit) and after it (round-off). */
  double gauss_pre;
  double gauss_post;
  double gauss_post_max; /* largest gauss_post over every solve so far */
  int64_t steps;         /* loop iterations run so far */
  int64_t sorts;         /* tile sorts of the particles run so far (sp_prestep and sp_step) */
} sp_diag;

typedef struct sp_ctx sp_ctx;
typedef struct sp_ensemble sp_ensemble;

/* Flds::new + Sim::new (src/flds/mod.rs:40-65, src/lib.rs:247-264): validates
 * the parameters like Config::new (src/lib.rs:68-117) and allocates every grid,
 * spectrum and FFT table on the current HIP device. */
int sp_create(const sp_params* params, sp_ctx** out);
int sp_destroy(sp_ctx* ctx);
/* Text of the last error on this context (or of the last failed sp_create
 * when ctx is NULL). */
const char* sp_last_error(const sp_ctx* ctx);

/* Prtl::new with caller-provided state (src/prtls.rs:60-84): charge, mass as
 * there (alpha, beta derived the same way); ix, iy are the reference's usize
 * cells as int32; dx..pz are Float arrays of length n.  gamma (psa) is
 * sqrt(1 + |p|^2/c^2), the expression of src/prtls.rs:248-249 and :440.
 * The arrays are copied; the caller keeps them. */
int sp_species_add(sp_ctx* ctx, double charge, double mass, int64_t n, const int32_t* ix,
                   const int32_t* iy, const void* dx, const void* dy, const void* px,
                   const void* py, const void* pz, int* id);
/* Prtl::new generated on the GPU (src/prtls.rs:191-261): positions exactly as
 * initialize_positions, momenta by the initialize_velocities recipe (Gaussian
 * of width vth*c, Zenitani flip, boost by gamma_inj) from a counter-based RNG
 * keyed by (seed, species id, particle index) instead of thread_rng. */
int sp_species_init(sp_ctx* ctx, double charge, double mass, double vth, uint64_t seed, int* id);
int sp_species_count(const sp_ctx* ctx, int id, int64_t* n);
/* Copies a species back in the order it was added; any output may be NULL
 * (psa is computed from the momenta whether or not px, py, pz are asked for). */
int sp_get_species(sp_ctx* ctx, int id, int64_t n, int32_t* ix, int32_t* iy, void* dx, void* dy,
                   void* px, void* py, void* pz, void* psa);
/* The species as it lies on the device (no reference counterpart; Prtl keeps
 * one order, src/prtls.rs:8-23): the cell-sorted order the particle passes
 * read, packed cells (iy << 16 | ix) and each particle's tag (its index in the
 * order the species was added).  Any output may be NULL. */
int sp_get_species_raw(sp_ctx* ctx, int id, int64_t n, int32_t* cell, int32_t* tag, void* dx, void* dy,
                       void* px, void* py, void* pz);

/* Padded real grids: count = (size_x+2)*(size_y+2) elements. */
int sp_get_grid(sp_ctx* ctx, int grid, void* out, size_t count);
int sp_set_grid(sp_ctx* ctx, int grid, const void* in, size_t count);
/* Spectra: interleaved (re, im), count = 2*size_x*size_y elements.  After a
 * step the J and rho spectra read back whole (the step keeps only their half
 * plane kx = 0..Nx/2; the getter fills the rest as conjugate mirrors); the E
 * and B buffers have been transformed back in place and hold no spectra, as
 * the reference's arrays after its in-place inverse (src/flds/mod.rs:280-290);
 * SP_BXW..SP_BZW persist. */
int sp_get_spectrum(sp_ctx* ctx, int spec, void* out, size_t count);
int sp_set_spectrum(sp_ctx* ctx, int spec, const void* in, size_t count);

/* Zero J, rho and dens including ghosts (src/lib.rs:157-167). */
int sp_zero_currents(sp_ctx* ctx);
/* Sim::move_and_deposit (src/lib.rs:478-489) for one species. */
int sp_move_and_deposit(sp_ctx* ctx, int id);
/* Field::deposit_ghosts on Jx, Jy, Jz, rho and dens (src/lib.rs:174-182,
 * src/flds/field.rs:259-373). */
int sp_deposit_ghosts(sp_ctx* ctx);
/* Flds::update (src/flds/mod.rs:87-309).  The seven forward spectra (Jx Jy
 * Jz rho Ex Ey Ez, transforms of real fields) are projected onto their
 * Hermitian part X[k] <- (X[k] + conj(X[-k])) / 2 in the same pass as the
 * spectral update -- exact in exact arithmetic; it keeps the persistent B_wrk
 * free of the FFT's anti-Hermitian rounding, which the update amplifies. */
int sp_fields_update(sp_ctx* ctx);
/* Prtl::boris_push (src/prtls.rs:262-443) for one species. */
int sp_boris_push(sp_ctx* ctx, int id);
/* `passes` calls of Field::binomial_filter_2_d on one grid
 * (src/flds/field.rs:375-426). */
int sp_filter(sp_ctx* ctx, int grid, int passes);
/* sp_filter without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  `passes` calls of
 * Field::binomial_filter_2_d on a caller's padded (size_y+2) x (size_x+2)
 * device grid of float (SP_F32) or double (SP_F64), in place, queued on
 * `stream` (a hipStream_t; NULL = the null stream) without waiting.  The
 * workspace is a device array inside the library (33,685,504 bytes per
 * device, held until the process ends): grids larger than that return
 * SP_EINVAL.  Calls that overlap on one device share the workspace and race.
 * It calls no hipMalloc or hipFree; errors read with sp_last_error(NULL). */
int sp_filter_noalloc(void* grid, int32_t size_x, int32_t size_y, int precision, int passes, void* stream);
/* Fft2D::fft / Fft2D::inv_fft on one spectrum (src/flds/fft_2d.rs:91-172). */
int sp_fft2d(sp_ctx* ctx, int spec, int inverse);
/* The pointwise spectral part of Flds::update (src/flds/mod.rs:105-278) on
 * spectra already holding the forward transforms, without the projection
 * (bit-identical to the reference's arithmetic). */
int sp_spectral_update(sp_ctx* ctx);

/* The untimed pre-loop of run(): move_and_deposit of every species, then the
 * ghost fold (src/lib.rs:130-142). */
int sp_prestep(sp_ctx* ctx);
/* n iterations of the loop body of run() (src/lib.rs:155-199): zero, move &
 * deposit every species, fold ghosts, Flds::update, Boris push every species.
 * Adjacent push / move&deposit pairs run as one fused particle pass. */
int sp_step(sp_ctx* ctx, int n);
/* Wait for the context's stream and report deferred invariant violations. */
int sp_sync(sp_ctx* ctx);
/* Phase timers (no reference counterpart beyond its stdout timers).  With
 * timing on, every phase is bracketed by HIP events on the context stream:
 * ms[0] "dep" (zero + move&deposit + ghost fold, src/lib.rs:169-184),
 * ms[1] "solve" (Flds::update, :187-189), ms[2] "push" (stand-alone Boris
 * pushes, :193-197) and ms[3] the particle passes alone (inside "dep"; a
 * fused push+move&deposit pass counts there).  Milliseconds accumulated since
 * the last reset; reset != 0 clears them. */
int sp_set_timing(sp_ctx* ctx, int on);
int sp_get_timers(sp_ctx* ctx, double* ms4, int reset);
/* Step-loop capture (no reference counterpart; the loop of src/lib.rs:155-199
 * is what gets captured).  With it on and timing off, sp_step runs whole
 * periods of 2 * sort_interval steps (2 without sorting), starting on a
 * period boundary with a push pending, as one replayed hipGraph: the buffer
 * swaps of a period cancel, so one captured graph serves every period.  It is
 * captured again when any buffer changed (species added, a stand-alone filter
 * call).  Results are those of the eager steps. */
int sp_set_graph(sp_ctx* ctx, int on);
/* Adaptive tile-sort schedule (no reference counterpart: the reference keeps
 * its particles in creation order).  The particle pass counts the particles
 * it takes through its general path (chunks whose cells spread past the LDS
 * windows, or that lie near a boundary).  With max_interval > 0 a step sorts
 * when at least min_interval steps have run since the last sort and either
 * max_interval have, or that count, as a fraction of all particles and read
 * two steps late, has risen by more than max_slow_rise since the step after
 * the last sort.  The counts depend on the particle state only, so the
 * schedule is the same in every run.  max_interval = 0 restores the fixed
 * schedule (every sort_interval steps, which must be > 0 either way).  The
 * fixed schedule applies while the step-loop graph is on and inside an
 * ensemble; after sp_checkpoint_load the adaptive count starts at the loaded
 * step. */
int sp_set_sort_policy(sp_ctx* ctx, int min_interval, int max_interval, double max_slow_rise);
/* FP64 diagnostics (no reference counterpart): energies, momentum, Gauss
 * residuals. */
int sp_get_diag(sp_ctx* ctx, sp_diag* out);
/* Restartable checkpoint.  save_output (src/save/mod.rs:25-81) writes species
 * 0 and the spatial fields only; a restart also needs every species and the
 * persistent spectral B_wrk (src/flds/mod.rs:31-33).  The file holds the
 * parameters, launch shapes, step count, every species in device order with
 * its tags, the 11 grids and the 14 spectra.  Save runs a pending push first.
 * Load goes into a context created with the same sp_params and no species;
 * stepping it continues the saved run exactly where it stopped (the learned
 * denoiser, if any, must be set again). */
int sp_checkpoint_save(sp_ctx* ctx, const char* path);
int sp_checkpoint_load(sp_ctx* ctx, const char* path);

/* ---- Launch shapes and the autotuner (no reference counterpart) ----
 * The knobs of one context's kernels.  They change speed only: cells, offsets
 * and momenta are the same for every tuning, and J/rho/dens only by the
 * summation order of the deposit. */
typedef struct sp_tuning {
  int32_t chunk;        /* particles per workgroup of the particle pass:
                           512, 1024 or 2048 (SP_F64: 1024) */
  int32_t tile_w_log;   /* sort tiles of 2^tile_w_log x 2^tile_h_log cells, */
  int32_t tile_h_log;   /*   each log in 1..5 (the deposit windows follow them) */
  int32_t fft_rows;     /* row transforms per workgroup, 1..64 (capped by LDS) */
  int32_t fft_cols;     /* column transforms per workgroup, 1..64 (capped by LDS) */
  int32_t fft_radix_x;  /* largest radix of the x / y transform plans: 8 or 4 */
  int32_t fft_radix_y;
  int32_t reserved;
} sp_tuning;
/* Defaults: chunk 2048 and 8x4 tiles, 16x2 from dens >= 16 (SP_F64: 1024 and 8x8), 1 row and 8
 * columns per workgroup, radix 8. */
int sp_set_tuning(sp_ctx* ctx, const sp_tuning* t);
int sp_get_tuning(const sp_ctx* ctx, sp_tuning* t);
/* Times the candidates of every knob on a scratch copy of the context (same
 * parameters, particles and fields; the context itself is not advanced),
 * `trials` timed repetitions each, and sets the fastest: chunk x tile shape
 * by the particle pass, radix x rows per workgroup by the row transforms,
 * radix x columns per workgroup by the column transforms.  The result is
 * kept per (grid size, dens, species count, precision, sort interval) for
 * the process; use_cache != 0 takes a kept result without measuring.
 * `out` may be NULL.  Needs at least one species. */
int sp_autotune(sp_ctx* ctx, int trials, int use_cache, sp_tuning* out);

/* ---- Ensembles of independent runs (no reference counterpart) ----
 * n contexts of any sizes stepped together on one GPU.  Each context joins
 * one of `branches` (1..16) concurrent branches, balanced by particle count;
 * every step of the ensemble forks into the branches and joins again, and
 * with the graph on (the default) `unit` steps of every member are captured
 * once as one hipGraph and replayed (captured again when a member's buffers,
 * push state, sort phase or tuning differ from every kept capture).  Members
 * keep their own results, getters and errors; they are launched on the
 * ensemble's stream while they belong to it, must outlive it (sp_destroy of
 * a member fails until then), and must have timing and their own graph off. */
int sp_ensemble_create(sp_ctx* const* members, int n, int branches, int unit, sp_ensemble** out);
int sp_ensemble_destroy(sp_ensemble* e);
const char* sp_ensemble_last_error(const sp_ensemble* e);
int sp_ensemble_set_graph(sp_ensemble* e, int on);
/* nsteps iterations of run()'s loop body for every member (as sp_step). */
int sp_ensemble_step(sp_ensemble* e, int nsteps);
/* Waits for the ensemble; reports the first member's deferred invariant
 * violation (its index in *bad_member, -1 when none; may be NULL). */
int sp_ensemble_sync(sp_ensemble* e, int* bad_member);
/* Captured graphs kept and captures made so far. */
int sp_ensemble_stats(const sp_ensemble* e, int* kept, int* captures);

/* ---- Learned current denoiser (no reference counterpart) ----
 * Five 3x3 convolutional layers (4 -> 32 -> 32 -> 32 -> 32 -> 4 channels,
 * GELU in its tanh form after each but the last) on BF16 MFMA with FP32
 * accumulation.  Input
 * x: Jx Jy Jz rho after the binomial filter, each divided by dens and a
 * channel scale; output x + r (residual).  Parameters, one FP32 vector of
 * SP_DN_NPARAMS: for each layer l = 0..4, w_l [cout][cin*9] then b_l [cout];
 * then scales [4] (w[co][ci*9 + ky*3 + kx]); the first SP_DN_NTRAIN are
 * trained. */
#define SP_DN_NTRAIN 30084
#define SP_DN_NPARAMS 30088
typedef struct sp_dn sp_dn;
/* A training context for batches of `batch` images of h x w, 4 planes each
 * ([batch][4][h][w], zero outside the image). */
int sp_dn_create(int batch, int h, int w, sp_dn** out);
int sp_dn_destroy(sp_dn* dn);
const char* sp_dn_last_error(const sp_dn* dn);
int sp_dn_set_params(sp_dn* dn, const float* params, int n);
int sp_dn_get_params(sp_dn* dn, float* params, int n);
/* Loss weight of each of the 4 channels (default 1): the loss is
 * sum_c w_c mean((x + r - t)_c^2). */
int sp_dn_set_loss_weights(sp_dn* dn, const float* w4);
/* y = x + r(x) for one batch (host arrays).  mode 0: zero outside the
 * images (training); 1: rows periodic; 2: rows periodic through the fused
 * one-launch kernel the simulation uses. */
int sp_dn_forward(sp_dn* dn, const float* x, float* y, int mode);
/* Mean squared error of x + r(x) against t and its gradient (SP_DN_NTRAIN
 * entries) for one batch. */
int sp_dn_gradient(sp_dn* dn, const float* x, const float* t, float* grad, double* loss);
/* `iters` Adam steps (lr, betas 0.9/0.999) on batches drawn with `seed` from
 * nsamples uploaded [4][h][w] pairs (xs, ts; NULL keeps the last upload);
 * loss_hist (iters entries, may be NULL) gets each batch's mean loss. */
int sp_dn_train(sp_dn* dn, const float* xs, const float* ts, int nsamples, int iters, float lr, uint64_t seed,
                double* loss_hist);
/* Apply a trained network in every field solve of the context (FP32 only),
 * after the binomial filter on its output, or in its place (see
 * sp_set_denoiser_input).  Fallback to the binomial-filtered currents when
 * |r| > max_rel |x| or r is not finite, when (max_gauss > 0) the previous
 * solve's relative pre-correction Gauss residual exceeds max_gauss, or when
 * the spectrum check of sp_set_denoiser_input fails.  params NULL disables. */
int sp_set_denoiser(sp_ctx* ctx, const float* params, int n, float max_rel, float max_gauss);
/* The network's input: raw = 0, the binomial-filtered currents (default);
 * raw = 1, the unfiltered deposit, the network then standing in for the
 * filter (src/flds/mod.rs:90-94) whose result remains the fallback (needs
 * n_pass <= 8).  max_hp > 0 adds a spectrum check on the current solve's
 * data: for every written channel, the high-pass energy sum |lap y|^2 of the
 * network's output y (5-point Laplacian, periodic) must not exceed max_hp
 * times that of the binomial-filtered grid; 0 switches it off. */
int sp_set_denoiser_input(sp_ctx* ctx, int raw, float max_hp);
/* Channels the network's output is written to when it is applied: bit c
 * for Jx, Jy, Jz, rho (default 0xF).  With rho left out, the Gauss-law
 * projection of the solve (src/flds/mod.rs:185-201) keeps the longitudinal
 * field tied to the deposited charge. */
int sp_set_denoiser_channels(sp_ctx* ctx, int mask);
/* Solves that applied the network / fell back since the denoiser was set. */
int sp_get_denoiser_stats(sp_ctx* ctx, int* applied, int* fallbacks);
/* Backoff (no reference counterpart): after a solve whose checks failed, the
 * next `solves` solves keep the binomial filter's result without running the
 * network (they count as fallbacks); 0 (the default) runs it in every solve.
 * The checks' verdict holds over long stretches of a run, so this saves the
 * inference that would only fall back again. */
int sp_set_denoiser_backoff(sp_ctx* ctx, int solves);
/* Solves skipped by the backoff since the denoiser was set. */
int sp_get_denoiser_skipped(sp_ctx* ctx, int* skipped);

#ifdef __cplusplus
}
#endif
#endif /* SPECTRAL_RS_H */
