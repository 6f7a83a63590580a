/*
 * tpm.h — C ABI of the MI355X TreePM force solver (libtpm.so).
 *
 * The reference (a Rust crate) has no FFI for this path; its "operator" is the method
 * Simulation::accelerations (src/integrator.rs:140-161) and the public functions it is made
 * of. Each entry point below names the reference function it replaces, so a Rust
 * `extern "C"` block can bind it beside the CPU path (see INTEGRATION.md for the binding).
 *
 * Conventions
 *  - Every call returns 0 on success and a negative TPM_E_* code on failure; nothing aborts
 *    across the ABI. tpm_last_error() returns a message for the last failure on a context.
 *  - Positions are AoS triples (x, y, z), the layout of Rust's Vec<[f32; 3]>; accelerations
 *    are written in the same layout and in the caller's particle order. Grids are n^3 values
 *    indexed (ix*n + iy)*n + iz, z fastest (src/mesh.rs:11,24-26).
 *  - Units are the reference's: H0 = 1, lengths in the box's units, accelerations dp/dt.
 *  - Host-pointer entry points copy in and out and return when the result is written.
 *    *_dev entry points take device pointers and enqueue on the given HIP stream
 *    (NULL = the context's stream) without synchronising.
 *  - A context owns all device memory, FFT tables and tree scratch (the PoissonSolver
 *    analogue, src/poisson.rs:23-29). Calls on one context must not overlap.
 */
#ifndef TPM_H
#define TPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TPM_OK 0
#define TPM_E_INVALID (-1)   /* bad argument (size, null pointer, mesh mismatch) */
#define TPM_E_DEVICE (-2)    /* HIP runtime error */
#define TPM_E_CAPACITY (-3)  /* more particles than the context was created for */
#define TPM_E_NONFINITE (-4) /* non-finite acceleration detected (TPM_FLAG_CHECK_FINITE) */
#define TPM_E_IO (-5)        /* file cannot be opened, written or read, or is not a snapshot */

/* Flags for tpm_params.flags. */
#define TPM_FLAG_TREEPM 1u           /* add the tree short-range force (split = Some(..)); off = pure PM */
#define TPM_FLAG_REF_LEAF_RULE 2u    /* walk a node whose octant-0 child is empty as a leaf, like
                                        the reference's children[0] test (src/tree.rs:280) */
#define TPM_FLAG_CHECK_FINITE 4u     /* scan accelerations for NaN/Inf and fail with TPM_E_NONFINITE */
#define TPM_FLAG_REF_WALK 8u         /* per-particle LIFO walk with the reference's per-particle opening
                                        test (src/tree.rs:264-310) instead of the group walk */

/*
 * Solver parameters: PoissonSolver::new(n, L) (src/poisson.rs:23) plus ForceSplit
 * (src/tree.rs:21-30) and the cosmology's Omega_m. r_split is used for the PM filter only
 * when TPM_FLAG_TREEPM is set (src/integrator.rs:142).
 */
typedef struct tpm_params {
    int32_t mesh_n;    /* PM cells per side, 2..1024 */
    uint32_t flags;    /* TPM_FLAG_* */
    double box_size;   /* periodic box side L */
    double omega_m;    /* matter density parameter */
    double r_split;    /* Gaussian split scale r_s */
    double r_cut;      /* short-range cutoff */
    double softening;  /* Plummer softening */
    double theta;      /* opening angle */
} tpm_params;

typedef struct tpm_ctx tpm_ctx;

/* Timing and work counters of the last force evaluation on a context. */
typedef struct tpm_stats {
    double ms_total;       /* device time of the whole evaluation */
    double ms_sort;        /* Morton keys + sort + gather */
    double ms_deposit;     /* fixed-point CIC deposit */
    double ms_fft;         /* r2c, Green multiply, inverse passes */
    double ms_interp;      /* CIC force interpolation */
    double ms_tree;        /* octree build + centres of mass */
    double ms_walk;        /* short-range walk */
    uint64_t interactions; /* particle-particle + particle-node interactions evaluated */
    uint64_t nodes;        /* octree nodes */
    uint64_t groups;       /* walk groups (wavefronts) */
    uint64_t listed;       /* particle sources listed by the group walks (before the box test) */
    uint64_t kept;         /* particle sources kept for evaluation (each costs one 64-lane step) */
    double ms_walk_kernel; /* the walk kernel alone (ms_walk also builds the walk groups) */
} tpm_stats;

/* Version string of the library. */
const char *tpm_version(void);

/* Creates a context for up to max_particles particles. Replaces Simulation::new's
 * PoissonSolver::new (src/integrator.rs:67, src/poisson.rs:23-29). */
int tpm_create(const tpm_params *params, size_t max_particles, tpm_ctx **out);

/* Releases every device allocation of the context. */
void tpm_destroy(tpm_ctx *ctx);

/* Message for the last failed call on ctx (or for tpm_create when ctx is NULL). */
const char *tpm_last_error(const tpm_ctx *ctx);

/* Updates the parameters of an existing context (mesh_n must not change). */
int tpm_set_params(tpm_ctx *ctx, const tpm_params *params);

/*
 * Full TreePM force evaluation: Simulation::accelerations (src/integrator.rs:140-161).
 * pos: 3n positions in [0, L); acc: 3n accelerations out; delta: n_mesh^3 density contrast
 * out (src/integrator.rs:117-126), or NULL. The f32 variant computes CIC weights in FP32 as
 * the reference does; the f64 variant computes them in FP64. Both walk the tree in FP32 on
 * positions kept in FP64.
 */
int tpm_accelerations_f32(tpm_ctx *ctx, const float *pos, size_t n, double a, float *acc, float *delta);
int tpm_accelerations_f64(tpm_ctx *ctx, const double *pos, size_t n, double a, float *acc, float *delta);

/* Device-pointer form of tpm_accelerations_f64, enqueued on `stream` (hipStream_t or NULL). */
int tpm_accelerations_dev_f64(tpm_ctx *ctx, const double *d_pos, size_t n, double a, float *d_acc,
                              float *d_delta, void *stream);

/*
 * Device-side failures of the evaluations enqueued on ctx since the status was last cleared
 * (no reference counterpart; the reference cannot fail this way). The *_dev entry points never
 * wait for the device, so they cannot return these; each evaluation ORs its flags into a sticky
 * word that this call reads after waiting for the context's stream and the stream of the last
 * evaluation. clear != 0 resets the word. The host-pointer entry points return such failures as
 * error codes as well.
 */
#define TPM_STATUS_WALK_OVERFLOW 1u  /* a walk's traversal stack overflowed: forces incomplete */
#define TPM_STATUS_TREE_CAPACITY 2u  /* the octree outgrew its node capacity: forces incomplete */
#define TPM_STATUS_NONFINITE 4u      /* a non-finite acceleration (only with TPM_FLAG_CHECK_FINITE) */
int tpm_get_device_status(tpm_ctx *ctx, uint32_t *status, int clear);

/* Statistics of the last force evaluation (device timings need TPM_FLAG timing enabled via
 * tpm_set_timing). */
int tpm_get_stats(const tpm_ctx *ctx, tpm_stats *out);
int tpm_set_timing(tpm_ctx *ctx, int enabled);
/* Interaction counting in the group walk (no reference counterpart): on by default; off, the walk
 * saves one vector instruction per particle pair and tpm_stats.interactions reads 0. */
int tpm_set_counting(tpm_ctx *ctx, int enabled);
/* The octree on a second stream beside the deposit, FFT and interpolation, for evaluations of more
 * than 2^22 particles (no reference counterpart): on by default; off, every stage runs on the
 * context's stream in turn (same forces, bit for bit; each kernel's time measured alone). */
int tpm_set_overlap(tpm_ctx *ctx, int enabled);

/*
 * The particle order of the octree: Tree::build's counting sort into octants (src/tree.rs:215-236)
 * applied at every level at once. order[k] is the caller index of the k-th particle in Morton
 * order (octant digits bit 0 = x, bit 1 = y, bit 2 = z, per axis floor(x / L * 2^21)), sorted
 * stably by the key digits of the context's octree depth (input order below it); keys (may be
 * NULL) receives the sorted particles' 63-bit Morton keys (for a depth of at most 10, the digits of
 * levels 11-21 read zero). Host memory.
 */
int tpm_morton_order(tpm_ctx *ctx, const double *pos, size_t n, uint32_t *order, uint64_t *keys);

/* Mesh::deposit (src/mesh.rs:62-69): cells (n_mesh^3) = CIC deposit of `mass` per particle,
 * accumulated in 2^-32 fixed point (bit-exact run to run). */
int tpm_deposit_f32(tpm_ctx *ctx, const float *pos, size_t n, float mass, float *cells);
int tpm_deposit_f64(tpm_ctx *ctx, const double *pos, size_t n, double mass, float *cells);

/*
 * Allocation-free variant of tpm_deposit_f64 (no counterpart in the reference): for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops that forbid
 * hipMalloc). No context: d_pos (3n FP64 positions) and d_cells (mesh_n^3 floats) are device
 * pointers; the work is enqueued on `stream` (hipStream_t, NULL = the null stream) without
 * synchronising. The same fixed-point CIC deposit, so d_cells equals tpm_deposit_f64's output bit
 * for bit. Its 2^-32 fixed-point grid is a 16 MiB array built into the library (one per GPU, held
 * until the process ends), so mesh_n is limited to [2, 128]; larger meshes fail with
 * TPM_E_INVALID before any device work. Calls that overlap on one GPU share that grid and race.
 * Errors: tpm_last_error(NULL). The first call on a device looks up the array's address there.
 */
int tpm_deposit_dev_f64_noalloc(const double *d_pos, size_t n, int32_t mesh_n, double box_size, double mass,
                                float *d_cells, void *stream);

/* Simulation::density_contrast (src/integrator.rs:117-126). */
int tpm_density_contrast_f32(tpm_ctx *ctx, const float *pos, size_t n, float *delta);

/* PoissonSolver::acceleration (src/poisson.rs:33-88): delta (n_mesh^3) -> gx, gy, gz grids,
 * filtered by exp(-k^2 r_split^2) with the given r_split (0 = unfiltered). */
int tpm_pm_accel(tpm_ctx *ctx, const float *delta, double a, double r_split, float *gx, float *gy, float *gz);

/* Mesh::interpolate (src/mesh.rs:72-77) of one grid at n positions. */
int tpm_interpolate_f32(tpm_ctx *ctx, const float *cells, const float *pos, size_t n, float *out);

/* tree::accelerations (src/tree.rs:314-323): short-range force only, coupling k given
 * explicitly as in the reference. Uses the context's r_split, r_cut, softening, theta. */
int tpm_short_range_f32(tpm_ctx *ctx, const float *pos, size_t n, double k, float *acc);
int tpm_short_range_f64(tpm_ctx *ctx, const double *pos, size_t n, double k, float *acc);

/*
 * FP64 TreePM judge (no reference counterpart; the error yardstick for force accuracy):
 * the full TreePM force in double precision (FP64 deposit, FFT, Green's function,
 * interpolation) with the per-particle reference walk at opening angle `theta`, evaluated
 * for the m particles listed in idx. acc: 3m doubles.
 */
int tpm_judge_f64(tpm_ctx *ctx, const double *pos, size_t n, const uint32_t *idx, size_t m, double a,
                  double theta, double *acc);

/*
 * Zel'dovich initial conditions on the mesh_n^3 particle lattice (the structure of the
 * reference's zeldovich, src/ics.rs:119-182): seeded counter-based white noise, power-law
 * spectrum P(k) = k^-2 exp(-k/k_c), displacement rescaled to a 3D RMS of disp_rms,
 * x = q + Psi wrapped into [0, L), p = mom_factor * Psi (mom_factor = a^2 f E,
 * src/ics.rs:157-160). pos, mom: 3 mesh_n^3 doubles in lattice order (ix, iy, iz); mom may
 * be NULL. Needs max_particles >= mesh_n^3.
 */
int tpm_zeldovich(tpm_ctx *ctx, uint64_t seed, double k_c, double disp_rms, double mom_factor, double *pos,
                  double *mom);
int tpm_zeldovich_dev(tpm_ctx *ctx, uint64_t seed, double k_c, double disp_rms, double mom_factor, double *d_pos,
                      double *d_mom, void *stream);

/*
 * Device memory helpers for the *_dev entry points (no reference counterpart: a host that keeps
 * particles resident on the GPU needs them; they use the same HIP runtime as the kernels).
 * kind: 0 host->device, 1 device->host, 2 device->device. tpm_sync waits for the context stream.
 */
int tpm_device_alloc(size_t bytes, void **out);
int tpm_device_free(void *ptr);
int tpm_memcpy(void *dst, const void *src, size_t bytes, int kind);
int tpm_sync(tpm_ctx *ctx);

/*
 * Clustered particle set (no reference counterpart; the tree walk's stress workload): n_halos
 * halos with masses from dN/dM ~ M^-slope on [mass_min, mass_max] (M_sun/h), NFW concentrations
 * uniform in [conc_min, conc_max], centres uniform in the box, particles placed isotropically
 * inside r200 (rho_crit = 2.775e11 h^2 M_sun/Mpc^3, lengths in the box's Mpc/h units) in numbers
 * proportional to mass, halo_fraction of all particles; the rest uniform. Counter-based
 * randomness: the output depends only on the spec and n.
 */
typedef struct tpm_nfw_spec {
    uint64_t seed;
    int32_t n_halos;
    double halo_fraction;
    double mass_min, mass_max, slope;
    double conc_min, conc_max;
} tpm_nfw_spec;

int tpm_nfw_halos(tpm_ctx *ctx, const tpm_nfw_spec *spec, size_t n, double *pos);
int tpm_nfw_halos_dev(tpm_ctx *ctx, const tpm_nfw_spec *spec, size_t n, double *d_pos, void *stream);

/*
 * Learned short-range correction (no reference counterpart; the exact walk is its fallback and
 * judge). A 64 -> 128 -> 128 -> 128 -> 4 SiLU MLP with BF16 weights and activations on MFMA (FP32
 * accumulation) predicts the short-range force from 33 features per particle (log(1 + delta) of its
 * cell and, relative to that, of the 26 cells around it; the PM force; its offset from the cell
 * centre in cells) and a
 * log-variance used as the error estimate. Forces are in the compressed units slog(F / (K/h^2)),
 * slog(x) = sign(x) log(1 + |x|), with K the coupling and h the cell size.
 * Parameters are FP32 masters (tpm_mlp_param_count() floats, layout: per layer Wt [out x in] then
 * bias, with in = 64, 128, 128, 128 and out = 128, 128, 128, 32; only 33 inputs and 4 outputs used).
 */
size_t tpm_mlp_param_count(void);
int tpm_mlp_init(tpm_ctx *ctx, uint64_t seed);
int tpm_mlp_get_params(tpm_ctx *ctx, float *params, size_t count);
int tpm_mlp_set_params(tpm_ctx *ctx, const float *params, size_t count);
/* Forward + backward (no update) on host samples: x rows x 33, t rows x 3 (compressed units);
 * grads: param_count. */
int tpm_mlp_gradients(tpm_ctx *ctx, const float *x, const float *t, size_t rows, float *grads, double *loss);
/* Forward on host samples: x rows x 33 -> out rows x 4 (force in K/h^2, log variance). */
int tpm_mlp_predict(tpm_ctx *ctx, const float *x, size_t rows, float *out);
/*
 * Trains on n_samples particles of d_pos (device, FP64), with the walk's short-range force as the
 * target, for `epochs` passes of Adam (batch rows, learning rate lr), then calibrates the gate on up
 * to 65,536 particles midway between training samples (their predicted and actual errors, see
 * tpm_mlp_error_scale). loss_history (epochs entries, may be NULL) receives each epoch's last-batch
 * loss.
 */
int tpm_mlp_train_dev(tpm_ctx *ctx, const double *d_pos, size_t n, double a, size_t n_samples, int epochs,
                      size_t batch, float lr, uint64_t seed, double *loss_history);
/*
 * TreePM force with the learned short-range term: PM force + the MLP's prediction wherever its error
 * estimate, kappa * sigma, is within tol * |a| (sigma: the network's predicted RMS error propagated to
 * force units; kappa: tpm_mlp_error_scale for this tol, from the calibration of tpm_mlp_train_dev);
 * every other particle is walked exactly (the context's walk: the group walk, or
 * with TPM_FLAG_REF_WALK the per-particle walk, whose result for a particle does not depend on which
 * others are walked). Features, network and gate run as one kernel. n_fallback (may be NULL)
 * receives the number of walked particles; d_walked (device, n words in caller order, may be NULL)
 * receives 1 for each walked particle and 0 for each predicted one.
 */
int tpm_accelerations_learned_dev(tpm_ctx *ctx, const double *d_pos, size_t n, double a, double tol, float *d_acc,
                                  uint64_t *n_fallback, uint32_t *d_walked);
/*
 * The gate's error-estimate multiplier kappa for tolerance tol. After training: tol / tau, with tau the
 * largest relative estimate sigma / |a| of the calibration particles such that 99% of those with
 * estimates up to tau (at least 100) have actual errors within tol * |a|; +inf when none qualify (the
 * gate then accepts nothing). Before any training: 3 sqrt(3).
 */
int tpm_mlp_error_scale(tpm_ctx *ctx, double tol, double *kappa);

/* Background cosmology for the KDK step factors (src/cosmology.rs:9-34, flat: Or + Om + OL = 1). */
typedef struct tpm_cosmology {
    double omega_r;
    double omega_m;
    double omega_l;
} tpm_cosmology;

/*
 * The reference's default IC spectrum (src/ics.rs:20-64): P(k) = A k^n_s T_BBKS(k)^2 with
 * q = k / (Omega_m h0 / 100) (src/cosmology.rs:93-103), A set so that sigma(8 Mpc/h) = sigma8 (top-hat,
 * Simpson in ln k over [1e-4, 1e3], 8192 intervals), all in FP64. No device work.
 * tpm_power_spectrum gives P(k) and A; tpm_power_sigma gives sigma(radius) of the normalised spectrum.
 */
int tpm_power_spectrum(const tpm_cosmology *cosmo, double h0, double n_s, double sigma8, double k, double *p,
                       double *amplitude);
int tpm_power_sigma(const tpm_cosmology *cosmo, double h0, double n_s, double sigma8, double radius, double *sigma);
/*
 * zeldovich (src/ics.rs:111-182) with that spectrum: mesh_n^3 lattice particles at scale factor a_init,
 * x = q + D(a) Psi wrapped into [0, L), p = a^2 f E D Psi, from the generator's own seeded noise.
 */
int tpm_zeldovich_bbks(tpm_ctx *ctx, const tpm_cosmology *cosmo, double h0, double n_s, double sigma8, uint64_t seed,
                       double a_init, double *pos, double *mom);
int tpm_zeldovich_bbks_dev(tpm_ctx *ctx, const tpm_cosmology *cosmo, double h0, double n_s, double sigma8,
                           uint64_t seed, double a_init, double *d_pos, double *d_mom, void *stream);

/*
 * Per-step outputs of the runner's snapshot (src/runner.rs:262-284) for FP64 device state, all
 * device pointers, asynchronous on `stream` (NULL: the context's): positions as float (3n),
 * speeds |p| / a (n), densities max(1 + delta(x), 0) CIC-sampled from the n_mesh^3 density-contrast
 * mesh d_delta (n), redshift-space shifts p / (a^2 E) (frame.rs:56-64, 3n). Any output may be NULL.
 */
int tpm_snapshot_dev(tpm_ctx *ctx, const double *d_pos, const double *d_mom, size_t n, double a, double e,
                     const float *d_delta, float *d_pos32, float *d_speed, float *d_density, float *d_rsd,
                     void *stream);
/*
 * downsample_grid (src/frame.rs:33-52): 1 + delta averaged over f^3 blocks, f = ceil(n_mesh / 128),
 * side m = n_mesh / f, x fastest. *side receives m; with d_out NULL nothing else is done.
 */
int tpm_downsample_grid_dev(tpm_ctx *ctx, const float *d_delta, float *d_out, int32_t *side, void *stream);

/*
 * Binary snapshot (planned in the reference, plan.md:160-163: positions, velocities, a, box size
 * and parameters, saved at chosen redshifts; the reference defines no format). File layout,
 * little-endian: this 128-byte header (magic "TPMSNAP1", version 1), then n positions and n
 * momenta, 3 doubles each, in caller order. The momentum is the reference's p (velocity = p / a).
 * tpm_snapshot_write sets magic, version and header_bytes itself. tpm_snapshot_read reads the
 * header and, when pos or mom is not NULL, the arrays (capacity >= n particles, else TPM_E_INVALID);
 * TPM_E_IO for a file that cannot be read or is not a snapshot. Host memory.
 */
typedef struct tpm_snapshot_header {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    uint64_t n;
    double a;
    double box_size;
    double omega_m, omega_l, omega_r;
    double h0;        /* Hubble parameter of the ICs' power spectrum, km/s/Mpc */
    double softening; /* force split (zeros for a PM-only run) */
    double r_split;
    double r_cut;
    double theta;
    int32_t mesh_n;
    int32_t step; /* KDK steps taken */
    uint64_t seed;
    uint64_t reserved;
} tpm_snapshot_header;

/*
 * Headless runner: RunSettings (src/runner.rs:29-82) and the worker loop of the `--no-viz` run
 * (src/runner.rs:186-260, src/main.rs:34-37) on the GPU. tpm_run_settings_default gives
 * RunSettings::default, tpm_run_settings_sanitized applies RunSettings::sanitized's clamps (FP32).
 * tpm_headless_run sanitizes, builds the ICs (BBKS/sigma_8 Zel'dovich at z_start, seed), runs
 * `steps` KDK steps evenly spaced in ln a on a mesh of twice the particle count per side (TreePM
 * with ForceSplit::with when treepm != 0) and calls on_frame after each step with the snapshot in
 * device memory (valid during the call; `snapshot` describes the run for tpm_snapshot_write with
 * d_state_pos and d_state_mom copied to the host); a non-zero return stops the run. The final positions and
 * momenta (3 N doubles each, N = particles_per_side^3 <= capacity) are copied out when given.
 */
typedef struct tpm_run_settings {
    double z_start;
    double z_end;
    int32_t steps;
    int32_t particles_per_side;
    uint64_t seed;
    int32_t treepm;
    double theta;
    double softening_scale;
    uint32_t pacing_ms; /* kept for the mirror; the headless loop does not pause */
} tpm_run_settings;

typedef struct tpm_frame {
    int32_t step;
    int32_t steps;
    double redshift;
    double max_speed; /* the run's blow-up check: max |p| / a */
    double elapsed_s;
    size_t n;
    const float *d_pos;     /* 3N: positions as float */
    const float *d_speed;   /* N: |p| / a */
    const float *d_density; /* N: max(1 + delta, 0) at the particles */
    const float *d_rsd;     /* 3N: p / (a^2 E) */
    const float *d_delta;   /* mesh_n^3 density contrast of the step's force evaluation */
    int32_t mesh_n;
    double a;                  /* scale factor after the step */
    const double *d_state_pos; /* 3N: the FP64 positions (device, valid during the callback) */
    const double *d_state_mom; /* 3N: the FP64 momenta */
    const tpm_snapshot_header *snapshot; /* this run's parameters, with a, n and step of this frame */
} tpm_frame;

typedef int (*tpm_frame_fn)(void *user, const tpm_frame *frame);

void tpm_run_settings_default(tpm_run_settings *settings);
int tpm_run_settings_sanitized(const tpm_run_settings *in, tpm_run_settings *out);
int tpm_headless_run(const tpm_run_settings *settings, tpm_frame_fn on_frame, void *user, double *final_pos,
                     double *final_mom, size_t capacity);

int tpm_snapshot_write(const char *path, const tpm_snapshot_header *header, const double *pos, const double *mom);
int tpm_snapshot_read(const char *path, tpm_snapshot_header *header, double *pos, double *mom, size_t capacity);

/*
 * Autotuning (no reference counterpart): times the candidates of the knobs below on the given input
 * and keeps the fastest (the default is a candidate) among those whose force error against the FP64
 * judge stays within 10% of the default's: the FFT kernels (wavefront transforms or LDS line
 * kernels), the cap on FFT lines per workgroup of the latter, the octree level of the walk groups'
 * cubes (as an offset from the default level), the octree depth limit and the small-node rule.
 */
typedef struct tpm_tuning {
    int32_t fft_cap;
    int32_t group_level_shift;
    int32_t fft_path; /* 0: wavefront transforms where available (mesh side 512), 1: LDS line kernels */
    int32_t tree_depth; /* octree depth limit, 8..21 (default 21): deeper nodes stay leaves and fewer key bits are
                           sorted; raised where the deposit cubes or the walk groups need finer sorted levels */
    int32_t small_node_direct; /* 0 (default): nodes of at most 16 particles that pass the opening criterion are
                                  monopoles; 1: they are summed directly too (more pairs, no monopole) */
    double ms_default; /* force evaluation with the default knobs */
    double ms_tuned;   /* with the chosen knobs */
    double err_default; /* force error with the default knobs: RMS |a - a_judge| / RMS |a_judge| against the
                           FP64 judge (theta 0.2) at up to 2048 evenly spaced particles */
    double err_tuned;   /* the same with the chosen knobs (candidates above 1.1 err_default are rejected) */
} tpm_tuning;

int tpm_autotune(tpm_ctx *ctx, const double *d_pos, size_t n, double a, float *d_acc, tpm_tuning *out);
int tpm_set_tuning(tpm_ctx *ctx, const tpm_tuning *tuning);

/*
 * Ensemble of independent periodic boxes (no reference counterpart): box b has sides[b]^3
 * particles on a sides[b]^3 mesh, the split of ForceSplit::with (src/tree.rs:40-50) with the given
 * softening (in cells) and theta. A run is the initial force and `steps` KDK steps of every box;
 * with use_graph it is captured once into a HIP graph (boxes forked onto their own streams) and
 * executed by one graph launch. ms receives the run's device time.
 */
typedef struct tpm_ensemble_params {
    double box_size;
    double omega_m;
    double softening_scale; /* softening in mesh cells */
    double theta;
} tpm_ensemble_params;

typedef struct tpm_ensemble tpm_ensemble;

int tpm_ensemble_create(const tpm_ensemble_params *params, int nboxes, const int32_t *sides, tpm_ensemble **out);
void tpm_ensemble_destroy(tpm_ensemble *ens);
const char *tpm_ensemble_last_error(const tpm_ensemble *ens);
/* Lattice Zel'dovich ICs per box (seeds[b]; displacement RMS in cells; momenta = mom_factor Psi). */
int tpm_ensemble_ics(tpm_ensemble *ens, const uint64_t *seeds, double k_c, double disp_cells, double mom_factor);
/* Autotunes each distinct box side on its first box; out (nboxes entries, may be NULL). */
int tpm_ensemble_autotune(tpm_ensemble *ens, double a, tpm_tuning *out);
int tpm_ensemble_run(tpm_ensemble *ens, const tpm_cosmology *cosmo, double a_start, double a_end, int steps,
                     int use_graph, double *ms);
/* Node counts of the captured run's HIP graph by type (kernels, memsets, other nodes such as event
 * records and waits); zeros before the first graph run. A capture that records a memset node is
 * refused (TPM_E_DEVICE): every operation of a run is a kernel. */
int tpm_ensemble_graph_nodes(tpm_ensemble *ens, int32_t *kernels, int32_t *memsets, int32_t *others);
/* Copies box `box`'s current positions, momenta (3 side^3 doubles each) and accelerations (floats). */
int tpm_ensemble_get(tpm_ensemble *ens, int box, double *pos, double *mom, float *acc);

/*
 * KDK step factors of Simulation::step_to (src/integrator.rs:174-186, 262-297) in FP64: the time
 * midpoint am, the half-kick integrals k1 = int_{a0}^{am} da/(aE), k2 = int_{am}^{a1} da/(aE) and
 * the drift integral d = int_{a0}^{a1} da/(a^3 E). No device work.
 */
int tpm_kdk_factors(const tpm_cosmology *cosmo, double a0, double a1, double *am, double *k1, double *d, double *k2);

/*
 * Linear growth D(a) (D(1) = 1), growth rate f = dlnD/dlna (Cosmology::growth, growth_rate:
 * src/cosmology.rs:41-91, the RK4 growth ODE in FP64) and E(a) = H/H0 (src/cosmology.rs:31-33).
 * Zel'dovich momenta are p = a^2 f E x displacement (src/ics.rs:155-160). No device work.
 */
int tpm_growth(const tpm_cosmology *cosmo, double a, double *d, double *f, double *e);

/*
 * One KDK step, Simulation::step_to (src/integrator.rs:174-186), on device arrays: p += acc k1;
 * x = (x + p d) mod L; acc = accelerations(a1) (written over d_acc); p += acc k2. d_acc must hold
 * the accelerations at a0 on entry (e.g. from tpm_accelerations_dev_f64). Positions and momenta
 * are FP64. d_delta (n_mesh^3) receives the density contrast at a1, or NULL.
 */
int tpm_step_to_dev(tpm_ctx *ctx, const tpm_cosmology *cosmo, double *d_pos, double *d_mom, size_t n, double a0,
                    double a1, float *d_acc, float *d_delta, void *stream);

/*
 * Simulation::new followed by Simulation::run (src/integrator.rs:54-79, 165-171): the initial
 * force at a_start, then `steps` KDK steps to a_end evenly spaced in ln a, on host arrays of FP64
 * positions and momenta updated in place.
 */
int tpm_run(tpm_ctx *ctx, const tpm_cosmology *cosmo, double *pos, double *mom, size_t n, double a_start,
            double a_end, int steps);

#ifdef __cplusplus
}
#endif

#endif /* TPM_H */
