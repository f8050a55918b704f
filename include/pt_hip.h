/*
 * pt_hip.h — C ABI of libpt_hip.so, the MI355X (gfx950) trajectory and
 * lag-statistics library for particle_transport.
 *
 * The reference has no plugin API: its Julia functions call CUDA.jl directly
 * (@cuda at phase_space_gpu_runner.jl:569, multimode_multienergy_cache_pipeline.jl:1474,1484,
 * compute_delta_mu2_curve.jl:203,355) and run the pair statistics on host
 * threads.  Each entry point below replaces one of those call sites; the
 * comment on each names the reference interface it replaces.
 *
 * Conventions
 *   - Every function returns int status: PT_OK (0) or a PT_ERR_* code.  No
 *     C++ exception crosses this boundary.  pt_last_error() returns the
 *     message of the last failure on the calling thread.
 *   - Arrays follow the reference's Julia layouts: column-major, first index
 *     fastest.  A Julia array (a,b,c) is the C array [c][b][a].
 *   - Lag steps are the reference's 1-based cached-step offsets (>= 1).
 *     Particle indices passed in selection lists are 0-based.
 *   - Host pointers belong to the caller and are only read/written during
 *     the call.  Device memory belongs to the handles.  "on_device" arguments
 *     say whether a pointer is a HIP device pointer (1) or host memory (0).
 *   - A context owns one HIP stream; every call on that context is ordered on
 *     it and returns after the stream has drained.  A context is not
 *     thread-safe: one host thread drives it, as one Julia task does today.
 */
#ifndef PT_HIP_H
#define PT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PT_ABI_VERSION 1

enum pt_status {
    PT_OK = 0,
    PT_ERR_INVALID = 1,  /* bad argument (shape, range, null pointer) */
    PT_ERR_HIP = 2,      /* HIP runtime failure */
    PT_ERR_MEMORY = 3,   /* device allocation failed or over budget */
    PT_ERR_STATE = 4     /* call not valid in the handle's current state */
};

enum pt_dtype { PT_F64 = 0, PT_F32 = 1 };

/* Field storage.  PLANAR = six separate (nx,ny,nz) column-major arrays, the
 * reference layout (phase_space_gpu_runner.jl:515-520).  BRICK = the six
 * components of a node interleaved in one 16-byte (BF16/FP16) or 32-byte
 * (FP32) record, nodes grouped in 8x8x8 bricks; BF16/FP16 bricks store
 * (value - offset_c) / scale_c per component c. */
enum pt_storage {
    PT_STORE_F64_PLANAR = 0,
    PT_STORE_F32_PLANAR = 1,
    PT_STORE_F32_BRICK = 2,
    PT_STORE_BF16_BRICK = 3,
    PT_STORE_F16_BRICK = 4
};

/* Pusher kinds.
 *  REF_F64:    FP64 state (position in metres, momentum in SI), reference
 *              operation order; needs a PT_STORE_F64_PLANAR field.
 *  PLAIN_F32:  reference structure in FP32 with split position (int32 cell +
 *              FP32 offset); state in HBM, six planar FP32 arrays, all 48
 *              corner values gathered every step, one launch per step.
 *  FAST:       FP32 pusher with split position; many steps per launch with
 *              state and the current cell's interpolation coefficients in
 *              registers; any BRICK field (or F32 planar). */
enum pt_pusher { PT_PUSH_REF_F64 = 0, PT_PUSH_PLAIN_F32 = 1, PT_PUSH_FAST = 2 };

/* Cadence output kinds (cache_mode in the reference's [run] section). */
enum pt_out { PT_OUT_NONE = 0, PT_OUT_PHASE = 1, PT_OUT_MU = 2 };

typedef struct pt_ctx pt_ctx;
typedef struct pt_field pt_field;
typedef struct pt_ens pt_ens;

/* ---- context ------------------------------------------------------------ */

int pt_abi_version(void);
const char *pt_last_error(void);

/* Replaces CUDA.jl's implicit device/stream setup (the reference uses the
 * default device and stream).  device: HIP ordinal. */
int pt_ctx_create(int device, pt_ctx **out);
int pt_ctx_destroy(pt_ctx *ctx);
/* Device bytes currently held by the context: fields, ensembles, buffers, its
 * scratch, and the per-call temporaries it keeps for reuse (at most 64 MiB each). */
int pt_ctx_bytes_in_use(const pt_ctx *ctx, int64_t *bytes);
/* Free and total memory of the context's device (hipMemGetInfo), for
 * monitoring a serving process.  No reference counterpart. */
int pt_device_memory(const pt_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes);

/* Device buffers owned by the context (cadence caches that stay in HBM and
 * feed the statistics entry points with on_device = 1).  No reference
 * counterpart: the reference holds these as CuArrays. direction: 0 host to
 * device, 1 device to host, 2 device to device. */
int pt_buffer_alloc(pt_ctx *ctx, int64_t bytes, void **dev_ptr);
int pt_buffer_free(pt_ctx *ctx, void *dev_ptr, int64_t bytes);
int pt_memcpy(pt_ctx *ctx, void *dst, const void *src, int64_t bytes, int direction);
/* Strided copy: `rows` rows of `width` bytes, row r from src + r * src_pitch to
 * dst + r * dst_pitch, by the copy engine (hipMemcpy2DAsync), then synchronised.
 * Replaces the host gather plus CuArray upload of a batch of particle columns
 * (read_particle_batch, compute_delta_mu2_dmumu_full.jl:411-437 and
 * compute_dpp_full.jl:200-220, called at :1137-1140 and :507): a contiguous run of particles of a cadence
 * cache in page-locked host memory goes to the device without the host reading it. */
int pt_memcpy_2d(pt_ctx *ctx, void *dst, int64_t dst_pitch, const void *src, int64_t src_pitch, int64_t width,
                 int64_t rows, int direction);
/* Page-locked host memory for fast cache downloads (hipHostMalloc). */
int pt_host_alloc(int64_t bytes, void **host_ptr);
int pt_host_free(void *host_ptr);

/* ---- fields ------------------------------------------------------------- */

/* Replaces the CuArray uploads of the six field components
 * (phase_space_gpu_runner.jl:515-520; multimode_multienergy_cache_pipeline.jl:1419-1424)
 * after load_static_fields (phase_space_gpu_runner.jl:95-123) has converted them to
 * SI.  comps = {Bx, By, Bz (tesla), ux, uy, uz (m/s)}, each Julia (nx,ny,nz).
 * box_length_m = box_length_pc * PC_TO_M; grid nodes x_i = range(0, L, length=n)
 * (:125-130).  The FP64 mean |B| of the source values (reference_B0_T,
 * time_units.jl:18-20) is computed on the device at creation. */
int pt_field_create(pt_ctx *ctx, int64_t nx, int64_t ny, int64_t nz, double box_length_m,
                    int storage, int src_dtype, int src_on_device, const void *const comps[6],
                    pt_field **out);
int pt_field_destroy(pt_field *field);

/* Cell table (no counterpart in the reference, which samples the field's
 * nodes every push): enable = 1 builds, for every cell of an FP32 planar or
 * brick field, the 6 x 8 FP32 trilinear coefficients that the production
 * pushers (PT_PUSH_FAST and the serving tick) otherwise rebuild from the
 * stored nodes whenever a particle changes cell, 192 bytes per cell in 8^3
 * cell bricks (25.8 GB for a 512^3-cell field); those pushers then load a
 * cell's record instead.  Trajectories are bit-identical with and without the
 * table; it pays where particles change cell every few tens of pushes
 * (10-100 TeV protons in the benchmark fields).  enable = 0 frees it;
 * pt_field_destroy frees it too.  A server's fields must all have one or all
 * have none. */
int pt_field_cell_table(pt_field *field, int enable);

/* reference_B0_T (time_units.jl:18-20): mean over all nodes of
 * sqrt(Bx^2+By^2+Bz^2), summed in FP64 on the device. */
int pt_field_mean_abs_b(const pt_field *field, double *B0_T);
/* Same reduction on caller arrays (total-field clock for decomposed modes,
 * README.md:132-148). */
int pt_mean_abs_b(pt_ctx *ctx, int64_t n, int src_dtype, int src_on_device, const void *bx,
                  const void *by, const void *bz, double *B0_T);

/* ---- ensembles ---------------------------------------------------------- */

/* Replaces the particle-state CuArray uploads and CUDA.fill(true, n) of
 * run_gpu_ensemble (phase_space_gpu_runner.jl:506-513) and run_mu_cache_ensemble
 * (multimode_multienergy_cache_pipeline.jl:1410-1417).  x/p: three host or device
 * arrays of n values each (positions in metres, momenta in kg m/s), as
 * init_particles returns them (phase_space_gpu_runner.jl:196-245).  The field
 * fixes the geometry and must outlive the ensemble.  periodic: boundary
 * "periodic" (1) or "open" (0). */
int pt_ens_create(pt_ctx *ctx, const pt_field *field, int pusher, int64_t n, int src_dtype,
                  int src_on_device, const void *const x[3], const void *const p[3], int periodic,
                  pt_ens **out);
int pt_ens_destroy(pt_ens *ens);

/* Injection settings of init_particles (phase_space_gpu_runner.jl:196-245),
 * drawn on the device.  The reference draws from MersenneTwister(seed), one
 * particle after another; here every particle draws from its own counters of
 * Philox4x64-10 (key (seed, 0), counter (particle index, block, 0, 0), four
 * 64-bit words per block, u = (word >> 11) * 2^-53), in the reference's
 * per-particle order: three position draws unless the position is fixed, then
 * z = 2u - 1 and phi = 2 pi u (rand_unit_vector, :61-66) or, for fixed mu,
 * phi = 2 pi u about the local B (fixed_mu_unit_vector, :170-194).  The state
 * of particle i does not depend on how the particles are split over calls. */
typedef struct {
    uint64_t seed;            /* cfg[:seed] */
    int32_t fixed_position;   /* injection_position_mode: 0 random, 1 fixed */
    int32_t fixed_mu;         /* injection_mode: 0 isotropic, 1 fixed_mu */
    double position_m[3];     /* fixed_injection_position (:145-168), metres */
    double mu0;               /* injection_mu0, in [-1, 1] */
    double p_magnitude;       /* pscale = gamma0 * M_P * v0 (:212), kg m/s */
} pt_inject_params;

/* Replaces the host loop of init_particles (phase_space_gpu_runner.jl:196-245):
 * writes the FP64 initial state of global particles first_index ..
 * first_index+n-1 into out = [x1 x2 x3 p1 p2 p3][n] (metres, kg m/s), host or
 * device memory.  Fixed-mu injection interpolates B (FP64 trilinear,
 * trilinear_sample :247-276) from the values the field stores and fails with
 * the reference's error where |B| is zero or non-finite. */
int pt_inject(pt_ctx *ctx, const pt_field *field, int64_t first_index, int64_t n,
              const pt_inject_params *params, int out_on_device, double *out);

/* pt_ens_create with the initial state drawn on the device by pt_inject
 * (first_index 0): init_particles and the state upload of run_gpu_ensemble
 * (phase_space_gpu_runner.jl:502-513) with no host O(n) work. */
int pt_ens_create_injected(pt_ctx *ctx, const pt_field *field, int pusher, int64_t n,
                           const pt_inject_params *params, int periodic, pt_ens **out);

typedef struct {
    double dt_s;          /* dt = trajectory_time.dt_s (phase_space_gpu_runner.jl:498) */
    double charge_C;      /* Q_E for the reference's protons (:9) */
    double mass_kg;       /* M_P (:10) */
    int64_t first_step;   /* 1-based si of the first step of this call */
    int64_t n_steps;      /* steps run by this call: si = first_step .. first_step+n_steps-1 */
    int64_t total_steps;  /* n_integration_steps; no push at si == total_steps (:568) */
    int64_t save_stride;  /* cadence: save at si with (si-1) % save_stride == 0 (:393-396) */
    int32_t out_kind;     /* PT_OUT_* */
    int32_t steps_per_launch; /* FAST pusher: steps per kernel launch (0 = library default) */
    int64_t sort_every_steps; /* split pushers: Morton re-sort before the first launch and
                                 every this many steps (0 = never) */
} pt_advance_params;

/* Replaces the step loop of run_gpu_ensemble (phase_space_gpu_runner.jl:562-591) and
 * run_mu_cache_ensemble (multimode_multienergy_cache_pipeline.jl:1467-1498): the
 * per-step advance_particles_kernel! launch, the per-step sum(dalive), the
 * reconstruct_mu_step_kernel! launch at save steps and the save-step D2H.
 *   out_a/out_b: PT_OUT_PHASE -> positions, momenta, each Julia (n,3,nsave_call)
 *                (C [save][3][n]); PT_OUT_MU -> mu in out_a, Julia (n,nsave_call)
 *                (C [save][n]); NaN for dead particles as the reference writes.
 *   out_dtype:   PT_F32 or PT_F64 (trajectory_output_precision /
 *                mu_cache_output_precision).
 *   out_capacity_saves: number of save slots the buffers hold.
 *   alive_counts (host, optional): n_steps values, sum(dalive) after each step.
 *   n_saved (optional): saves written by this call. */
int pt_ens_advance(pt_ens *ens, const pt_field *field, const pt_advance_params *params,
                   void *out_a, void *out_b, int out_dtype, int out_on_device,
                   int64_t out_capacity_saves, int64_t *alive_counts, int64_t *n_saved);

/* Download the current state (positions in metres, momenta in kg m/s) and
 * alive flags.  No reference counterpart (the reference never stores the
 * final state, exact_final_state_stored=false, phase_space_gpu_runner.jl:436); used
 * by parity tests and checkpointing. */
int pt_ens_get_state(pt_ens *ens, double *x[3], double *p[3], uint8_t *alive);

/* Reorder the particles of a split-precision ensemble (PLAIN_F32, FAST) by
 * the Morton key of their 8^3 brick so neighbouring lanes gather neighbouring
 * field records.  Outputs, pt_ens_get_state and alive counts keep the
 * original particle order.  No reference counterpart (the reference keeps
 * injection order). */
int pt_ens_sort(pt_ens *ens);

/* Kernel time (ms) of the last pt_ens_advance, measured with HIP events on
 * the context's stream around the launches, and the number of launches. */
int pt_ens_last_timing(const pt_ens *ens, double *kernel_ms, int64_t *launches);

/* ---- mu reconstruction -------------------------------------------------- */

/* Replaces reconstruct_mu_chunk_gpu / reconstruct_mu_kernel!
 * (compute_delta_mu2_curve.jl:151-225): mu for every (step, particle) of a
 * phase-space chunk, with B trilinear from a PLANAR field, in the field's
 * precision (compute_precision).  positions/momenta: Julia (chunk,3,nsave)
 * (C [s][3][chunk]) of src_dtype.  mu_out: Julia (nsave, chunk)
 * (time fastest) of out_dtype. */
int pt_mu_from_phase(pt_ctx *ctx, const pt_field *field, int64_t chunk, int64_t nsave,
                     int src_dtype, int src_on_device, const void *positions,
                     const void *momenta, int out_dtype, int out_on_device, void *mu_out);
/* Same reconstruction with a choice of output layout: sample_major = 0 is
 * pt_mu_from_phase's (the reference's mu_chunk, each particle's series
 * contiguous); sample_major = 1 writes C [nsave][chunk], the layout in which
 * the pair walk (pt_dmumu_accumulate with stride_s = chunk, stride_p = 1) reads the series
 * coalesced.  No reference counterpart: run_combined_full's device path uses
 * it between the two steps. */
int pt_mu_from_phase_layout(pt_ctx *ctx, const pt_field *field, int64_t chunk, int64_t nsave,
                            int src_dtype, int src_on_device, const void *positions,
                            const void *momenta, int out_dtype, int out_on_device, void *mu_out,
                            int sample_major);

/* ---- lag statistics ----------------------------------------------------- */

/* Replaces process_mu_chunk_dmumu_cpu! (compute_delta_mu2_dmumu_full.jl:656-704) and
 * its two bodies, process_mu_chunk_full_pairs_cpu! (:457-560, start_mode 0 =
 * sliding) and process_mu_chunk_injection_pairs_cpu! (:562-654, start_mode 1 =
 * injection).  mu is float32; element (s, j) of the chunk (save s, selected
 * particle j) is at mu[s*stride_s + col_j*stride_p], where col_j =
 * particle_index[j] if particle_index is non-null, else j.  The reference's
 * chunk is Julia (nsave, chunk): stride_s = 1, stride_p = nsave.  A device
 * mu cache written by pt_ens_advance is [save][n]: stride_s = n, stride_p = 1.
 * Accumulators are host arrays in the reference's shapes and are updated in
 * place: counts/bin sums added, the per-lag (count, mean, M2) merged with
 * Chan's formula (combine_welford_stats!, :439-455).
 *   per lag (nlags): particle_counts, particle_means, particle_m2s,
 *                    pair_sum_squares, pair_counts
 *   Julia (n_bins, n_lags): dmumu_counts, dmumu_sum_delta, dmumu_sum_delta2
 * mu_edges: n_bins+1 FP64 edges (build_mu_edges_full, :289-302). */
int pt_dmumu_accumulate(pt_ctx *ctx, const float *mu, int mu_on_device, int64_t nsave,
                        int64_t n_sel, int64_t stride_s, int64_t stride_p,
                        const int64_t *particle_index, const int64_t *lag_steps, int64_t nlags,
                        const double *mu_edges, int64_t n_bins, int start_mode, int bin_abs,
                        int64_t *particle_counts, double *particle_means, double *particle_m2s,
                        double *pair_sum_squares, int64_t *pair_counts, int64_t *dmumu_counts,
                        double *dmumu_sum_delta, double *dmumu_sum_delta2);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the pass of pt_dmumu_accumulate over a mu cache already
 * in device memory (mu must be a device pointer), with every device work area
 * in a fixed 32 MiB array the library holds on each GPU it runs on (looked up
 * at pt_ctx_create) and its host staging in a fixed static area.  It calls no
 * allocator.  Same arguments, accumulators and results as
 * pt_dmumu_accumulate (bitwise).  Limits: at most 4096 lags and 65536
 * (lag, bin) cells, and about 40 nlags (n_bins + ceil(n_sel / 256) + 1) + 4 n_sel
 * (+ 8 n_sel with particle_index) bytes within the 32 MiB; beyond them the
 * call fails with PT_ERR_MEMORY before any kernel runs.  Calls that overlap on
 * one GPU share the array and race; calls from several host threads are
 * serialised on the host staging area. */
int pt_dmumu_accumulate_noalloc(pt_ctx *ctx, const float *mu, int64_t nsave, int64_t n_sel, int64_t stride_s,
                                int64_t stride_p, const int64_t *particle_index, const int64_t *lag_steps,
                                int64_t nlags, const double *mu_edges, int64_t n_bins, int start_mode, int bin_abs,
                                int64_t *particle_counts, double *particle_means, double *particle_m2s,
                                double *pair_sum_squares, int64_t *pair_counts, int64_t *dmumu_counts,
                                double *dmumu_sum_delta, double *dmumu_sum_delta2);

/* Replaces process_momenta_chunk_dpp_cpu! (compute_dpp_full.jl:247-288).
 * momenta: float32, Julia (total,3,nsave) (C [s][3][total]) with the chunk's
 * selected columns given by particle_index (or 0..n_sel-1).  |p| in FP64
 * (momentum_magnitude, :232-234); sums in FP64; updated in place.  The
 * p0-normalised sums are (sum Delta p)/p0 and (sum Delta p^2)/p0^2, equal to
 * the reference's per-pair sums of Delta p/p0 up to the rounding of the terms.
 * Deterministic: the same inputs give the same bits on every call. */
int pt_dpp_accumulate(pt_ctx *ctx, const float *momenta, int on_device, int64_t nsave,
                      int64_t total, int64_t n_sel, const int64_t *particle_index,
                      const int64_t *lag_steps, int64_t nlags, double p0, int64_t *counts,
                      double *sum_delta_p, double *sum_delta_p2, double *sum_delta_p_norm,
                      double *sum_delta_p_norm2);

/* process_momenta_chunk_dpp_cpu! (compute_dpp_full.jl:247-288) over a cache of
 * |p| (kg m/s, FP64) laid out [save][total], as a serving request with
 * PT_OUT_PMAG writes it; same accumulators and selection as pt_dpp_accumulate. */
int pt_dpp_accumulate_pmag(pt_ctx *ctx, const double *pmag, int on_device, int64_t nsave,
                           int64_t total, int64_t n_sel, const int64_t *particle_index,
                           const int64_t *lag_steps, int64_t nlags, double p0, int64_t *counts,
                           double *sum_delta_p, double *sum_delta_p2, double *sum_delta_p_norm,
                           double *sum_delta_p_norm2);

/* The same pairs over a relative cache delta = |p|/p_ref - 1 (float32,
 * [save][total]), as a serving request with PT_OUT_DPREL writes it: Delta p =
 * p_ref (delta(t+tau) - delta(t)), the difference taken in FP64.  Half the
 * bytes of the FP64 |p| cache; the rounding of delta to float32 is |delta|
 * 6e-8 relative, |delta| times smaller than the rounding of a Float32
 * phase-space cache that the reference's D_pp reads (its |p| carries 6e-8
 * relative).  Same accumulators (kg m/s and p0-normalised) and selection. */
int pt_dpp_accumulate_dprel(pt_ctx *ctx, const float *dprel, int on_device, int64_t nsave, int64_t total,
                            int64_t n_sel, const int64_t *particle_index, const int64_t *lag_steps, int64_t nlags,
                            double p_ref, double p0, int64_t *counts, double *sum_delta_p, double *sum_delta_p2,
                            double *sum_delta_p_norm, double *sum_delta_p_norm2);

/* Replaces fill_energy_snapshots! (compute_dpp_full.jl:290-299): kinetic energy
 * in GeV of each selected particle at each snapshot save index (0-based),
 * written into out, Julia (k, out_cols) column-major, at column col0 + j. */
int pt_energy_snapshots(pt_ctx *ctx, const float *momenta, int on_device, int64_t nsave,
                        int64_t total, int64_t n_sel, const int64_t *particle_index,
                        const int64_t *snapshot_idx, int64_t k, double *out, int64_t col0);

/* ---- serving: continuous batching of transport campaigns ---------------- */

/* No counterpart in the reference, which runs one (campaign, energy) job at a
 * time from run_mode_campaigns (multimode_multienergy_cache_pipeline.jl:2924)
 * through run_energy_pipeline! (:1857).  A server keeps fields resident,
 * admits requests against an HBM budget (FIFO), and advances every admitted
 * request in one fused multi-campaign push launch per tick, each with its own
 * field, dt, charge, mass, duration and cadence.  Requests that arrive while
 * others run are admitted at the next tick that has memory for them.  All
 * fields of one server share one BRICK storage kind. */

typedef struct pt_serve pt_serve;

enum pt_serve_state { PT_REQ_QUEUED = 0, PT_REQ_RUNNING = 1, PT_REQ_DONE = 2, PT_REQ_RELEASED = 3 };

/* Extra cadence outputs for serving, for D_pp: |p| (kg m/s, FP64), or the
 * relative momentum |p|/p_ref - 1 (float32, p_ref = the request's out_p_ref). */
#define PT_OUT_PMAG 3
#define PT_OUT_DPREL 4

typedef struct {
    int32_t field_id;        /* from pt_serve_add_field */
    int32_t out_kind;        /* PT_OUT_MU (float32), PT_OUT_PMAG (float64) or PT_OUT_DPREL (float32) cache */
    int64_t n;               /* particles */
    double dt_s, charge_C, mass_kg;
    int64_t total_steps;     /* n_integration_steps: steps si = 1..total_steps */
    int64_t save_stride;     /* cadence (trajectory_time_stride) */
    int32_t periodic;
    int32_t inject_fixed_mu; /* 1: momenta from mu0 about local B (fixed_mu_unit_vector,
                                phase_space_gpu_runner.jl:170-194) on the device */
    double mu0;              /* injection_mu0 for inject_fixed_mu */
    double p_magnitude;      /* |p| (kg m/s) for inject_fixed_mu */
    const double *x[3];      /* host positions (m); unused when inject_on_device */
    const double *p[3];      /* host momenta (kg m/s); unused when inject_fixed_mu or inject_on_device */
    const double *phase;     /* host gyrophases 2*pi*rand for inject_fixed_mu */
    int32_t inject_on_device; /* 1: the initial state is drawn on the device at admission, as pt_inject
                                 draws particles inject_first_index .. + n - 1 with `inject` (init_particles,
                                 phase_space_gpu_runner.jl:196-245); no host arrays, no host O(n) work.
                                 inject_fixed_mu, mu0, p_magnitude and the arrays above are then unused. */
    int32_t reserved0;
    int64_t inject_first_index;
    pt_inject_params inject;
    double out_p_ref;         /* p_ref (kg m/s) of PT_OUT_DPREL */
} pt_serve_request;

/* Reserves one device arena of the budget plus 12.5 % (for best-fit
 * fragmentation) for the server's lifetime; requests' state and caches are
 * placed in it at admission and returned at pt_serve_release. */
int pt_serve_create(pt_ctx *ctx, int64_t hbm_budget_bytes, int32_t steps_per_tick,
                    int32_t max_active, pt_serve **out);
int pt_serve_destroy(pt_serve *srv);
int pt_serve_add_field(pt_serve *srv, const pt_field *field, int32_t *field_id);
/* Copies the initial state; the request waits in the FIFO until admitted. */
int pt_serve_submit(pt_serve *srv, const pt_serve_request *req, int64_t *request_id);
/* Completes the previous tick (requests that finished become DONE), admits
 * queued requests that fit, and enqueues the next tick without waiting for
 * it.  Counts are after admission. */
int pt_serve_step(pt_serve *srv, int64_t *n_running, int64_t *n_queued, int64_t *n_done);
/* Waits for the enqueued tick, if any, and completes it. */
int pt_serve_drain(pt_serve *srv);
/* Ids of the requests that became DONE since the previous call, oldest first
 * (at most `capacity`; the rest stay for the next call), so an event loop
 * need not poll every outstanding request. */
int pt_serve_take_done(pt_serve *srv, int64_t *ids, int64_t capacity, int64_t *count);
int pt_serve_poll(pt_serve *srv, int64_t request_id, int32_t *state);
/* For a DONE request: the device cadence cache [save][n] (float32 mu or
 * float64 |p|), its shape and the number of particles alive at the end. */
int pt_serve_fetch(pt_serve *srv, int64_t request_id, const void **cache_device, int64_t *nsave,
                   int64_t *n, int64_t *alive_final);
/* Copies a DONE request's cadence cache to host memory (bytes must equal
 * nsave * n * element size). */
int pt_serve_copy_cache(pt_serve *srv, int64_t request_id, void *host_dst, int64_t bytes);
/* Returns a DONE request's memory to the arena (or drops a QUEUED one).
 * Call it after the last use of the request's cache has completed. */
int pt_serve_release(pt_serve *srv, int64_t request_id);
/* Budget bytes held by admitted requests, particle-steps pushed so far and
 * the summed kernel time of all ticks (HIP events on the server's stream). */
int pt_serve_stats(pt_serve *srv, int64_t *bytes_reserved, int64_t *particle_steps,
                   double *kernel_ms);

#ifdef __cplusplus
}
#endif

#endif /* PT_HIP_H */
