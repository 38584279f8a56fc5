/*
 * teuk_hip.h -- C ABI of the MI355X-native Teukolsky stage operator and
 * RK4 stepper (libteuk_hip.so).
 *
 * This is the drop-in boundary for the reference's spatial-operator / RK4
 * path.  The reference (HengruiZhu99/teuk-kokkos) exposes that path as the
 * C++ class `teuk::SpatialPipeline` (include/teuk/spatial_pipeline.hpp:
 * 144-620) driven by `teuk::run_solver` (include/teuk/solver_driver.hpp:
 * 86-353) from a schema-v1 configuration (include/teuk/config.hpp:
 * 194-713).  Each entry point below names the reference interface it
 * replaces.
 *
 * Conventions
 *  - Every call returns TEUK_OK (0) or an error code; TEUK_EINVAL mirrors the
 *    reference's std::invalid_argument (contract / schema violations) and
 *    TEUK_ERUNTIME its std::runtime_error (numerical, IO or device failures).
 *    The message of the last failure is teuk_last_error(plan), or
 *    teuk_config_last_error() for calls without a plan.  Failures happen
 *    before any state mutation, as in the reference.
 *  - State buffers use the reference interchange layout: interleaved
 *    complex128 (re, im) in (mode, field = 13, radial, theta) order with
 *    theta fastest (include/teuk/pipeline_storage.hpp:18-33,
 *    include/teuk/io.hpp:96-99).
 *  - A plan owns all its device memory, streams and events.  It allocates
 *    device memory and creates streams and events only in teuk_plan_create*,
 *    the teuk_plan_set_* tuning calls and teuk_initialize (its one-time
 *    initial-data tables); every other call allocates nothing (kernel timing
 *    aside, whose events are made while it is on).  teuk_allocation_count
 *    shows it, as the reference's tests check its own allocation counts
 *    (tests/test_spatial_pipeline.cpp:102-125).
 *  - Work is enqueued on the plan's stream, or on the caller's stream after
 *    teuk_plan_set_stream (the reference's execution-space argument).  Calls
 *    that take device pointers return without synchronising; teuk_sync
 *    waits.  A plan is not thread-safe.
 *  - A batch plan holds several independent instances (each its own
 *    configuration) that advance together, one launch per RK stage.
 */
#ifndef TEUK_HIP_H
#define TEUK_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  TEUK_OK = 0,
  TEUK_EINVAL = 1,   /* std::invalid_argument in the reference */
  TEUK_ERUNTIME = 2  /* std::runtime_error in the reference */
};

typedef struct teuk_plan teuk_plan;

/* Endpoint waveform record (include/teuk/waveform_output.hpp:23-33). */
typedef struct teuk_waveform_record {
  uint64_t step;
  double time;
  int boundary; /* 0 = scri (R = 0), 1 = horizon (R = R_H) */
  int order;    /* 1 = FirstPsi, 2 = SecondPsi */
  int m;
  int ell;
  double rescaled_re, rescaled_im; /* modal F coefficient */
  double endpoint_re, endpoint_im; /* L^2 F at scri, R_H F at the horizon */
} teuk_waveform_record;

/* Latched second-order source activation state
 * (include/teuk/source_activation.hpp, spatial_pipeline.hpp:386-424). */
typedef struct teuk_source_activation {
  int active;
  int consecutive_passes;
  double activation_time;
  double last_eligibility_time;
} teuk_source_activation;

/* ---------------------------------------------------------------------
 * Configuration (no GPU needed)
 * ------------------------------------------------------------------- */

/* parse_run_configuration_text + resolved_configuration_text
 * (config.hpp:686-831).  Writes the resolved text (NUL-terminated) into
 * out[out_size]; *needed receives the required size including the NUL. */
int teuk_config_resolve(const char* cfg_text, const char* const* overrides, int n_overrides,
                        char* out, size_t out_size, size_t* needed);
const char* teuk_config_last_error(void);

/* ---------------------------------------------------------------------
 * Plans: SpatialPipeline construction (spatial_pipeline.hpp:146-363) plus
 * the run_solver set-up checks (solver_driver.hpp:86-136: plus2 gate, CFL).
 * ------------------------------------------------------------------- */
int teuk_plan_create(const char* cfg_text, const char* const* overrides, int n_overrides,
                     teuk_plan** out);
/* n_instances configuration texts; the overrides apply to every instance. */
int teuk_plan_create_batch(const char* const* cfg_texts, int n_instances,
                           const char* const* overrides, int n_overrides, teuk_plan** out);
void teuk_plan_destroy(teuk_plan* plan);
const char* teuk_last_error(const teuk_plan* plan);

int teuk_plan_instance_count(const teuk_plan* plan);
/* SpatialPipelineStorage extents (pipeline_storage.hpp:104-109): stored
 * modes, fields (13), radial points, theta nodes. */
int teuk_plan_shape(const teuk_plan* plan, int instance, int* modes, int* fields, int* nr,
                    int* ntheta);
/* Configured time step final_time / steps (solver_driver.hpp:117-118). */
double teuk_plan_time_step(const teuk_plan* plan, int instance);
/* Simulation time and completed steps of an instance. */
int teuk_plan_progress(const teuk_plan* plan, int instance, double* time, uint64_t* steps);
/* sum over instances of modes * nr * ntheta: the grid points one step
 * updates (solver_driver.hpp:323-341); and the number of retained modal
 * coefficients (first-order modes * nr * n_ell) the kernels evolve. */
double teuk_plan_grid_points(const teuk_plan* plan);
double teuk_plan_modal_points(const teuk_plan* plan);

/* ---------------------------------------------------------------------
 * State
 * ------------------------------------------------------------------- */
/* initialize_pipeline_state for every instance (initial_data_factory.hpp:
 * 82-110): Gaussian data (pipeline_initial_data.hpp:211-340), or the
 * checkpoint named by initial_data.checkpoint_directory. */
int teuk_initialize(teuk_plan* plan);
/* write_pipeline_checkpoint (pipeline_checkpoint.hpp:782-857): a new
 * directory with metadata.txt and state.bin (format version 5) holding the
 * instance's state, latch and progress; published with one rename, an
 * existing destination is refused. */
int teuk_checkpoint_write(teuk_plan* plan, int instance, const char* directory);
/* load_pipeline_checkpoint (pipeline_checkpoint.hpp:869-943): every metadata,
 * registry, configuration, geometry, size, checksum and band check passes
 * before the instance's state, latch and progress change. */
int teuk_checkpoint_load(teuk_plan* plan, int instance, const char* directory);
/* import_pipeline_state (pipeline_bands.hpp:211-232): host buffer of
 * 2 * modes * 13 * nr * ntheta doubles, rejected with TEUK_ERUNTIME when
 * its off-band relative content exceeds relative_tolerance (reference
 * default 5e-11).  *offband receives the measured value (may be NULL). */
int teuk_state_import(teuk_plan* plan, int instance, const double* host,
                      double relative_tolerance, double* offband);
/* Copy of storage().state() in the interchange layout. */
int teuk_state_export(teuk_plan* plan, int instance, double* host);

/* ---------------------------------------------------------------------
 * Operator and stepper
 * ------------------------------------------------------------------- */
/* SpatialPipeline::evaluate_rhs_at_time (spatial_pipeline.hpp:458-515,
 * Disabled source policy): d_stage and d_out are device buffers in the
 * interchange layout; fields 3..12 and non-first-order modes of the output
 * are written as zero, fields 0..2 are band-projected. */
int teuk_evaluate_rhs(teuk_plan* plan, int instance, const void* d_stage, void* d_out,
                      double time);
/* No counterpart in the reference.  teuk_evaluate_rhs for callers that
 * cannot allocate device memory around the call (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * kernels, with the modal staging in a fixed per-device array of
 * 2 x 2^20 complex doubles (32 MiB, held until the process ends) instead of
 * the plan's buffers.  Linear plans only.  The staging is indexed like the
 * plan's own buffers, so an instance is refused with TEUK_EINVAL when its
 * modal lines end past 2^20 values in the plan's modal order (in a batch
 * plan the lines of the instances before it count).  Calls that overlap on
 * one device share the array and race. */
int teuk_evaluate_rhs_noalloc(teuk_plan* plan, int instance, const void* d_stage, void* d_out,
                              double time);
/* SpatialPipeline::step (spatial_pipeline.hpp:590-620): one classical RK4
 * step of every instance with the given time and step. */
int teuk_step(teuk_plan* plan, double time, double dt);
/* The driver's step loop (solver_driver.hpp:289-317): n_steps steps of every
 * instance with its configured step, captured as a hipGraph when enabled. */
int teuk_advance(teuk_plan* plan, int n_steps);
int teuk_sync(teuk_plan* plan);

/* EndpointWaveformSampler::sample (waveform_output.hpp:79-160): records in
 * the reference order (first order scri, horizon; second order scri,
 * horizon; modes ascending, ell ascending). */
int teuk_sample_waveform(teuk_plan* plan, int instance, teuk_waveform_record* out,
                         int capacity, int* count);

int teuk_get_source_activation(const teuk_plan* plan, int instance,
                               teuk_source_activation* state);
/* restore_source_activation (spatial_pipeline.hpp:395-400). */
int teuk_set_source_activation(teuk_plan* plan, int instance,
                               const teuk_source_activation* state);
/* source_constraint_norms (spatial_pipeline.hpp:378-380): the norms of the
 * last constraint-aware activation check of a second-order instance,
 * SourceConstraintNorms (source_activation.hpp:24-37) as 14 doubles: for
 * psi3 Bianchi, psi2 Bianchi and h_ll reality in turn the absolute maximum,
 * normalized maximum, weighted RMS and normalized weighted norm, then the
 * controlling normalized value and 1.0 / 0.0 for all_finite. */
int teuk_get_source_constraint_norms(const teuk_plan* plan, int instance, double* norms14);

/* The sampled diagnostics of the reference driver (solver_driver.hpp:205-242)
 * for one instance: evaluate_rhs_at_time of the current state at `time`,
 * then PipelineDiagnostics::sample_pipeline (pipeline_diagnostics.hpp:
 * 145-305), PipelineTransportEquationDiagnostics::sample
 * (pipeline_reconstruction_diagnostics.hpp:43-190),
 * PipelineIndependentReconstructionDiagnostics::sample
 * (pipeline_independent_reconstruction_diagnostics.hpp:33-98) and the horizon
 * maxima of HorizonTransverseDiagnostics (horizon_diagnostics.hpp:113-160).
 * Writes TEUK_DIAG_COUNT doubles in the layout below.  Fails with
 * TEUK_ERUNTIME and the reference's message when a sampled value is not
 * finite.  Call between steps: it overwrites the plan's RHS scratch. */
enum {
  TEUK_DIAG_STATE_RMS = 0,            /* 13 fields */
  TEUK_DIAG_STATE_MAX = 13,
  TEUK_DIAG_RHS_RMS = 26,
  TEUK_DIAG_RHS_MAX = 39,
  TEUK_DIAG_FIRST_CONSTRAINT = 52,    /* rms, max of Q - D Psi */
  TEUK_DIAG_SECOND_CONSTRAINT = 54,
  TEUK_DIAG_SOURCE = 56,              /* rms, max of source_over_r3 */
  TEUK_DIAG_FORCING = 58,             /* rms, max of the latched forcing */
  TEUK_DIAG_INDEPENDENT = 60,         /* rms, max: psi3, psi2, h_ll, combined */
  TEUK_DIAG_TRANSPORT = 68,           /* rms, max: G, Lambda, H, B, Pi, C, U, combined */
  TEUK_DIAG_HORIZON_FIRST = 84,       /* max |d^k FirstPsi/dR^k| at R_H, k = 0..4 */
  TEUK_DIAG_HORIZON_SECOND = 89,      /* the same for SecondPsi */
  TEUK_DIAG_SOURCE_CONSTRAINT_MAX = 94,
  TEUK_DIAG_SOURCE_ACTIVE = 95,
  TEUK_DIAG_ACTIVATION_TIME = 96,
  TEUK_DIAG_SCRI_FINITE = 97,
  TEUK_DIAG_HORIZON_FINITE = 98,
  TEUK_DIAG_ALL_FINITE = 99,
  TEUK_DIAG_COUNT = 100
};
int teuk_sample_diagnostics(teuk_plan* plan, int instance, double time, double* out);
/* The per-pair source report of the last teuk_sample_diagnostics call on a
 * second-order instance (solver_driver.hpp:251-280): for every ordered pair
 * in registry order, 7 doubles: m1, m2, target, RMS and maximum of |D|, RMS
 * and maximum of |T| over (r, theta).  *count receives the pair count;
 * out may be null to query it. */
int teuk_source_pairs(const teuk_plan* plan, int instance, double* out, int capacity,
                      int* count);

/* ---------------------------------------------------------------------
 * Execution tuning and measurement (no reference counterpart: they replace
 * the reference's SpatialPipelineTiming, spatial_pipeline.hpp:27-38)
 * ------------------------------------------------------------------- */
/* Graph mode (default off): one RK4 step of every instance is captured as a
 * hipGraph now, and teuk_advance replays it; linear plans also capture eight
 * steps as one graph (each execution group's stage chain joined once per
 * eight steps) and replay that while eight or more steps remain, with the
 * same results bit for bit.  Second-order plans replay the one-step graph
 * for every step in which no instance splits at its source start and no
 * latch decision needs a constraint evaluation (the Unrestricted and
 * Disabled policies decide on the host); other steps launch directly, with
 * the same results bit for bit.  teuk_plan_graph_replays counts the steps
 * taken as replays. */
int teuk_plan_set_graph(teuk_plan* plan, int enable);
int teuk_plan_graph_replays(const teuk_plan* plan, long long* replays);
/* The execution-space argument of SpatialPipeline::evaluate_rhs_at_time and
 * step (spatial_pipeline.hpp:458-462, 590): every later call enqueues on
 * `stream` (a hipStream_t the caller owns and keeps alive while the plan
 * uses it); NULL returns to the plan's own stream.  Waits for the work
 * already enqueued on the previous stream. */
int teuk_plan_set_stream(teuk_plan* plan, void* stream);
/* Device allocations, streams and events made by the library so far (all
 * plans and teuk_device_alloc). */
int teuk_allocation_count(unsigned long long* count);
/* LDS budget per workgroup (bytes) and fixed tile rows (0 = automatic);
 * both re-tile the plan. */
int teuk_plan_set_lds_budget(teuk_plan* plan, int bytes);
int teuk_plan_set_tile_rows(teuk_plan* plan, int rows);
/* Stage kernels specialised per line variant (default 1); 0 runs every line
 * through the general kernel, which gives the same results bit for bit (with
 * teuk_plan_set_mixed_gram on, the general kernel keeps the FP32 Gram
 * correction: the two then agree to 1e-13).  Re-tiles the plan. */
int teuk_plan_set_specialized(teuk_plan* plan, int enable);
/* Mixed precision (default 0, no reference counterpart): 1 takes the Gram
 * correction E X + N Psi of bands whose E = A S - I and N are below 2e-11
 * (ell 24) on BF16 MFMA in the specialised Kerr-Gram kernel, where 0 keeps
 * it on FP32 MFMA, as accurate as the FP64 product.  The correction is
 * 6e-12 of the RHS there; BF16 holds it to 4e-3 of itself.  Changes no
 * tiling and allocates nothing; takes effect from the next launch or replay. */
int teuk_plan_set_mixed_gram(teuk_plan* plan, int enable);
int teuk_plan_tiles(const teuk_plan* plan, int* tiles, int* max_lds_bytes);
/* HIP-event timing of every stage-kernel launch on the plan's stream
 * (eager mode only); reset by enabling. */
int teuk_plan_set_kernel_timing(teuk_plan* plan, int enable);
int teuk_plan_kernel_timing(const teuk_plan* plan, double* total_ms, long long* launches);
/* GPU time of whole RK4 steps (all execution groups, which run on streams of
 * their own between a fork from and a join into the plan's stream). */
int teuk_plan_step_timing(const teuk_plan* plan, double* total_ms, long long* steps);
/* Execution groups: the tiles of one radial scheme, ell-tile class
 * (ell_tiles = 1, 2 or 4 tiles of 16 held by the MFMA routines) and line
 * variant (Kerr lines without dissipation in free-damped mode, with and
 * without the Gram projection, run kernels specialised for them), one
 * stage-kernel launch per group and RK stage.  kernel_ms: the group's timed
 * launches since teuk_plan_set_kernel_timing(plan, 1). */
int teuk_plan_group_count(const teuk_plan* plan, int* count);
int teuk_plan_group_info(const teuk_plan* plan, int group, int* scheme, int* ell_tiles,
                         int* tiles, int* lds_bytes, double* kernel_ms);
/* The group's line variant: 0 general, 1 Kerr with G = I, 2 Kerr with the
 * Gram projection (both without dissipation, free-damped). */
int teuk_plan_group_variant(const teuk_plan* plan, int group, int* variant);
/* Tile shape of the execution group keyed (scheme, ell_tiles, variant), now
 * and after later re-tiling: LDS budget per workgroup and tile rows (0 = the
 * plan's setting).  Re-tiles the plan; results do not depend on the shape.
 * teuk_plan_group_shape reports the shape set for group `group`. */
int teuk_plan_set_group_shape(teuk_plan* plan, int scheme, int ell_tiles, int variant,
                              int lds_budget, int tile_rows);
int teuk_plan_group_shape(const teuk_plan* plan, int group, int* lds_budget, int* tile_rows);

/* ---------------------------------------------------------------------
 * Device buffers for callers without a HIP runtime of their own (no
 * reference counterpart: the reference's Kokkos::View allocations).
 * kind: 0 host->device, 1 device->host, 2 device->device; synchronous.
 * ------------------------------------------------------------------- */
int teuk_device_alloc(size_t bytes, void** ptr);
int teuk_device_free(void* ptr);
int teuk_memcpy(void* dst, const void* src, size_t bytes, int kind);

/* ---------------------------------------------------------------------
 * Host-side tables (no GPU needed; used by the CPU tests)
 * ------------------------------------------------------------------- */
/* scheme: 0 d4-2, 1 d8-4, 2 d10-5.  h*D(row, col) and H(i)/h of an n-point
 * grid (sbp.hpp, sbp84.hpp, sbp105.hpp). */
double teuk_sbp_derivative_entry(int scheme, int n, int row, int col);
double teuk_sbp_norm_weight(int scheme, int n, int index);
/* S (ntheta x n_ell) and A (n_ell x ntheta) of one fixed-(s, m) band
 * (angular.hpp:329-355); returns n_ell or -1. */
int teuk_band_matrices(int spin, int m, int ell_max, int ntheta, double* S, double* A);
/* ModeRegistry (modes.hpp:32-213) of a configuration: sorted stored modes,
 * sharp indices, and ordered pairs (m1, m2, target) target-major. */
int teuk_config_modes(const char* cfg_text, int* modes, int* sharp, int capacity, int* count,
                      int* pairs, int pair_capacity, int* pair_count);

#ifdef __cplusplus
}
#endif

#endif /* TEUK_HIP_H */
