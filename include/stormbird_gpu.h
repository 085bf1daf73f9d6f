/*
 * stormbird_gpu.h — C ABI of the MI355X (gfx950) dynamic lifting-line library.
 *
 * Drop-in for Stormbird's `Simulation` (stormbird 0.9.0, Rust) on its dynamic
 * and quasi-steady lifting-line paths: the same JSON setup string, the same step semantics and the same
 * result fields, with the wake state resident in GPU memory.  The convention
 * follows the reference's own C interface (interfaces/cstormbird): opaque
 * handle built from a NUL-terminated JSON string (NULL on error), explicit
 * destroy, caller-owned output arrays, negative integer error codes, no
 * panics across the ABI.  Vectors cross the ABI as double[3*n] (x,y,z per
 * point), the layout of Rust's `[f64; 3]` / `SpatialVector`.
 *
 * Error codes: 0 ok; -1 null argument; -2 zero capacity; -3 size mismatch;
 * -4 unsupported or invalid request; -5 GPU runtime error.
 * Non-convergence of the fixed-point solve is not an error (as in the
 * reference): it shows as iterations == max_iterations_per_time_step.
 *
 * Threading: one handle is used by one thread at a time; distinct handles may
 * run concurrently (each owns its own HIP stream).
 */
#ifndef STORMBIRD_GPU_H
#define STORMBIRD_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sb_sim sb_sim;

/* Per-step outputs. Any array pointer may be NULL (that output is skipped).
 * Mirrors SimulationResult (stormbird/src/common_utils/results/simulation.rs:27-40):
 * force_input.{circulation_strength, angles_of_attack, velocity},
 * solver_input_ctrl_points_velocity, ctrl_points, sectional_forces.*,
 * integrated_forces[*].*, integrated_moments[*].*, iterations, residual. */
typedef struct sb_step_result {
    double time;
    long iterations;
    double residual;
    double* circulation_strength;  /* [N]   */
    double* angles_of_attack;      /* [N]   */
    double* ctrl_points_velocity;  /* [3N]  force_input.velocity            */
    double* solver_input_velocity; /* [3N]  solver_input_ctrl_points_velocity */
    double* ctrl_points;           /* [3N]  */
    double* sectional_forces;      /* [6][3N] circulatory, viscous_lift, sectional_drag,
                                      added_mass, gyroscopic, total */
    double* integrated_forces;     /* [6][3W] same order */
    double* integrated_moments;    /* [6][3W] same order */
} sb_step_result;

/* Arithmetic of the Biot-Savart sums.
 * SB_MODE_EXACT: the reference's operation order (no fused multiply-add,
 *   IEEE division and square root, ascending panel order per target); results
 *   equal the f64 CPU restatement bit for bit.
 * SB_MODE_FAST: fused multiply-add and reciprocal square root in the pair
 *   kernels.  Measured for one step of four sails x 40 strips from a rolled-up
 *   wake of 2,000 rows against EXACT from the same state: circulation 1.9e-14
 *   and integrated forces 1.8e-15 relative; wake-point velocities up to 3.5e-10
 *   relative to the largest (targets close to a ring edge, where the viscous-core
 *   distance cancels).  A free wake amplifies such differences over many steps.
 * SB_MODE_FP32_KAHAN: pair terms in fp32 (relative vectors formed in f64, the
 *   near/far switch decided in f64) with compensated (Kahan) fp32 per-target
 *   accumulation of the free-wake self-induction; the fixed velocities fold the
 *   fp32 terms in f64; unit factors, solver, geometry and wake convection stay
 *   f64.  Not bit-exact: its deviation from the f64 path is measured and
 *   reported.  Single simulations only (sb_batch_create takes EXACT or FAST). */
enum { SB_MODE_EXACT = 0, SB_MODE_FAST = 1, SB_MODE_FP32_KAHAN = 2 };

/* Message of the last error on this thread. */
const char* sb_last_error(void);

/* Replaces Simulation::new_from_string (stormbird/src/lifting_line/simulation.rs:33-39).
 * Returns NULL on a parse error or an unsupported setup (see sb_last_error).
 * Both SimulationSettings variants are taken: Dynamic (shed free wake) and
 * QuasiSteady (one horseshoe vortex per strip, no wake points: the freestream
 * is given at the control points only), which is also the default when
 * simulation_settings is absent (simulation_builder.rs:51-64).  Solvers:
 * SimpleIterative (with start_with_linearized_solution) and Linearized
 * (linearized.rs:54-222; a singular equation system fails the step with -4
 * where the reference's unwrap panics). */
sb_sim* sb_sim_create(const char* setup_json);
/* Parses and checks a setup string without touching a GPU (what
 * SimulationBuilder's deserialization accepts, simulation_builder.rs:42-66, plus
 * this library's limits).  out[0..7): strips, wings, span points, wake rows,
 * freestream points, wake panels, 1 for quasi-steady.  No counterpart in the
 * reference (there, building the simulation is the check). */
int sb_setup_info(const char* setup_json, long* out, size_t cap);
/* Same, on GPU `device` with a Biot-Savart arithmetic mode (SB_MODE_*). */
sb_sim* sb_sim_create_ex(const char* setup_json, int device, int mode);
/* Mirrors complete_sail_model_drop (interfaces/cstormbird/src/lifting_line/complete_sail_model.rs:59-66). */
void sb_sim_destroy(sb_sim* sim);

/* Sizes: LineForceModel::nr_span_lines / nr_wings (line_force_model/data_access.rs:13-20),
 * WakeIndices::nr_points / nr_panels (dynamic_wake/settings.rs:34-42). */
long sb_sim_nr_ctrl_points(const sb_sim* sim);
long sb_sim_nr_wings(const sb_sim* sim);
long sb_sim_nr_span_points(const sb_sim* sim);
long sb_sim_nr_wake_points(const sb_sim* sim);
long sb_sim_nr_panels(const sb_sim* sim);
long sb_sim_nr_freestream_points(const sb_sim* sim);

/* Replaces Simulation::get_freestream_velocity_points (simulation.rs:45-63):
 * ctrl points, then all wake points. Returns the count written (<= cap). */
long sb_sim_get_freestream_points(const sb_sim* sim, double* xyz, size_t cap);

/* Replaces Simulation::do_step (simulation.rs:107-203). `freestream_xyz` holds
 * the freestream at every point of sb_sim_get_freestream_points, host memory. */
int sb_sim_do_step(sb_sim* sim, double time, double time_step, const double* freestream_xyz,
                   size_t n_points, sb_step_result* out);
/* Same with `freestream_xyz` already in device memory (HBM-resident input),
 * read after the work queued so far on the null stream (where the caller wrote
 * it, e.g. torch's default stream). */
int sb_sim_do_step_device(sb_sim* sim, double time, double time_step,
                          const double* freestream_xyz_device, size_t n_points,
                          sb_step_result* out);

/* Replaces Simulation::induced_velocities (simulation.rs:207-216); -4 for a
 * quasi-steady simulation, where the reference panics. */
int sb_sim_induced_velocities(sb_sim* sim, const double* points_xyz, size_t n, double* out_xyz);

/* The warm-start hook: Simulation::previous_circulation_strength (simulation.rs:26,165,189). */
int sb_sim_get_previous_circulation(const sb_sim* sim, double* out, size_t n);
int sb_sim_set_previous_circulation(sb_sim* sim, const double* gamma, size_t n);

/* LineForceModel::set_translation_and_rotation_with_finite_difference_for_the_velocity
 * (line_force_model/data_update.rs:188-205). */
/* Wake export in the reference's formats (wake/dynamic_wake/export.rs:30-210):
 * DynamicWake::write_wake_to_obj_file, ::write_wake_to_vtk_file, and the panel
 * corner indices [nr_panels][4] export_to_plotly_mesh builds its triangles from
 * (panel_point_indices, dynamic_wake/mod.rs:82-91).  Numbers print as Rust's
 * `{}` prints an f64.  With `write_wake_data_to_file` set, every step writes
 * `{wake_files_folder_path}/wake_{(time / time_step) as usize}.vtp`
 * (simulation.rs:181-186). */
int sb_sim_write_wake_obj(const sb_sim* sim, const char* path);
int sb_sim_write_wake_vtk(const sb_sim* sim, const char* path);
int sb_sim_get_panel_point_indices(const sb_sim* sim, long* indices, size_t cap);
/* Rust `{}` formatting of one f64 into buf (returns its length); test hook. */
int sb_debug_format_f64(double value, char* buf, size_t cap);
/* Back to the state of a newly built simulation: straight initial wake, zero
 * strengths, no completed step (the line-force model's motion is kept).  No
 * counterpart in the reference (which builds a new Simulation); it lets one
 * handle serve many independent first-step solves. */
int sb_sim_reset(sb_sim* sim);
/* The solver's public settings (SimpleIterative: max_iterations_per_time_step,
 * damping_factor, residual_tolerance_absolute, strength_difference_tolerance;
 * solvers/simple_iterative.rs:59-76), used from the next step on. */
int sb_sim_set_solver_settings(sb_sim* sim, long max_iterations, double damping, double residual_tol,
                               double strength_tol);
int sb_sim_set_translation_and_rotation_fd(sb_sim* sim, double time_step,
                                           const double* translation3, const double* rotation3);
/* LineForceModel::set_local_wing_angles (data_update.rs:171-177). */
int sb_sim_set_local_wing_angles(sb_sim* sim, const double* angles, size_t n);
/* LineForceModel::local_wing_angles (line_force_model/mod.rs:92), n = number of wings. */
int sb_sim_get_local_wing_angles(const sb_sim* sim, double* out, size_t n);

/* LineForceModel::wing_indices (line_force_model/mod.rs:69) as [start, end) pairs
 * [2 * nr_wings]; -3 when cap < 2 * nr_wings. */
int sb_sim_get_wing_ranges(const sb_sim* sim, long* ranges, size_t cap);
/* LineForceModel::rigid_body_motion (line_force_model/mod.rs:88; stormath/src/rigid_body_motion.rs:22-29):
 * translation, rotation, velocity_linear, velocity_angular [12]. */
int sb_sim_get_rigid_body_motion(const sb_sim* sim, double* motion12);

/* DynamicWake public fields points / velocity_at_points / strengths /
 * number_of_time_steps_completed (dynamic_wake/mod.rs:41-64) and
 * Simulation::first_time_step_completed (simulation.rs:27). */
int sb_sim_get_wake_state(const sb_sim* sim, double* points, double* velocities,
                          double* strengths, long* steps_completed, int* first_step_done);
int sb_sim_set_wake_state(sb_sim* sim, const double* points, const double* velocities,
                          const double* strengths, long steps_completed, int first_step_done);
/* Same with the length of each array: n_points points [3 n_points] for points and
 * velocities, n_strengths panels for strengths (NULL arrays skip the field and
 * their length is not checked); -3 when a length differs from the simulation's
 * sb_sim_nr_wake_points / sb_sim_nr_panels.  The unsized form above trusts the
 * caller's arrays to be that long. */
int sb_sim_set_wake_state_ex(sb_sim* sim, const double* points, size_t n_points, const double* velocities,
                             size_t n_velocities, const double* strengths, size_t n_strengths,
                             long steps_completed, int first_step_done);
/* Panel records [nr_panels][22]: 4 corners, center, normal, far-field length^2,
 * viscous core length, area/(4 pi), strength (singularity_elements/panel.rs:16-23). */
int sb_sim_get_panels(const sb_sim* sim, double* records);
/* FrozenWake::fixed_velocities [3N] and variable_velocity_factors [N][N][3]
 * row-major (wake/frozen_wake.rs:31-47). */
int sb_sim_get_frozen_wake(const sb_sim* sim, double* fixed, double* factors);

/* Per-kernel interval timers (no counterpart in the reference): with on != 0,
 * every later step records HIP events between its kernels on the handle's
 * stream; off (the default) it records none, since the seven records cost about
 * 25 us of a 0.18 ms step.  Returns 0, or -1 for a null handle. */
int sb_sim_set_kernel_timers(sb_sim* sim, int on);
/* Per-kernel timing of the last step (milliseconds, HIP events on the
 * handle's stream). names: "wake_update", "fixed", "factors", "solve",
 * "forces", "self_induction".  Returns -2 for an unknown name and -3 when the
 * last step ran with the timers off (sb_sim_set_kernel_timers). */
double sb_sim_last_kernel_ms(const sb_sim* sim, const char* name);
/* Pair evaluations with nonzero panel strength in the last step:
 * [0] self-induction, [1] fixed velocities, [2] unit-strength factor pairs. */
int sb_sim_last_pair_counts(const sb_sim* sim, double* counts3);

/* Order in which the free-wake self-induction visits its targets (no counterpart
 * in the reference; results do not depend on it, since every target keeps its
 * ascending panel order): waves of `span_points_per_wave` consecutive span points
 * of 64 / span_points_per_wave consecutive wake rows (a divisor of 64), or 0 for
 * the flat point order (the default, and the fastest measured). */
int sb_sim_set_target_tiling(sb_sim* sim, int span_points_per_wave);
int sb_sim_get_target_tiling(const sb_sim* sim);

/* Near-field (vortex-ring branch) and total pair counts, with nonzero panel
 * strength, of the free-wake self-induction sum at the current wake state
 * (update_velocity_at_points, dynamic_wake/update_data.rs:125-141). */
int sb_sim_count_self_induction_pairs(sb_sim* sim, double* near_total2);

/* Standalone Biot-Savart sum (kernel-level entry for parity tests): induced
 * velocity at `n_targets` points from `n_panels` panels given as records
 * [n_panels][22] (layout of sb_sim_get_panels), symmetry 0 none, 1 X, 2 Y,
 * 3 Z. Host arrays; mirrors DynamicWake::induced_velocities
 * (dynamic_wake/induced_velocity_calc.rs:18-26, 111-126). */
int sb_wake_induced_velocities(const double* panel_records, size_t n_panels,
                               const double* targets_xyz, size_t n_targets, int symmetry,
                               int mode, double* out_xyz);

/* Allocation-free form of sb_wake_induced_velocities.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * records [n_panels][22], targets [3 n_targets] and output [3 n_targets] are in
 * device memory on the calling thread's current device; the work is queued on
 * `stream` (a hipStream_t; NULL: the null stream) and not waited for.  The
 * panel records are staged in a fixed 32 MiB device array that the library
 * holds on each device it runs on until the process ends; no memory is
 * allocated or freed.  Limit: n_panels <= sb_noalloc_max_panels() (-4 with a
 * message beyond it).  Every call on a device reuses that device's array, so
 * calls that overlap on one GPU (other streams or threads) race.  Results
 * equal sb_wake_induced_velocities's bit for bit.  mode: SB_MODE_EXACT or
 * SB_MODE_FAST. */
int sb_wake_induced_velocities_noalloc(const double* panel_records_device, size_t n_panels,
                                       const double* targets_xyz_device, size_t n_targets, int symmetry,
                                       int mode, double* out_xyz_device, void* stream);
/* The largest n_panels sb_wake_induced_velocities_noalloc takes. */
long sb_noalloc_max_panels(void);

/* Frozen-wake matrix-vector product (kernel-level entry for parity tests):
 * out[r] = fixed[r] + sum_c factors[r][c] * strengths[c], the x, y and z sums
 * each taken over the columns in ascending order, as
 * FrozenWake::update_induced_velocities_at_control_points
 * (wake/frozen_wake.rs:268-301).  Host arrays: fixed [3n], factors [n][n][3]
 * row-major (layout of sb_sim_get_frozen_wake), strengths [n], out [3n];
 * n <= 65536. */
int sb_frozen_matvec(const double* fixed, const double* factors, const double* strengths, size_t n,
                     double* out);

/* Device-side elementary functions, for the host/device bit-identity test:
 * fn 0 sin, 1 cos, 2 exp, 3 acos, 4 pow(x, y), 5 log, 6 atan. */
int sb_debug_libm_device(int fn, const double* x, const double* y, double* out, size_t n);
/* Same functions evaluated by the library's host code. */
int sb_debug_libm_host(int fn, const double* x, const double* y, double* out, size_t n);
/* 1 when the Linearized solve of n_strips keeps its augmented matrix in the
 * workgroup's shared memory on the current device, 0 when in global memory,
 * -1 for n_strips outside [1, 1024]. */
int sb_debug_linearized_in_shared_memory(long n_strips);
/* The one-workgroup Gaussian elimination with partial pivoting that the
 * Linearized solve runs (stormath matrix/linalg.rs:41-104), on a given system:
 * A row-major n x n, b [n], x [n]; variant 1 (256 threads, n <= 64), 2 (256,
 * n <= 127), 3 (256, n <= 128) or 4 (512, n <= 139).  Returns 0, k + 1 when the pivot of step k is
 * below 1e-12 in magnitude (where linalg.rs:67-71 returns its error; x is then
 * zero), -1 when n does not fit the variant, -2 on a device error. */
int sb_debug_gauss_device(const double* A, const double* b, size_t n, double* x, int variant);
/* The same elimination in the blocked form the Linearized solve takes for larger
 * systems (16-column panels, trailing updates, one-wave back substitution), the
 * matrix in global memory: n <= 1024.  Returns as sb_debug_gauss_device. */
int sb_debug_gauss_blocked_device(const double* A, const double* b, size_t n, double* x);

/* ---- Batched simulations ------------------------------------------------
 * Many independent simulations (each built from its own setup string, with
 * its own strips, wings and wake length) stepped together: one call advances
 * every instance by one time step with the kernels of sb_sim_do_step launched
 * once over all instances, the pair sums as persistent workgroups pulling
 * (instance, target block) jobs from a queue, and, from the second step on,
 * the step replayed from a captured HIP graph.  Each instance's results are
 * bit-identical to stepping it alone with sb_sim_do_step (same mode).
 * No counterpart in the reference, whose callers loop over Simulation values
 * (interfaces/pystormbird/src/lifting_line/simulation.rs:131-149 per instance);
 * it replaces such a loop. */
typedef struct sb_batch sb_batch;

/* n setup strings; NULL on error (sb_last_error). mode: SB_MODE_EXACT or SB_MODE_FAST. */
sb_batch* sb_batch_create(const char* const* setup_jsons, long n, int device, int mode);
void sb_batch_destroy(sb_batch* b);
long sb_batch_size(const sb_batch* b);
/* Borrowed handle of instance i (valid until sb_batch_destroy): the sb_sim_*
 * queries and state setters apply to it; step it only through the batch. */
sb_sim* sb_batch_instance(sb_batch* b, long i);
/* offsets[i] = first freestream point of instance i in the concatenated
 * freestream array (instance order, each instance's own point order);
 * offsets[n] = total.  Returns n + 1, or -1 / -2. */
long sb_batch_freestream_offsets(const sb_batch* b, long* offsets, size_t cap);
/* One time step of every instance.  times/dts: [n] (times may be NULL);
 * fs: the concatenated freestream, n_points points (host memory, or device
 * memory for the _device variant, read after the work queued so far on the
 * null stream); results: [n] or NULL. */
int sb_batch_do_step(sb_batch* b, const double* times, const double* dts, const double* fs, size_t n_points,
                     sb_step_result* results);
int sb_batch_do_step_device(sb_batch* b, const double* times, const double* dts, const double* fs,
                            size_t n_points, sb_step_result* results);
/* Wind steps (no counterpart in the reference for a batch; per instance they
 * are CompleteSailModel::freestream_velocity, complete_sail_model/mod.rs:215-255,
 * followed by the step, :174-191, without a controller): a wind environment
 * (environment.rs:21-34 serde form; NULL or "{}" for the defaults) and one wind
 * condition per instance (wind_condition/mod.rs:18-39 serde form, with its
 * discretized gust spectra), parsed once; replaced by the next call.  -3 for a
 * quasi-steady batch, -4 for a string that does not parse. */
int sb_batch_set_wind(sb_batch* b, const char* environment_json, const char* const* wind_conditions_json);
/* One time step of every instance in the wind set by sb_batch_set_wind, at
 * times[i] with ship_velocities[i] (along the environment's zero direction):
 * the unsteady apparent wind at the control points (with the inflow
 * corrections, on the host) and at every wake point where it is before the
 * step (on the device, straight into the batch's freestream). */
int sb_batch_do_step_wind(sb_batch* b, const double* times, const double* dts, const double* ship_velocities,
                          sb_step_result* results);
/* The freestream of the last batched step, n_points points (the concatenation
 * that sb_batch_freestream_offsets describes), into host memory. */
int sb_batch_get_freestream(const sb_batch* b, double* out, size_t n_points);
/* Device time of the last batched step (ms, HIP events around the step). */
double sb_batch_last_step_ms(const sb_batch* b);
/* sb_sim_reset on every instance. */
int sb_batch_reset(sb_batch* b);

/* Kernel variants of the batched pair sums (no counterpart in the reference):
 * fixed_per = panels per compute wave per block of the fixed-velocity kernel
 * (1, 2, 3), targets_tile = panel records per LDS tile of the self-induction
 * kernel (64, 128, 256).  Results do not depend on them (every target keeps its
 * ascending panel order). */
int sb_batch_set_tuning(sb_batch* b, int fixed_per, int targets_tile);
int sb_batch_get_tuning(const sb_batch* b, int* fixed_per, int* targets_tile);
/* Autotuner: times every variant on the batch's last-step state (a step must
 * have been taken; the state is left unchanged) and applies the fastest.
 * flags bit 0: use a choice stored in cache_path for the same device and shape
 * class (instance count, pair counts and largest strip count, in powers of
 * two) without timing; bit 1: store the result in cache_path (a JSON file).
 * Each variant runs once untimed, then `reps` timed launches (the fastest
 * counts); max_jobs > 0 times only the first max_jobs jobs of each kernel's
 * queue (the fixed-velocity queue is longest first).  timings [6] (may be
 * NULL): ms of the fixed-velocity kernel for 1, 2, 3 panels per wave, then of
 * the self-induction kernel for tiles of 64, 128, 256 (-1: no work).  Returns
 * 1 when the choice came from the cache, 0 when timed. */
int sb_batch_autotune(sb_batch* b, const char* cache_path, int flags, int reps, long max_jobs, double* timings,
                      size_t cap);

/* ---- Learned warm start ---------------------------------------------------
 * A per-strip network (64 -> 256 -> 256 -> 256 -> 256 -> 1, GELU; bf16 matrix
 * -core GEMMs with fp32 accumulation, fp32 master weights, Adam) predicts
 * circulation / (c |u| / 2) from strip features; attached to a batch it
 * supplies the solver's initial circulation, and the unchanged fixed-point
 * solve finishes from it to the same tolerances.  An instance whose
 * prediction is not finite or exceeds 10 in magnitude keeps the default guess
 * (previous circulation, or zero at the first step) and counts as a fallback.
 * No counterpart in the reference.  Parameters: fp32, for each layer W[out][in]
 * then b[out]. */
typedef struct sb_mlp sb_mlp;
sb_mlp* sb_mlp_create(int device, unsigned long long seed);
void sb_mlp_destroy(sb_mlp* m);
long sb_mlp_nr_params(const sb_mlp* m);
int sb_mlp_get_params(const sb_mlp* m, float* params, size_t cap);
int sb_mlp_set_params(sb_mlp* m, const float* params, size_t n);
/* Device pointers: x [n][64] fp32, y/t [n] fp32.  The calls read their device
 * inputs after the work queued so far on the null stream (where the caller
 * wrote them, e.g. torch's default stream), and return when their outputs are
 * written. */
int sb_mlp_forward(sb_mlp* m, const float* x, long n, float* y);
/* Mean squared error of the predictions against t and its gradient with
 * respect to every parameter (host array, parameter layout). */
int sb_mlp_gradients(sb_mlp* m, const float* x, const float* t, long n, float* grads, float* loss);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on the mean squared error. */
int sb_mlp_train_step(sb_mlp* m, const float* x, const float* t, long n, float learning_rate, float* loss);

/* mode 0: off; 1: at the first step; 2: at every step; 3: at every step,
 * judged: per instance the solver's first evaluation keeps whichever of the
 * default guess (previous circulation) and the prediction has the smaller
 * max |Gamma_est - Gamma|.  Plus 4 (modes 5, 6, 7): the network predicts the
 * change from the previous circulation (guess = previous + prediction x scale)
 * instead of the circulation.  The network must outlive its use by the batch. */
int sb_batch_set_warm_start(sb_batch* b, sb_mlp* m, int mode);
/* Record the per-strip features at every step (also when no network is set). */
int sb_batch_set_record_features(sb_batch* b, int on);
long sb_batch_nr_strips(const sb_batch* b);
/* Features [strips][64] and scale = c |u| / 2 [strips] of the last step, in
 * instance order. */
int sb_batch_get_features(const sb_batch* b, float* features, float* scale, size_t cap_strips);
/* Instances that kept their default initial guess in the last step because
 * the prediction was not usable (not finite, or |prediction| > 10). */
long sb_batch_last_fallbacks(const sb_batch* b);
/* Instances whose last solve started from the network's guess. */
long sb_batch_last_warm_taken(const sb_batch* b);

/* ---- Complete sail model ---------------------------------------------------
 * CompleteSailModel (stormbird/src/lifting_line/complete_sail_model/mod.rs:38-330):
 * a lifting-line simulation, a WindEnvironment (wind/environment.rs:21-360)
 * and a Controller (controller/mod.rs:22-60), built from the JSON of
 * CompleteSailModelBuilder (complete_sail_model/builder.rs:14-20):
 *   {"lifting_line_simulation": <SimulationBuilder>, "wind_environment": {..},
 *    "controller": <ControllerBuilder>}.
 * Wind conditions cross the ABI as the JSON of WindCondition
 * (wind/wind_condition/mod.rs:18-29).  The freestream of a step is evaluated
 * on the device for the wake points and on the host for the control points
 * (inflow corrections and controller measurements).  Where the reference
 * panics (a controller with time_steps_between_updates 0, set points for more
 * sails than the model has, an empty measurement range, interpolation tables
 * of unequal length) these calls return -4 with a message.
 * The reference's own C interface (cstormbird) is in cstormbird_compat.h. */
typedef struct sb_csm sb_csm;
sb_csm* sb_csm_create_ex(const char* setup_json, int device, int mode);  /* new_from_string */
/* Parses and checks a CompleteSailModelBuilder string without touching a GPU:
 * out[0..7) as sb_setup_info for its simulation, out[7] controller set points,
 * out[8] 1 with inflow corrections (cap >= 9). */
int sb_csm_setup_info(const char* setup_json, long* out, size_t cap);
void sb_csm_destroy(sb_csm* m);
/* The model's simulation (owned by the model): counts, wake state, settings. */
sb_sim* sb_csm_simulation(sb_csm* m);
/* do_step (mod.rs:174-191): freestream from the wind condition, then one step. */
int sb_csm_do_step(sb_csm* m, double time, double time_step, const char* wind_condition_json,
                   double ship_velocity, sb_step_result* out);
/* The wake starts afresh at the next step, as do_multiple_steps does (mod.rs:152). */
int sb_csm_restart(sb_csm* m);
/* apply_controller_based_on_wind_condition (mod.rs:193-211). */
int sb_csm_apply_controller_wind(sb_csm* m, double time, double time_step, const char* wind_condition_json,
                                 double ship_velocity, double loading);
/* apply_controller_based_on_simulation_result (mod.rs:290-318): the result's
 * angles of attack [N], force-input velocity [3N], solver input velocity [3N]
 * and control points [3N]. */
int sb_csm_apply_controller_result(sb_csm* m, double time, double time_step, double loading,
                                   const double* angles_of_attack, const double* velocity,
                                   const double* input_velocity, const double* ctrl_points);
/* simulate_steady_state_condition (mod.rs:99-122). */
int sb_csm_simulate_steady_state(sb_csm* m, const char* wind_condition_json, double ship_velocity,
                                 double loading, sb_step_result* out);
/* simulate_condition_optimal_controller_loading (mod.rs:58-97): the result of
 * the loading (of linspace(0.1, 1, n)) with the largest delivered power. */
int sb_csm_simulate_optimal_loading(sb_csm* m, const char* wind_condition_json, double ship_velocity,
                                    long nr_loadings, sb_step_result* out, double* best_loading);
/* freestream_velocity (mod.rs:213-255) at the control points and wake points
 * [3 * sb_sim_nr_freestream_points]. */
int sb_csm_freestream(sb_csm* m, const char* wind_condition_json, double ship_velocity, double time,
                      double* out, size_t n);
/* WindEnvironment queries at n locations [3n] (environment.rs:66-190), with
 * the gust signals at `time` and linear_velocity [3] (NULL: zero):
 * kind 0 unsteady_apparent_wind_velocity_vector_at_location, 1 steady apparent,
 * 2 unsteady_true_wind_velocity_vector_at_location, 3 steady true vector
 * (vectors, out [3n]); 4 steady_true_wind_velocity_at_location, 5
 * unsteady_parallel_true_wind_velocity_at_location (speeds, out [n]). */
int sb_wind_velocity(const char* environment_json, const char* wind_condition_json, int kind,
                     const double* locations, size_t n, const double* linear_velocity, double time, double* out);
/* apparent_wind_direction_from_condition_and_linear_velocity (environment.rs:258-278). */
int sb_wind_apparent_direction(const char* environment_json, const char* wind_condition_json,
                               const double* linear_velocity, double height, double* out);
/* apparent_wind_velocity_vectors_at_ctrl_points_with_corrections_applied
 * (environment.rs:192-223): wing_ranges [2 * nr_wings] as (start, end). */
int sb_wind_ctrl_points_velocity(const char* environment_json, const char* wind_condition_json,
                                 const double* ctrl_points, size_t n, const double* linear_velocity, double time,
                                 const long* wing_ranges, size_t nr_wings, double* out);
/* WindCondition values (wind_condition/mod.rs:37-61) at n heights: which 0
 * steady_true_wind_velocity_at_height, 1 unsteady_parallel_true_wind_velocity_at_height,
 * 2 unsteady_perpendicular_true_wind_velocity, 3 unsteady_vertical_true_wind_velocity
 * (2 and 3 ignore the height), 4 businger_dyer_unscaled_correction (0 unless
 * the profile is a LogarithmicModel). */
int sb_wind_condition_values(const char* wind_condition_json, int which, const double* heights, size_t n,
                             double time, double* out);

#ifdef __cplusplus
}
#endif
#endif
