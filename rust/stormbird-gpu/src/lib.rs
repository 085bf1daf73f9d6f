//! Rust binding of libstormbird_amd: Stormbird's dynamic lifting-line step on one MI355X GPU.
//!
//! [`GpuSimulation`] is a drop-in for `stormbird::lifting_line::simulation::Simulation`
//! (stormbird/src/lifting_line/simulation.rs:20-216): it is built from the same
//! `SimulationBuilder` JSON string, takes the same `do_step(time, time_step, freestream)`
//! arguments and returns the reference's own `SimulationResult`, with the wake state, the
//! Biot-Savart sums and the circulation solve resident on the GPU.  In the default EXACT
//! arithmetic the results equal the reference f64 path to the last bit wherever both use
//! the same elementary functions (the library's own, shared by its host and device code).
//!
//! [`ffi`] holds the raw C ABI of include/stormbird_gpu.h, one declaration per exported
//! function; [`GpuBatch`] steps many simulations in one launch.
//!
//! Threading follows the reference: a handle is used by one thread at a time (`Send`, not
//! `Sync`); distinct handles may run concurrently.

use std::ffi::{CStr, CString};
use std::fmt;
use std::os::raw::{c_char, c_int, c_long};
use std::ptr::NonNull;

use stormath::rigid_body_motion::RigidBodyMotion;
use stormath::spatial_vector::SpatialVector;
use stormath::type_aliases::Float;
use stormbird::common_utils::flow_derivatives::FlowDerivatives;
use stormbird::common_utils::forces_and_moments::{
    CoordinateSystem, IntegratedValues, SectionalForces, SectionalForcesInput,
};
use stormbird::common_utils::results::simulation::SimulationResult;
use stormbird::error::Error;
use stormbird::lifting_line::simulation_builder::{SimulationBuilder, SimulationSettings};

// The GPU path computes in f64; stormbird built with `single_precision` cannot share its types.
const _: () = assert!(std::mem::size_of::<Float>() == 8);

/// The raw C ABI (include/stormbird_gpu.h).  Vectors cross it as `[x, y, z]` f64 triples;
/// every output array is owned by the caller; negative return values are error codes
/// (-1 null argument, -2 zero capacity, -3 size mismatch, -4 unsupported request,
/// -5 GPU runtime error), with the message in `sb_last_error()`.
#[allow(non_camel_case_types)]
pub mod ffi {
    use std::os::raw::{c_char, c_int, c_long, c_void};

    /// Opaque simulation handle (one `Simulation`).
    #[repr(C)]
    pub struct sb_sim {
        _private: [u8; 0],
    }
    /// Opaque batch of simulations stepped together.
    #[repr(C)]
    pub struct sb_batch {
        _private: [u8; 0],
    }
    /// Opaque warm-start network.
    #[repr(C)]
    pub struct sb_mlp {
        _private: [u8; 0],
    }
    /// Opaque complete sail model.
    #[repr(C)]
    pub struct sb_csm {
        _private: [u8; 0],
    }

    /// Per-step outputs (`sb_step_result`); a null array pointer skips that output.
    #[repr(C)]
    pub struct sb_step_result {
        pub time: f64,
        pub iterations: c_long,
        pub residual: f64,
        /// [N]
        pub circulation_strength: *mut f64,
        /// [N]
        pub angles_of_attack: *mut f64,
        /// [3N] force_input.velocity
        pub ctrl_points_velocity: *mut f64,
        /// [3N] solver_input_ctrl_points_velocity
        pub solver_input_velocity: *mut f64,
        /// [3N]
        pub ctrl_points: *mut f64,
        /// [6][3N]: circulatory, viscous_lift, sectional_drag, added_mass, gyroscopic, total
        pub sectional_forces: *mut f64,
        /// [6][3W] same order
        pub integrated_forces: *mut f64,
        /// [6][3W] same order
        pub integrated_moments: *mut f64,
    }

    pub const SB_MODE_EXACT: c_int = 0;
    pub const SB_MODE_FAST: c_int = 1;
    pub const SB_MODE_FP32_KAHAN: c_int = 2;

    extern "C" {
        pub fn sb_last_error() -> *const c_char;
        pub fn sb_sim_create(setup_json: *const c_char) -> *mut sb_sim;
        pub fn sb_setup_info(setup_json: *const c_char, out: *mut c_long, cap: usize) -> c_int;
        pub fn sb_sim_create_ex(setup_json: *const c_char, device: c_int, mode: c_int) -> *mut sb_sim;
        pub fn sb_sim_destroy(sim: *mut sb_sim);
        pub fn sb_sim_nr_ctrl_points(sim: *const sb_sim) -> c_long;
        pub fn sb_sim_nr_wings(sim: *const sb_sim) -> c_long;
        pub fn sb_sim_nr_span_points(sim: *const sb_sim) -> c_long;
        pub fn sb_sim_nr_wake_points(sim: *const sb_sim) -> c_long;
        pub fn sb_sim_nr_panels(sim: *const sb_sim) -> c_long;
        pub fn sb_sim_nr_freestream_points(sim: *const sb_sim) -> c_long;
        pub fn sb_sim_get_freestream_points(sim: *const sb_sim, xyz: *mut f64, cap: usize) -> c_long;
        pub fn sb_sim_do_step(sim: *mut sb_sim, time: f64, time_step: f64, freestream_xyz: *const f64, n_points: usize, out: *mut sb_step_result) -> c_int;
        pub fn sb_sim_do_step_device(sim: *mut sb_sim, time: f64, time_step: f64, freestream_xyz_device: *const f64, n_points: usize, out: *mut sb_step_result) -> c_int;
        pub fn sb_sim_induced_velocities(sim: *mut sb_sim, points_xyz: *const f64, n: usize, out_xyz: *mut f64) -> c_int;
        pub fn sb_sim_get_previous_circulation(sim: *const sb_sim, out: *mut f64, n: usize) -> c_int;
        pub fn sb_sim_set_previous_circulation(sim: *mut sb_sim, gamma: *const f64, n: usize) -> c_int;
        pub fn sb_sim_write_wake_obj(sim: *const sb_sim, path: *const c_char) -> c_int;
        pub fn sb_sim_write_wake_vtk(sim: *const sb_sim, path: *const c_char) -> c_int;
        pub fn sb_sim_get_panel_point_indices(sim: *const sb_sim, indices: *mut c_long, cap: usize) -> c_int;
        pub fn sb_debug_format_f64(value: f64, buf: *mut c_char, cap: usize) -> c_int;
        pub fn sb_sim_reset(sim: *mut sb_sim) -> c_int;
        pub fn sb_sim_set_solver_settings(sim: *mut sb_sim, max_iterations: c_long, damping: f64, residual_tol: f64, strength_tol: f64) -> c_int;
        pub fn sb_sim_set_translation_and_rotation_fd(sim: *mut sb_sim, time_step: f64, translation3: *const f64, rotation3: *const f64) -> c_int;
        pub fn sb_sim_set_local_wing_angles(sim: *mut sb_sim, angles: *const f64, n: usize) -> c_int;
        pub fn sb_sim_get_local_wing_angles(sim: *const sb_sim, out: *mut f64, n: usize) -> c_int;
        pub fn sb_sim_get_wing_ranges(sim: *const sb_sim, ranges: *mut c_long, cap: usize) -> c_int;
        pub fn sb_sim_get_rigid_body_motion(sim: *const sb_sim, motion12: *mut f64) -> c_int;
        pub fn sb_sim_get_wake_state(sim: *const sb_sim, points: *mut f64, velocities: *mut f64, strengths: *mut f64, steps_completed: *mut c_long, first_step_done: *mut c_int) -> c_int;
        pub fn sb_sim_set_wake_state(sim: *mut sb_sim, points: *const f64, velocities: *const f64, strengths: *const f64, steps_completed: c_long, first_step_done: c_int) -> c_int;
        pub fn sb_sim_set_wake_state_ex(sim: *mut sb_sim, points: *const f64, n_points: usize, velocities: *const f64, n_velocities: usize, strengths: *const f64, n_strengths: usize, steps_completed: c_long, first_step_done: c_int) -> c_int;
        pub fn sb_sim_get_panels(sim: *const sb_sim, records: *mut f64) -> c_int;
        pub fn sb_sim_get_frozen_wake(sim: *const sb_sim, fixed: *mut f64, factors: *mut f64) -> c_int;
        pub fn sb_sim_set_kernel_timers(sim: *mut sb_sim, on: c_int) -> c_int;
        pub fn sb_sim_last_kernel_ms(sim: *const sb_sim, name: *const c_char) -> f64;
        pub fn sb_sim_last_pair_counts(sim: *const sb_sim, counts3: *mut f64) -> c_int;
        pub fn sb_sim_set_target_tiling(sim: *mut sb_sim, span_points_per_wave: c_int) -> c_int;
        pub fn sb_sim_get_target_tiling(sim: *const sb_sim) -> c_int;
        pub fn sb_sim_count_self_induction_pairs(sim: *mut sb_sim, near_total2: *mut f64) -> c_int;
        pub fn sb_wake_induced_velocities(panel_records: *const f64, n_panels: usize, targets_xyz: *const f64, n_targets: usize, symmetry: c_int, mode: c_int, out_xyz: *mut f64) -> c_int;
        pub fn sb_wake_induced_velocities_noalloc(panel_records_device: *const f64, n_panels: usize, targets_xyz_device: *const f64, n_targets: usize, symmetry: c_int, mode: c_int, out_xyz_device: *mut f64, stream: *mut c_void) -> c_int;
        pub fn sb_noalloc_max_panels() -> c_long;
        pub fn sb_frozen_matvec(fixed: *const f64, factors: *const f64, strengths: *const f64, n: usize, out: *mut f64) -> c_int;
        pub fn sb_debug_libm_device(func: c_int, x: *const f64, y: *const f64, out: *mut f64, n: usize) -> c_int;
        pub fn sb_debug_libm_host(func: c_int, x: *const f64, y: *const f64, out: *mut f64, n: usize) -> c_int;
        pub fn sb_debug_linearized_in_shared_memory(n_strips: c_long) -> c_int;
        pub fn sb_debug_gauss_device(a: *const f64, b: *const f64, n: usize, x: *mut f64, variant: c_int) -> c_int;
        pub fn sb_debug_gauss_blocked_device(a: *const f64, b: *const f64, n: usize, x: *mut f64) -> c_int;
        pub fn sb_batch_create(setup_jsons: *const *const c_char, n: c_long, device: c_int, mode: c_int) -> *mut sb_batch;
        pub fn sb_batch_destroy(b: *mut sb_batch);
        pub fn sb_batch_size(b: *const sb_batch) -> c_long;
        pub fn sb_batch_instance(b: *mut sb_batch, i: c_long) -> *mut sb_sim;
        pub fn sb_batch_freestream_offsets(b: *const sb_batch, offsets: *mut c_long, cap: usize) -> c_long;
        pub fn sb_batch_do_step(b: *mut sb_batch, times: *const f64, dts: *const f64, fs: *const f64, n_points: usize, results: *mut sb_step_result) -> c_int;
        pub fn sb_batch_do_step_device(b: *mut sb_batch, times: *const f64, dts: *const f64, fs: *const f64, n_points: usize, results: *mut sb_step_result) -> c_int;
        pub fn sb_batch_last_step_ms(b: *const sb_batch) -> f64;
        pub fn sb_batch_reset(b: *mut sb_batch) -> c_int;
        pub fn sb_batch_set_wind(b: *mut sb_batch, environment_json: *const c_char, wind_conditions_json: *const *const c_char) -> c_int;
        pub fn sb_batch_do_step_wind(b: *mut sb_batch, times: *const f64, dts: *const f64, ship_velocities: *const f64, results: *mut sb_step_result) -> c_int;
        pub fn sb_batch_get_freestream(b: *const sb_batch, out: *mut f64, n_points: usize) -> c_int;
        pub fn sb_batch_set_tuning(b: *mut sb_batch, fixed_per: c_int, targets_tile: c_int) -> c_int;
        pub fn sb_batch_get_tuning(b: *const sb_batch, fixed_per: *mut c_int, targets_tile: *mut c_int) -> c_int;
        pub fn sb_batch_autotune(b: *mut sb_batch, cache_path: *const c_char, flags: c_int, reps: c_int, max_jobs: c_long, timings: *mut f64, cap: usize) -> c_int;
        pub fn sb_mlp_create(device: c_int, seed: u64) -> *mut sb_mlp;
        pub fn sb_mlp_destroy(m: *mut sb_mlp);
        pub fn sb_mlp_nr_params(m: *const sb_mlp) -> c_long;
        pub fn sb_mlp_get_params(m: *const sb_mlp, params: *mut f32, cap: usize) -> c_int;
        pub fn sb_mlp_set_params(m: *mut sb_mlp, params: *const f32, n: usize) -> c_int;
        pub fn sb_mlp_forward(m: *mut sb_mlp, x: *const f32, n: c_long, y: *mut f32) -> c_int;
        pub fn sb_mlp_gradients(m: *mut sb_mlp, x: *const f32, t: *const f32, n: c_long, grads: *mut f32, loss: *mut f32) -> c_int;
        pub fn sb_mlp_train_step(m: *mut sb_mlp, x: *const f32, t: *const f32, n: c_long, learning_rate: f32, loss: *mut f32) -> c_int;
        pub fn sb_batch_set_warm_start(b: *mut sb_batch, m: *mut sb_mlp, mode: c_int) -> c_int;
        pub fn sb_batch_set_record_features(b: *mut sb_batch, on: c_int) -> c_int;
        pub fn sb_batch_nr_strips(b: *const sb_batch) -> c_long;
        pub fn sb_batch_get_features(b: *const sb_batch, features: *mut f32, scale: *mut f32, cap_strips: usize) -> c_int;
        pub fn sb_batch_last_fallbacks(b: *const sb_batch) -> c_long;
        pub fn sb_batch_last_warm_taken(b: *const sb_batch) -> c_long;
        pub fn sb_csm_create_ex(setup_json: *const c_char, device: c_int, mode: c_int) -> *mut sb_csm;
        pub fn sb_csm_setup_info(setup_json: *const c_char, out: *mut c_long, cap: usize) -> c_int;
        pub fn sb_csm_destroy(m: *mut sb_csm);
        pub fn sb_csm_simulation(m: *mut sb_csm) -> *mut sb_sim;
        pub fn sb_csm_do_step(m: *mut sb_csm, time: f64, time_step: f64, wind_condition_json: *const c_char, ship_velocity: f64, out: *mut sb_step_result) -> c_int;
        pub fn sb_csm_restart(m: *mut sb_csm) -> c_int;
        pub fn sb_csm_apply_controller_wind(m: *mut sb_csm, time: f64, time_step: f64, wind_condition_json: *const c_char, ship_velocity: f64, loading: f64) -> c_int;
        pub fn sb_csm_apply_controller_result(m: *mut sb_csm, time: f64, time_step: f64, loading: f64, angles_of_attack: *const f64, velocity: *const f64, input_velocity: *const f64, ctrl_points: *const f64) -> c_int;
        pub fn sb_csm_simulate_steady_state(m: *mut sb_csm, wind_condition_json: *const c_char, ship_velocity: f64, loading: f64, out: *mut sb_step_result) -> c_int;
        pub fn sb_csm_simulate_optimal_loading(m: *mut sb_csm, wind_condition_json: *const c_char, ship_velocity: f64, nr_loadings: c_long, out: *mut sb_step_result, best_loading: *mut f64) -> c_int;
        pub fn sb_csm_freestream(m: *mut sb_csm, wind_condition_json: *const c_char, ship_velocity: f64, time: f64, out: *mut f64, n: usize) -> c_int;
        pub fn sb_wind_velocity(environment_json: *const c_char, wind_condition_json: *const c_char, kind: c_int, locations: *const f64, n: usize, linear_velocity: *const f64, time: f64, out: *mut f64) -> c_int;
        pub fn sb_wind_apparent_direction(environment_json: *const c_char, wind_condition_json: *const c_char, linear_velocity: *const f64, height: f64, out: *mut f64) -> c_int;
        pub fn sb_wind_ctrl_points_velocity(environment_json: *const c_char, wind_condition_json: *const c_char, ctrl_points: *const f64, n: usize, linear_velocity: *const f64, time: f64, wing_ranges: *const c_long, nr_wings: usize, out: *mut f64) -> c_int;
        pub fn sb_wind_condition_values(wind_condition_json: *const c_char, which: c_int, heights: *const f64, n: usize, time: f64, out: *mut f64) -> c_int;
    }
}

/// Arithmetic of the Biot-Savart sums (`SB_MODE_*`).
#[derive(Clone, Copy, Debug, Default, PartialEq, Eq)]
pub enum Arithmetic {
    /// The reference's operation order, IEEE division and square root, no fused
    /// multiply-add: equal to the f64 path bit for bit.
    #[default]
    Exact,
    /// Fused multiply-add and refined reciprocal square roots in the pair sums.
    Fast,
    /// fp32 pair terms with compensated (Kahan) sums; single simulations only.
    Fp32Kahan,
}

impl Arithmetic {
    fn code(self) -> c_int {
        match self {
            Arithmetic::Exact => ffi::SB_MODE_EXACT,
            Arithmetic::Fast => ffi::SB_MODE_FAST,
            Arithmetic::Fp32Kahan => ffi::SB_MODE_FP32_KAHAN,
        }
    }
}

/// An error code of the C ABI with the library's message.
#[derive(Debug, Clone)]
pub struct GpuError {
    pub code: i32,
    pub message: String,
}

impl fmt::Display for GpuError {
    fn fmt(&self, f: &mut fmt::Formatter<'_>) -> fmt::Result {
        write!(f, "stormbird GPU error {}: {}", self.code, self.message)
    }
}

impl std::error::Error for GpuError {}

impl From<GpuError> for Error {
    fn from(e: GpuError) -> Self {
        Error::CustomStringError(e.to_string())
    }
}

fn last_error() -> String {
    // SAFETY: sb_last_error returns a NUL-terminated string owned by the library (thread local).
    unsafe {
        let p = ffi::sb_last_error();
        if p.is_null() {
            String::new()
        } else {
            CStr::from_ptr(p).to_string_lossy().into_owned()
        }
    }
}

fn check(code: c_int) -> Result<(), GpuError> {
    if code == 0 {
        Ok(())
    } else {
        Err(GpuError { code, message: last_error() })
    }
}

fn to_xyz(v: &[SpatialVector]) -> Vec<f64> {
    let mut out = Vec::with_capacity(3 * v.len());
    for p in v {
        out.extend_from_slice(&[p[0], p[1], p[2]]);
    }
    out
}

fn from_xyz(a: &[f64]) -> Vec<SpatialVector> {
    a.chunks_exact(3).map(|c| SpatialVector::from([c[0], c[1], c[2]])).collect()
}

fn vec3(a: &[f64]) -> SpatialVector {
    SpatialVector::from([a[0], a[1], a[2]])
}

/// Host arrays one step's outputs are written into (N strips, W wings).
struct StepArrays {
    n: usize,
    w: usize,
    gamma: Vec<f64>,
    alpha: Vec<f64>,
    velocity: Vec<f64>,
    input: Vec<f64>,
    ctrl: Vec<f64>,
    sectional: Vec<f64>,
    forces: Vec<f64>,
    moments: Vec<f64>,
}

impl StepArrays {
    fn new(n: usize, w: usize) -> Self {
        StepArrays {
            n,
            w,
            gamma: vec![0.0; n],
            alpha: vec![0.0; n],
            velocity: vec![0.0; 3 * n],
            input: vec![0.0; 3 * n],
            ctrl: vec![0.0; 3 * n],
            sectional: vec![0.0; 18 * n],
            forces: vec![0.0; 18 * w],
            moments: vec![0.0; 18 * w],
        }
    }

    fn raw(&mut self) -> ffi::sb_step_result {
        ffi::sb_step_result {
            time: 0.0,
            iterations: 0,
            residual: 0.0,
            circulation_strength: self.gamma.as_mut_ptr(),
            angles_of_attack: self.alpha.as_mut_ptr(),
            ctrl_points_velocity: self.velocity.as_mut_ptr(),
            solver_input_velocity: self.input.as_mut_ptr(),
            ctrl_points: self.ctrl.as_mut_ptr(),
            sectional_forces: self.sectional.as_mut_ptr(),
            integrated_forces: self.forces.as_mut_ptr(),
            integrated_moments: self.moments.as_mut_ptr(),
        }
    }

    fn sectional(&self, cs: CoordinateSystem) -> SectionalForces {
        let k = |i: usize| from_xyz(&self.sectional[3 * self.n * i..3 * self.n * (i + 1)]);
        SectionalForces {
            circulatory: k(0),
            viscous_lift: k(1),
            sectional_drag: k(2),
            added_mass: k(3),
            gyroscopic: k(4),
            total: k(5),
            coordinate_system: cs,
        }
    }

    fn integrated(&self, a: &[f64]) -> Vec<IntegratedValues> {
        let at = |kind: usize, w: usize| vec3(&a[3 * (self.w * kind + w)..]);
        (0..self.w)
            .map(|w| IntegratedValues {
                circulatory: at(0, w),
                viscous_lift: at(1, w),
                sectional_drag: at(2, w),
                added_mass: at(3, w),
                gyroscopic: at(4, w),
                total: at(5, w),
            })
            .collect()
    }
}

/// What the result of one step needs besides the step's own arrays.
struct ResultContext<'a> {
    coordinate_system: CoordinateSystem,
    wing_indices: &'a [std::ops::Range<usize>],
    motion: RigidBodyMotion,
    acceleration: Vec<SpatialVector>,
}

fn simulation_result(raw: &ffi::sb_step_result, a: &StepArrays, cx: ResultContext) -> SimulationResult {
    let velocity = from_xyz(&a.velocity);
    // force_calculations.rs:336-350: the rotation velocity in the output coordinate system
    let rotation_velocity = match cx.coordinate_system {
        CoordinateSystem::Global => cx.motion.velocity_angular,
        CoordinateSystem::Body => cx.motion.vector_in_body_fixed_coordinate_system(cx.motion.velocity_angular),
    };
    SimulationResult {
        time: raw.time,
        ctrl_points: from_xyz(&a.ctrl),
        solver_input_ctrl_points_velocity: from_xyz(&a.input),
        force_input: SectionalForcesInput {
            circulation_strength: a.gamma.clone(),
            velocity,
            angles_of_attack: a.alpha.clone(),
            acceleration: cx.acceleration,
            rotation_velocity,
            coordinate_system: cx.coordinate_system,
        },
        sectional_forces: a.sectional(cx.coordinate_system),
        integrated_forces: a.integrated(&a.forces),
        integrated_moments: a.integrated(&a.moments),
        // the library takes the NoPower input-power model only, whose power is zero
        input_power: vec![0.0; a.w],
        iterations: raw.iterations as usize,
        residual: raw.residual,
        wing_indices: cx.wing_indices.to_vec(),
        rigid_body_motion: cx.motion,
    }
}

/// `Simulation` on the GPU (stormbird/src/lifting_line/simulation.rs:20-216).
pub struct GpuSimulation {
    raw: NonNull<ffi::sb_sim>,
    n: usize,
    w: usize,
    quasi_steady: bool,
    coordinate_system: CoordinateSystem,
    wing_indices: Vec<std::ops::Range<usize>>,
    /// The reference's FlowDerivatives over the returned (global) control-point velocities,
    /// for `force_input.acceleration`; the GPU step applies the same acceleration to the
    /// added-mass force itself.
    flow_derivatives: Option<FlowDerivatives>,
    arrays: StepArrays,
}

// One thread at a time, as the reference's `&mut self` API.
unsafe impl Send for GpuSimulation {}

impl GpuSimulation {
    /// `Simulation::new_from_string` (simulation.rs:33-39) on GPU 0 in EXACT arithmetic.
    /// The string is first read by the reference's own `SimulationBuilder`, so that the same
    /// setups are accepted and refused with the same `Error`.
    pub fn new_from_string(setup_string: &str) -> Result<Self, Error> {
        Self::new_from_string_on(setup_string, 0, Arithmetic::Exact)
    }

    /// The same on GPU `device` with the given arithmetic.
    pub fn new_from_string_on(setup_string: &str, device: i32, arithmetic: Arithmetic) -> Result<Self, Error> {
        let builder = SimulationBuilder::new_from_string(setup_string)?;
        let quasi_steady = matches!(builder.simulation_settings, SimulationSettings::QuasiSteady(_));
        let coordinate_system = builder.line_force_model.output_coordinate_system;
        let c = CString::new(setup_string).map_err(|e| Error::CustomStringError(e.to_string()))?;
        // SAFETY: c is NUL-terminated and outlives the call; a null return reports an error.
        let raw = unsafe { ffi::sb_sim_create_ex(c.as_ptr(), device, arithmetic.code()) };
        let raw = NonNull::new(raw).ok_or_else(|| Error::CustomStringError(last_error()))?;
        // SAFETY: raw is a live handle.
        let (n, w) = unsafe { (ffi::sb_sim_nr_ctrl_points(raw.as_ptr()), ffi::sb_sim_nr_wings(raw.as_ptr())) };
        let (n, w) = (n as usize, w as usize);
        let mut ranges = vec![0 as c_long; 2 * w];
        // SAFETY: ranges holds 2 * w longs.
        check(unsafe { ffi::sb_sim_get_wing_ranges(raw.as_ptr(), ranges.as_mut_ptr(), ranges.len()) })?;
        let wing_indices = ranges.chunks_exact(2).map(|r| r[0] as usize..r[1] as usize).collect();
        Ok(GpuSimulation {
            raw,
            n,
            w,
            quasi_steady,
            coordinate_system,
            wing_indices,
            flow_derivatives: None,
            arrays: StepArrays::new(n, w),
        })
    }

    pub fn nr_ctrl_points(&self) -> usize {
        self.n
    }

    pub fn nr_wings(&self) -> usize {
        self.w
    }

    /// `Simulation::get_freestream_velocity_points` (simulation.rs:45-63): control points,
    /// then every wake point.
    pub fn get_freestream_velocity_points(&self) -> Vec<SpatialVector> {
        // SAFETY: live handle; the buffer holds the count the library reports.
        unsafe {
            let total = ffi::sb_sim_nr_freestream_points(self.raw.as_ptr()) as usize;
            let mut xyz = vec![0.0; 3 * total];
            let got = ffi::sb_sim_get_freestream_points(self.raw.as_ptr(), xyz.as_mut_ptr(), total);
            assert!(got >= 0, "{}", last_error());
            xyz.truncate(3 * got as usize);
            from_xyz(&xyz)
        }
    }

    /// `Simulation::do_step` (simulation.rs:107-203).  Panics where the reference panics
    /// (a singular Linearized system) and on a GPU runtime error; `try_do_step` returns them.
    pub fn do_step(&mut self, time: Float, time_step: Float, freestream_velocity: &[SpatialVector]) -> SimulationResult {
        self.try_do_step(time, time_step, freestream_velocity).unwrap_or_else(|e| panic!("{e}"))
    }

    pub fn try_do_step(
        &mut self,
        time: Float,
        time_step: Float,
        freestream_velocity: &[SpatialVector],
    ) -> Result<SimulationResult, GpuError> {
        let fs = to_xyz(freestream_velocity);
        let mut raw = self.arrays.raw();
        // SAFETY: fs holds freestream_velocity.len() points; raw points into self.arrays,
        // sized for this simulation's N and W.
        check(unsafe {
            ffi::sb_sim_do_step(self.raw.as_ptr(), time, time_step, fs.as_ptr(), freestream_velocity.len(), &mut raw)
        })?;
        let motion = self.rigid_body_motion()?;
        let acceleration = self.acceleration(&freestream_velocity[..self.n], time_step, &motion);
        let cx = ResultContext {
            coordinate_system: self.coordinate_system,
            wing_indices: &self.wing_indices,
            motion,
            acceleration,
        };
        Ok(simulation_result(&raw, &self.arrays, cx))
    }

    /// flow_derivatives.rs:30-62 over the step's output velocities (global), reset to the
    /// control-point freestream at the first step (simulation.rs:65-94).
    fn acceleration(&mut self, ctrl_freestream: &[SpatialVector], dt: Float, m: &RigidBodyMotion) -> Vec<SpatialVector> {
        let returned = from_xyz(&self.arrays.velocity);
        let global: Vec<SpatialVector> = match self.coordinate_system {
            CoordinateSystem::Global => returned,
            CoordinateSystem::Body => returned
                .iter()
                .map(|v| v.from_rotated_to_global_system(m.rotation, m.rotation_type))
                .collect(),
        };
        let fd = self.flow_derivatives.get_or_insert_with(|| FlowDerivatives::new(ctrl_freestream));
        let acc = fd.acceleration(&global, dt);
        fd.update(&global);
        match self.coordinate_system {
            CoordinateSystem::Global => acc,
            CoordinateSystem::Body => acc.iter().map(|a| m.vector_in_body_fixed_coordinate_system(*a)).collect(),
        }
    }

    /// `Simulation::induced_velocities` (simulation.rs:207-216); panics for a quasi-steady
    /// simulation, as the reference does.
    pub fn induced_velocities(&self, points: &[SpatialVector]) -> Vec<SpatialVector> {
        if self.quasi_steady {
            panic!("Induced velocities not implemented for pure quasi-steady simulations yet");
        }
        let xyz = to_xyz(points);
        let mut out = vec![0.0; xyz.len()];
        // SAFETY: both arrays hold points.len() points.
        let rc = unsafe { ffi::sb_sim_induced_velocities(self.raw.as_ptr(), xyz.as_ptr(), points.len(), out.as_mut_ptr()) };
        assert!(rc == 0, "{}", last_error());
        from_xyz(&out)
    }

    /// `Simulation::previous_circulation_strength` (simulation.rs:26): the solver's initial guess.
    pub fn previous_circulation_strength(&self) -> Vec<Float> {
        let mut g = vec![0.0; self.n];
        // SAFETY: g holds n values.
        let rc = unsafe { ffi::sb_sim_get_previous_circulation(self.raw.as_ptr(), g.as_mut_ptr(), self.n) };
        assert!(rc == 0, "{}", last_error());
        g
    }

    pub fn set_previous_circulation_strength(&mut self, gamma: &[Float]) -> Result<(), GpuError> {
        // SAFETY: gamma.len() values are read.
        check(unsafe { ffi::sb_sim_set_previous_circulation(self.raw.as_ptr(), gamma.as_ptr(), gamma.len()) })
    }

    /// `LineForceModel::set_translation_and_rotation_with_finite_difference_for_the_velocity`
    /// (line_force_model/data_update.rs:188-205).
    pub fn set_translation_and_rotation_with_finite_difference_for_the_velocity(
        &mut self,
        time_step: Float,
        translation: SpatialVector,
        rotation: SpatialVector,
    ) -> Result<(), GpuError> {
        let t = [translation[0], translation[1], translation[2]];
        let r = [rotation[0], rotation[1], rotation[2]];
        // SAFETY: t and r hold three values each.
        check(unsafe { ffi::sb_sim_set_translation_and_rotation_fd(self.raw.as_ptr(), time_step, t.as_ptr(), r.as_ptr()) })
    }

    /// `LineForceModel::set_local_wing_angles` (line_force_model/data_update.rs:171-177).
    pub fn set_local_wing_angles(&mut self, angles: &[Float]) -> Result<(), GpuError> {
        // SAFETY: angles.len() values are read.
        check(unsafe { ffi::sb_sim_set_local_wing_angles(self.raw.as_ptr(), angles.as_ptr(), angles.len()) })
    }

    /// `LineForceModel::rigid_body_motion` (line_force_model/mod.rs:88).
    pub fn rigid_body_motion(&self) -> Result<RigidBodyMotion, GpuError> {
        let mut m = [0.0; 12];
        // SAFETY: m holds 12 values.
        check(unsafe { ffi::sb_sim_get_rigid_body_motion(self.raw.as_ptr(), m.as_mut_ptr()) })?;
        Ok(RigidBodyMotion {
            translation: vec3(&m[0..]),
            rotation: vec3(&m[3..]),
            velocity_linear: vec3(&m[6..]),
            velocity_angular: vec3(&m[9..]),
            ..Default::default()
        })
    }

    /// The underlying handle, for the calls of [`ffi`] this wrapper does not cover.
    pub fn as_raw(&self) -> *mut ffi::sb_sim {
        self.raw.as_ptr()
    }
}

impl Drop for GpuSimulation {
    fn drop(&mut self) {
        // SAFETY: the handle was created by sb_sim_create_ex and is destroyed once.
        unsafe { ffi::sb_sim_destroy(self.raw.as_ptr()) }
    }
}

/// Many simulations stepped together in one launch per kernel (`sb_batch_*`); each
/// instance's results equal stepping it alone.  No counterpart in the reference, whose
/// callers loop over `Simulation` values.
pub struct GpuBatch {
    raw: NonNull<ffi::sb_batch>,
    offsets: Vec<usize>,
    shapes: Vec<(usize, usize, CoordinateSystem, Vec<std::ops::Range<usize>>)>,
}

unsafe impl Send for GpuBatch {}

impl GpuBatch {
    /// One simulation per setup string; `arithmetic` is Exact or Fast.
    pub fn new(setup_strings: &[&str], device: i32, arithmetic: Arithmetic) -> Result<Self, Error> {
        let mut owned = Vec::with_capacity(setup_strings.len());
        let mut systems = Vec::with_capacity(setup_strings.len());
        for s in setup_strings {
            let b = SimulationBuilder::new_from_string(s)?;
            systems.push(b.line_force_model.output_coordinate_system);
            owned.push(CString::new(*s).map_err(|e| Error::CustomStringError(e.to_string()))?);
        }
        let ptrs: Vec<*const c_char> = owned.iter().map(|c| c.as_ptr()).collect();
        // SAFETY: ptrs holds setup_strings.len() NUL-terminated strings alive for the call.
        let raw = unsafe { ffi::sb_batch_create(ptrs.as_ptr(), ptrs.len() as c_long, device, arithmetic.code()) };
        let raw = NonNull::new(raw).ok_or_else(|| Error::CustomStringError(last_error()))?;
        let n = setup_strings.len();
        let mut offsets = vec![0 as c_long; n + 1];
        // SAFETY: offsets holds n + 1 longs.
        let got = unsafe { ffi::sb_batch_freestream_offsets(raw.as_ptr(), offsets.as_mut_ptr(), offsets.len()) };
        if got < 0 {
            return Err(Error::CustomStringError(last_error()));
        }
        let mut shapes = Vec::with_capacity(n);
        for (i, cs) in systems.into_iter().enumerate() {
            // SAFETY: instance i of a live batch; ranges holds 2 * wings longs.
            unsafe {
                let s = ffi::sb_batch_instance(raw.as_ptr(), i as c_long);
                let (nc, nw) = (ffi::sb_sim_nr_ctrl_points(s) as usize, ffi::sb_sim_nr_wings(s) as usize);
                let mut ranges = vec![0 as c_long; 2 * nw];
                check(ffi::sb_sim_get_wing_ranges(s, ranges.as_mut_ptr(), ranges.len()))?;
                let wi = ranges.chunks_exact(2).map(|r| r[0] as usize..r[1] as usize).collect();
                shapes.push((nc, nw, cs, wi));
            }
        }
        Ok(GpuBatch { raw, offsets: offsets.iter().map(|&o| o as usize).collect(), shapes })
    }

    pub fn len(&self) -> usize {
        self.shapes.len()
    }

    pub fn is_empty(&self) -> bool {
        self.shapes.is_empty()
    }

    /// Instance i's freestream points are `offsets[i]..offsets[i + 1]` of the concatenated array.
    pub fn freestream_offsets(&self) -> &[usize] {
        &self.offsets
    }

    /// One step of every instance: `times`, `time_steps` per instance, the concatenated
    /// freestream of all instances (each in its own point order).  `force_input.acceleration`
    /// is left empty here (the step applies it to the added-mass force).
    pub fn do_step(
        &mut self,
        times: &[Float],
        time_steps: &[Float],
        freestream_velocity: &[SpatialVector],
    ) -> Result<Vec<SimulationResult>, GpuError> {
        let n = self.len();
        if times.len() != n || time_steps.len() != n {
            return Err(GpuError { code: -3, message: "one time and one time step per instance".into() });
        }
        let fs = to_xyz(freestream_velocity);
        let mut arrays: Vec<StepArrays> = self.shapes.iter().map(|s| StepArrays::new(s.0, s.1)).collect();
        let mut raws: Vec<ffi::sb_step_result> = arrays.iter_mut().map(|a| a.raw()).collect();
        // SAFETY: every array is sized for its instance; fs holds freestream_velocity.len() points.
        check(unsafe {
            ffi::sb_batch_do_step(self.raw.as_ptr(), times.as_ptr(), time_steps.as_ptr(), fs.as_ptr(),
                                  freestream_velocity.len(), raws.as_mut_ptr())
        })?;
        let mut out = Vec::with_capacity(n);
        for (i, (raw, a)) in raws.iter().zip(arrays.iter()).enumerate() {
            let mut m = [0.0; 12];
            // SAFETY: instance i of a live batch; m holds 12 values.
            unsafe {
                let s = ffi::sb_batch_instance(self.raw.as_ptr(), i as c_long);
                check(ffi::sb_sim_get_rigid_body_motion(s, m.as_mut_ptr()))?;
            }
            let motion = RigidBodyMotion {
                translation: vec3(&m[0..]),
                rotation: vec3(&m[3..]),
                velocity_linear: vec3(&m[6..]),
                velocity_angular: vec3(&m[9..]),
                ..Default::default()
            };
            let cx = ResultContext {
                coordinate_system: self.shapes[i].2,
                wing_indices: &self.shapes[i].3,
                motion,
                acceleration: Vec::new(),
            };
            out.push(simulation_result(raw, a, cx));
        }
        Ok(out)
    }
}

impl Drop for GpuBatch {
    fn drop(&mut self) {
        // SAFETY: created by sb_batch_create, destroyed once.
        unsafe { ffi::sb_batch_destroy(self.raw.as_ptr()) }
    }
}
