//! The GPU simulation against the reference's own CPU `Simulation`, built from one setup
//! string and stepped with one freestream.
//!
//! Bar: 1e-10 relative in circulation and integrated forces.  The GPU's EXACT arithmetic
//! keeps the reference's operation order; the two builds differ only in the elementary
//! functions (the library's own sin/exp/acos/pow against the Rust build's libm, <= 1 ulp
//! apart), whose differences the undamped frozen-wake iteration amplifies over tens of
//! steps, so the trajectory is compared over its first 40 steps.

use std::f64::consts::PI;

use stormath::spatial_vector::SpatialVector;
use stormbird::common_utils::results::simulation::SimulationResult;
use stormbird::lifting_line::simulation::Simulation;
use stormbird_gpu::GpuSimulation;

/// One 30 m x 10 m wing sail, 20 cosine-spaced strips, a frozen-shape wake of 200 rows.
fn single_wing_sail(dynamic: bool) -> String {
    let points: Vec<String> = (0..=20)
        .map(|k| format!(r#"{{"x": 0.0, "y": 0.0, "z": {:?}}}"#, 15.0 * (1.0 - (k as f64 * PI / 20.0).cos())))
        .collect();
    let chords = vec![r#"{"x": 10.0, "y": 0.0, "z": 0.0}"#; 21].join(", ");
    let settings = if dynamic {
        r#"{"Dynamic": {"wake": {"nr_panels_per_line_element": 200, "viscous_core_length": {"Absolute": 1.0},
            "first_panel_relative_length": 0.25}}}"#
    } else {
        r#"{"QuasiSteady": {}}"#
    };
    format!(
        r#"{{"line_force_model": {{"nr_sections": 12, "wing_builders": [{{
            "section_points": [{}], "chord_vectors": [{}],
            "section_model": {{"Foil": {{"cl_initial_slope": {:?}, "mean_positive_stall_angle": {:?},
                "mean_negative_stall_angle": {:?}, "cd_min": 0.01, "cd_second_order_factor": 0.05}}}}}}]}},
            "simulation_settings": {}}}"#,
        points.join(", "),
        chords,
        2.0 * PI,
        15.0f64.to_radians(),
        15.0f64.to_radians(),
        settings
    )
}

fn rel_dev(a: &[f64], b: &[f64]) -> f64 {
    let scale = b.iter().fold(1e-300f64, |m, x| m.max(x.abs()));
    a.iter().zip(b).fold(0.0f64, |m, (x, y)| m.max((x - y).abs())) / scale
}

fn forces(r: &SimulationResult) -> Vec<f64> {
    r.integrated_forces.iter().flat_map(|v| [v.total[0], v.total[1], v.total[2]]).collect()
}

fn compare(setup: &str, steps: usize) {
    let mut cpu = Simulation::new_from_string(setup).expect("reference build");
    let mut gpu = GpuSimulation::new_from_string(setup).expect("GPU build");
    let cpu_points = cpu.get_freestream_velocity_points();
    let gpu_points = gpu.get_freestream_velocity_points();
    assert_eq!(cpu_points.len(), gpu_points.len());
    let u = SpatialVector::from([10.0 * 10.0f64.to_radians().cos(), 10.0 * 10.0f64.to_radians().sin(), 0.0]);
    let freestream = vec![u; cpu_points.len()];
    let dt = 0.25;
    for k in 0..steps {
        let t = k as f64 * dt;
        let rc = cpu.do_step(t, dt, &freestream);
        let rg = gpu.do_step(t, dt, &freestream);
        let dg = rel_dev(&rg.force_input.circulation_strength, &rc.force_input.circulation_strength);
        let df = rel_dev(&forces(&rg), &forces(&rc));
        assert!(dg <= 1e-10 && df <= 1e-10, "step {k}: circulation {dg:.3e}, forces {df:.3e}");
        assert_eq!(rg.iterations, rc.iterations, "step {k}");
        assert_eq!(rg.wing_indices, rc.wing_indices);
    }
    assert_eq!(gpu.previous_circulation_strength().len(), cpu.previous_circulation_strength.len());
}

#[test]
fn dynamic_single_wing_sail_matches_cpu() {
    compare(&single_wing_sail(true), 40);
}

#[test]
fn quasi_steady_single_wing_sail_matches_cpu() {
    compare(&single_wing_sail(false), 5);
}

#[test]
fn induced_velocities_match_cpu() {
    let setup = single_wing_sail(true);
    let mut cpu = Simulation::new_from_string(&setup).unwrap();
    let mut gpu = GpuSimulation::new_from_string(&setup).unwrap();
    let n = cpu.get_freestream_velocity_points().len();
    let freestream = vec![SpatialVector::from([10.0, 1.0, 0.0]); n];
    for k in 0..5 {
        cpu.do_step(k as f64 * 0.25, 0.25, &freestream);
        gpu.do_step(k as f64 * 0.25, 0.25, &freestream);
    }
    let points: Vec<SpatialVector> = (0..50).map(|i| SpatialVector::from([i as f64, 2.0, 10.0])).collect();
    let a: Vec<f64> = gpu.induced_velocities(&points).iter().flat_map(|v| [v[0], v[1], v[2]]).collect();
    let b: Vec<f64> = cpu.induced_velocities(&points).iter().flat_map(|v| [v[0], v[1], v[2]]).collect();
    assert!(rel_dev(&a, &b) <= 1e-12);
}

#[test]
fn bad_setups_fail_like_the_reference() {
    assert!(GpuSimulation::new_from_string("not json").is_err());
    assert!(Simulation::new_from_string("not json").is_err());
}
