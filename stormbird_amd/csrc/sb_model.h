// Host-side model of a lifting-line simulation: setup parsing, line-force-model
// geometry and solver/wake settings.  The per-step O(N) geometry (rigid-body
// transform of span lines and chord vectors) runs here and is uploaded; all
// O(N*M) and O(N^2) work runs on the GPU.
#pragma once

#include <string>
#include <vector>

#include "sb_common.h"

namespace sb {

enum class RotType { XYZ, YawPitchRoll };

// stormath/src/rigid_body_motion.rs:22-103
struct RigidBody {
    V3 translation{0, 0, 0}, rotation{0, 0, 0}, velocity_linear{0, 0, 0}, velocity_angular{0, 0, 0};
    RotType rot = RotType::XYZ;
    V3 rotate(V3 v) const;
    V3 transform_point(V3 p) const { return rotate(p) + translation; }
    V3 transform_vector(V3 v) const { return rotate(v); }
    V3 body_fixed(V3 v) const;
    V3 velocity_at_point(V3 p) const { return velocity_linear + cross(velocity_angular, p - translation); }
    void update_translation_fd(V3 t, double dt);
    void update_rotation_fd(V3 r, double dt);
};

V3 rotate_around_axis(V3 v, double angle, V3 axis);

struct SpanLine {
    V3 start, end;
    V3 rel() const { return end - start; }
    double length() const { return len(rel()); }
    V3 dir() const { return normalize(rel()); }
    V3 ctrl() const { return 0.5 * (start + end); }
};

struct LineModel {
    std::vector<SpanLine> local, global;
    std::vector<V3> chord_local_not_rotated, chord_local, chord_global, chord_global_at_span;
    std::vector<double> chord_len;
    std::vector<char> is_virtual;
    std::vector<FoilParams> foils;  // one per wing
    std::vector<int> wing_start, wing_end, wing_of;
    std::vector<double> wing_angles;
    std::vector<V3> span_points, ctrl_points;
    double density = 1.225;
    bool output_body = false;
    bool viscous_lift_in_circulation = false;
    RigidBody rbm;

    int n() const { return (int)local.size(); }
    int nw() const { return (int)wing_start.size(); }
    void update_global();  // data_update.rs:13-90
    template <class T>
    std::vector<T> span_values(const std::vector<T>& cv) const;  // value_mapping.rs:87-137
};

struct WakeConfig {
    int M = 100;
    double first_rel = 0.75, last_rel = 0.25, damping = 0.0, far_ratio = 5.0;
    int first_dir = 0;  // 0 chord, 1 freestream, 2 actual velocity
    int symmetry = 0;   // 0 none, 1 X, 2 Y, 3 Z
    bool neglect_self = false;
    long end_index = 0;
    double rep_chord = 0.0;
    double initial_relative_wake_length = 50.0;
    std::vector<double> core;  // per panel, [M][N]
    bool write_files = false;  // write_wake_data_to_file
    std::string files_folder;  // wake_files_folder_path
};

struct SolverConfig {
    int kind = 0;  // 0 SimpleIterative, 1 Linearized (solvers/mod.rs:21-24)
    int max_iterations = 40;
    double damping = 0.05, residual_tol = 1e-4, strength_tol = 1e-6;
    int velocity_correction = 0;  // 0 none, 1 max ratio, 2 fixed magnitude
    double velocity_correction_ratio = 0.0;
    double correction_factor_initial = 1.0;
    bool start_linearized = false;  // start_with_linearized_solution
    // Linearized settings (linearized.rs:26-48); the velocity correction above applies
    bool disable_viscous_corrections = false;
    int induced_velocity_correction = 1;  // 0 NoCorrection, 1 FullCorrection
};

// Quasi-steady wake (wake/settings.rs:33-58): one horseshoe vortex per strip.
struct QuasiConfig {
    double wake_length_factor = 100.0;
    int symmetry = 0;     // 0 none, 1 X, 2 Y, 3 Z
    int core_kind = 0;    // 0 Relative, 1 Absolute, 2 NoViscousCore
    double core_value = 0.1;
};

struct Setup {
    LineModel lfm;
    bool quasi = false;   // SimulationSettings::QuasiSteady (no shed wake)
    QuasiConfig qs;
    WakeConfig wake;
    SolverConfig solver;
};

// Parses a SimulationBuilder JSON string (simulation_builder.rs:34-143). Throws
// std::runtime_error for malformed input and for settings outside this path.
Setup parse_setup(const char* json);

// ---- wind and the complete sail model (stormbird/src/wind, controller,
// lifting_line/complete_sail_model) --------------------------------------------
struct Spectrum {  // DiscretizedSpectrum (wind_condition/discretized_spectrum.rs:12-17)
    std::vector<double> frequencies, amplitudes, phases;
    double value_at_time(double t) const;  // :37-45
};

struct WindConditionCfg {  // wind_condition/mod.rs:18-29
    double direction_coming_from = 0.0;
    int profile = 0;  // VelocityProfile (sb_wind.h)
    double p0 = 0.0, p1 = 0.0, p2 = 0.0;
    bool has_obukhov = false;
    double obukhov = 0.0, stable = 6.0, unstable = 19.3;
    bool has_gust[3] = {false, false, false};  // parallel, perpendicular, vertical
    Spectrum gust[3];
};

struct InflowCorrectionsSingle {  // inflow_corrections.rs:13-19
    std::vector<double> heights, magnitude, angle;
    std::vector<std::pair<long, long>> wing_ranges;
};

struct WindEnvironmentCfg {  // environment.rs:21-34
    V3 up{0.0, 0.0, 1.0}, axis{0.0, 0.0, 1.0}, zero_dir{1.0, 0.0, 0.0};
    double water_plane_height = 0.0;
    bool has_inflow = false;
    std::vector<double> inflow_directions;
    std::vector<InflowCorrectionsSingle> inflow;
};

struct MeasurementCfg {  // controller/measurements.rs:20-30
    int type = 0;  // 0 Mean, 1 Max, 2 Min
    long start_index = 0, end_offset = 0;
};

struct SetPointsCfg {  // controller/set_points.rs:52-68
    std::vector<double> directions;
    bool has_aoa = false, has_state = false;
    std::vector<double> aoa, state;
    bool spin_ratio = false;  // InternalStateType::SpinRatio
    double diameter = 0.0, max_rps = 0.0;
    bool use_effective = false;
    bool has_max_angle_rate = false, has_max_state_rate = false;
    double max_angle_rate = 0.0, max_state_rate = 0.0;
};

struct ControllerCfg {  // controller/builder.rs:14-31
    std::vector<SetPointsCfg> set_points;
    MeasurementCfg aoa, direction, velocity;
    long steps_between_updates = 1;
    double start_time = 0.0;
    bool use_input_velocity = false;
};

struct CompleteSailModelCfg {  // complete_sail_model/builder.rs:14-20
    Setup sim;
    WindEnvironmentCfg env;
    ControllerCfg controller;
};

CompleteSailModelCfg parse_complete_sail_model(const char* json);
WindConditionCfg parse_wind_condition(const char* json);
WindEnvironmentCfg parse_wind_environment(const char* json);

// stormath/src/interpolation.rs:17-100 (clamped piecewise-linear, binary search)
double interpolate_scalar(double x, const double* xs, const double* ys, size_t n);
V3 interpolate_vector(double x, const double* xs, const V3* ys, size_t n);

}  // namespace sb
