/*
 * ORACLE — test infrastructure only.
 *
 * C ABI of the CPU f64 restatement of Stormbird's dynamic lifting-line path
 * (stormbird 0.9.0, Rust).  Only tests/, __graft_entry__.smoke() and the
 * bench's cpu_baseline leg load this library, and only as the checker or the
 * timed CPU baseline.  The product library (stormbird_amd) never links it.
 *
 * The entry points mirror the product's C ABI (include/stormbird_gpu.h) one to
 * one, with an `orc_` prefix, so the same test code can drive both.
 */
#ifndef STORMBIRD_ORACLE_H
#define STORMBIRD_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

/* Same layout as sb_step_result in include/stormbird_gpu.h. Any array pointer
 * may be NULL, in which case that output is skipped. */
typedef struct orc_step_result {
    double time;
    long iterations;
    double residual;
    double* circulation_strength;  /* [N]            */
    double* angles_of_attack;      /* [N]            */
    double* ctrl_points_velocity;  /* [3N]  force_input.velocity               */
    double* solver_input_velocity; /* [3N]  solver_input_ctrl_points_velocity */
    double* ctrl_points;           /* [3N]  */
    double* sectional_forces;      /* [6][3N]: circulatory, viscous_lift,
                                      sectional_drag, added_mass, gyroscopic, total */
    double* integrated_forces;     /* [6][3W] same order            */
    double* integrated_moments;    /* [6][3W] same order            */
} orc_step_result;

const char* orc_last_error(void);
/* Gaussian elimination with partial pivoting (stormath matrix/linalg.rs:41-104);
 * A row-major n x n. Returns 0, or -1 for a (nearly) singular pivot. */
int orc_solve_gaussian_elimination(const double* A, const double* b, size_t n, double* x);
/* Unit-strength velocity of one horseshoe vortex (horseshoe_vortex.rs:23-103):
 * bound segment start->end, trailing legs start+wake_start->start and end->end+wake_end. */
void orc_horseshoe_unit_velocity(const double* start, const double* end, const double* wake_start,
                                 const double* wake_end, const double* x, double core, double* out);

orc_sim* orc_sim_create(const char* setup_json);
void orc_sim_destroy(orc_sim* sim);
orc_sim* orc_sim_clone(const orc_sim* sim);

long orc_sim_nr_ctrl_points(const orc_sim* sim);
long orc_sim_nr_wings(const orc_sim* sim);
long orc_sim_nr_span_points(const orc_sim* sim);
long orc_sim_nr_wake_points(const orc_sim* sim);
long orc_sim_nr_panels(const orc_sim* sim);
long orc_sim_nr_freestream_points(const orc_sim* sim);
long orc_sim_get_freestream_points(const orc_sim* sim, double* xyz, size_t cap);

int orc_sim_do_step(orc_sim* sim, double time, double time_step,
                    const double* freestream_xyz, size_t n_points,
                    orc_step_result* out);

int orc_sim_induced_velocities(const orc_sim* sim, const double* pts, size_t n,
                               double* out);

int orc_sim_get_previous_circulation(const orc_sim* sim, double* out, size_t n);
int orc_sim_set_previous_circulation(orc_sim* sim, const double* g, size_t n);

int orc_sim_set_translation_and_rotation_fd(orc_sim* sim, double time_step,
                                           const double* translation3,
                                           const double* rotation3);
int orc_sim_set_local_wing_angles(orc_sim* sim, const double* angles, size_t n);

/* Line-force-model state the controller restatement reads: global span lines
 * [N][2][3], chord_vectors_global [N][3], chord_vectors_local_not_rotated
 * [N][3], wing index ranges [W][2], local wing angles [W], rigid body
 * [translation, rotation, velocity_linear, velocity_angular][3].  Any pointer
 * may be NULL. */
int orc_sim_get_line_model(const orc_sim* sim, double* span_lines, double* chord_global,
                           double* chord_local_not_rotated, long* wing_ranges, double* wing_angles,
                           double* rigid_body);
/* first_time_step_completed = false (complete_sail_model/mod.rs:152). */
int orc_sim_restart(orc_sim* sim);

/* Wake state access (parity tests start both sides from one state). */
int orc_sim_get_wake_state(const orc_sim* sim, double* points, double* velocities,
                           double* strengths, long* steps_completed,
                           int* first_step_done);
int orc_sim_set_wake_state(orc_sim* sim, const double* points,
                           const double* velocities, const double* strengths,
                           long steps_completed, int first_step_done);
/* Rebuilds panels from the current points (update_panel_data). */
int orc_sim_update_panels(orc_sim* sim);
int orc_sim_get_panels(const orc_sim* sim, double* rec /* [nr_panels][22] */);
/* FrozenWake internals after the last step. */
int orc_sim_get_frozen_wake(const orc_sim* sim, double* fixed /*3N*/,
                            double* factors /*N*N*3 row-major*/);

/* Kernel-level access for closed-form tests. */
void orc_segment_unit_velocity(const double* p0, const double* p1,
                               const double* x, double core, double* out);
void orc_panel_unit_velocity(const double* corners12, double far_ratio,
                             double core, const double* x, int mode,
                             double* out);
double orc_foil_cl(const char* foil_json, double alpha, int which);
double orc_sigmoid(double x, double x0, double range);
/* stormath transformations (spatial_vector/transformations.rs:29-180);
 * rotation_type 0 XYZ, 1 YawPitchRoll. */
void orc_rotate(const double* v, const double* rotation, int rotation_type, double* out);
void orc_rotate_around_axis(const double* v, double angle, const double* axis, double* out);
void orc_in_rotated_coordinate_system(const double* v, const double* rotation, int rotation_type,
                                      double* out);
void orc_moment_in_new_coordinate_system(const double* moment, const double* rotation,
                                         const double* translation, const double* force,
                                         int rotation_type, double* out);
/* RigidBodyMotion (stormath/src/rigid_body_motion.rs:22-103); motion12 =
 * translation, rotation, velocity_linear, velocity_angular. */
void orc_rigid_body_transform_point(const double* motion12, int rotation_type, const double* p,
                                    double* out);
void orc_rigid_body_velocity_at_point(const double* motion12, int rotation_type, const double* p,
                                      double* out);
/* update_translation_/update_rotation_with_velocity_using_finite_difference, in place. */
void orc_rigid_body_update_fd(double* motion12, const double* translation, const double* rotation,
                              double time_step);
/* VelocityCorrections::corrected_velocity (solvers/velocity_corrections.rs:28-100):
 * kind 1 MaxInducedVelocityMagnitudeRatio(ratio), 2 FixedMagnitudeEqualToFreestream,
 * 0 none (freestream + induced). fs, ui, out: [n][3]. */
void orc_corrected_velocity(int kind, double ratio, const double* fs, const double* ui, size_t n,
                            double* out);
/* FrozenWake::update_induced_velocities_at_control_points (frozen_wake.rs:268-301):
 * fixed [3n], factors [n][n][3] row-major, strengths [n], out [3n]. */
void orc_frozen_matvec(const double* fixed, const double* factors, const double* strengths, size_t n,
                       double* out);
/* OpenMP threads of the per-target loops (n > 0 sets it); returns the current count. */
int orc_set_threads(int n);
/* 0: host libm, 1: portable libm shared with the device code */
int orc_libm_variant(void);
/* fn: 0 sin, 1 cos, 2 exp, 3 acos, 4 pow(x, y), 5 log, 6 atan */
void orc_libm_eval(int fn, const double* x, const double* y, double* out, size_t n);

#ifdef __cplusplus
}
#endif
#endif
