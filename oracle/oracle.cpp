// ORACLE — test infrastructure only.  Never linked by the product library.
//
// CPU f64 restatement of Stormbird 0.9.0's dynamic lifting-line path
// (Simulation::do_step with DynamicWake + FrozenWake + SimpleIterative and
// the Foil section model).  Each function follows the Rust source line by
// line; the file:line it restates is cited beside it (paths relative to the
// stormbird repository root).  Arithmetic is written in the reference's
// evaluation order and compiled with -ffp-contract=off so no FMA is formed;
// transcendental functions come from the host libm, as the Rust build does.
//
// Multithreading: only the per-target induced-velocity sum runs in parallel
// (OpenMP), exactly where the reference uses rayon
// (stormbird/src/lifting_line/wake/dynamic_wake/induced_velocity_calc.rs:108-126).
// Each target's fold is sequential in ascending panel order, so the result
// does not depend on the thread count.

#include "oracle.h"

#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "json_lite.h"

#ifdef _OPENMP
#include <omp.h>
#endif

// Elementary functions: the host libm by default (what the Rust build calls on
// Linux); with ORC_PORTABLE_LIBM the portable implementation that the device
// code also uses, so that whole free-wake trajectories can be compared bit for
// bit.  The two variants differ by at most an ulp per call (tests/test_libm.py).
#ifdef ORC_PORTABLE_LIBM
#include "../stormbird_amd/csrc/sb_libm.h"
#define ORC_SIN sbm::sin
#define ORC_COS sbm::cos
#define ORC_EXP sbm::exp
#define ORC_ACOS sbm::acos
#define ORC_POW sbm::pow
#define ORC_LOG sbm::log
#define ORC_ATAN sbm::atan
#else
#define ORC_SIN std::sin
#define ORC_COS std::cos
#define ORC_EXP std::exp
#define ORC_ACOS std::acos
#define ORC_POW std::pow
#define ORC_LOG std::log
#define ORC_ATAN std::atan
#endif

namespace {

using orc_json::Value;

constexpr double PI = 3.14159265358979323846;
constexpr double TAU = 6.28318530717958647692;
constexpr double MIN_POSITIVE = DBL_MIN;             // stormath/src/consts.rs:26
constexpr double FOUR_PI_INVERSE = 1.0 / (4.0 * PI);  // vortex_line.rs:11, panel.rs:11
constexpr double CLOSENESS_ERROR = 1.0e-10;          // vortex_line.rs:13

thread_local std::string g_err;

// ---------------------------------------------------------------------------
// SpatialVector  (stormath/src/spatial_vector/{mod,operators,vector_math}.rs)
// ---------------------------------------------------------------------------
struct V3 {
    double v[3] = {0.0, 0.0, 0.0};
    V3() = default;
    V3(double x, double y, double z) { v[0] = x; v[1] = y; v[2] = z; }
    double& operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
};
inline V3 operator+(const V3& a, const V3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 operator-(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3& operator+=(V3& a, const V3& b) { a = a + b; return a; }
inline V3& operator-=(V3& a, const V3& b) { a = a - b; return a; }
// SpatialVector * Float and Float * SpatialVector both compute v[i] * s (operators.rs:74-123)
inline V3 operator*(const V3& a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
inline V3 operator*(double s, const V3& a) { return {a[0] * s, a[1] * s, a[2] * s}; }
inline V3 operator/(const V3& a, double s) { return {a[0] / s, a[1] / s, a[2] / s}; }
inline V3 operator-(const V3& a) { return {-a[0], -a[1], -a[2]}; }
// vector_math.rs:11-110
inline double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double length_squared(const V3& a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }
inline double length(const V3& a) { return std::sqrt(length_squared(a)); }
inline V3 normalize(const V3& a) {
    double l = length(a);
    if (l > 0.0) return {a[0] / l, a[1] / l, a[2] / l};
    return V3();
}
inline V3 cross(const V3& a, const V3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline double rmax(double a, double b) { return std::fmax(a, b); }
inline double rmin(double a, double b) { return std::fmin(a, b); }
inline double absolute_angle_between(const V3& a, const V3& b) {  // vector_math.rs:57-71
    double sa = length_squared(a), sb = length_squared(b);
    if (sa == 0.0 || sb == 0.0) return 0.0;
    double c = dot(a, b) / std::sqrt(sa * sb);
    c = rmin(rmax(c, -1.0), 1.0);
    return ORC_ACOS(c);
}
inline double signed_angle_between(const V3& a, const V3& b, const V3& axis) {  // :77-89
    double triple = dot(a, cross(b, axis));
    double ang = absolute_angle_between(a, b);
    return triple > 0.0 ? ang : -ang;
}
inline V3 project(const V3& a, const V3& onto) {  // :93-97
    V3 n = normalize(onto);
    return n * dot(a, n);
}
// Rust f64::signum: +1 for +0.0 and positives, -1 for -0.0 and negatives.
inline double signum(double x) {
    if (std::isnan(x)) return x;
    return std::signbit(x) ? -1.0 : 1.0;
}

// transformations.rs:29-97 (XYZ) and :101-114 (yaw/pitch/roll)
inline V3 rotate_around_axis(const V3& self, double angle, const V3& axis) {
    V3 an = normalize(axis);
    double c = ORC_COS(angle), s = ORC_SIN(angle);
    V3 t1 = self * c;
    V3 t2 = cross(an, self) * s;
    V3 t3 = an * dot(an, self) * (1.0 - c);
    return t1 + t2 + t3;
}
inline V3 rotate_xyz(const V3& self, const V3& r) {
    V3 nv = self;
    if (r[0] != 0.0) {
        double ty = nv[1], tz = nv[2];
        nv[1] = ty * ORC_COS(r[0]) - tz * ORC_SIN(r[0]);
        nv[2] = ty * ORC_SIN(r[0]) + tz * ORC_COS(r[0]);
    }
    if (r[1] != 0.0) {
        double tx = nv[0], tz = nv[2];
        nv[0] = tx * ORC_COS(r[1]) + tz * ORC_SIN(r[1]);
        nv[2] = -tx * ORC_SIN(r[1]) + tz * ORC_COS(r[1]);
    }
    if (r[2] != 0.0) {
        double tx = nv[0], ty = nv[1];
        nv[0] = tx * ORC_COS(r[2]) - ty * ORC_SIN(r[2]);
        nv[1] = tx * ORC_SIN(r[2]) + ty * ORC_COS(r[2]);
    }
    return nv;
}
inline V3 rotate_ypr(const V3& self, const V3& r) {
    V3 yaw(0, 0, 1);
    V3 pitch = rotate_around_axis(V3(0, 1, 0), r[2], yaw);
    V3 roll = rotate_around_axis(rotate_around_axis(V3(1, 0, 0), r[2], yaw), r[1], pitch);
    return rotate_around_axis(rotate_around_axis(rotate_around_axis(self, r[2], yaw), r[1], pitch),
                              r[0], roll);
}
enum class RotationType { XYZ, YawPitchRoll };
inline V3 rotate(const V3& self, const V3& r, RotationType t) {
    return t == RotationType::XYZ ? rotate_xyz(self, r) : rotate_ypr(self, r);
}
inline V3 in_rotated_coordinate_system(const V3& self, const V3& r, RotationType t) {  // :147-155
    V3 ax = rotate(V3(1, 0, 0), r, t), ay = rotate(V3(0, 1, 0), r, t), az = rotate(V3(0, 0, 1), r, t);
    return {dot(self, ax), dot(self, ay), dot(self, az)};
}
// :157-180: rotate the moment, then add the moment of the force about the new origin (arm = -translation)
inline V3 moment_in_new_coordinate_system(const V3& self, const V3& r, const V3& translation, const V3& force,
                                          RotationType t) {
    V3 rotated = in_rotated_coordinate_system(self, r, t);
    V3 arm = -translation;
    V3 from_force = cross(arm, force);
    return rotated + in_rotated_coordinate_system(from_force, r, t);
}

// ---------------------------------------------------------------------------
// Rigid body motion (stormath/src/rigid_body_motion.rs:22-103)
// ---------------------------------------------------------------------------
struct RigidBodyMotion {
    V3 translation, rotation, velocity_linear, velocity_angular;
    RotationType rotation_type = RotationType::XYZ;
    V3 transform_point(const V3& p) const { return rotate(p, rotation, rotation_type) + translation; }
    V3 transform_vector(const V3& v) const { return rotate(v, rotation, rotation_type); }
    V3 body_fixed(const V3& v) const { return in_rotated_coordinate_system(v, rotation, rotation_type); }
    V3 velocity_at_point(const V3& p) const {
        return velocity_linear + cross(velocity_angular, p - translation);
    }
    void update_translation_fd(const V3& t, double dt) {
        V3 old = translation;
        translation = t;
        velocity_linear = (translation - old) / dt;
    }
    void update_rotation_fd(const V3& r, double dt) {
        V3 old = rotation;
        rotation = r;
        V3 d = rotation - old;
        for (int i = 0; i < 3; ++i) {  // both branches add TAU (reference behaviour, :94-100)
            if (d[i] > PI) d[i] = d[i] + TAU;
            else if (d[i] < -PI) d[i] = d[i] + TAU;
        }
        velocity_angular = d / dt;
    }
};

// ---------------------------------------------------------------------------
// Singularity elements
// ---------------------------------------------------------------------------
// vortex_line.rs:59-84
inline double normal_distance_squared(const V3& l0, const V3& l1, const V3& p, double r1sq,
                                      double r2sq) {
    V3 rel_line = l1 - l0;
    V3 rel_point = p - l0;
    double line_len_sq = length_squared(rel_line);
    double line_len = std::sqrt(line_len_sq);
    V3 dir = rel_line / line_len;
    double par = dot(rel_point, dir);
    if (par < 0.0) return r1sq;
    if (par > line_len) return r2sq;
    double par_sq = par * par;
    return length_squared(rel_point) - par_sq;
}
// vortex_line.rs:90-106; rc.powi(4) is lowered to (rc*rc)*(rc*rc)
inline double viscous_core_term(const V3& l0, const V3& l1, const V3& p, double rc, double r1sq,
                                double r2sq) {
    double d2 = normal_distance_squared(l0, l1, p, r1sq, r2sq);
    double rc2 = rc * rc;
    double den = std::sqrt(rc2 * rc2 + d2 * d2);
    return den > 0.0 ? d2 / den : 0.0;
}
// vortex_line.rs:18-55
inline V3 segment_unit_velocity(const V3& l0, const V3& l1, const V3& p, double rc) {
    V3 r1 = p - l0;
    V3 r2 = p - l1;
    double r1sq = length_squared(r1), r2sq = length_squared(r2);
    double r1l = std::sqrt(r1sq), r2l = std::sqrt(r2sq);
    double r1r2 = r1l * r2l;
    double den = r1r2 * (r1r2 + dot(r1, r2));
    if (std::fabs(den) > CLOSENESS_ERROR) {
        double core = (rc == 0.0) ? 1.0 : viscous_core_term(l0, l1, p, rc, r1sq, r2sq);
        double k = (r1l + r2l) / den;
        return (cross(r1, r2) * core) * (k * FOUR_PI_INVERSE);
    }
    return V3();
}

// panel.rs:16-118 and geometry_functions.rs:11-30
struct Panel {
    V3 points[4];
    V3 center, normal;
    double far_field_length_squared = 0.0, viscous_core_length = 0.0, area_term = 0.0;

    static Panel make(const V3 pts[4], double far_ratio, double rc) {
        Panel P;
        for (int k = 0; k < 4; ++k) P.points[k] = pts[k];
        P.center = 0.25 * (pts[0] + pts[1] + pts[2] + pts[3]);
        double a = length(pts[1] - pts[0]);
        double b = length(pts[2] - pts[1]);
        double c = length(pts[3] - pts[2]);
        double d = length(pts[0] - pts[3]);
        double s = (a + b + c + d) * 0.5;
        double area = std::sqrt((s - a) * (s - b) * (s - c) * (s - d));
        P.normal = normalize(cross(pts[2] - pts[0], pts[3] - pts[1]));
        double bound = length(pts[1] - pts[0]);
        double stream = length(pts[2] - pts[1]);
        double rep = rmax(bound, stream);
        double f = rep * far_ratio;
        P.far_field_length_squared = f * f;
        P.viscous_core_length = rc;
        P.area_term = area * FOUR_PI_INVERSE;
        return P;
    }
    bool near(const V3& x) const { return length_squared(x - center) <= far_field_length_squared; }
    V3 doublet(const V3& x) const {  // panel.rs:82-96
        V3 t = x - center;
        double d2 = length_squared(t);
        double d5 = d2 * d2 * std::sqrt(d2);
        double nh = dot(t, normal);
        V3 num = t * (3.0 * nh) - normal * d2;
        return (num * area_term) / d5;
    }
    V3 rings(const V3& x) const {  // panel.rs:99-118
        V3 u;
        for (int k = 0; k < 4; ++k) {
            const V3& a = points[k];
            const V3& b = points[(k + 1) & 3];
            u += segment_unit_velocity(a, b, x, viscous_core_length);
        }
        return u;
    }
    V3 unit_velocity(const V3& x) const { return near(x) ? rings(x) : doublet(x); }
};

enum class Symmetry { None, X, Y, Z };
inline bool mirrored_point(Symmetry s, const V3& p, V3& m) {  // symmetry_condition.rs:27-34
    switch (s) {
        case Symmetry::None: return false;
        case Symmetry::X: m = V3(-p[0], p[1], p[2]); return true;
        case Symmetry::Y: m = V3(p[0], -p[1], p[2]); return true;
        case Symmetry::Z: m = V3(p[0], p[1], -p[2]); return true;
    }
    return false;
}
inline V3 corrected_velocity(Symmetry s, const V3& u, const V3& m) {  // :36-43
    switch (s) {
        case Symmetry::None: return u;
        case Symmetry::X: return {u[0] - m[0], u[1] + m[1], u[2] + m[2]};
        case Symmetry::Y: return {u[0] + m[0], u[1] - m[1], u[2] + m[2]};
        case Symmetry::Z: return {u[0] + m[0], u[1] + m[1], u[2] - m[2]};
    }
    return u;
}

// ---------------------------------------------------------------------------
// Foil section model (stormbird/src/section_models/foil.rs)
// ---------------------------------------------------------------------------
inline double sigmoid_zero_to_one(double x, double x0, double range) {  // transition_functions.rs:20-28
    double slope = 4.5951212 / range;
    double xp = slope * (x - x0);
    return 1.0 / (1.0 + ORC_EXP(-xp));
}
inline double to_radians(double deg) { return deg * (PI / 180.0); }

struct Foil {
    double cl_zero_angle = 0.0, cl_initial_slope = TAU;
    double cl_high_order_factor_positive = 0.0, cl_high_order_factor_negative = 0.0;
    double cl_high_order_power = 0.0, cl_max_after_stall = 1.0;
    double cd_min = 0.0, angle_cd_min = 0.0, cd_second_order_factor = 0.0;
    double cd_max_after_stall = 1.0, cd_power_after_stall = 1.6, cdi_correction_factor = 0.0;
    double mean_positive_stall_angle = to_radians(20.0), mean_negative_stall_angle = to_radians(20.0);
    double stall_range = to_radians(6.0), cd_bump_during_stall = 0.0, cd_stall_angle_offset = 0.0;
    double added_mass_factor = 0.0;

    static double stall_angle(double a) {  // foil.rs:119-129
        double e = std::fabs(a);
        while (e > PI) e -= PI;
        e *= signum(a);
        return e;
    }
    double cl_linear(double a) const { return cl_zero_angle + cl_initial_slope * ORC_SIN(a); }
    double cl_pre_raw(double a) const {  // :166-181
        double hp = cl_high_order_power > 0.0 ? ORC_POW(std::fabs(a), cl_high_order_power) * signum(a) : 0.0;
        double f = a >= 0.0 ? cl_high_order_factor_positive : cl_high_order_factor_negative;
        return cl_linear(a) + f * hp;
    }
    double amount_of_stall(double a) const {  // :256-270
        double e = a + angle_cd_min;
        double m = e >= 0.0 ? std::fabs(mean_positive_stall_angle) : std::fabs(mean_negative_stall_angle);
        return sigmoid_zero_to_one(std::fabs(e), m, stall_range);
    }
    double cl_pre_dropoff(double a) const {  // :184-190
        double s = amount_of_stall(a);
        double raw = cl_pre_raw(a);
        return raw * (1.0 - s);
    }
    double cl_post_raw(double a) const { return cl_max_after_stall * ORC_SIN(2.0 * stall_angle(a)); }
    double cl_post_weighted(double a) const {  // :200-206
        double raw = cl_post_raw(a);
        double s = amount_of_stall(a);
        return raw * s;
    }
    double cl(double a) const {  // :152-158
        double pre = cl_pre_dropoff(a);
        double post = cl_post_weighted(a);
        return post + pre;
    }
    double cd(double a) const {  // :212-235
        double sa = stall_angle(a);
        double pe = std::fabs(a + angle_cd_min);
        double pre = cd_min + cd_second_order_factor * (pe * pe);
        double post = cd_max_after_stall * ORC_POW(std::fabs(ORC_SIN(sa)), cd_power_after_stall);
        double at = a + cd_stall_angle_offset * signum(a);
        double s = amount_of_stall(at);
        double raw = pre * (1.0 - s) + s * post;
        double s2 = amount_of_stall(at);
        double during = cd_bump_during_stall * ((1.0 - ORC_COS(s2 * TAU)) * 0.5);
        if (cdi_correction_factor != 0.0) {
            double c = std::fabs(cl(a));
            double corr = cdi_correction_factor * (c * c);
            return raw + corr + during;
        }
        return raw + during;
    }
    double added_mass_coefficient(double acc) const { return added_mass_factor * acc; }
};

// ---------------------------------------------------------------------------
// Line force model (stormbird/src/line_force_model/*)
// ---------------------------------------------------------------------------
struct SpanLine {
    V3 start, end;
    V3 rel() const { return end - start; }
    double len() const { return length(rel()); }
    V3 dir() const { return normalize(rel()); }
    V3 ctrl() const { return 0.5 * (start + end); }
};

enum class CoordSys { Global, Body };

struct Range { size_t start, end; };

struct LineForceModel {
    std::vector<SpanLine> span_lines_local, span_lines_global;
    std::vector<V3> chord_local_not_rotated, chord_local, chord_global, chord_global_at_span_points;
    std::vector<double> chord_lengths;
    std::vector<bool> is_virtual;
    std::vector<Foil> foils;  // one per wing
    std::vector<Range> wing_indices;
    double density = 1.225;
    CoordSys output_cs = CoordSys::Global;
    RigidBodyMotion rbm;
    std::vector<double> local_wing_angles;
    std::vector<V3> span_points_global, ctrl_points_global;
    bool include_viscous_lift_in_circulation = false;

    size_t n() const { return span_lines_local.size(); }
    size_t nw() const { return wing_indices.size(); }
    size_t wing_of(size_t i) const {  // mod.rs:210-218
        for (size_t w = 0; w < wing_indices.size(); ++w)
            if (i >= wing_indices[w].start && i < wing_indices[w].end) return w;
        throw std::runtime_error("Wing index not found");
    }
    template <class T>
    std::vector<T> span_point_values(const std::vector<T>& cv) const {  // value_mapping.rs:87-137 (no extrapolation)
        std::vector<T> out;
        for (auto& r : wing_indices) {
            out.push_back(cv[r.start]);
            for (size_t i = r.start; i < r.end; ++i) {
                if (i == r.end - 1) out.push_back(cv[i]);
                else out.push_back((cv[i] + cv[i + 1]) * 0.5);
            }
        }
        return out;
    }
    void update_global() {  // data_update.rs:13-90
        // update_chord_vectors
        for (size_t i = 0; i < n(); ++i) {
            size_t w = wing_of(i);
            V3 axis = span_lines_local[wing_indices[w].start].rel();
            chord_local[i] = rotate_around_axis(chord_local_not_rotated[i], local_wing_angles[w], axis);
            chord_global[i] = rbm.transform_vector(chord_local[i]);
        }
        chord_global_at_span_points = span_point_values(chord_global);
        for (size_t i = 0; i < n(); ++i) {
            span_lines_global[i].start = rbm.transform_point(span_lines_local[i].start);
            span_lines_global[i].end = rbm.transform_point(span_lines_local[i].end);
        }
        span_points_global.clear();
        for (auto& r : wing_indices) {
            for (size_t i = r.start; i < r.end; ++i) span_points_global.push_back(span_lines_global[i].start);
            span_points_global.push_back(span_lines_global[r.end - 1].end);
        }
        ctrl_points_global.resize(n());
        for (size_t i = 0; i < n(); ++i) ctrl_points_global[i] = span_lines_global[i].ctrl();
    }
    const std::vector<SpanLine>& lines(CoordSys cs) const {
        return cs == CoordSys::Global ? span_lines_global : span_lines_local;
    }
    std::vector<V3> felt_ctrl_velocity(const V3* fs) const {  // force_calculations.rs:50-69
        std::vector<V3> out(n());
        for (size_t i = 0; i < n(); ++i) out[i] = fs[i] - rbm.velocity_at_point(ctrl_points_global[i]);
        return out;
    }
    std::vector<V3> remove_span_velocity(const std::vector<V3>& v) const {  // mod.rs:235-254 (Global)
        std::vector<V3> out(n());
        for (size_t i = 0; i < n(); ++i) out[i] = v[i] - project(v[i], span_lines_global[i].rel());
        return out;
    }
    std::vector<double> angles_of_attack(const std::vector<V3>& v) const {  // force_calculations.rs:79-107
        std::vector<double> a(v.size());
        for (size_t i = 0; i < v.size(); ++i)
            a[i] = length(v[i]) > MIN_POSITIVE
                       ? signed_angle_between(chord_global[i], v[i], span_lines_global[i].dir())
                       : 0.0;
        return a;
    }
    std::vector<double> cl_pre(const std::vector<double>& a) const {
        std::vector<double> c(n());
        for (size_t i = 0; i < n(); ++i) c[i] = foils[wing_of(i)].cl_pre_dropoff(a[i]);
        return c;
    }
    std::vector<double> circulation_strength(const std::vector<double>& a,
                                             const std::vector<V3>& v) const {  // :271-292
        std::vector<double> cl = cl_pre(a);
        if (include_viscous_lift_in_circulation)
            for (size_t i = 0; i < n(); ++i) cl[i] += foils[wing_of(i)].cl_post_weighted(a[i]);
        std::vector<double> g(v.size());
        for (size_t i = 0; i < v.size(); ++i) g[i] = -0.5 * chord_lengths[i] * length(v[i]) * cl[i];
        return g;
    }
    std::vector<V3> circulatory_forces(const std::vector<double>& g,
                                       const std::vector<V3>& v) const {  // :390-411
        const auto& sl = lines(output_cs);
        std::vector<V3> f(n());
        for (size_t i = 0; i < n(); ++i)
            f[i] = length(v[i]) == 0.0 ? V3() : (g[i] * cross(v[i], sl[i].rel())) * density;
        return f;
    }
    double average_residual_absolute(const std::vector<double>& g, const std::vector<double>& a,
                                     const std::vector<V3>& v) const {  // :552-627
        std::vector<V3> f = circulatory_forces(g, v);
        std::vector<double> cd = cl_pre(a);
        double sum = 0.0;
        for (size_t i = 0; i < n(); ++i) {
            double sign = -1.0 * signum(g[i]);
            double area = chord_lengths[i] * span_lines_local[i].len();
            double ff = 0.5 * area * density * length_squared(v[i]);
            double clg = ff == 0.0 ? 0.0 : length(f[i]) * sign / ff;
            sum += std::fabs(clg - cd[i]);
        }
        return sum / (double)n();
    }
};

struct SectionalForces {
    std::vector<V3> k[6];  // circulatory, viscous_lift, sectional_drag, added_mass, gyroscopic, total
    CoordSys cs = CoordSys::Global;
};

// ---------------------------------------------------------------------------
// Dynamic wake (stormbird/src/lifting_line/wake/dynamic_wake/*)
// ---------------------------------------------------------------------------
enum class FirstDir { Chord, Freestream, ActualVelocity };

struct DynamicWake {
    size_t P = 0, N = 0, M = 0;  // points along span, panels along span, panels per line element
    std::vector<V3> points, velocity_at_points;
    std::vector<double> strengths, core;
    std::vector<Panel> panels;
    std::vector<Range> wing_indices;
    double first_rel = 0.75, last_rel = 0.25, damping = 0.0, far_ratio = 5.0;
    FirstDir first_dir = FirstDir::Chord;
    size_t end_index = 0;
    bool neglect_self = false;
    Symmetry sym = Symmetry::None;
    size_t steps_completed = 0;
    double rep_chord = 0.0;

    size_t nr_points() const { return P * (M + 1); }
    size_t nr_panels() const { return N * M; }
    size_t pidx(size_t s, size_t j) const { return s * P + j; }
    size_t wing_index(size_t span_index) const {  // mod.rs:68-76
        for (size_t w = 0; w < wing_indices.size(); ++w)
            if (span_index >= wing_indices[w].start && span_index < wing_indices[w].end) return w;
        throw std::runtime_error("Span index not found in any wing");
    }
    void panel_points(size_t s, size_t i, V3 out[4]) const {  // mod.rs:82-103
        size_t w = wing_index(i);
        out[0] = points[pidx(s, i + w)];
        out[1] = points[pidx(s, i + 1 + w)];
        out[2] = points[pidx(s + 1, i + 1 + w)];
        out[3] = points[pidx(s + 1, i + w)];
    }
    void update_panel_data() {  // update_data.rs:102-115
        for (size_t k = 0; k < nr_panels(); ++k) {
            V3 pts[4];
            panel_points(k / N, k % N, pts);
            panels[k] = Panel::make(pts, far_ratio, core[k]);
        }
    }
    V3 unit_velocity(size_t k, const V3& x) const {  // induced_velocity_calc.rs:182-198
        V3 u = panels[k].unit_velocity(x);
        V3 m;
        if (mirrored_point(sym, x, m)) return corrected_velocity(sym, u, panels[k].unit_velocity(m));
        return u;
    }
    V3 from_panel(size_t k, const V3& x) const {  // :202-210
        if (strengths[k] == 0.0) return V3();
        return strengths[k] * unit_velocity(k, x);
    }
    std::vector<V3> induced_local(const std::vector<V3>& pts, size_t s, size_t e, bool neglect) const {
        std::vector<V3> res(pts.size());
        if (neglect) {  // :128-150
            for (size_t k = s; k < e; ++k) {
                size_t wp = wing_index(k % N);
                for (size_t i = 0; i < pts.size(); ++i)
                    if (wp != wing_index(i)) res[i] += from_panel(k, pts[i]);
            }
        } else {  // :111-126, rayon over targets, ascending panel fold per target
            long np = (long)pts.size();
#pragma omp parallel for schedule(dynamic, 16)
            for (long i = 0; i < np; ++i) {
                V3 acc;
                for (size_t k = s; k < e; ++k) acc = acc + from_panel(k, pts[i]);
                res[i] = acc;
            }
        }
        return res;
    }
    std::vector<V3> induced_velocities(const std::vector<V3>& pts) const {
        return induced_local(pts, 0, strengths.size(), false);
    }

    // update_data.rs:90-99
    void sync_first(const LineForceModel& lfm) {
        for (size_t i = 0; i < lfm.span_points_global.size(); ++i) points[i] = lfm.span_points_global[i];
    }
    std::vector<V3> first_dirs(const LineForceModel& lfm, const V3* felt_span) const {  // :152-195
        std::vector<V3> d(P);
        for (size_t i = 0; i < P; ++i) {
            switch (first_dir) {
                case FirstDir::Chord: d[i] = normalize(lfm.chord_global_at_span_points[i]); break;
                case FirstDir::Freestream: d[i] = normalize(felt_span[i]); break;
                case FirstDir::ActualVelocity:
                    if (steps_completed > 2)
                        d[i] = normalize(0.5 * (velocity_at_points[i] + velocity_at_points[P + i]));
                    else
                        d[i] = normalize(felt_span[i]);
                    break;
            }
        }
        return d;
    }
    void stream_free(double dt) {  // :295-316
        for (size_t s = M; s >= 2; --s) {
            for (size_t j = 0; j < P; ++j) {
                size_t prev = pidx(s - 1, j), cur = pidx(s, j);
                V3 integ = points[prev] + dt * velocity_at_points[prev];
                if (damping > 0.0) points[cur] = points[cur] * damping + integ * (1.0 - damping);
                else points[cur] = integ;
            }
        }
    }
    void move_first(const LineForceModel& lfm, const V3* felt_span) {  // :200-252
        std::vector<V3> old(points.begin() + P, points.begin() + 2 * P);
        std::vector<V3> d = first_dirs(lfm, felt_span);
        double len = rep_chord * first_rel;
        for (size_t i = 0; i < P; ++i) {
            V3 est = points[i] + d[i] * len;
            if (damping > 0.0) points[i + P] = old[i] * damping + est * (1.0 - damping);
            else points[i + P] = est;
        }
    }
    void move_last(const V3* felt_span) {  // :262-292
        size_t n = points.size();
        size_t last = n - P, prev = last - P;
        std::vector<V3> dirs(P);
        for (size_t i = 0; i < P; ++i) {
            if (M > 1) dirs[i] = normalize(points[prev + i] - points[prev - P + i]);
            else dirs[i] = normalize(felt_span[i]);
        }
        for (size_t i = 0; i < P; ++i) {
            V3 cv = last_rel * rep_chord * dirs[i];
            points[last + i] = points[prev + i] + cv;
        }
    }
    void update_points_before_solving(double dt, const LineForceModel& lfm, const V3* felt_span) {  // :136-153
        sync_first(lfm);
        stream_free(dt);
        move_first(lfm, felt_span);
        move_last(felt_span);
    }
    void stream_strengths() {  // :319-328
        for (size_t s = M - 1; s >= 1; --s)
            for (size_t i = 0; i < N; ++i) strengths[s * N + i] = strengths[(s - 1) * N + i];
    }
    void update_before_solving(double dt, const LineForceModel& lfm, const std::vector<V3>& felt_span) {
        update_points_before_solving(dt, lfm, felt_span.data());
        update_panel_data();
        stream_strengths();
    }
    void update_velocity_at_points(const V3* wake_fs) {  // :125-141
        for (size_t i = 0; i < points.size(); ++i) velocity_at_points[i] = wake_fs[i];
        size_t end = std::min(end_index, points.size());
        if (end > 0 && steps_completed > 2) {
            std::vector<V3> tp(points.begin(), points.begin() + end);
            std::vector<V3> u = induced_velocities(tp);
            for (size_t i = 0; i < end; ++i) velocity_at_points[i] += u[i];
        }
    }
    void update_after_solving(const std::vector<double>& g, const V3* wake_fs) {  // :49-59
        for (size_t i = 0; i < g.size(); ++i) strengths[i] = g[i];
        update_velocity_at_points(wake_fs);
        steps_completed += 1;
    }
    void initialize_with_velocity(const LineForceModel& lfm, const V3& vel, double dt) {  // initialization.rs:12-41
        std::fill(strengths.begin(), strengths.end(), 0.0);
        std::fill(velocity_at_points.begin(), velocity_at_points.end(), vel);
        std::vector<V3> wfs(points.size(), vel);
        sync_first(lfm);
        for (size_t k = 0; k < M + 1; ++k) update_points_before_solving(dt, lfm, wfs.data());
        update_panel_data();
    }
    void initialize_based_on_chord(const LineForceModel& lfm, double factor) {  // :46-71
        V3 sum;
        for (auto& c : lfm.chord_global_at_span_points) sum = sum + c;
        V3 avg = sum / (double)lfm.chord_global_at_span_points.size();
        double wl = factor * length(avg);
        V3 vel = (wl / (double)M) * avg;
        initialize_with_velocity(lfm, vel, 1.0);
    }
};

// ---------------------------------------------------------------------------
// Frozen wake (stormbird/src/lifting_line/wake/frozen_wake.rs:31-301)
// ---------------------------------------------------------------------------
struct FrozenWake {
    std::vector<V3> fixed, induced;
    std::vector<V3> factors;  // N x N row-major
    size_t N = 0;
    void init(size_t n) {
        N = n;
        fixed.assign(n, V3());
        induced.assign(n, V3());
        factors.assign(n * n, V3());
    }
    void update_from_full_wake(const std::vector<V3>& ctrl, const DynamicWake& w) {  // :189-261
        fixed = w.induced_local(ctrl, w.N, w.strengths.size(), w.neglect_self);
        for (size_t j = 0; j < N; ++j)
            for (size_t i = 0; i < N; ++i) {
                if (w.neglect_self && w.wing_index(i) == w.wing_index(j)) factors[i * N + j] = V3();
                else factors[i * N + j] = w.unit_velocity(j, ctrl[i]);
            }
    }
    void matvec(const std::vector<double>& g) {  // :268-301
        induced = fixed;
        for (size_t r = 0; r < N; ++r) {
            double sx = 0.0, sy = 0.0, sz = 0.0;
            for (size_t c = 0; c < N; ++c) {
                const V3& f = factors[r * N + c];
                sx += f[0] * g[c];
                sy += f[1] * g[c];
                sz += f[2] * g[c];
            }
            induced[r][0] += sx;
            induced[r][1] += sy;
            induced[r][2] += sz;
        }
    }
};

// Quasi-steady wake (wake/settings.rs:33-58) and the steady frozen wake built
// from horseshoe vortices (frozen_wake.rs:68-129, horseshoe_vortex.rs:23-103).
struct QuasiSteadyWake {
    double wake_length_factor = 100.0;
    Symmetry sym = Symmetry::None;
    int core_kind = 0;  // 0 Relative, 1 Absolute, 2 NoViscousCore
    double core_value = 0.1;
};

struct Horseshoe {
    V3 b0, b1, s0, s1, e0, e1;  // bound, start trailing, end trailing (two points each)
    double core;
    V3 unit(const V3& x) const {  // horseshoe_vortex.rs:23-51
        V3 u;
        u += segment_unit_velocity(b0, b1, x, core);
        u += segment_unit_velocity(s0, s1, x, core);
        u += segment_unit_velocity(e0, e1, x, core);
        return u;
    }
};

void update_as_steady(FrozenWake& fw, const LineForceModel& lfm, const std::vector<V3>& span_vel,
                      const QuasiSteadyWake& w) {  // frozen_wake.rs:68-129
    const size_t n = lfm.n();
    double csum = 0.0;
    for (double c : lfm.chord_lengths) csum += c;
    double avg_chord = csum / (double)n;
    double wake_length = w.wake_length_factor * avg_chord;
    std::vector<V3> wv(span_vel.size());
    for (size_t i = 0; i < span_vel.size(); ++i) wv[i] = normalize(span_vel[i]) * wake_length;
    double core = 0.0;
    if (w.core_kind == 0) {
        double lsum = 0.0;
        for (size_t i = 0; i < n; ++i) lsum += lfm.span_lines_global[i].len();
        core = w.core_value * (lsum / (double)n);
    } else if (w.core_kind == 1) {
        core = w.core_value;
    }
    std::vector<Horseshoe> hs;  // horseshoe_vortex.rs:55-103
    size_t k = 0;
    for (size_t wi = 0; wi < lfm.nw(); ++wi) {
        for (size_t li = lfm.wing_indices[wi].start; li < lfm.wing_indices[wi].end; ++li) {
            const SpanLine& L = lfm.span_lines_global[li];
            hs.push_back({L.start, L.end, L.start + wv[k], L.start, L.end, L.end + wv[k + 1], core});
            k += 1;
        }
        k += 1;
    }
    for (size_t r = 0; r < n; ++r) {
        const V3& x = lfm.ctrl_points_global[r];
        for (size_t c = 0; c < n; ++c) {
            V3 u = hs[c].unit(x);
            V3 m;
            if (mirrored_point(w.sym, x, m)) u = corrected_velocity(w.sym, u, hs[c].unit(m));
            fw.factors[r * n + c] = u;
        }
    }
}

// ---------------------------------------------------------------------------
// Solver (stormbird/src/lifting_line/solvers/{simple_iterative,velocity_corrections}.rs)
// ---------------------------------------------------------------------------
enum class VelCorr { None, MaxRatio, FixedMagnitude };

struct SimpleIterative {
    size_t max_iterations = 40;
    double damping = 0.05, residual_tol = 1e-4, strength_tol = 1e-6;
    VelCorr vc = VelCorr::None;
    double vc_ratio = 0.0;
    bool use_raw = false, start_linearized = false;
    double correction_factor_initial = 1.0;
};

struct SolverResult {
    std::vector<V3> input_velocity, output_velocity;
    std::vector<double> gamma;
    size_t iterations = 0;
    double residual = 0.0;
};

std::vector<V3> corrected_velocity(VelCorr vc, double vc_ratio, const std::vector<V3>& fs,
                                   const std::vector<V3>& ui) {  // velocity_corrections.rs:18-100
    std::vector<V3> out(fs.size());
    if (vc == VelCorr::MaxRatio) {
        for (size_t i = 0; i < fs.size(); ++i) {
            double um = length(ui[i]);
            double fm = length(fs[i]);
            double mx = vc_ratio * fm;
            V3 c = um > mx ? normalize(ui[i]) * mx : ui[i];
            out[i] = fs[i] + c;
        }
    } else {
        for (size_t i = 0; i < fs.size(); ++i) out[i] = fs[i] + ui[i];
        for (size_t i = 0; i < fs.size(); ++i) out[i] = normalize(out[i]) * length(fs[i]);
    }
    return out;
}

// Gaussian elimination with partial pivoting (stormath/src/matrix/linalg.rs:41-104);
// A row-major n x n.  Throws on a (nearly) singular pivot, where the reference's
// Linearized solver unwraps an error (linearized.rs:133).
std::vector<double> solve_gaussian_elimination(const std::vector<double>& A, const std::vector<double>& b,
                                               size_t n) {
    const size_t m = n + 1;
    std::vector<double> aug(n * m);
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = 0; j < n; ++j) aug[i * m + j] = A[i * n + j];
        aug[i * m + n] = b[i];
    }
    for (size_t k = 0; k < n; ++k) {
        size_t max_row = k;
        for (size_t i = k + 1; i < n; ++i)
            if (std::fabs(aug[i * m + k]) > std::fabs(aug[max_row * m + k])) max_row = i;
        if (max_row != k)
            for (size_t j = 0; j <= n; ++j) std::swap(aug[k * m + j], aug[max_row * m + j]);
        if (std::fabs(aug[k * m + k]) < 1e-12)
            throw std::runtime_error("Matrix is singular or nearly singular at row " + std::to_string(k));
        for (size_t i = k + 1; i < n; ++i) {
            double factor = aug[i * m + k] / aug[k * m + k];
            for (size_t j = k; j <= n; ++j) aug[i * m + j] -= factor * aug[k * m + j];
        }
    }
    std::vector<double> x(n, 0.0);
    for (size_t ii = n; ii-- > 0;) {
        double sum = 0.0;
        for (size_t j = ii + 1; j < n; ++j) sum += aug[ii * m + j] * x[j];
        x[ii] = (aug[ii * m + n] - sum) / aug[ii * m + ii];
    }
    return x;
}

enum class IvcMethod { NoCorrection, FullCorrection };

struct Linearized {  // solvers/linearized.rs:17-48
    VelCorr vc = VelCorr::None;
    double vc_ratio = 0.0;
    bool disable_viscous = false;
    IvcMethod ivc = IvcMethod::FullCorrection;
};

// linearized.rs:54-222.  check_for_zero_linear_lift (:224-266) starts its flag at
// true and so never ends the solve early; it has no other effect.
SolverResult solve_linearized(const Linearized& s, const LineForceModel& lfm, const std::vector<V3>& felt,
                              FrozenWake& fw) {
    const size_t n = lfm.n();
    std::vector<V3> u = felt;
    for (size_t i = 0; i < n; ++i) u[i] += fw.fixed[i];
    u = lfm.remove_span_velocity(u);
    std::vector<V3> vdir(n), ndir(n);
    for (size_t i = 0; i < n; ++i) {
        V3 un = normalize(u[i]);
        V3 sd = normalize(lfm.span_lines_global[i].rel());
        vdir[i] = un;
        ndir[i] = cross(un, sd);
    }
    std::vector<double> a = lfm.angles_of_attack(u);
    std::vector<double> cl_lin(n), dcl(n);
    for (size_t i = 0; i < n; ++i) {
        cl_lin[i] = lfm.foils[lfm.wing_of(i)].cl_linear(a[i]);  // force_calculations.rs:188-210
        dcl[i] = lfm.foils[lfm.wing_of(i)].cl_initial_slope;    // :212-229
    }
    std::vector<double> A(n * n), rhs(n);
    for (size_t r = 0; r < n; ++r) {
        double ut = dot(u[r], vdir[r]);
        for (size_t c = 0; c < n; ++c) {
            double uin = dot(fw.factors[r * n + c], ndir[r]);
            A[r * n + c] = -dcl[r] * uin / ut;
            if (r == c) A[r * n + c] += 2.0 / (lfm.chord_lengths[r] * ut);
        }
        rhs[r] = -cl_lin[r];
    }
    std::vector<double> g = solve_gaussian_elimination(A, rhs, n);
    fw.matvec(g);
    auto felt_plus_induced = [&]() {
        if (s.vc != VelCorr::None) {
            u = corrected_velocity(s.vc, s.vc_ratio, felt, fw.induced);
        } else {
            for (size_t i = 0; i < n; ++i) u[i] = felt[i] + fw.induced[i];
        }
    };
    felt_plus_induced();
    if (!s.disable_viscous) {
        a = lfm.angles_of_attack(u);
        std::vector<double> cl_full = lfm.cl_pre(a);
        for (size_t i = 0; i < n; ++i) {
            double cll = lfm.foils[lfm.wing_of(i)].cl_linear(a[i]);
            if (std::fabs(cll) > MIN_POSITIVE) g[i] *= cl_full[i] / cll;
        }
        if (s.ivc == IvcMethod::FullCorrection) fw.matvec(g);
        felt_plus_induced();
    }
    a = lfm.angles_of_attack(u);
    SolverResult r;
    r.residual = lfm.average_residual_absolute(g, a, u);
    r.input_velocity = felt;
    r.gamma = g;
    r.output_velocity = u;
    r.iterations = 1;
    return r;
}

SolverResult solve(const SimpleIterative& s, const LineForceModel& lfm, const std::vector<V3>& felt,
                   FrozenWake& fw, const std::vector<double>& initial) {  // simple_iterative.rs:86-212
    size_t n = lfm.n();
    std::vector<double> g = initial;
    if (std::fabs(s.correction_factor_initial - 1.0) > MIN_POSITIVE)
        for (auto& x : g) x *= s.correction_factor_initial;
    if (s.start_linearized) g = solve_linearized(Linearized(), lfm, felt, fw).gamma;  // :103-113
    std::vector<V3> u(n);
    std::vector<double> a(n, 0.0);
    double residual = lfm.average_residual_absolute(g, a, u);
    size_t it = 0;
    bool converged = false;
    while (it < s.max_iterations && !converged) {
        it += 1;
        fw.matvec(g);
        if (s.vc != VelCorr::None) u = corrected_velocity(s.vc, s.vc_ratio, felt, fw.induced);
        else
            for (size_t i = 0; i < n; ++i) u[i] = felt[i] + fw.induced[i];
        u = lfm.remove_span_velocity(u);
        a = lfm.angles_of_attack(u);
        std::vector<double> est = lfm.circulation_strength(a, u);
        residual = lfm.average_residual_absolute(g, a, u);
        if (residual < s.residual_tol) converged = true;
        double maxd = 0.0;
        for (size_t i = 0; i < n; ++i) {
            double d = est[i] - g[i];
            double ad = std::fabs(d);
            if (ad > maxd) maxd = ad;
            g[i] += s.damping * d;
        }
        if (maxd < s.strength_tol) converged = true;
    }
    SolverResult r;
    r.gamma = lfm.circulation_strength(a, u);
    r.input_velocity = felt;
    r.output_velocity = u;
    r.iterations = it;
    r.residual = residual;
    return r;
}

// ---------------------------------------------------------------------------
// Flow derivatives (stormbird/src/common_utils/flow_derivatives.rs:16-63)
// ---------------------------------------------------------------------------
struct FlowDerivatives {
    std::vector<V3> h0, h1;
    size_t count = 0;
    void reset(const std::vector<V3>& v) { h0 = v; h1 = v; count = 0; }
    std::vector<V3> acceleration(const std::vector<V3>& cur, double dt) const {
        std::vector<V3> a(cur.size());
        if (count < 2) return a;
        for (size_t i = 0; i < cur.size(); ++i)
            a[i] = (cur[i] * 3.0 - h1[i] * 4.0 + h0[i]) / (2.0 * dt);  // finite_difference.rs:33-42
        return a;
    }
    void update(const std::vector<V3>& cur) { h0 = h1; h1 = cur; count += 1; }
};

// ---------------------------------------------------------------------------
// JSON setup (simulation_builder.rs, dynamic_wake/builder.rs, line_force_model/builder/*)
// ---------------------------------------------------------------------------
void check_keys(const Value* o, std::initializer_list<const char*> allowed, const char* ctx) {
    if (!o || o->kind != Value::Object) throw std::runtime_error(std::string(ctx) + ": expected object");
    for (auto& kv : o->obj) {
        bool ok = false;
        for (auto* a : allowed) ok = ok || kv.first == a;
        if (!ok) throw std::runtime_error(std::string(ctx) + ": unknown field `" + kv.first + "`");
    }
}
double num(const Value* v, const char* ctx) {
    if (!v || v->kind != Value::Number) throw std::runtime_error(std::string(ctx) + ": expected number");
    return v->num;
}
V3 vec3(const Value* v, const char* ctx) {  // serde_implementations.rs: {x,y,z} or [x,y,z]
    if (v && v->kind == Value::Array && v->arr.size() == 3)
        return V3(num(v->arr[0].get(), ctx), num(v->arr[1].get(), ctx), num(v->arr[2].get(), ctx));
    if (v && v->kind == Value::Object) {
        const Value *x = v->get("x"), *y = v->get("y"), *z = v->get("z");
        if (x && y && z) return V3(num(x, ctx), num(y, ctx), num(z, ctx));
    }
    throw std::runtime_error(std::string(ctx) + ": expected vector");
}
// Externally tagged enum: "Unit" or {"Variant": payload}
std::string enum_tag(const Value* v, const Value** payload, const char* ctx) {
    *payload = nullptr;
    if (v && v->kind == Value::String) return v->str;
    if (v && v->kind == Value::Object && v->obj.size() == 1) {
        *payload = v->obj[0].second.get();
        return v->obj[0].first;
    }
    throw std::runtime_error(std::string(ctx) + ": expected enum");
}

Foil parse_foil(const Value* o) {
    check_keys(o, {"cl_zero_angle", "cl_initial_slope", "cl_high_order_factor_positive",
                   "cl_high_order_factor_negative", "cl_high_order_power", "cl_max_after_stall",
                   "cd_min", "angle_cd_min", "cd_second_order_factor", "cd_max_after_stall",
                   "cd_power_after_stall", "cdi_correction_factor", "mean_positive_stall_angle",
                   "mean_negative_stall_angle", "stall_range", "cd_bump_during_stall",
                   "cd_stall_angle_offset", "added_mass_factor"},
               "Foil");
    Foil f;
    auto rd = [&](const char* k, double& d) { if (auto* v = o->get(k)) d = num(v, k); };
    rd("cl_zero_angle", f.cl_zero_angle);
    rd("cl_initial_slope", f.cl_initial_slope);
    rd("cl_high_order_factor_positive", f.cl_high_order_factor_positive);
    rd("cl_high_order_factor_negative", f.cl_high_order_factor_negative);
    rd("cl_high_order_power", f.cl_high_order_power);
    rd("cl_max_after_stall", f.cl_max_after_stall);
    rd("cd_min", f.cd_min);
    rd("angle_cd_min", f.angle_cd_min);
    rd("cd_second_order_factor", f.cd_second_order_factor);
    rd("cd_max_after_stall", f.cd_max_after_stall);
    rd("cd_power_after_stall", f.cd_power_after_stall);
    rd("cdi_correction_factor", f.cdi_correction_factor);
    rd("mean_positive_stall_angle", f.mean_positive_stall_angle);
    rd("mean_negative_stall_angle", f.mean_negative_stall_angle);
    rd("stall_range", f.stall_range);
    rd("cd_bump_during_stall", f.cd_bump_during_stall);
    rd("cd_stall_angle_offset", f.cd_stall_angle_offset);
    rd("added_mass_factor", f.added_mass_factor);
    return f;
}
Foil parse_section_model(const Value* v) {
    const Value* p;
    std::string tag = enum_tag(v, &p, "SectionModel");
    if (tag == "Foil") return parse_foil(p);
    throw std::runtime_error("SectionModel `" + tag + "` is not supported by this path");
}

// linear_interpolation (stormath/src/interpolation.rs:17-100)
template <class T>
T lin2(double x, double x0, double x1, const T& y0, const T& y1) {
    if (x <= x0) return y0;
    if (x >= x1) return y1;
    T dy = y1 - y0;
    double dx = x1 - x0;
    return y0 + dy * ((x - x0) / dx);
}
template <class T>
T linear_interpolation(double x, const std::vector<double>& xd, const std::vector<T>& yd) {
    if (xd.size() == 1) return yd[0];
    if (xd.size() == 2) return lin2(x, xd[0], xd[1], yd[0], yd[1]);
    size_t lo = 0, hi = xd.size() - 1;
    while (hi - lo > 1) {
        size_t mid = (lo + hi) / 2;
        if (xd[mid] > x) hi = mid;
        else lo = mid;
    }
    if (lo == xd.size() - 1) return yd[lo];
    return lin2(x, xd[lo], xd[lo + 1], yd[lo], yd[lo + 1]);
}

void add_wing(LineForceModel& lfm, const Value* wb, size_t default_nr_sections) {  // single_wing.rs:73-150, mod.rs:164-205
    check_keys(wb, {"section_points", "chord_vectors", "section_model", "line_segment_is_virtual",
                    "non_zero_circulation_at_ends", "nr_sections", "input_power_model"},
               "WingBuilder");
    std::vector<V3> sp, cv;
    const Value* a = wb->get("section_points");
    if (!a || a->kind != Value::Array) throw std::runtime_error("WingBuilder: section_points");
    for (auto& e : a->arr) sp.push_back(vec3(e.get(), "section_points"));
    a = wb->get("chord_vectors");
    if (!a || a->kind != Value::Array) throw std::runtime_error("WingBuilder: chord_vectors");
    for (auto& e : a->arr) cv.push_back(vec3(e.get(), "chord_vectors"));
    if (sp.size() < 2 || cv.size() != sp.size()) throw std::runtime_error("WingBuilder: sizes");
    Foil foil = parse_section_model(wb->get("section_model"));
    std::vector<bool> seg_virtual;
    if (auto* v = wb->get("line_segment_is_virtual"))
        if (v->kind == Value::Array)
            for (auto& e : v->arr) seg_virtual.push_back(e->kind == Value::Bool && e->b);
    size_t nsec = default_nr_sections;
    if (auto* v = wb->get("nr_sections"))
        if (v->kind == Value::Number) nsec = (size_t)v->num;
    if (auto* v = wb->get("input_power_model")) {
        const Value* p;
        if (enum_tag(v, &p, "InputPowerModel") != "NoPower")
            throw std::runtime_error("InputPowerModel other than NoPower is not supported");
    }
    std::vector<double> dist;
    for (size_t i = 0; i < sp.size(); ++i) {
        if (dist.empty()) dist.push_back(0.0);
        else dist.push_back(dist.back() + length(sp[i] - sp[i - 1]));
    }
    double total = dist.back();
    double ds_avg = total / (double)nsec;
    size_t nseg = sp.size() - 1;
    std::vector<size_t> nloc(nseg);
    std::vector<double> dsloc(nseg);
    for (size_t i = 0; i < nseg; ++i) {
        double segd = dist[i + 1] - dist[i];
        nloc[i] = (size_t)std::ceil(segd / ds_avg);
        dsloc[i] = segd / (double)nloc[i];
    }
    size_t start = lfm.wing_indices.empty() ? 0 : lfm.wing_indices.back().end;
    size_t count = 0;
    for (size_t seg = 0; seg < nseg; ++seg) {
        double ds = dsloc[seg];
        double s0 = dist[seg];
        bool virt = seg_virtual.empty() ? false : (bool)seg_virtual.at(seg);
        for (size_t i = 0; i < nloc[seg]; ++i) {
            double sd = s0 + (double)i * ds;
            double ed = sd + ds;
            double cd = 0.5 * (sd + ed);
            SpanLine L{linear_interpolation(sd, dist, sp), linear_interpolation(ed, dist, sp)};
            V3 ch = linear_interpolation(cd, dist, cv);
            lfm.span_lines_local.push_back(L);
            lfm.span_lines_global.push_back(L);
            lfm.chord_local_not_rotated.push_back(ch);
            lfm.chord_local.push_back(ch);
            lfm.chord_global.push_back(ch);
            lfm.chord_lengths.push_back(length(ch));
            lfm.is_virtual.push_back(virt);
            ++count;
        }
    }
    lfm.wing_indices.push_back({start, start + count});
    lfm.foils.push_back(foil);
    lfm.local_wing_angles.push_back(0.0);
    lfm.update_global();
}

struct Setup {
    LineForceModel lfm;
    DynamicWake wake;
    SimpleIterative solver;
};

LineForceModel build_lfm(const Value* o) {  // line_force_model/builder/mod.rs:79-113
    check_keys(o, {"wing_builders", "nr_sections", "density", "circulation_correction",
                   "angle_of_attack_correction", "output_coordinate_system", "rotation_type",
                   "local_wing_angles", "rotation", "translation", "force_calculation_settings"},
               "LineForceModelBuilder");
    LineForceModel lfm;
    size_t nsec = (size_t)num(o->get("nr_sections"), "nr_sections");
    if (auto* v = o->get("density")) lfm.density = num(v, "density");
    const Value* wbs = o->get("wing_builders");
    if (!wbs || wbs->kind != Value::Array || wbs->arr.empty())
        throw std::runtime_error("LineForceModelBuilder: wing_builders");
    for (auto& wb : wbs->arr) add_wing(lfm, wb.get(), nsec);
    if (auto* v = o->get("circulation_correction")) {
        const Value* p;
        if (enum_tag(v, &p, "CirculationCorrectionBuilder") != "None")
            throw std::runtime_error("circulation corrections are not supported by this path");
    }
    if (auto* v = o->get("angle_of_attack_correction")) {
        const Value* p;
        if (enum_tag(v, &p, "AngleOfAttackCorrection") != "None")
            throw std::runtime_error("angle of attack corrections are not supported");
    }
    if (auto* v = o->get("output_coordinate_system")) {
        const Value* p;
        std::string t = enum_tag(v, &p, "CoordinateSystem");
        if (t == "Global") lfm.output_cs = CoordSys::Global;
        else if (t == "Body") lfm.output_cs = CoordSys::Body;
        else throw std::runtime_error("bad CoordinateSystem");
    }
    if (auto* v = o->get("rotation_type")) {  // parsed but not applied by the reference builder
        const Value* p;
        std::string t = enum_tag(v, &p, "RotationType");
        if (t != "XYZ" && t != "YawPitchRoll") throw std::runtime_error("bad RotationType");
    }
    if (auto* v = o->get("local_wing_angles")) {
        if (v->kind != Value::Array) throw std::runtime_error("local_wing_angles");
        if (!v->arr.empty()) {
            if (v->arr.size() != lfm.nw())
                throw std::runtime_error("The number of local wing angles does not match the number of wings.");
            for (size_t i = 0; i < v->arr.size(); ++i) lfm.local_wing_angles[i] = num(v->arr[i].get(), "angle");
        }
    }
    if (auto* v = o->get("force_calculation_settings")) {
        check_keys(v, {"include_viscous_lift_in_the_circulation"}, "ForceCalculationSettings");
        if (auto* b = v->get("include_viscous_lift_in_the_circulation"))
            lfm.include_viscous_lift_in_circulation = b->kind == Value::Bool && b->b;
    }
    if (auto* v = o->get("translation")) lfm.rbm.translation = vec3(v, "translation");
    if (auto* v = o->get("rotation")) lfm.rbm.rotation = vec3(v, "rotation");
    lfm.update_global();
    return lfm;
}

double core_from(const Value* v, double strip_len) {  // dynamic_wake/builder.rs:214-252
    const Value* p;
    std::string t = enum_tag(v, &p, "ViscousCoreLength");
    if (t == "Relative") return num(p, "Relative") * strip_len;
    if (t == "Absolute") return num(p, "Absolute");
    if (t == "NoViscousCore") return MIN_POSITIVE;
    throw std::runtime_error("bad ViscousCoreLength");
}

void build_wake(DynamicWake& w, const Value* o, const LineForceModel& lfm) {  // builder.rs:121-198
    if (o) check_keys(o, {"nr_panels_per_line_element", "viscous_core_length",
                          "viscous_core_length_evolution", "first_panel_relative_length",
                          "last_panel_relative_length", "first_wake_points_direction",
                          "symmetry_condition", "ratio_of_wake_affected_by_induced_velocities",
                          "far_field_ratio", "shape_damping_factor", "neglect_self_induced_velocities",
                          "initial_relative_wake_length", "write_wake_data_to_file",
                          "wake_files_folder_path"},
                      "DynamicWakeBuilder");
    auto get = [&](const char* k) -> const Value* { return o ? o->get(k) : nullptr; };
    w.M = 100;
    if (auto* v = get("nr_panels_per_line_element")) w.M = (size_t)num(v, "nr_panels");
    if (w.M < 1) throw std::runtime_error("nr_panels_per_line_element must be >= 1");
    w.N = lfm.n();
    w.P = lfm.span_points_global.size();
    w.wing_indices = lfm.wing_indices;
    if (auto* v = get("first_panel_relative_length")) w.first_rel = num(v, "first");
    if (auto* v = get("last_panel_relative_length")) w.last_rel = num(v, "last");
    if (auto* v = get("first_wake_points_direction")) {
        const Value* p;
        std::string t = enum_tag(v, &p, "FirstWakePointsDirection");
        if (t == "Chord") w.first_dir = FirstDir::Chord;
        else if (t == "Freestream") w.first_dir = FirstDir::Freestream;
        else if (t == "ActualVelocity") w.first_dir = FirstDir::ActualVelocity;
        else throw std::runtime_error("bad FirstWakePointsDirection");
    }
    if (auto* v = get("symmetry_condition")) {
        const Value* p;
        std::string t = enum_tag(v, &p, "SymmetryCondition");
        if (t == "NoSymmetry") w.sym = Symmetry::None;
        else if (t == "X") w.sym = Symmetry::X;
        else if (t == "Y") w.sym = Symmetry::Y;
        else if (t == "Z") w.sym = Symmetry::Z;
        else throw std::runtime_error("bad SymmetryCondition");
    }
    double ratio = 0.0;
    if (auto* v = get("ratio_of_wake_affected_by_induced_velocities")) ratio = num(v, "ratio");
    if (auto* v = get("far_field_ratio")) w.far_ratio = num(v, "far");
    if (auto* v = get("shape_damping_factor")) w.damping = num(v, "damping");
    if (auto* v = get("neglect_self_induced_velocities")) w.neglect_self = v->kind == Value::Bool && v->b;
    double init_len = 50.0;
    if (auto* v = get("initial_relative_wake_length")) init_len = num(v, "init");
    w.end_index = (size_t)std::ceil(ratio * (double)(w.nr_points() - 1));
    w.points.assign(w.nr_points(), V3());
    for (size_t s = 0; s <= w.M; ++s)
        for (size_t j = 0; j < w.P; ++j) w.points[w.pidx(s, j)] = lfm.span_points_global[j];
    w.velocity_at_points.assign(w.nr_points(), V3());
    w.strengths.assign(w.nr_panels(), 0.0);
    w.panels.assign(w.nr_panels(), Panel());
    // viscous core length per panel
    std::vector<double> slen(w.N);
    for (size_t i = 0; i < w.N; ++i) slen[i] = lfm.span_lines_global[i].len();
    const Value* cvl = get("viscous_core_length");
    const Value* evo = get("viscous_core_length_evolution");
    const Value* evo_p = nullptr;
    std::string evo_t = "Constant";
    if (evo) evo_t = enum_tag(evo, &evo_p, "ViscousCoreLengthEvolution");
    double evo_factor = 1.0;
    const Value* last_val = nullptr;
    if (evo_t == "SinIncrease") {
        check_keys(evo_p, {"last_panel_value", "evolution_length_factor"}, "SinIncrease");
        last_val = evo_p->get("last_panel_value");
        if (!last_val) throw std::runtime_error("SinIncrease: last_panel_value");
        if (auto* f = evo_p->get("evolution_length_factor")) evo_factor = num(f, "factor");
    } else if (evo_t != "Constant") {
        throw std::runtime_error("bad ViscousCoreLengthEvolution");
    }
    w.core.assign(w.nr_panels(), 0.0);
    orc_json::Value def_core;
    def_core.kind = Value::Object;
    auto rel = std::make_shared<Value>();
    rel->kind = Value::Number;
    rel->num = 0.1;
    def_core.obj.emplace_back("Relative", rel);
    if (!cvl) cvl = &def_core;
    for (size_t s = 0; s < w.M; ++s)
        for (size_t i = 0; i < w.N; ++i) {
            double raw = core_from(cvl, slen[i]);
            double val = raw;
            if (last_val) {
                double fin = core_from(last_val, slen[i]);
                double lb = (double)(w.M - 1) * evo_factor;
                double rs = rmax(rmin((double)s / lb, 1.0), 0.0);
                double si = 0.5 * rs * PI;
                double sf = ORC_SIN(si);
                val = raw * (1.0 - sf) + fin * sf;
            }
            w.core[s * w.N + i] = val;
        }
    double csum = 0.0;
    for (double c : lfm.chord_lengths) csum += c;
    w.rep_chord = csum / (double)lfm.chord_lengths.size();
    w.initialize_based_on_chord(lfm, init_len);
}

VelCorr parse_vc(const Value* x, double& ratio) {  // velocity_corrections.rs:9-16
    const Value* q;
    std::string vt = enum_tag(x, &q, "VelocityCorrections");
    if (vt == "NoCorrection") return VelCorr::None;
    if (vt == "MaxInducedVelocityMagnitudeRatio") {
        ratio = num(q, "ratio");
        return VelCorr::MaxRatio;
    }
    if (vt == "FixedMagnitudeEqualToFreestream") return VelCorr::FixedMagnitude;
    throw std::runtime_error("bad VelocityCorrections");
}

struct SolverSpec {  // solvers/mod.rs:21-55
    bool linearized = false;
    SimpleIterative si;
    Linearized lin;
};

Linearized parse_linearized(const Value* p) {  // linearized.rs:26-48
    Linearized l;
    if (!p) return l;
    check_keys(p, {"velocity_corrections", "disable_viscous_corrections", "induced_velocity_correction_method"},
               "Linearized");
    if (auto* x = p->get("velocity_corrections")) l.vc = parse_vc(x, l.vc_ratio);
    if (auto* x = p->get("disable_viscous_corrections")) l.disable_viscous = x->kind == Value::Bool && x->b;
    if (auto* x = p->get("induced_velocity_correction_method")) {
        const Value* q;
        std::string t = enum_tag(x, &q, "InducedVelocityCorrectionMethod");
        if (t == "NoCorrection") l.ivc = IvcMethod::NoCorrection;
        else if (t == "FullCorrection") l.ivc = IvcMethod::FullCorrection;
        else throw std::runtime_error("bad InducedVelocityCorrectionMethod");
    }
    return l;
}

// Solver (dynamic settings, simple_iterative.rs:59-84) or QuasiSteadySolverBuilder
// (quasi_steady: simple_iterative.rs:14-45 with 1000 iterations and damping 0.04;
// default Linearized, solvers/mod.rs:31-36).
SolverSpec parse_solver(const Value* v, bool quasi) {
    SolverSpec spec;
    if (quasi) {
        spec.si.max_iterations = 1000;
        spec.si.damping = 0.04;
        if (!v) {
            spec.linearized = true;
            return spec;
        }
    }
    if (!v) return spec;
    const Value* p;
    std::string t = enum_tag(v, &p, "Solver");
    if (t == "Linearized") {
        spec.linearized = true;
        spec.lin = parse_linearized(p);
        return spec;
    }
    if (t != "SimpleIterative") throw std::runtime_error("Solver `" + t + "` is not supported by this path");
    SimpleIterative& s = spec.si;
    if (quasi)
        check_keys(p, {"max_iterations_per_time_step", "damping_factor", "residual_tolerance_absolute",
                       "strength_difference_tolerance", "velocity_corrections", "start_with_linearized_solution",
                       "correction_factor_initial_solution"},
                   "QuasiSteadySimpleIterativeBuilder");
    else
        check_keys(p, {"max_iterations_per_time_step", "damping_factor", "residual_tolerance_absolute",
                       "strength_difference_tolerance", "velocity_corrections",
                       "use_raw_circulation_during_iterations", "start_with_linearized_solution",
                       "correction_factor_initial_solution"},
                   "SimpleIterative");
    if (auto* x = p->get("max_iterations_per_time_step")) s.max_iterations = (size_t)num(x, "max_it");
    if (auto* x = p->get("damping_factor")) s.damping = num(x, "damping");
    if (auto* x = p->get("residual_tolerance_absolute")) s.residual_tol = num(x, "rtol");
    if (auto* x = p->get("strength_difference_tolerance")) s.strength_tol = num(x, "stol");
    if (auto* x = p->get("use_raw_circulation_during_iterations")) s.use_raw = x->kind == Value::Bool && x->b;
    if (auto* x = p->get("start_with_linearized_solution")) s.start_linearized = x->kind == Value::Bool && x->b;
    if (auto* x = p->get("correction_factor_initial_solution")) s.correction_factor_initial = num(x, "cf");
    if (auto* x = p->get("velocity_corrections")) s.vc = parse_vc(x, s.vc_ratio);
    return spec;
}

QuasiSteadyWake parse_quasi_wake(const Value* o) {  // wake/settings.rs:33-58
    QuasiSteadyWake w;
    if (!o) return w;
    check_keys(o, {"wake_length_factor", "symmetry_condition", "viscous_core_length"}, "QuasiSteadyWakeSettings");
    if (auto* x = o->get("wake_length_factor")) w.wake_length_factor = num(x, "wake_length_factor");
    if (auto* x = o->get("symmetry_condition")) {
        const Value* q;
        std::string t = enum_tag(x, &q, "SymmetryCondition");
        if (t == "NoSymmetry") w.sym = Symmetry::None;
        else if (t == "X") w.sym = Symmetry::X;
        else if (t == "Y") w.sym = Symmetry::Y;
        else if (t == "Z") w.sym = Symmetry::Z;
        else throw std::runtime_error("bad SymmetryCondition");
    }
    if (auto* x = o->get("viscous_core_length")) {
        const Value* q;
        std::string t = enum_tag(x, &q, "ViscousCoreLength");
        if (t == "Relative") { w.core_kind = 0; w.core_value = num(q, "Relative"); }
        else if (t == "Absolute") { w.core_kind = 1; w.core_value = num(q, "Absolute"); }
        else if (t == "NoViscousCore") { w.core_kind = 2; w.core_value = 0.0; }
        else throw std::runtime_error("bad ViscousCoreLength");
    }
    return w;
}

}  // namespace

// ---------------------------------------------------------------------------
// Simulation (stormbird/src/lifting_line/simulation.rs:20-216)
// ---------------------------------------------------------------------------
struct orc_sim {
    LineForceModel lfm;
    FlowDerivatives fd;
    DynamicWake wake;
    FrozenWake frozen;
    SolverSpec solver;
    bool quasi = false;      // WakeData::QuasiSteady
    QuasiSteadyWake qs;
    std::vector<double> prev_gamma;
    bool first_done = false;

    std::vector<V3> freestream_points() const {  // simulation.rs:45-63
        std::vector<V3> p = lfm.ctrl_points_global;
        if (!quasi) p.insert(p.end(), wake.points.begin(), wake.points.end());
        return p;
    }
    void initialize(const V3* ctrl_fs, double dt) {  // simulation.rs:65-94
        if (!quasi) {
            V3 sum;
            for (size_t i = 0; i < lfm.n(); ++i) sum = sum + ctrl_fs[i];
            V3 avg = sum / (double)lfm.n();
            for (auto& v : wake.velocity_at_points) v = avg;
            wake.initialize_with_velocity(lfm, avg, dt);
        }
        fd.reset(std::vector<V3>(ctrl_fs, ctrl_fs + lfm.n()));
        first_done = true;
        prev_gamma.assign(lfm.n(), 0.0);
    }
    void do_step(double time, double dt, const V3* fs, size_t nfs, orc_step_result* out) {  // :107-203
        (void)nfs;
        lfm.update_global();
        size_t n = lfm.n();
        std::vector<V3> felt = lfm.felt_ctrl_velocity(fs);
        std::vector<V3> felt_span = lfm.span_point_values(felt);
        if (!first_done) initialize(fs, dt);
        if (quasi) {
            update_as_steady(frozen, lfm, felt_span, qs);
        } else {
            wake.update_before_solving(dt, lfm, felt_span);
            frozen.update_from_full_wake(lfm.ctrl_points_global, wake);
        }
        SolverResult sr = solver.linearized ? solve_linearized(solver.lin, lfm, felt, frozen)
                                            : solve(solver.si, lfm, felt, frozen, prev_gamma);
        if (!quasi) wake.update_after_solving(sr.gamma, fs + n);
        prev_gamma = sr.gamma;
        std::vector<V3> acc = fd.acceleration(sr.output_velocity, dt);
        fd.update(sr.output_velocity);
        write_result(time, sr, acc, out);
    }
    void write_result(double time, const SolverResult& sr, const std::vector<V3>& acc_in,
                      orc_step_result* out) const {  // force_calculations.rs:322-380, 669-698
        size_t n = lfm.n(), nw = lfm.nw();
        std::vector<double> alpha = lfm.angles_of_attack(sr.output_velocity);
        std::vector<V3> vel = sr.output_velocity, acc = acc_in;
        V3 rot_vel = lfm.rbm.velocity_angular;
        if (lfm.output_cs == CoordSys::Body) {
            for (size_t i = 0; i < n; ++i) {
                vel[i] = lfm.rbm.body_fixed(vel[i]);
                acc[i] = lfm.rbm.body_fixed(acc[i]);
            }
            rot_vel = lfm.rbm.body_fixed(rot_vel);
        }
        SectionalForces F;
        F.cs = lfm.output_cs;
        F.k[0] = lfm.circulatory_forces(sr.gamma, vel);
        for (int k = 1; k < 6; ++k) F.k[k].assign(n, V3());
        if (!lfm.include_viscous_lift_in_circulation) {  // :413-433
            for (size_t i = 0; i < n; ++i) {
                double clv = lfm.foils[lfm.wing_of(i)].cl_post_weighted(alpha[i]);
                V3 dir = normalize(cross(lfm.span_lines_local[i].rel(), vel[i]));
                double area = lfm.chord_lengths[i] * lfm.span_lines_local[i].len();
                double ff = 0.5 * area * lfm.density * length_squared(vel[i]);
                F.k[1][i] = dir * clv * ff;
            }
        }
        for (size_t i = 0; i < n; ++i) {  // :438-452
            double cd = lfm.foils[lfm.wing_of(i)].cd(alpha[i]);
            V3 dir = normalize(vel[i]);
            double area = lfm.chord_lengths[i] * lfm.span_lines_local[i].len();
            double ff = 0.5 * area * lfm.density * length_squared(vel[i]);
            F.k[2][i] = dir * cd * ff;
        }
        {  // added mass :467-513
            const auto& sl = lfm.lines(lfm.output_cs);
            const auto& ch = lfm.output_cs == CoordSys::Global ? lfm.chord_global : lfm.chord_local;
            for (size_t i = 0; i < n; ++i) {
                const Foil& fo = lfm.foils[lfm.wing_of(i)];
                double area = lfm.chord_lengths[i] * sl[i].len();
                V3 ra = acc[i];
                ra -= project(ra, sl[i].dir());
                ra -= project(ra, ch[i]);
                if (length(ra) == 0.0) F.k[3][i] = V3();
                else {
                    double c = fo.added_mass_coefficient(length(ra));
                    F.k[3][i] = (c * lfm.density * area) * normalize(ra);
                }
            }
        }
        // gyroscopic: zero for foils (:520-548)
        for (size_t i = 0; i < n; ++i) {  // forces_and_moments.rs:163-172
            V3 t = F.k[0][i];
            t += F.k[1][i];
            t += F.k[2][i];
            t += F.k[3][i];
            t += F.k[4][i];
            F.k[5][i] = t;
        }
        if (!out) return;
        out->time = time;
        out->iterations = (long)sr.iterations;
        out->residual = sr.residual;
        auto put = [](double* dst, const std::vector<V3>& v) {
            if (!dst) return;
            for (size_t i = 0; i < v.size(); ++i)
                for (int c = 0; c < 3; ++c) dst[3 * i + c] = v[i][c];
        };
        if (out->circulation_strength)
            for (size_t i = 0; i < n; ++i) out->circulation_strength[i] = sr.gamma[i];
        if (out->angles_of_attack)
            for (size_t i = 0; i < n; ++i) out->angles_of_attack[i] = alpha[i];
        put(out->ctrl_points_velocity, vel);
        put(out->solver_input_velocity, sr.input_velocity);
        put(out->ctrl_points, lfm.ctrl_points_global);
        if (out->sectional_forces)
            for (int k = 0; k < 6; ++k) put(out->sectional_forces + (size_t)k * 3 * n, F.k[k]);
        // integrate forces / moments (forces_and_moments.rs:177-245)
        const auto& msl = lfm.lines(F.cs);
        for (int k = 0; k < 6; ++k) {
            for (size_t w = 0; w < nw; ++w) {
                V3 fs, ms;
                for (size_t i = lfm.wing_indices[w].start; i < lfm.wing_indices[w].end; ++i) {
                    if (lfm.is_virtual[i]) continue;
                    fs += F.k[k][i];
                    ms += cross(msl[i].ctrl(), F.k[k][i]);
                }
                for (int c = 0; c < 3; ++c) {
                    if (out->integrated_forces) out->integrated_forces[(size_t)k * 3 * nw + 3 * w + c] = fs[c];
                    if (out->integrated_moments) out->integrated_moments[(size_t)k * 3 * nw + 3 * w + c] = ms[c];
                }
            }
        }
    }
};

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_solve_gaussian_elimination(const double* A, const double* b, size_t n, double* x) {
    try {
        std::vector<double> a(A, A + n * n), r(b, b + n);
        std::vector<double> s = solve_gaussian_elimination(a, r, n);
        for (size_t i = 0; i < n; ++i) x[i] = s[i];
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void orc_horseshoe_unit_velocity(const double* start, const double* end, const double* wake_start,
                                 const double* wake_end, const double* x, double core, double* out) {
    V3 a(start[0], start[1], start[2]), b(end[0], end[1], end[2]);
    V3 ws(wake_start[0], wake_start[1], wake_start[2]), we(wake_end[0], wake_end[1], wake_end[2]);
    Horseshoe h{a, b, a + ws, a, b, b + we, core};
    V3 u = h.unit(V3(x[0], x[1], x[2]));
    for (int c = 0; c < 3; ++c) out[c] = u[c];
}

orc_sim* orc_sim_create(const char* json) {
    if (!json) { g_err = "null setup string"; return nullptr; }
    try {
        auto root = orc_json::parse(json);
        check_keys(root.get(), {"line_force_model", "simulation_settings"}, "SimulationBuilder");
        auto sim = std::make_unique<orc_sim>();
        const Value* lfmv = root->get("line_force_model");
        if (!lfmv) throw std::runtime_error("missing field `line_force_model`");
        sim->lfm = build_lfm(lfmv);
        const Value* ss = root->get("simulation_settings");  // simulation_builder.rs:21-49 (default QuasiSteady)
        const Value* p = nullptr;
        std::string t = ss ? enum_tag(ss, &p, "SimulationSettings") : std::string("QuasiSteady");
        if (t == "QuasiSteady") {
            if (p) check_keys(p, {"solver", "wake"}, "QuasiSteadySettings");
            sim->quasi = true;
            sim->solver = parse_solver(p ? p->get("solver") : nullptr, true);
            sim->qs = parse_quasi_wake(p ? p->get("wake") : nullptr);
        } else if (t == "Dynamic") {
            if (p) check_keys(p, {"solver", "wake"}, "DynamicSettings");
            sim->solver = parse_solver(p ? p->get("solver") : nullptr, false);
            build_wake(sim->wake, p ? p->get("wake") : nullptr, sim->lfm);
        } else {
            throw std::runtime_error("SimulationSettings `" + t + "` is not supported by this path");
        }
        sim->frozen.init(sim->lfm.n());
        sim->prev_gamma.assign(sim->lfm.n(), 0.0);
        sim->fd.reset(std::vector<V3>(sim->lfm.n(), V3()));
        return sim.release();
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_sim_destroy(orc_sim* s) { delete s; }
orc_sim* orc_sim_clone(const orc_sim* s) { return s ? new orc_sim(*s) : nullptr; }

long orc_sim_nr_ctrl_points(const orc_sim* s) { return s ? (long)s->lfm.n() : -1; }
long orc_sim_nr_wings(const orc_sim* s) { return s ? (long)s->lfm.nw() : -1; }
long orc_sim_nr_span_points(const orc_sim* s) { return s ? (long)s->wake.P : -1; }
long orc_sim_nr_wake_points(const orc_sim* s) { return s ? (long)s->wake.nr_points() : -1; }
long orc_sim_nr_panels(const orc_sim* s) { return s ? (long)s->wake.nr_panels() : -1; }
long orc_sim_nr_freestream_points(const orc_sim* s) {
    return s ? (long)(s->lfm.n() + s->wake.nr_points()) : -1;
}
long orc_sim_get_freestream_points(const orc_sim* s, double* xyz, size_t cap) {
    if (!s || !xyz) return -1;
    if (cap == 0) return -2;
    auto p = s->freestream_points();
    size_t n = std::min(cap, p.size());
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) xyz[3 * i + c] = p[i][c];
    return (long)n;
}

int orc_sim_do_step(orc_sim* s, double time, double dt, const double* fs, size_t n,
                    orc_step_result* out) {
    if (!s || !fs) return -1;
    if (n != s->lfm.n() + s->wake.nr_points()) { g_err = "freestream length mismatch"; return -3; }
    try {
        std::vector<V3> v(n);
        for (size_t i = 0; i < n; ++i) v[i] = V3(fs[3 * i], fs[3 * i + 1], fs[3 * i + 2]);
        s->do_step(time, dt, v.data(), n, out);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -4;
    }
}

int orc_sim_induced_velocities(const orc_sim* s, const double* pts, size_t n, double* out) {
    if (!s || (!pts && n) || (!out && n)) return -1;
    std::vector<V3> p(n);
    for (size_t i = 0; i < n; ++i) p[i] = V3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
    auto u = s->wake.induced_velocities(p);
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) out[3 * i + c] = u[i][c];
    return 0;
}

int orc_sim_get_previous_circulation(const orc_sim* s, double* out, size_t n) {
    if (!s || !out) return -1;
    if (n != s->prev_gamma.size()) return -3;
    for (size_t i = 0; i < n; ++i) out[i] = s->prev_gamma[i];
    return 0;
}
int orc_sim_set_previous_circulation(orc_sim* s, const double* g, size_t n) {
    if (!s || !g) return -1;
    if (n != s->prev_gamma.size()) return -3;
    for (size_t i = 0; i < n; ++i) s->prev_gamma[i] = g[i];
    return 0;
}
int orc_sim_set_translation_and_rotation_fd(orc_sim* s, double dt, const double* t, const double* r) {
    if (!s || !t || !r) return -1;
    s->lfm.rbm.update_translation_fd(V3(t[0], t[1], t[2]), dt);  // data_update.rs:188-205
    s->lfm.rbm.update_rotation_fd(V3(r[0], r[1], r[2]), dt);
    s->lfm.update_global();
    return 0;
}
int orc_sim_set_local_wing_angles(orc_sim* s, const double* a, size_t n) {
    if (!s || !a) return -1;
    if (n != s->lfm.nw()) return -3;
    for (size_t i = 0; i < n; ++i) s->lfm.local_wing_angles[i] = a[i];
    s->lfm.update_global();
    return 0;
}

int orc_sim_get_line_model(const orc_sim* s, double* span_lines, double* chord_global,
                           double* chord_local_not_rotated, long* wing_ranges, double* wing_angles,
                           double* rigid_body) {
    if (!s) return -1;
    const auto& m = s->lfm;
    for (size_t i = 0; i < m.n(); ++i)
        for (int c = 0; c < 3; ++c) {
            if (span_lines) {
                span_lines[6 * i + c] = m.span_lines_global[i].start[c];
                span_lines[6 * i + 3 + c] = m.span_lines_global[i].end[c];
            }
            if (chord_global) chord_global[3 * i + c] = m.chord_global[i][c];
            if (chord_local_not_rotated) chord_local_not_rotated[3 * i + c] = m.chord_local_not_rotated[i][c];
        }
    for (size_t w = 0; w < m.nw(); ++w) {
        if (wing_ranges) {
            wing_ranges[2 * w] = (long)m.wing_indices[w].start;
            wing_ranges[2 * w + 1] = (long)m.wing_indices[w].end;
        }
        if (wing_angles) wing_angles[w] = m.local_wing_angles[w];
    }
    if (rigid_body)
        for (int c = 0; c < 3; ++c) {
            rigid_body[c] = m.rbm.translation[c];
            rigid_body[3 + c] = m.rbm.rotation[c];
            rigid_body[6 + c] = m.rbm.velocity_linear[c];
            rigid_body[9 + c] = m.rbm.velocity_angular[c];
        }
    return 0;
}

int orc_sim_restart(orc_sim* s) {  // complete_sail_model/mod.rs:152
    if (!s) return -1;
    s->first_done = false;
    return 0;
}

int orc_sim_get_wake_state(const orc_sim* s, double* pts, double* vel, double* str, long* steps,
                           int* first) {
    if (!s) return -1;
    const auto& w = s->wake;
    for (size_t i = 0; i < w.nr_points(); ++i)
        for (int c = 0; c < 3; ++c) {
            if (pts) pts[3 * i + c] = w.points[i][c];
            if (vel) vel[3 * i + c] = w.velocity_at_points[i][c];
        }
    if (str)
        for (size_t k = 0; k < w.nr_panels(); ++k) str[k] = w.strengths[k];
    if (steps) *steps = (long)w.steps_completed;
    if (first) *first = s->first_done ? 1 : 0;
    return 0;
}
int orc_sim_set_wake_state(orc_sim* s, const double* pts, const double* vel, const double* str,
                           long steps, int first) {
    if (!s) return -1;
    auto& w = s->wake;
    for (size_t i = 0; i < w.nr_points(); ++i)
        for (int c = 0; c < 3; ++c) {
            if (pts) w.points[i][c] = pts[3 * i + c];
            if (vel) w.velocity_at_points[i][c] = vel[3 * i + c];
        }
    if (str)
        for (size_t k = 0; k < w.nr_panels(); ++k) w.strengths[k] = str[k];
    w.steps_completed = (size_t)steps;
    if (first && !s->first_done) {
        s->first_done = true;
        s->prev_gamma.assign(s->lfm.n(), 0.0);
    }
    w.update_panel_data();
    return 0;
}
int orc_sim_update_panels(orc_sim* s) {
    if (!s) return -1;
    s->wake.update_panel_data();
    return 0;
}
int orc_sim_get_panels(const orc_sim* s, double* rec) {
    if (!s || !rec) return -1;
    const auto& w = s->wake;
    for (size_t k = 0; k < w.nr_panels(); ++k) {
        const Panel& P = w.panels[k];
        double* r = rec + 22 * k;
        for (int q = 0; q < 4; ++q)
            for (int c = 0; c < 3; ++c) r[3 * q + c] = P.points[q][c];
        for (int c = 0; c < 3; ++c) { r[12 + c] = P.center[c]; r[15 + c] = P.normal[c]; }
        r[18] = P.far_field_length_squared;
        r[19] = P.viscous_core_length;
        r[20] = P.area_term;
        r[21] = w.strengths[k];
    }
    return 0;
}
int orc_sim_get_frozen_wake(const orc_sim* s, double* fixed, double* factors) {
    if (!s) return -1;
    size_t n = s->frozen.N;
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c)
            if (fixed) fixed[3 * i + c] = s->frozen.fixed[i][c];
    if (factors)
        for (size_t k = 0; k < n * n; ++k)
            for (int c = 0; c < 3; ++c) factors[3 * k + c] = s->frozen.factors[k][c];
    return 0;
}

void orc_segment_unit_velocity(const double* p0, const double* p1, const double* x, double core,
                               double* out) {
    V3 u = segment_unit_velocity(V3(p0[0], p0[1], p0[2]), V3(p1[0], p1[1], p1[2]),
                                 V3(x[0], x[1], x[2]), core);
    for (int c = 0; c < 3; ++c) out[c] = u[c];
}
void orc_panel_unit_velocity(const double* c12, double far_ratio, double core, const double* x,
                             int mode, double* out) {
    V3 pts[4];
    for (int q = 0; q < 4; ++q) pts[q] = V3(c12[3 * q], c12[3 * q + 1], c12[3 * q + 2]);
    Panel P = Panel::make(pts, far_ratio, core);
    V3 xx(x[0], x[1], x[2]);
    V3 u = mode == 0 ? P.unit_velocity(xx) : (mode == 1 ? P.rings(xx) : P.doublet(xx));
    for (int c = 0; c < 3; ++c) out[c] = u[c];
}
double orc_foil_cl(const char* foil_json, double alpha, int which) {
    try {
        auto v = orc_json::parse(foil_json);
        Foil f = parse_foil(v.get());
        switch (which) {
            case 0: return f.cl(alpha);
            case 1: return f.cl_pre_dropoff(alpha);
            case 2: return f.cd(alpha);
            case 3: return f.amount_of_stall(alpha);
            default: return f.cl_post_weighted(alpha);
        }
    } catch (const std::exception& e) {
        g_err = e.what();
        return std::nan("");
    }
}
double orc_sigmoid(double x, double x0, double r) { return sigmoid_zero_to_one(x, x0, r); }

// stormath transformations (spatial_vector/transformations.rs) and rigid-body motion
// (rigid_body_motion.rs:22-103); rotation_type 0 XYZ, 1 YawPitchRoll.
static V3 v3(const double* p) { return V3(p[0], p[1], p[2]); }
static void put3(double* out, const V3& v) {
    for (int c = 0; c < 3; ++c) out[c] = v[c];
}
static RotationType rtype(int t) { return t == 1 ? RotationType::YawPitchRoll : RotationType::XYZ; }
static RigidBodyMotion motion_of(const double* m12, int t) {
    RigidBodyMotion m;
    m.translation = v3(m12);
    m.rotation = v3(m12 + 3);
    m.velocity_linear = v3(m12 + 6);
    m.velocity_angular = v3(m12 + 9);
    m.rotation_type = rtype(t);
    return m;
}
void orc_rotate(const double* v, const double* rotation, int rotation_type, double* out) {
    put3(out, rotate(v3(v), v3(rotation), rtype(rotation_type)));
}
void orc_rotate_around_axis(const double* v, double angle, const double* axis, double* out) {
    put3(out, rotate_around_axis(v3(v), angle, v3(axis)));
}
void orc_in_rotated_coordinate_system(const double* v, const double* rotation, int rotation_type, double* out) {
    put3(out, in_rotated_coordinate_system(v3(v), v3(rotation), rtype(rotation_type)));
}
void orc_moment_in_new_coordinate_system(const double* moment, const double* rotation, const double* translation,
                                         const double* force, int rotation_type, double* out) {
    put3(out, moment_in_new_coordinate_system(v3(moment), v3(rotation), v3(translation), v3(force),
                                              rtype(rotation_type)));
}
void orc_rigid_body_transform_point(const double* motion12, int rotation_type, const double* p, double* out) {
    put3(out, motion_of(motion12, rotation_type).transform_point(v3(p)));
}
void orc_rigid_body_velocity_at_point(const double* motion12, int rotation_type, const double* p, double* out) {
    put3(out, motion_of(motion12, rotation_type).velocity_at_point(v3(p)));
}
void orc_rigid_body_update_fd(double* motion12, const double* translation, const double* rotation,
                              double time_step) {
    RigidBodyMotion m = motion_of(motion12, 0);
    m.update_translation_fd(v3(translation), time_step);
    m.update_rotation_fd(v3(rotation), time_step);
    put3(motion12, m.translation);
    put3(motion12 + 3, m.rotation);
    put3(motion12 + 6, m.velocity_linear);
    put3(motion12 + 9, m.velocity_angular);
}
// FrozenWake::update_induced_velocities_at_control_points (frozen_wake.rs:268-301) on given arrays.
void orc_frozen_matvec(const double* fixed, const double* factors, const double* strengths, size_t n,
                       double* out) {
    FrozenWake fw;
    fw.init(n);
    for (size_t r = 0; r < n; ++r) fw.fixed[r] = v3(fixed + 3 * r);
    for (size_t k = 0; k < n * n; ++k) fw.factors[k] = v3(factors + 3 * k);
    fw.matvec(std::vector<double>(strengths, strengths + n));
    for (size_t r = 0; r < n; ++r) put3(out + 3 * r, fw.induced[r]);
}
// VelocityCorrections::corrected_velocity (velocity_corrections.rs:28-100); kind 1
// MaxInducedVelocityMagnitudeRatio(ratio), 2 FixedMagnitudeEqualToFreestream, else freestream + induced.
void orc_corrected_velocity(int kind, double ratio, const double* fs, const double* ui, size_t n, double* out) {
    std::vector<V3> f(n), u(n);
    for (size_t i = 0; i < n; ++i) {
        f[i] = v3(fs + 3 * i);
        u[i] = v3(ui + 3 * i);
    }
    std::vector<V3> r(n);
    if (kind == 1) r = corrected_velocity(VelCorr::MaxRatio, ratio, f, u);
    else if (kind == 2) r = corrected_velocity(VelCorr::FixedMagnitude, ratio, f, u);
    else
        for (size_t i = 0; i < n; ++i) r[i] = f[i] + u[i];
    for (size_t i = 0; i < n; ++i) put3(out + 3 * i, r[i]);
}

int orc_set_threads(int n) {  // OpenMP threads of the target loop (0: the runtime default)
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int orc_libm_variant(void) {
#ifdef ORC_PORTABLE_LIBM
    return 1;
#else
    return 0;
#endif
}
// fn: 0 sin, 1 cos, 2 exp, 3 acos, 4 pow(x, y), 5 log
void orc_libm_eval(int fn, const double* x, const double* y, double* out, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        switch (fn) {
            case 0: out[i] = ORC_SIN(x[i]); break;
            case 1: out[i] = ORC_COS(x[i]); break;
            case 2: out[i] = ORC_EXP(x[i]); break;
            case 3: out[i] = ORC_ACOS(x[i]); break;
            case 4: out[i] = ORC_POW(x[i], y[i]); break;
            case 6: out[i] = ORC_ATAN(x[i]); break;
            default: out[i] = ORC_LOG(x[i]); break;
        }
    }
}

}  // extern "C"
