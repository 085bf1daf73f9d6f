// Host/device 3-vector arithmetic and device data layout of the lifting-line
// step.
//
// All vector operations evaluate in the order Stormbird's SpatialVector does
// (stormath/src/spatial_vector/{operators,vector_math}.rs): component-wise
// +,-,*,/ ; dot = (a0*b0 + a1*b1) + a2*b2 ; Float * vector multiplies each
// component by the scalar.  Every translation unit is compiled with
// -ffp-contract=off, so no multiply-add is fused unless written as fma().
#pragma once

#include <stdint.h>

#include "sb_libm.h"

#if defined(__HIPCC__) || defined(__HIP__)
#define SB_DEV __device__ inline
#else
#define SB_DEV inline
#endif

namespace sb {

constexpr double PI = 3.14159265358979323846;
constexpr double TAU = 6.28318530717958647692;
constexpr double MIN_POSITIVE = 2.2250738585072014e-308;  // f64::MIN_POSITIVE
constexpr double FOUR_PI_INVERSE = 1.0 / (4.0 * PI);
constexpr double CLOSENESS_ERROR = 1.0e-10;

struct V3 {
    double x, y, z;
};
SB_HD V3 v3(double x, double y, double z) { return V3{x, y, z}; }
SB_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
SB_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
SB_HD V3 operator*(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }
SB_HD V3 operator*(double s, V3 a) { return V3{a.x * s, a.y * s, a.z * s}; }
SB_HD V3 operator/(V3 a, double s) { return V3{a.x / s, a.y / s, a.z / s}; }
SB_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
SB_HD double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SB_HD double len2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
SB_HD double len(V3 a) { return sbm::sqrt_(len2(a)); }
SB_HD V3 normalize(V3 a) {
    double l = len(a);
    if (l > 0.0) return V3{a.x / l, a.y / l, a.z / l};
    return V3{0.0, 0.0, 0.0};
}
SB_HD V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
SB_HD double fmax_(double a, double b) { return __builtin_fmax(a, b); }
SB_HD double fmin_(double a, double b) { return __builtin_fmin(a, b); }
// Rust f64::signum: +1 for +0.0 and positive values, -1 for -0.0 and negative values.
SB_HD double signum(double x) {
    if (x != x) return x;
    return __builtin_signbit(x) ? -1.0 : 1.0;
}
SB_HD V3 project(V3 a, V3 onto) {
    V3 n = normalize(onto);
    return n * dot(a, n);
}
SB_HD double absolute_angle_between(V3 a, V3 b) {
    double sa = len2(a), sb2 = len2(b);
    if (sa == 0.0 || sb2 == 0.0) return 0.0;
    double c = dot(a, b) / sbm::sqrt_(sa * sb2);
    c = fmin_(fmax_(c, -1.0), 1.0);
    return sbm::acos(c);
}
SB_HD double signed_angle_between(V3 a, V3 b, V3 axis) {
    double triple = dot(a, cross(b, axis));
    double ang = absolute_angle_between(a, b);
    return triple > 0.0 ? ang : -ang;
}
SB_HD V3 load3(const double* p) { return V3{p[0], p[1], p[2]}; }
SB_HD void store3(double* p, V3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}

// Foil section-model parameters (section_models/foil.rs:43-117), one per wing.
struct FoilParams {
    double cl_zero_angle, cl_initial_slope, cl_high_order_factor_positive,
        cl_high_order_factor_negative, cl_high_order_power, cl_max_after_stall, cd_min,
        angle_cd_min, cd_second_order_factor, cd_max_after_stall, cd_power_after_stall,
        cdi_correction_factor, mean_positive_stall_angle, mean_negative_stall_angle, stall_range,
        cd_bump_during_stall, cd_stall_angle_offset, added_mass_factor;
};

// transition_functions.rs:20-28
SB_HD double sigmoid_zero_to_one(double x, double x0, double range) {
    double slope = 4.5951212 / range;
    double xp = slope * (x - x0);
    return 1.0 / (1.0 + sbm::exp(-xp));
}
// foil.rs:119-129
SB_HD double foil_stall_angle(double a) {
    double e = sbm::fabs_(a);
    while (e > PI) e -= PI;
    e *= signum(a);
    return e;
}
SB_HD double foil_amount_of_stall(const FoilParams& f, double a) {  // foil.rs:256-270
    double e = a + f.angle_cd_min;
    double m = e >= 0.0 ? sbm::fabs_(f.mean_positive_stall_angle) : sbm::fabs_(f.mean_negative_stall_angle);
    return sigmoid_zero_to_one(sbm::fabs_(e), m, f.stall_range);
}
SB_HD double foil_cl_linear(const FoilParams& f, double a) {  // foil.rs:161-163
    return f.cl_zero_angle + f.cl_initial_slope * sbm::sin(a);
}
SB_HD double foil_cl_pre_raw(const FoilParams& f, double a) {  // foil.rs:161-181
    double hp = f.cl_high_order_power > 0.0 ? sbm::pow(sbm::fabs_(a), f.cl_high_order_power) * signum(a) : 0.0;
    double k = a >= 0.0 ? f.cl_high_order_factor_positive : f.cl_high_order_factor_negative;
    double lin = f.cl_zero_angle + f.cl_initial_slope * sbm::sin(a);
    return lin + k * hp;
}
SB_HD double foil_cl_pre_dropoff(const FoilParams& f, double a) {  // foil.rs:184-190
    double s = foil_amount_of_stall(f, a);
    double raw = foil_cl_pre_raw(f, a);
    return raw * (1.0 - s);
}
SB_HD double foil_cl_post_weighted(const FoilParams& f, double a) {  // foil.rs:193-206
    double raw = f.cl_max_after_stall * sbm::sin(2.0 * foil_stall_angle(a));
    double s = foil_amount_of_stall(f, a);
    return raw * s;
}
SB_HD double foil_cl(const FoilParams& f, double a) {  // foil.rs:152-158
    double pre = foil_cl_pre_dropoff(f, a);
    double post = foil_cl_post_weighted(f, a);
    return post + pre;
}
SB_HD double foil_cd(const FoilParams& f, double a) {  // foil.rs:212-243
    double sa = foil_stall_angle(a);
    double pe = sbm::fabs_(a + f.angle_cd_min);
    double pre = f.cd_min + f.cd_second_order_factor * (pe * pe);
    double post = f.cd_max_after_stall * sbm::pow(sbm::fabs_(sbm::sin(sa)), f.cd_power_after_stall);
    double at = a + f.cd_stall_angle_offset * signum(a);
    double s = foil_amount_of_stall(f, at);
    double raw = pre * (1.0 - s) + s * post;
    double s2 = foil_amount_of_stall(f, at);
    double during = f.cd_bump_during_stall * ((1.0 - sbm::cos(s2 * TAU)) * 0.5);
    if (f.cdi_correction_factor != 0.0) {
        double c = sbm::fabs_(foil_cl(f, a));
        double corr = f.cdi_correction_factor * (c * c);
        return raw + corr + during;
    }
    return raw + during;
}

// Record of one wake panel, the unit the pair kernels stream through LDS.
// (panel.rs:16-23; the corners live in the wake point array and are gathered
// only for near pairs.)
struct PanelRec {
    double cx, cy, cz;   // center
    double nx, ny, nz;   // unit normal
    double far2;         // (max(|p1-p0|, |p2-p1|) * far_field_ratio)^2
    double area_term;    // area / (4 pi)
    double strength;     // circulation of the panel
    double cbase;        // flat index of corner 0 in the wake point array (s*P + i + wing), or -1
};
static_assert(sizeof(PanelRec) == 80, "PanelRec layout");

}  // namespace sb
