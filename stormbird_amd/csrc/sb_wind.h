// Wind field of one wind condition at one instant, evaluated identically on the
// host and on the device (stormbird/src/wind/environment.rs:60-190,
// wind/wind_condition/mod.rs:33-61, velocity_variation/*.rs).
//
// The host resolves everything that does not depend on the location (the
// direction vectors, the three gust signals at this time, the ship's linear
// velocity); a location then costs one dot product, one velocity-profile
// evaluation and three scaled vectors.
#pragma once

#include "sb_common.h"

namespace sb {

enum VelocityProfile { VP_CONSTANT = 0, VP_POWER = 1, VP_LOGARITHMIC = 2 };

struct WindField {
    V3 up;                 // WindEnvironment::up_direction
    V3 parallel_dir;       // zero direction rotated by direction_coming_from about the wind axis
    V3 perpendicular_dir;  // up x parallel_dir
    V3 linear;             // ship velocity * zero direction
    double water_plane_height;
    int profile;           // VelocityProfile
    double p0, p1, p2;     // constant: p0 | power: reference velocity, height, factor |
                           // logarithmic: friction velocity, surface roughness, von Karman constant
    int has_obukhov;       // logarithmic: Monin-Obukhov correction present
    double obukhov, stable, unstable;
    int has_parallel_gust;
    double parallel_gust, perpendicular, vertical;  // gust values at this time (0 when absent)
};

// environment.rs:66-71
SB_HD double wind_height(const WindField& w, V3 location) {
    return fmax_(dot(location, w.up) - w.water_plane_height, 0.0);
}

// power_model.rs:28-40
SB_HD double power_profile(const WindField& w, double height) {
    double f = w.p2 > 0.0 ? sbm::pow(height / w.p1, w.p2) : 1.0;
    return w.p0 * f;
}

// logarithmic_model.rs:52-74: Businger-Dyer stability correction, unscaled
SB_HD double businger_dyer_unscaled(const WindField& w, double height) {
    if (!w.has_obukhov) return 0.0;
    double zeta = height / w.obukhov;
    if (w.obukhov > 0.0) return -w.stable * zeta;
    double x = sbm::pow(1.0 - w.unstable * zeta, 0.25);
    double first = 2.0 * sbm::log((1.0 + x) / 2.0);
    double second = sbm::log((1.0 + x * x) / 2.0);
    double third = -2.0 * sbm::atan(x) + PI / 2.0;
    return first + second + third;
}

// logarithmic_model.rs:26-50, 76-80: neutral profile minus the scaled correction
SB_HD double logarithmic_profile(const WindField& w, double height) {
    const double scale = w.p0 / w.p2;
    double neutral = 0.0;
    if (!(w.p1 <= 0.0 || height <= 0.0)) neutral = scale * sbm::log(height / w.p1);
    return neutral - businger_dyer_unscaled(w, height) * scale;
}

// velocity_variation/mod.rs:22-28
SB_HD double profile_at_height(const WindField& w, double height) {
    if (w.profile == VP_POWER) return power_profile(w, height);
    if (w.profile == VP_LOGARITHMIC) return logarithmic_profile(w, height);
    return w.p0;
}

// wind_condition/mod.rs:41-48
SB_HD double unsteady_parallel_at_height(const WindField& w, double height) {
    double u = profile_at_height(w, height);
    if (w.has_parallel_gust) u += w.parallel_gust;
    return u;
}

// environment.rs:127-149: unsteady true wind vector
SB_HD V3 true_wind_at(const WindField& w, V3 location) {
    double u = unsteady_parallel_at_height(w, wind_height(w, location));
    V3 t = u * w.parallel_dir + w.perpendicular * w.perpendicular_dir;
    return t + w.vertical * w.up;
}

// environment.rs:182-190: unsteady apparent wind (true wind plus the ship's linear velocity)
SB_HD V3 apparent_wind_at(const WindField& w, V3 location) { return true_wind_at(w, location) + w.linear; }

}  // namespace sb
