// Host-side setup parsing and line-force-model geometry.
//
// Reference: stormbird/src/line_force_model/{builder/mod.rs, builder/single_wing.rs,
// data_update.rs, value_mapping.rs}, stormbird/src/lifting_line/{simulation_builder.rs,
// wake/dynamic_wake/builder.rs, solvers/simple_iterative.rs}, section_models/foil.rs,
// stormath/src/{rigid_body_motion.rs, spatial_vector/transformations.rs, interpolation.rs}.
#include "sb_model.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "sb_json.h"

namespace sb {

using json::Node;

// ----------------------------------------------------------------------------
// rotations (transformations.rs:29-155)
// ----------------------------------------------------------------------------
V3 rotate_around_axis(V3 v, double angle, V3 axis) {
    V3 a = normalize(axis);
    double c = sbm::cos(angle), s = sbm::sin(angle);
    V3 t1 = v * c;
    V3 t2 = cross(a, v) * s;
    V3 t3 = a * dot(a, v) * (1.0 - c);
    return t1 + t2 + t3;
}

static V3 rotate_xyz(V3 v, V3 r) {
    if (r.x != 0.0) {
        double y = v.y, z = v.z;
        v.y = y * sbm::cos(r.x) - z * sbm::sin(r.x);
        v.z = y * sbm::sin(r.x) + z * sbm::cos(r.x);
    }
    if (r.y != 0.0) {
        double x = v.x, z = v.z;
        v.x = x * sbm::cos(r.y) + z * sbm::sin(r.y);
        v.z = -x * sbm::sin(r.y) + z * sbm::cos(r.y);
    }
    if (r.z != 0.0) {
        double x = v.x, y = v.y;
        v.x = x * sbm::cos(r.z) - y * sbm::sin(r.z);
        v.y = x * sbm::sin(r.z) + y * sbm::cos(r.z);
    }
    return v;
}

static V3 rotate_ypr(V3 v, V3 r) {
    V3 yaw = v3(0, 0, 1);
    V3 pitch = rotate_around_axis(v3(0, 1, 0), r.z, yaw);
    V3 roll = rotate_around_axis(rotate_around_axis(v3(1, 0, 0), r.z, yaw), r.y, pitch);
    return rotate_around_axis(rotate_around_axis(rotate_around_axis(v, r.z, yaw), r.y, pitch), r.x, roll);
}

V3 RigidBody::rotate(V3 v) const { return rot == RotType::XYZ ? rotate_xyz(v, rotation) : rotate_ypr(v, rotation); }

V3 RigidBody::body_fixed(V3 v) const {
    RigidBody tmp = *this;
    V3 ax = tmp.rotate(v3(1, 0, 0)), ay = tmp.rotate(v3(0, 1, 0)), az = tmp.rotate(v3(0, 0, 1));
    return v3(dot(v, ax), dot(v, ay), dot(v, az));
}

void RigidBody::update_translation_fd(V3 t, double dt) {
    V3 old = translation;
    translation = t;
    velocity_linear = (translation - old) / dt;
}

void RigidBody::update_rotation_fd(V3 r, double dt) {
    V3 old = rotation;
    rotation = r;
    V3 d = rotation - old;
    double* c[3] = {&d.x, &d.y, &d.z};
    for (int i = 0; i < 3; ++i) {  // rigid_body_motion.rs:94-100 adds TAU in both branches
        if (*c[i] > PI) *c[i] = *c[i] + TAU;
        else if (*c[i] < -PI) *c[i] = *c[i] + TAU;
    }
    velocity_angular = d / dt;
}

// ----------------------------------------------------------------------------
// line force model geometry
// ----------------------------------------------------------------------------
template <class T>
std::vector<T> LineModel::span_values(const std::vector<T>& cv) const {
    std::vector<T> out;
    out.reserve(cv.size() + wing_start.size());
    for (int w = 0; w < nw(); ++w) {
        out.push_back(cv[wing_start[w]]);
        for (int i = wing_start[w]; i < wing_end[w]; ++i) {
            if (i == wing_end[w] - 1) out.push_back(cv[i]);
            else out.push_back((cv[i] + cv[i + 1]) * 0.5);
        }
    }
    return out;
}
template std::vector<V3> LineModel::span_values<V3>(const std::vector<V3>&) const;

void LineModel::update_global() {
    for (int i = 0; i < n(); ++i) {
        int w = wing_of[i];
        V3 axis = local[wing_start[w]].rel();
        chord_local[i] = rotate_around_axis(chord_local_not_rotated[i], wing_angles[w], axis);
        chord_global[i] = rbm.transform_vector(chord_local[i]);
    }
    chord_global_at_span = span_values(chord_global);
    for (int i = 0; i < n(); ++i) {
        global[i].start = rbm.transform_point(local[i].start);
        global[i].end = rbm.transform_point(local[i].end);
    }
    span_points.clear();
    for (int w = 0; w < nw(); ++w) {
        for (int i = wing_start[w]; i < wing_end[w]; ++i) span_points.push_back(global[i].start);
        span_points.push_back(global[wing_end[w] - 1].end);
    }
    ctrl_points.resize(n());
    for (int i = 0; i < n(); ++i) ctrl_points[i] = global[i].ctrl();
}

// ----------------------------------------------------------------------------
// JSON helpers (serde conventions: externally tagged enums, deny_unknown_fields)
// ----------------------------------------------------------------------------
static void allow_only(const Node* o, std::initializer_list<const char*> names, const char* what) {
    if (!o || o->t != Node::Obj) throw std::runtime_error(std::string(what) + ": expected an object");
    for (const auto& k : o->keys) {
        bool known = false;
        for (const char* n : names) known = known || k == n;
        if (!known) throw std::runtime_error(std::string(what) + ": unknown field `" + k + "`");
    }
}
static double number(const Node* n, const char* what) {
    if (!n || n->t != Node::Num) throw std::runtime_error(std::string(what) + ": expected a number");
    return n->n;
}
// usize fields: serde reads them from non-negative integer literals only.
static long count(const Node* n, const char* what) {
    if (!n || n->t != Node::Num || !n->is_int || n->n < 0.0 || n->n > 9.0e15)
        throw std::runtime_error(std::string(what) + ": expected a non-negative integer");
    return (long)n->n;
}
static bool boolean(const Node* n, const char* what) {
    if (!n || n->t != Node::Bool) throw std::runtime_error(std::string(what) + ": expected a boolean");
    return n->b;
}
static V3 vector(const Node* n, const char* what) {
    if (n && n->t == Node::Arr && n->items.size() == 3)
        return v3(number(n->items[0].get(), what), number(n->items[1].get(), what), number(n->items[2].get(), what));
    if (n && n->t == Node::Obj && n->find("x") && n->find("y") && n->find("z"))
        return v3(number(n->find("x"), what), number(n->find("y"), what), number(n->find("z"), what));
    throw std::runtime_error(std::string(what) + ": expected a vector");
}
static std::string variant(const Node* n, const Node** payload, const char* what) {
    *payload = nullptr;
    if (n && n->t == Node::Str) return n->s;
    if (n && n->t == Node::Obj && n->items.size() == 1) {
        *payload = n->items[0].get();
        return n->keys[0];
    }
    throw std::runtime_error(std::string(what) + ": expected an enum variant");
}

static FoilParams default_foil() {
    FoilParams f;
    f.cl_zero_angle = 0.0;
    f.cl_initial_slope = TAU;
    f.cl_high_order_factor_positive = 0.0;
    f.cl_high_order_factor_negative = 0.0;
    f.cl_high_order_power = 0.0;
    f.cl_max_after_stall = 1.0;
    f.cd_min = 0.0;
    f.angle_cd_min = 0.0;
    f.cd_second_order_factor = 0.0;
    f.cd_max_after_stall = 1.0;
    f.cd_power_after_stall = 1.6;
    f.cdi_correction_factor = 0.0;
    f.mean_positive_stall_angle = 20.0 * (PI / 180.0);  // f64::to_radians
    f.mean_negative_stall_angle = 20.0 * (PI / 180.0);
    f.stall_range = 6.0 * (PI / 180.0);
    f.cd_bump_during_stall = 0.0;
    f.cd_stall_angle_offset = 0.0;
    f.added_mass_factor = 0.0;
    return f;
}

static FoilParams parse_foil(const Node* o) {
    struct Field { const char* name; double FoilParams::*m; };
    static const Field fields[] = {
        {"cl_zero_angle", &FoilParams::cl_zero_angle},
        {"cl_initial_slope", &FoilParams::cl_initial_slope},
        {"cl_high_order_factor_positive", &FoilParams::cl_high_order_factor_positive},
        {"cl_high_order_factor_negative", &FoilParams::cl_high_order_factor_negative},
        {"cl_high_order_power", &FoilParams::cl_high_order_power},
        {"cl_max_after_stall", &FoilParams::cl_max_after_stall},
        {"cd_min", &FoilParams::cd_min},
        {"angle_cd_min", &FoilParams::angle_cd_min},
        {"cd_second_order_factor", &FoilParams::cd_second_order_factor},
        {"cd_max_after_stall", &FoilParams::cd_max_after_stall},
        {"cd_power_after_stall", &FoilParams::cd_power_after_stall},
        {"cdi_correction_factor", &FoilParams::cdi_correction_factor},
        {"mean_positive_stall_angle", &FoilParams::mean_positive_stall_angle},
        {"mean_negative_stall_angle", &FoilParams::mean_negative_stall_angle},
        {"stall_range", &FoilParams::stall_range},
        {"cd_bump_during_stall", &FoilParams::cd_bump_during_stall},
        {"cd_stall_angle_offset", &FoilParams::cd_stall_angle_offset},
        {"added_mass_factor", &FoilParams::added_mass_factor},
    };
    if (!o || o->t != Node::Obj) throw std::runtime_error("Foil: expected an object");
    FoilParams f = default_foil();
    for (size_t k = 0; k < o->keys.size(); ++k) {
        bool known = false;
        for (const auto& fd : fields)
            if (o->keys[k] == fd.name) {
                f.*(fd.m) = number(o->items[k].get(), fd.name);
                known = true;
            }
        if (!known) throw std::runtime_error("Foil: unknown field `" + o->keys[k] + "`");
    }
    return f;
}

// stormath/src/interpolation.rs:17-100
template <class T>
static T lerp2(double x, double x0, double x1, const T& y0, const T& y1) {
    if (x <= x0) return y0;
    if (x >= x1) return y1;
    T dy = y1 - y0;
    double dx = x1 - x0;
    return y0 + dy * ((x - x0) / dx);
}
template <class T>
static T interpolate(double x, const std::vector<double>& xs, const std::vector<T>& ys) {
    if (xs.size() == 1) return ys[0];
    if (xs.size() == 2) return lerp2(x, xs[0], xs[1], ys[0], ys[1]);
    size_t lo = 0, hi = xs.size() - 1;
    while (hi - lo > 1) {
        size_t mid = (lo + hi) / 2;
        if (xs[mid] > x) hi = mid;
        else lo = mid;
    }
    if (lo == xs.size() - 1) return ys[lo];
    return lerp2(x, xs[lo], xs[lo + 1], ys[lo], ys[lo + 1]);
}

// builder/single_wing.rs:73-150 and line_force_model/mod.rs:164-205
static void add_wing(LineModel& m, const Node* wb, int default_sections) {
    allow_only(wb, {"section_points", "chord_vectors", "section_model", "line_segment_is_virtual",
                    "non_zero_circulation_at_ends", "nr_sections", "input_power_model"},
               "WingBuilder");
    std::vector<V3> pts, chords;
    const Node* a = wb->find("section_points");
    if (!a || a->t != Node::Arr) throw std::runtime_error("WingBuilder: missing section_points");
    for (auto& e : a->items) pts.push_back(vector(e.get(), "section_points"));
    a = wb->find("chord_vectors");
    if (!a || a->t != Node::Arr) throw std::runtime_error("WingBuilder: missing chord_vectors");
    for (auto& e : a->items) chords.push_back(vector(e.get(), "chord_vectors"));
    if (pts.size() < 2 || chords.size() != pts.size())
        throw std::runtime_error("WingBuilder: section_points and chord_vectors must have equal length >= 2");
    const Node* pl;
    std::string sm = variant(wb->find("section_model"), &pl, "SectionModel");
    if (sm != "Foil") throw std::runtime_error("SectionModel `" + sm + "` is not supported on the GPU path");
    FoilParams foil = parse_foil(pl);
    std::vector<char> seg_virtual;  // Option<Vec<bool>>: null is None
    if (const Node* v = wb->find("line_segment_is_virtual")) {
        if (v->t == Node::Arr)
            for (auto& e : v->items) seg_virtual.push_back(boolean(e.get(), "line_segment_is_virtual"));
        else if (v->t != Node::Null)
            throw std::runtime_error("line_segment_is_virtual: expected a list of booleans or null");
    }
    if (const Node* v = wb->find("non_zero_circulation_at_ends")) {  // [bool; 2]
        if (v->t != Node::Arr || v->items.size() != 2)
            throw std::runtime_error("non_zero_circulation_at_ends: expected two booleans");
        for (auto& e : v->items) boolean(e.get(), "non_zero_circulation_at_ends");
    }
    int nsec = default_sections;  // Option<usize>: null is None
    if (const Node* v = wb->find("nr_sections"))
        if (v->t != Node::Null) nsec = (int)count(v, "WingBuilder.nr_sections");
    if (nsec < 1) throw std::runtime_error("nr_sections must be at least 1");
    if (const Node* v = wb->find("input_power_model")) {
        const Node* p;
        if (variant(v, &p, "InputPowerModel") != "NoPower")
            throw std::runtime_error("InputPowerModel other than NoPower is not supported on the GPU path");
    }
    std::vector<double> dist;
    for (size_t i = 0; i < pts.size(); ++i)
        dist.push_back(dist.empty() ? 0.0 : dist.back() + len(pts[i] - pts[i - 1]));
    double ds_avg = dist.back() / (double)nsec;
    int start = m.wing_start.empty() ? 0 : m.wing_end.back();
    int w = (int)m.wing_start.size();
    int count = 0;
    for (size_t seg = 0; seg + 1 < pts.size(); ++seg) {
        double seg_len = dist[seg + 1] - dist[seg];
        int pieces = (int)std::ceil(seg_len / ds_avg);
        double ds = seg_len / (double)pieces;
        char virt = seg_virtual.empty() ? 0 : seg_virtual.at(seg);
        for (int i = 0; i < pieces; ++i) {
            double s0 = dist[seg] + (double)i * ds;
            double s1 = s0 + ds;
            double sc = 0.5 * (s0 + s1);
            SpanLine L{interpolate(s0, dist, pts), interpolate(s1, dist, pts)};
            V3 ch = interpolate(sc, dist, chords);
            m.local.push_back(L);
            m.global.push_back(L);
            m.chord_local_not_rotated.push_back(ch);
            m.chord_local.push_back(ch);
            m.chord_global.push_back(ch);
            m.chord_len.push_back(len(ch));
            m.is_virtual.push_back(virt);
            m.wing_of.push_back(w);
            ++count;
        }
    }
    m.wing_start.push_back(start);
    m.wing_end.push_back(start + count);
    m.foils.push_back(foil);
    m.wing_angles.push_back(0.0);
    m.update_global();
}

static LineModel parse_line_model(const Node* o) {  // builder/mod.rs:79-113
    allow_only(o, {"wing_builders", "nr_sections", "density", "circulation_correction",
                   "angle_of_attack_correction", "output_coordinate_system", "rotation_type",
                   "local_wing_angles", "rotation", "translation", "force_calculation_settings"},
               "LineForceModelBuilder");
    LineModel m;
    if (!o->find("nr_sections")) throw std::runtime_error("LineForceModelBuilder: missing field `nr_sections`");
    int nsec = (int)count(o->find("nr_sections"), "LineForceModelBuilder.nr_sections");
    if (const Node* v = o->find("density")) m.density = number(v, "density");
    const Node* wbs = o->find("wing_builders");
    if (!wbs || wbs->t != Node::Arr || wbs->items.empty())
        throw std::runtime_error("LineForceModelBuilder: wing_builders must be a non-empty list");
    for (auto& wb : wbs->items) add_wing(m, wb.get(), nsec);
    const Node* p;
    if (const Node* v = o->find("circulation_correction"))
        if (variant(v, &p, "CirculationCorrectionBuilder") != "None")
            throw std::runtime_error("circulation corrections are not supported on the GPU path");
    if (const Node* v = o->find("angle_of_attack_correction"))
        if (variant(v, &p, "AngleOfAttackCorrection") != "None")
            throw std::runtime_error("angle of attack corrections are not supported on the GPU path");
    if (const Node* v = o->find("output_coordinate_system")) {
        std::string t = variant(v, &p, "CoordinateSystem");
        if (t == "Body") m.output_body = true;
        else if (t != "Global") throw std::runtime_error("unknown CoordinateSystem `" + t + "`");
    }
    if (const Node* v = o->find("rotation_type")) {  // parsed; the reference builder does not apply it
        std::string t = variant(v, &p, "RotationType");
        if (t != "XYZ" && t != "YawPitchRoll") throw std::runtime_error("unknown RotationType `" + t + "`");
    }
    if (const Node* v = o->find("local_wing_angles")) {
        if (v->t != Node::Arr) throw std::runtime_error("local_wing_angles: expected a list");
        if (!v->items.empty()) {
            if ((int)v->items.size() != m.nw())
                throw std::runtime_error("The number of local wing angles does not match the number of wings.");
            for (size_t i = 0; i < v->items.size(); ++i) m.wing_angles[i] = number(v->items[i].get(), "local_wing_angles");
        }
    }
    if (const Node* v = o->find("force_calculation_settings")) {
        // ForceCalculationSettings has no deny_unknown_fields (force_calculations.rs:41-45):
        // other keys are ignored
        if (v->t != Node::Obj) throw std::runtime_error("ForceCalculationSettings: expected an object");
        if (const Node* b = v->find("include_viscous_lift_in_the_circulation"))
            m.viscous_lift_in_circulation = boolean(b, "include_viscous_lift_in_the_circulation");
    }
    if (const Node* v = o->find("translation")) m.rbm.translation = vector(v, "translation");
    if (const Node* v = o->find("rotation")) m.rbm.rotation = vector(v, "rotation");
    m.update_global();
    return m;
}

static double core_length(const Node* v, double strip_length) {  // dynamic_wake/builder.rs:222-226
    const Node* p;
    std::string t = variant(v, &p, "ViscousCoreLength");
    if (t == "Relative") return number(p, "Relative") * strip_length;
    if (t == "Absolute") return number(p, "Absolute");
    if (t == "NoViscousCore") return MIN_POSITIVE;
    throw std::runtime_error("unknown ViscousCoreLength `" + t + "`");
}

static void parse_wake(WakeConfig& w, const Node* o, const LineModel& m) {  // builder.rs:58-255
    if (o)
        allow_only(o, {"nr_panels_per_line_element", "viscous_core_length", "viscous_core_length_evolution",
                       "first_panel_relative_length", "last_panel_relative_length", "first_wake_points_direction",
                       "symmetry_condition", "ratio_of_wake_affected_by_induced_velocities", "far_field_ratio",
                       "shape_damping_factor", "neglect_self_induced_velocities", "initial_relative_wake_length",
                       "write_wake_data_to_file", "wake_files_folder_path"},
                   "DynamicWakeBuilder");
    auto f = [&](const char* k) { return o ? o->find(k) : nullptr; };
    const Node* p;
    if (const Node* v = f("nr_panels_per_line_element")) w.M = (int)count(v, "nr_panels_per_line_element");
    if (w.M < 1) throw std::runtime_error("nr_panels_per_line_element must be at least 1");
    if (const Node* v = f("first_panel_relative_length")) w.first_rel = number(v, "first_panel_relative_length");
    if (const Node* v = f("last_panel_relative_length")) w.last_rel = number(v, "last_panel_relative_length");
    if (const Node* v = f("first_wake_points_direction")) {
        std::string t = variant(v, &p, "FirstWakePointsDirection");
        if (t == "Chord") w.first_dir = 0;
        else if (t == "Freestream") w.first_dir = 1;
        else if (t == "ActualVelocity") w.first_dir = 2;
        else throw std::runtime_error("unknown FirstWakePointsDirection `" + t + "`");
    }
    if (const Node* v = f("symmetry_condition")) {
        std::string t = variant(v, &p, "SymmetryCondition");
        if (t == "NoSymmetry") w.symmetry = 0;
        else if (t == "X") w.symmetry = 1;
        else if (t == "Y") w.symmetry = 2;
        else if (t == "Z") w.symmetry = 3;
        else throw std::runtime_error("unknown SymmetryCondition `" + t + "`");
    }
    double ratio = 0.0;
    if (const Node* v = f("ratio_of_wake_affected_by_induced_velocities")) ratio = number(v, "ratio");
    if (const Node* v = f("far_field_ratio")) w.far_ratio = number(v, "far_field_ratio");
    if (const Node* v = f("shape_damping_factor")) w.damping = number(v, "shape_damping_factor");
    if (const Node* v = f("neglect_self_induced_velocities")) w.neglect_self = boolean(v, "neglect_self_induced_velocities");
    if (const Node* v = f("initial_relative_wake_length")) w.initial_relative_wake_length = number(v, "initial_relative_wake_length");
    if (const Node* v = f("write_wake_data_to_file")) w.write_files = boolean(v, "write_wake_data_to_file");
    if (const Node* v = f("wake_files_folder_path")) {
        if (v->t != Node::Str) throw std::runtime_error("wake_files_folder_path: expected a string");
        w.files_folder = v->s;
    }
    int N = m.n();
    long P = (long)m.span_points.size();
    long nr_points = P * (long)(w.M + 1);
    w.end_index = (long)std::ceil(ratio * (double)(nr_points - 1));
    // per-panel viscous core length
    std::vector<double> strip_len(N);
    for (int i = 0; i < N; ++i) strip_len[i] = m.global[i].length();
    const Node* core = f("viscous_core_length");
    const Node* evo = f("viscous_core_length_evolution");
    const Node* last = nullptr;
    double evo_factor = 1.0;
    if (evo) {
        const Node* ep;
        std::string t = variant(evo, &ep, "ViscousCoreLengthEvolution");
        if (t == "SinIncrease") {
            allow_only(ep, {"last_panel_value", "evolution_length_factor"}, "SinIncreasedViscousCoreLength");
            last = ep->find("last_panel_value");
            if (!last) throw std::runtime_error("SinIncrease: missing last_panel_value");
            if (const Node* e = ep->find("evolution_length_factor")) evo_factor = number(e, "evolution_length_factor");
        } else if (t != "Constant") {
            throw std::runtime_error("unknown ViscousCoreLengthEvolution `" + t + "`");
        }
    }
    w.core.assign((size_t)w.M * N, 0.0);
    for (int s = 0; s < w.M; ++s)
        for (int i = 0; i < N; ++i) {
            double raw = core ? core_length(core, strip_len[i]) : 0.1 * strip_len[i];
            double val = raw;
            if (last) {
                double fin = core_length(last, strip_len[i]);
                double lb = (double)(w.M - 1) * evo_factor;
                double rs = fmax_(fmin_((double)s / lb, 1.0), 0.0);
                double sf = sbm::sin(0.5 * rs * PI);
                val = raw * (1.0 - sf) + fin * sf;
            }
            w.core[(size_t)s * N + i] = val;
        }
    double csum = 0.0;
    for (double c : m.chord_len) csum += c;
    w.rep_chord = csum / (double)m.chord_len.size();
}

static void parse_velocity_correction(SolverConfig& s, const Node* x) {  // velocity_corrections.rs:9-16
    const Node* q;
    std::string vt = variant(x, &q, "VelocityCorrections");
    if (vt == "NoCorrection") s.velocity_correction = 0;
    else if (vt == "MaxInducedVelocityMagnitudeRatio") {
        s.velocity_correction = 1;
        s.velocity_correction_ratio = number(q, "MaxInducedVelocityMagnitudeRatio");
    } else if (vt == "FixedMagnitudeEqualToFreestream") s.velocity_correction = 2;
    else throw std::runtime_error("unknown VelocityCorrections `" + vt + "`");
}

static void parse_linearized(SolverConfig& s, const Node* p) {  // linearized.rs:26-48
    s.kind = 1;
    if (!p) return;
    allow_only(p, {"velocity_corrections", "disable_viscous_corrections", "induced_velocity_correction_method"},
               "Linearized");
    if (const Node* x = p->find("velocity_corrections")) parse_velocity_correction(s, x);
    if (const Node* x = p->find("disable_viscous_corrections"))
        s.disable_viscous_corrections = boolean(x, "disable_viscous_corrections");
    if (const Node* x = p->find("induced_velocity_correction_method")) {
        const Node* q;
        std::string t = variant(x, &q, "InducedVelocityCorrectionMethod");
        if (t == "NoCorrection") s.induced_velocity_correction = 0;
        else if (t == "FullCorrection") s.induced_velocity_correction = 1;
        else throw std::runtime_error("unknown InducedVelocityCorrectionMethod `" + t + "`");
    }
}

// Solver (solvers/mod.rs:21-30, simple_iterative.rs:59-84) or, for quasi-steady
// settings, QuasiSteadySolverBuilder (mod.rs:32-57): default Linearized, and a
// SimpleIterative builder with 1000 iterations and damping 0.04 by default
// (simple_iterative.rs:14-45).
static void parse_solver(SolverConfig& s, const Node* v, bool quasi) {
    if (quasi) {
        s.max_iterations = 1000;
        s.damping = 0.04;
        if (!v) {
            s.kind = 1;
            return;
        }
    }
    if (!v) return;
    const Node* p;
    std::string t = variant(v, &p, "Solver");
    if (t == "Linearized") {
        parse_linearized(s, p);
        return;
    }
    if (t != "SimpleIterative") throw std::runtime_error("unknown Solver `" + t + "`");
    if (quasi)
        allow_only(p, {"max_iterations_per_time_step", "damping_factor", "residual_tolerance_absolute",
                       "strength_difference_tolerance", "velocity_corrections", "start_with_linearized_solution",
                       "correction_factor_initial_solution"},
                   "QuasiSteadySimpleIterativeBuilder");
    else
        allow_only(p, {"max_iterations_per_time_step", "damping_factor", "residual_tolerance_absolute",
                       "strength_difference_tolerance", "velocity_corrections", "use_raw_circulation_during_iterations",
                       "start_with_linearized_solution", "correction_factor_initial_solution"},
                   "SimpleIterative");
    if (const Node* x = p->find("max_iterations_per_time_step"))
        s.max_iterations = (int)std::min<long>(count(x, "max_iterations_per_time_step"), 2000000000L);
    if (const Node* x = p->find("damping_factor")) s.damping = number(x, "damping_factor");
    if (const Node* x = p->find("residual_tolerance_absolute")) s.residual_tol = number(x, "residual_tolerance_absolute");
    if (const Node* x = p->find("strength_difference_tolerance")) s.strength_tol = number(x, "strength_difference_tolerance");
    if (const Node* x = p->find("use_raw_circulation_during_iterations")) boolean(x, "use_raw_circulation_during_iterations");
    if (const Node* x = p->find("start_with_linearized_solution"))
        s.start_linearized = boolean(x, "start_with_linearized_solution");
    if (const Node* x = p->find("correction_factor_initial_solution")) s.correction_factor_initial = number(x, "correction_factor_initial_solution");
    if (const Node* x = p->find("velocity_corrections")) parse_velocity_correction(s, x);
}

static void parse_quasi_wake(QuasiConfig& q, const Node* o) {  // wake/settings.rs:15-58
    if (!o) return;
    allow_only(o, {"wake_length_factor", "symmetry_condition", "viscous_core_length"}, "QuasiSteadyWakeSettings");
    const Node* p;
    if (const Node* v = o->find("wake_length_factor")) q.wake_length_factor = number(v, "wake_length_factor");
    if (const Node* v = o->find("symmetry_condition")) {
        std::string t = variant(v, &p, "SymmetryCondition");
        if (t == "NoSymmetry") q.symmetry = 0;
        else if (t == "X") q.symmetry = 1;
        else if (t == "Y") q.symmetry = 2;
        else if (t == "Z") q.symmetry = 3;
        else throw std::runtime_error("unknown SymmetryCondition `" + t + "`");
    }
    if (const Node* v = o->find("viscous_core_length")) {
        std::string t = variant(v, &p, "ViscousCoreLength");
        if (t == "Relative") {
            q.core_kind = 0;
            q.core_value = number(p, "Relative");
        } else if (t == "Absolute") {
            q.core_kind = 1;
            q.core_value = number(p, "Absolute");
        } else if (t == "NoViscousCore") {
            q.core_kind = 2;
            q.core_value = 0.0;
        } else {
            throw std::runtime_error("unknown ViscousCoreLength `" + t + "`");
        }
    }
}

static Setup parse_setup_node(const Node* root) {
    allow_only(root, {"line_force_model", "simulation_settings"}, "SimulationBuilder");
    const Node* lfm = root->find("line_force_model");
    if (!lfm) throw std::runtime_error("SimulationBuilder: missing field `line_force_model`");
    Setup s;
    s.lfm = parse_line_model(lfm);
    // simulation_builder.rs:42-64: absent settings are QuasiSteadySettings::default()
    const Node* ss = root->find("simulation_settings");
    const Node* p = nullptr;
    std::string t = ss ? variant(ss, &p, "SimulationSettings") : std::string("QuasiSteady");
    if (t == "QuasiSteady") {
        if (p) allow_only(p, {"solver", "wake"}, "QuasiSteadySettings");
        s.quasi = true;
        parse_solver(s.solver, p ? p->find("solver") : nullptr, true);
        parse_quasi_wake(s.qs, p ? p->find("wake") : nullptr);
        s.wake.M = 0;
        double csum = 0.0;
        for (double c : s.lfm.chord_len) csum += c;
        s.wake.rep_chord = csum / (double)s.lfm.chord_len.size();
        return s;
    }
    if (t != "Dynamic") throw std::runtime_error("unknown SimulationSettings `" + t + "`");
    if (p) allow_only(p, {"solver", "wake"}, "DynamicSettings");
    parse_solver(s.solver, p ? p->find("solver") : nullptr, false);
    parse_wake(s.wake, p ? p->find("wake") : nullptr, s.lfm);
    return s;
}

Setup parse_setup(const char* text) {
    if (!text) throw std::runtime_error("null setup string");
    auto root = json::parse(text);
    return parse_setup_node(root.get());
}

// ---------------------------------------------------------------------------
// Wind, controller and complete sail model settings
// ---------------------------------------------------------------------------
static std::vector<double> numbers(const Node* n, const char* what) {
    if (!n || n->t != Node::Arr) throw std::runtime_error(std::string(what) + ": expected a list of numbers");
    std::vector<double> v;
    for (auto& e : n->items) v.push_back(number(e.get(), what));
    return v;
}
static const Node* required(const Node* o, const char* key, const char* what) {
    const Node* v = o->find(key);
    if (!v) throw std::runtime_error(std::string(what) + ": missing field `" + key + "`");
    return v;
}
static void require_object(const Node* o, const char* what) {
    if (!o || o->t != Node::Obj) throw std::runtime_error(std::string(what) + ": expected an object");
}

double Spectrum::value_at_time(double t) const {  // discretized_spectrum.rs:37-45
    double out = 0.0;
    for (size_t i = 0; i < frequencies.size(); ++i)
        out += amplitudes[i] * sbm::sin(TAU * frequencies[i] * t + phases[i]);
    return out;
}

static Spectrum parse_spectrum(const Node* o) {  // no deny_unknown_fields
    require_object(o, "DiscretizedSpectrum");
    Spectrum s;
    s.frequencies = numbers(required(o, "frequencies", "DiscretizedSpectrum"), "frequencies");
    s.amplitudes = numbers(required(o, "amplitudes", "DiscretizedSpectrum"), "amplitudes");
    s.phases = numbers(required(o, "phases", "DiscretizedSpectrum"), "phases");
    if (s.amplitudes.size() < s.frequencies.size() || s.phases.size() < s.frequencies.size())
        throw std::runtime_error("DiscretizedSpectrum: fewer amplitudes or phases than frequencies");
    return s;
}

static WindConditionCfg parse_wind_condition_node(const Node* o) {  // wind_condition/mod.rs:18-29
    require_object(o, "WindCondition");  // no deny_unknown_fields
    WindConditionCfg w;
    w.direction_coming_from = number(required(o, "direction_coming_from", "WindCondition"), "direction_coming_from");
    const Node* p;
    std::string t = variant(required(o, "velocity_variation", "WindCondition"), &p, "VelocityVariation");
    if (t == "Constant") {
        w.profile = 0;
        w.p0 = number(p, "Constant");
    } else if (t == "PowerModel") {  // power_model.rs:10-23, no deny_unknown_fields
        require_object(p, "PowerModel");
        w.profile = 1;
        w.p0 = number(required(p, "reference_velocity", "PowerModel"), "reference_velocity");
        w.p1 = 10.0;
        w.p2 = 1.0 / 9.0;
        if (const Node* x = p->find("reference_height")) w.p1 = number(x, "reference_height");
        if (const Node* x = p->find("power_factor")) w.p2 = number(x, "power_factor");
    } else if (t == "LogarithmicModel") {  // logarithmic_model.rs:10-28, no deny_unknown_fields
        require_object(p, "LogarithmicModel");
        w.profile = 2;
        w.p0 = number(required(p, "friction_velocity", "LogarithmicModel"), "friction_velocity");
        w.p1 = number(required(p, "surface_roughness", "LogarithmicModel"), "surface_roughness");
        w.p2 = 0.41;
        if (const Node* x = p->find("obukhov_length"))
            if (x->t != Node::Null) {
                w.has_obukhov = true;
                w.obukhov = number(x, "obukhov_length");
            }
        if (const Node* x = p->find("von_karman_constant")) w.p2 = number(x, "von_karman_constant");
        if (const Node* x = p->find("stable_coefficient")) w.stable = number(x, "stable_coefficient");
        if (const Node* x = p->find("unstable_coefficient")) w.unstable = number(x, "unstable_coefficient");
    } else {
        throw std::runtime_error("unknown VelocityVariation `" + t + "`");
    }
    const char* gusts[3] = {"parallel_gust", "perpendicular_gust", "vertical_gust"};
    for (int k = 0; k < 3; ++k)
        if (const Node* g = o->find(gusts[k]))
            if (g->t != Node::Null) {
                w.has_gust[k] = true;
                w.gust[k] = parse_spectrum(g);
            }
    return w;
}

WindConditionCfg parse_wind_condition(const char* text) {
    if (!text) throw std::runtime_error("null wind condition string");
    auto root = json::parse(text);
    return parse_wind_condition_node(root.get());
}

static std::pair<long, long> parse_range(const Node* n) {  // serde's Range<usize>: {start, end} or [start, end]
    if (n && n->t == Node::Arr && n->items.size() == 2)
        return {count(n->items[0].get(), "Range.start"), count(n->items[1].get(), "Range.end")};
    allow_only(n, {"start", "end"}, "Range");
    return {count(required(n, "start", "Range"), "Range.start"), count(required(n, "end", "Range"), "Range.end")};
}

static WindEnvironmentCfg parse_wind_environment_node(const Node* o) {  // environment.rs:21-34
    allow_only(o, {"up_direction", "wind_rotation_axis", "zero_direction_vector", "water_plane_height",
                   "inflow_corrections"},
               "WindEnvironment");
    WindEnvironmentCfg e;
    if (const Node* v = o->find("up_direction")) e.up = vector(v, "up_direction");
    if (const Node* v = o->find("wind_rotation_axis")) e.axis = vector(v, "wind_rotation_axis");
    if (const Node* v = o->find("zero_direction_vector")) e.zero_dir = vector(v, "zero_direction_vector");
    if (const Node* v = o->find("water_plane_height")) e.water_plane_height = number(v, "water_plane_height");
    if (const Node* v = o->find("inflow_corrections"))
        if (v->t != Node::Null) {  // inflow_corrections.rs:49-53, no deny_unknown_fields
            require_object(v, "InflowCorrections");
            e.has_inflow = true;
            e.inflow_directions = numbers(required(v, "apparent_wind_directions", "InflowCorrections"),
                                          "apparent_wind_directions");
            const Node* cs = required(v, "corrections", "InflowCorrections");
            if (cs->t != Node::Arr) throw std::runtime_error("InflowCorrections.corrections: expected a list");
            for (auto& c : cs->items) {  // :13-19, no deny_unknown_fields
                require_object(c.get(), "InflowCorrectionsSingleDirection");
                InflowCorrectionsSingle one;
                one.heights = numbers(required(c.get(), "height_values", "InflowCorrectionsSingleDirection"),
                                      "height_values");
                one.magnitude = numbers(required(c.get(), "magnitude_corrections", "InflowCorrectionsSingleDirection"),
                                        "magnitude_corrections");
                one.angle = numbers(required(c.get(), "angle_corrections", "InflowCorrectionsSingleDirection"),
                                    "angle_corrections");
                const Node* wi = required(c.get(), "wing_indices", "InflowCorrectionsSingleDirection");
                if (wi->t != Node::Arr) throw std::runtime_error("wing_indices: expected a list of ranges");
                for (auto& r : wi->items) one.wing_ranges.push_back(parse_range(r.get()));
                e.inflow.push_back(std::move(one));
            }
        }
    return e;
}

WindEnvironmentCfg parse_wind_environment(const char* text) {
    if (!text) throw std::runtime_error("null wind environment string");
    auto root = json::parse(text);
    return parse_wind_environment_node(root.get());
}

static MeasurementCfg parse_measurement(const Node* o) {  // measurements.rs:20-30, no deny_unknown_fields
    MeasurementCfg m;
    if (!o) return m;
    require_object(o, "MeasurementSettings");
    if (const Node* v = o->find("measurement_type")) {
        const Node* p;
        std::string t = variant(v, &p, "MeasurementType");
        if (t == "Mean") m.type = 0;
        else if (t == "Max") m.type = 1;
        else if (t == "Min") m.type = 2;
        else throw std::runtime_error("unknown MeasurementType `" + t + "`");
    }
    if (const Node* v = o->find("start_index")) m.start_index = count(v, "start_index");
    if (const Node* v = o->find("end_offset")) m.end_offset = count(v, "end_offset");
    return m;
}

static SetPointsCfg parse_set_points(const Node* o) {  // set_points.rs:52-68
    allow_only(o, {"apparent_wind_directions_data", "angle_of_attack_data", "section_model_internal_state_data",
                   "internal_state_type", "use_effective_angle_of_attack", "max_local_wing_angle_change_rate",
                   "max_internal_section_state_change_rate"},
               "ControllerSetPoints");
    SetPointsCfg c;
    c.directions = numbers(required(o, "apparent_wind_directions_data", "ControllerSetPoints"),
                           "apparent_wind_directions_data");
    if (const Node* v = o->find("angle_of_attack_data"))
        if (v->t != Node::Null) {
            c.has_aoa = true;
            c.aoa = numbers(v, "angle_of_attack_data");
        }
    if (const Node* v = o->find("section_model_internal_state_data"))
        if (v->t != Node::Null) {
            c.has_state = true;
            c.state = numbers(v, "section_model_internal_state_data");
        }
    if (const Node* v = o->find("internal_state_type")) {
        const Node* p;
        std::string t = variant(v, &p, "InternalStateType");
        if (t == "SpinRatio") {  // SpinRatioConversion: no deny_unknown_fields
            require_object(p, "SpinRatioConversion");
            c.spin_ratio = true;
            c.diameter = number(required(p, "diameter", "SpinRatioConversion"), "diameter");
            c.max_rps = number(required(p, "max_rps", "SpinRatioConversion"), "max_rps");
        } else if (t != "Generic") {
            throw std::runtime_error("unknown InternalStateType `" + t + "`");
        }
    }
    if (const Node* v = o->find("use_effective_angle_of_attack")) c.use_effective = boolean(v, "use_effective_angle_of_attack");
    if (const Node* v = o->find("max_local_wing_angle_change_rate"))
        if (v->t != Node::Null) {
            c.has_max_angle_rate = true;
            c.max_angle_rate = number(v, "max_local_wing_angle_change_rate");
        }
    if (const Node* v = o->find("max_internal_section_state_change_rate"))
        if (v->t != Node::Null) {
            c.has_max_state_rate = true;
            c.max_state_rate = number(v, "max_internal_section_state_change_rate");
        }
    return c;
}

static ControllerCfg parse_controller(const Node* o) {  // controller/builder.rs:14-31
    allow_only(o, {"set_points", "flow_measurement_settings", "time_steps_between_updates", "start_time",
                   "moving_average_window_size", "use_input_velocity_for_apparent_wind_direction"},
               "ControllerBuilder");
    ControllerCfg c;
    const Node* sp = required(o, "set_points", "ControllerBuilder");
    if (sp->t != Node::Arr) throw std::runtime_error("ControllerBuilder.set_points: expected a list");
    for (auto& e : sp->items) c.set_points.push_back(parse_set_points(e.get()));
    if (const Node* v = o->find("flow_measurement_settings")) {  // measurements.rs:32-41
        allow_only(v, {"angle_of_attack", "wind_direction", "wind_velocity"}, "FlowMeasurementSettings");
        c.aoa = parse_measurement(v->find("angle_of_attack"));
        c.direction = parse_measurement(v->find("wind_direction"));
        c.velocity = parse_measurement(v->find("wind_velocity"));
    }
    if (const Node* v = o->find("time_steps_between_updates")) c.steps_between_updates = count(v, "time_steps_between_updates");
    if (const Node* v = o->find("start_time")) c.start_time = number(v, "start_time");
    if (const Node* v = o->find("moving_average_window_size"))  // parsed; the built controller does not use it
        if (v->t != Node::Null) count(v, "moving_average_window_size");
    if (const Node* v = o->find("use_input_velocity_for_apparent_wind_direction"))
        c.use_input_velocity = boolean(v, "use_input_velocity_for_apparent_wind_direction");
    return c;
}

CompleteSailModelCfg parse_complete_sail_model(const char* text) {  // complete_sail_model/builder.rs:14-20
    if (!text) throw std::runtime_error("null setup string");
    auto root = json::parse(text);
    allow_only(root.get(), {"lifting_line_simulation", "wind_environment", "controller"}, "CompleteSailModelBuilder");
    CompleteSailModelCfg c;
    c.sim = parse_setup_node(required(root.get(), "lifting_line_simulation", "CompleteSailModelBuilder"));
    c.env = parse_wind_environment_node(required(root.get(), "wind_environment", "CompleteSailModelBuilder"));
    c.controller = parse_controller(required(root.get(), "controller", "CompleteSailModelBuilder"));
    return c;
}

double interpolate_scalar(double x, const double* xs, const double* ys, size_t n) {
    return interpolate(x, std::vector<double>(xs, xs + n), std::vector<double>(ys, ys + n));
}
V3 interpolate_vector(double x, const double* xs, const V3* ys, size_t n) {
    return interpolate(x, std::vector<double>(xs, xs + n), std::vector<V3>(ys, ys + n));
}

}  // namespace sb
