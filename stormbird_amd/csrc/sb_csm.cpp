// The complete sail model: a lifting-line simulation driven by a wind
// environment and a controller (stormbird/src/lifting_line/complete_sail_model/mod.rs:38-330,
// wind/environment.rs:60-360, wind/inflow_corrections.rs, controller/*.rs), with
// the C interface of the reference's cstormbird crate
// (interfaces/cstormbird/src/lifting_line/complete_sail_model.rs:17-109) and a
// fuller one (sb_csm_*) for the rest of the model's methods.
//
// A step's freestream is evaluated where its points live: the control points
// (with the inflow corrections and the controller's measurements) on the host,
// the wake points (up to hundreds of thousands) by k_wind_points straight into
// the simulation's freestream buffer, so no freestream crosses PCIe.
#include <cstdint>

#include "sb_host.h"

#include "../../include/cstormbird_compat.h"

namespace {


// The wind field of condition c at time t (environment.rs:73-190): direction
// vectors, gust values and the ship's linear velocity resolved once.
WindField wind_field(const WindEnvironmentCfg& e, const WindConditionCfg& c, double ship_velocity, double time) {
    WindField w{};
    w.up = e.up;
    w.parallel_dir = rotate_around_axis(e.zero_dir, c.direction_coming_from, e.axis);  // :73-79
    w.perpendicular_dir = cross(e.up, w.parallel_dir);
    w.linear = ship_velocity * e.zero_dir;  // complete_sail_model/mod.rs:223
    w.water_plane_height = e.water_plane_height;
    w.profile = c.profile;
    w.p0 = c.p0;
    w.p1 = c.p1;
    w.p2 = c.p2;
    w.has_obukhov = c.has_obukhov ? 1 : 0;
    w.obukhov = c.obukhov;
    w.stable = c.stable;
    w.unstable = c.unstable;
    w.has_parallel_gust = c.has_gust[0] ? 1 : 0;
    w.parallel_gust = c.has_gust[0] ? c.gust[0].value_at_time(time) : 0.0;  // wind_condition/mod.rs:41-61
    w.perpendicular = c.has_gust[1] ? c.gust[1].value_at_time(time) : 0.0;
    w.vertical = c.has_gust[2] ? c.gust[2].value_at_time(time) : 0.0;
    return w;
}

// environment.rs:151-158: steady true wind (profile along the wind direction only)
V3 steady_true_wind(const WindField& w, V3 location) {
    return profile_at_height(w, wind_height(w, location)) * w.parallel_dir;
}

// environment.rs:258-278
double apparent_wind_direction(const WindEnvironmentCfg& e, const WindField& w, double height) {
    V3 location = height * e.up;
    V3 apparent = steady_true_wind(w, location) + w.linear;
    return signed_angle_between(e.zero_dir, apparent, e.axis);
}

template <class T>
std::vector<T> slice(const std::vector<T>& v, std::pair<long, long> r, const char* what) {
    if (r.first < 0 || r.second < r.first || r.second > (long)v.size())
        throw std::runtime_error(std::string(what) + ": wing index range outside the table");
    return std::vector<T>(v.begin() + r.first, v.begin() + r.second);
}

double interp(double x, const std::vector<double>& xs, const std::vector<double>& ys, const char* what) {
    if (xs.empty() || xs.size() != ys.size())
        throw std::runtime_error(std::string(what) + ": interpolation tables must be non-empty and of equal length");
    return interpolate_scalar(x, xs.data(), ys.data(), xs.size());
}

// inflow_corrections.rs:21-93
V3 corrected_velocity(const WindEnvironmentCfg& e, int wing, double direction, double height, V3 v) {
    if (e.inflow_directions.empty()) throw std::runtime_error("InflowCorrections: no apparent wind directions");
    if (e.inflow.size() < e.inflow_directions.size())
        throw std::runtime_error("InflowCorrections: fewer correction tables than apparent wind directions");
    std::vector<V3> per_direction;
    for (size_t k = 0; k < e.inflow_directions.size(); ++k) {
        const InflowCorrectionsSingle& one = e.inflow[k];
        if (wing >= (int)one.wing_ranges.size())
            throw std::runtime_error("InflowCorrections: no wing index range for wing " + std::to_string(wing));
        std::pair<long, long> r = one.wing_ranges[wing];
        std::vector<double> h = slice(one.heights, r, "height_values");
        double m = interp(height, h, slice(one.magnitude, r, "magnitude_corrections"), "magnitude_corrections");
        double a = interp(height, h, slice(one.angle, r, "angle_corrections"), "angle_corrections");
        per_direction.push_back(m * rotate_around_axis(v, a, e.up));
    }
    return interpolate_vector(direction, e.inflow_directions.data(), per_direction.data(), per_direction.size());
}

// environment.rs:225-254: corrections per sail at the control-point heights
void apply_inflow_corrections(const WindEnvironmentCfg& e, double direction, V3* v, const V3* ctrl,
                              const std::vector<std::pair<long, long>>& wing_ranges) {
    if (!e.has_inflow) return;
    for (size_t wing = 0; wing < wing_ranges.size(); ++wing)
        for (long i = wing_ranges[wing].first; i < wing_ranges[wing].second; ++i)
            v[i] = corrected_velocity(e, (int)wing, direction, dot(ctrl[i], e.up), v[i]);
}

// controller/measurements.rs:43-77 (usize arithmetic of the slice end kept)
std::vector<double> measure(const std::vector<double>& values, const LineModel& m, const MeasurementCfg& s) {
    std::vector<double> out(m.nw(), 0.0);
    for (int w = 0; w < m.nw(); ++w) {
        const unsigned long long n = (unsigned long long)(m.wing_end[w] - m.wing_start[w]);
        const unsigned long long start = (unsigned long long)s.start_index;
        unsigned long long end = n - (unsigned long long)s.end_offset - 1ull;  // wraps as usize does
        end = std::max(end, start + 1);
        end = std::min(end, n - 1);
        if (start >= end || end > n)
            throw std::runtime_error("flow measurement: empty strip range for wing " + std::to_string(w));
        const double* x = values.data() + m.wing_start[w] + start;
        const size_t len = (size_t)(end - start);
        double r = x[0];
        if (s.type == 0) {  // stormath/src/statistics.rs:11-24
            for (size_t i = 1; i < len; ++i) r = r + x[i];
            r = r / (double)len;
        } else if (s.type == 1) {  // :43-53
            for (size_t i = 1; i < len; ++i)
                if (x[i] > r) r = x[i];
        } else {  // :56-66
            for (size_t i = 1; i < len; ++i)
                if (x[i] < r) r = x[i];
        }
        out[w] = r;
    }
    return out;
}

struct ControllerInput {  // controller/input.rs:21-30
    double loading, wing_angle, internal_state, aoa, velocity, direction;
};

// set_points.rs:28-46
double limit_value(double old_value, double raw_new_value, double max_change) {
    double diff = raw_new_value - old_value;
    return fabs(diff) > max_change ? old_value + max_change * signum(diff) : raw_new_value;
}
double wrap_angle(double a) {  // set_points.rs:188-200
    while (a > PI) a -= TAU;
    while (a < -PI) a += TAU;
    return a;
}

// set_points.rs:70-186
void controller_output(const SetPointsCfg& sp, const ControllerInput& in, double dt, double& angle, double& state) {
    angle = 0.0;
    if (sp.has_aoa) {
        double set_point = in.loading * interp(in.direction, sp.directions, sp.aoa, "angle_of_attack_data");
        if (sp.use_effective) {
            double err = wrap_angle(in.aoa - set_point);
            angle = wrap_angle(in.wing_angle + err);
        } else {
            angle = in.direction - set_point;
        }
    }
    if (sp.has_max_angle_rate) angle = limit_value(in.wing_angle, angle, sp.max_angle_rate * dt);
    state = 0.0;
    if (sp.has_state) {
        double raw = in.loading * interp(in.direction, sp.directions, sp.state, "section_model_internal_state_data");
        state = raw;
        if (sp.spin_ratio) {  // :12-25
            double circumference = PI * sp.diameter;
            double rps = raw * in.velocity / circumference;
            state = fabs(rps) > sp.max_rps ? sp.max_rps * signum(rps) : rps;
        }
    }
    if (sp.has_max_state_rate) state = limit_value(in.internal_state, state, sp.max_state_rate * dt);
}

}  // namespace

namespace sb_host {

WindField freestream_at_ctrl_points(const WindEnvironmentCfg& e, const WindConditionCfg& c, double ship_velocity,
                                    double time, const LineModel& m, V3* out) {
    WindField w = wind_field(e, c, ship_velocity, time);
    for (int i = 0; i < m.n(); ++i) out[i] = apparent_wind_at(w, m.ctrl_points[i]);
    if (e.has_inflow) {
        std::vector<std::pair<long, long>> ranges;
        for (int wing = 0; wing < m.nw(); ++wing) ranges.push_back({m.wing_start[wing], m.wing_end[wing]});
        apply_inflow_corrections(e, apparent_wind_direction(e, w, 10.0), out, m.ctrl_points.data(), ranges);
    }
    return w;
}

}  // namespace sb_host

struct sb_csm {
    std::unique_ptr<sb_sim> sim;
    WindEnvironmentCfg env;
    ControllerCfg ctrl;
    std::vector<V3> fs_ctrl;

    void create(const char* json, int dev, int md) {
        CompleteSailModelCfg c = parse_complete_sail_model(json);
        env = c.env;
        ctrl = c.controller;
        sim = std::make_unique<sb_sim>();
        sim->create(std::move(c.sim), dev, md);
        fs_ctrl.assign(sim->N, V3{0.0, 0.0, 0.0});
    }

    const LineModel& lfm() const { return sim->setup.lfm; }

    // Freestream of every point into the simulation's buffer (on its stream).
    void freestream(const WindConditionCfg& wc, double ship_velocity, double time) {
        sim->setup.lfm.update_global();
        WindField w = freestream_at_ctrl_points(env, wc, ship_velocity, time, lfm(), fs_ctrl.data());
        check(hipSetDevice(sim->device), "hipSetDevice");
        check(hipMemcpyAsync(sim->d_fs, fs_ctrl.data(), sizeof(V3) * fs_ctrl.size(), hipMemcpyHostToDevice,
                             sim->stream),
              "upload freestream");
        check(launch_wind_points(sim->inst(), w, sim->stream), "wind at wake points");
    }

    // complete_sail_model/mod.rs:174-191
    void step(double time, double dt, const WindConditionCfg& wc, double ship_velocity, sb_step_result* out) {
        freestream(wc, ship_velocity, time);
        sim->step(time, dt, nullptr, (size_t)(sim->N + sim->nr_points), true, out);
    }

    // environment.rs:297-327 for strip i's velocity v
    double strip_direction(int i, V3 v) const {
        const LineModel& m = lfm();
        int first = m.wing_start[m.wing_of[i]];
        V3 axis = normalize(m.global[first].rel());
        V3 chord = m.rbm.transform_vector(m.chord_local_not_rotated[first]);
        return signed_angle_between(chord, v, axis);
    }

    // controller/input.rs:91-157
    std::vector<ControllerInput> input_from_velocity(double loading, const std::vector<V3>& v) const {
        const LineModel& m = lfm();
        const int n = m.n();
        std::vector<double> speed(n), aoa(n), dir(n);
        for (int i = 0; i < n; ++i) {
            speed[i] = len(v[i]);
            aoa[i] = len(v[i]) > MIN_POSITIVE ? signed_angle_between(m.chord_global[i], v[i], m.global[i].dir()) : 0.0;
            dir[i] = strip_direction(i, v[i]);
        }
        return inputs(loading, measure(aoa, m, ctrl.aoa), measure(speed, m, ctrl.velocity),
                      measure(dir, m, ctrl.direction));
    }

    // controller/input.rs:33-88 with results/simulation.rs:118-149
    std::vector<ControllerInput> input_from_result(double loading, const double* aoa, const double* velocity,
                                                   const double* input_velocity, const double* ctrl_points) const {
        const LineModel& m = lfm();
        const int n = m.n();
        std::vector<double> a(aoa, aoa + n), speed(n), dir(n);
        const double* rel = ctrl.use_input_velocity ? input_velocity : velocity;
        for (int i = 0; i < n; ++i) {
            speed[i] = len(v3(velocity[3 * i], velocity[3 * i + 1], velocity[3 * i + 2]));
            V3 p = v3(ctrl_points[3 * i], ctrl_points[3 * i + 1], ctrl_points[3 * i + 2]);
            V3 r = v3(rel[3 * i], rel[3 * i + 1], rel[3 * i + 2]) + cross(m.rbm.velocity_angular, p - m.rbm.translation);
            dir[i] = strip_direction(i, r);
        }
        return inputs(loading, measure(a, m, ctrl.aoa), measure(speed, m, ctrl.velocity),
                      measure(dir, m, ctrl.direction));
    }

    std::vector<ControllerInput> inputs(double loading, const std::vector<double>& aoa, const std::vector<double>& speed,
                                        const std::vector<double>& dir) const {
        const LineModel& m = lfm();
        std::vector<ControllerInput> in(m.nw());
        for (int w = 0; w < m.nw(); ++w)  // Foil sections have no internal state (data_access.rs:101-117)
            in[w] = ControllerInput{loading, m.wing_angles[w], 0.0, aoa[w], speed[w], dir[w]};
        return in;
    }

    // controller/mod.rs:25-60 and data_update.rs:160-177.  The controller's
    // time-step index stays 0 (update takes &self), so every call at or after
    // start_time updates and the first-step branch never runs.
    void update(double time, double dt, const std::vector<ControllerInput>& in) {
        if (ctrl.steps_between_updates == 0)
            throw std::runtime_error("controller: time_steps_between_updates is 0 (the reference divides by it)");
        if (!(time >= ctrl.start_time)) return;
        LineModel& m = sim->setup.lfm;
        std::vector<double> angles;
        for (size_t k = 0; k < ctrl.set_points.size(); ++k) {
            if (k >= in.size()) throw std::runtime_error("controller: more set points than sails");
            double angle, state;
            controller_output(ctrl.set_points[k], in[k], dt, angle, state);
            angles.push_back(angle);
        }
        for (size_t k = 0; k < angles.size(); ++k) m.wing_angles[k] = angles[k];
        m.update_global();
    }

    void apply_controller_wind(double time, double dt, const WindConditionCfg& wc, double ship_velocity,
                               double loading) {  // complete_sail_model/mod.rs:193-211, 257-288
        sim->setup.lfm.update_global();
        freestream_at_ctrl_points(env, wc, ship_velocity, time, lfm(), fs_ctrl.data());
        update(time, dt, input_from_velocity(loading, fs_ctrl));
    }

    // complete_sail_model/mod.rs:99-122
    void steady_state(const WindConditionCfg& wc, double ship_velocity, double loading, sb_step_result* out) {
        apply_controller_wind(0.0, 1.0, wc, ship_velocity, loading);
        step(0.0, 1.0, wc, ship_velocity, out);
    }

    // complete_sail_model/mod.rs:58-97
    double optimal_loading(const WindConditionCfg& wc, double ship_velocity, long n, sb_step_result* out) {
        if (n < 1) throw std::invalid_argument("at least one loading to test");
        const double dx = (1.0 - 0.1) / (double)(n - 1);  // stormath array_generation::linspace
        double x = 0.1;
        std::vector<std::vector<double>> kept;
        double best = -__builtin_inf();
        long best_i = 0;
        double best_loading = 0.1;
        for (long i = 0; i < n; ++i) {
            steady_state(wc, ship_velocity, x, nullptr);
            const double* h = sim->h_out.data();
            double fx = 0.0;
            for (int w = 0; w < sim->W; ++w) fx += h[OutLayout::integ_f(sim->N) + 3L * (5L * sim->W + w)];
            double thrust = -fx;
            double effective = thrust * ship_velocity - 0.0;  // input power of Foil sections is zero
            if (effective > best) {
                best = effective;
                best_i = i;
                best_loading = x;
            }
            kept.emplace_back(sim->h_out.begin(), sim->h_out.end());
            x += dx;
        }
        sim->h_out.assign(kept[best_i].begin(), kept[best_i].end());
        sim->write_result(0.0, out);
        return best_loading;
    }
};

namespace {

WindConditionCfg cstormbird_condition(const WindCondition& c) {  // interfaces/cstormbird/src/wind.rs:13-33
    WindConditionCfg w;
    w.direction_coming_from = c.direction_coming_from;
    w.profile = VP_POWER;
    w.p0 = c.velocity;
    w.p1 = 10.0;
    w.p2 = 1.0 / 9.0;
    return w;
}

}  // namespace

extern "C" {

// ---- the reference's C interface (cstormbird) --------------------------------
CompleteSailModel* complete_sail_model_new(const char* setup_string) {
    if (!setup_string) return nullptr;
    auto m = std::make_unique<sb_csm>();
    if (guarded([&] { m->create(setup_string, 0, MODE_EXACT); }) != 0) return nullptr;
    return reinterpret_cast<CompleteSailModel*>(m.release());
}

int32_t complete_sail_model_get_number_of_sails(CompleteSailModel* sail_model) {
    if (!sail_model) return -1;
    return (int32_t)reinterpret_cast<sb_csm*>(sail_model)->sim->W;
}

void complete_sail_model_drop(CompleteSailModel* sail_model) { delete reinterpret_cast<sb_csm*>(sail_model); }

int32_t complete_sail_model_simulate_steady_state_condition(CompleteSailModel* sail_model, WindCondition wind_condition,
                                                            double ship_velocity, double controller_loading,
                                                            SingleSailResult* results_out, size_t results_out_length) {
    if (!sail_model || !results_out) return -1;
    if (results_out_length == 0) return -2;
    sb_csm* m = reinterpret_cast<sb_csm*>(sail_model);
    int rc = guarded([&] {
        m->steady_state(cstormbird_condition(wind_condition), ship_velocity, controller_loading, nullptr);
    });
    if (rc != 0) return rc;
    const int W = m->sim->W, N = m->sim->N;
    const double* h = m->sim->h_out.data();
    size_t count = std::min((size_t)W, results_out_length);
    for (size_t w = 0; w < count; ++w) {  // results/simulation.rs:150-166
        for (int c = 0; c < 3; ++c) {
            results_out[w].force[c] = h[OutLayout::integ_f(N) + 3L * (5L * W + w) + c];
            results_out[w].moment[c] = h[OutLayout::integ_m(N, W) + 3L * (5L * W + w) + c];
        }
        results_out[w].input_power = 0.0;
    }
    return (int32_t)count;
}

// ---- the rest of CompleteSailModel ------------------------------------------
sb_csm* sb_csm_create_ex(const char* setup_json, int device, int mode) {
    if (!setup_json) {
        g_error = "null setup string";
        return nullptr;
    }
    if (mode != MODE_EXACT && mode != MODE_FAST && mode != MODE_FP32_KAHAN) {
        g_error = "unknown mode";
        return nullptr;
    }
    auto m = std::make_unique<sb_csm>();
    if (guarded([&] { m->create(setup_json, device, mode); }) != 0) return nullptr;
    return m.release();
}

int sb_csm_setup_info(const char* setup_json, long* out, size_t cap) {
    if (!setup_json || !out) return -1;
    if (cap < 9) return -3;
    return guarded([&] {
        CompleteSailModelCfg c = parse_complete_sail_model(setup_json);
        setup_counts(c.sim, out);
        out[7] = (long)c.controller.set_points.size();
        out[8] = c.env.has_inflow ? 1 : 0;
    });
}

void sb_csm_destroy(sb_csm* m) { delete m; }

sb_sim* sb_csm_simulation(sb_csm* m) { return m ? m->sim.get() : nullptr; }

int sb_csm_do_step(sb_csm* m, double time, double time_step, const char* wind_condition_json, double ship_velocity,
                   sb_step_result* out) {
    if (!m || !wind_condition_json) return -1;
    return guarded([&] { m->step(time, time_step, parse_wind_condition(wind_condition_json), ship_velocity, out); });
}

int sb_csm_restart(sb_csm* m) {
    if (!m) return -1;
    m->sim->first_done = false;  // complete_sail_model/mod.rs:152
    return 0;
}

int sb_csm_apply_controller_wind(sb_csm* m, double time, double time_step, const char* wind_condition_json,
                                 double ship_velocity, double loading) {
    if (!m || !wind_condition_json) return -1;
    return guarded([&] {
        m->apply_controller_wind(time, time_step, parse_wind_condition(wind_condition_json), ship_velocity, loading);
    });
}

int sb_csm_apply_controller_result(sb_csm* m, double time, double time_step, double loading,
                                   const double* angles_of_attack, const double* velocity,
                                   const double* input_velocity, const double* ctrl_points) {
    if (!m || !angles_of_attack || !velocity || !input_velocity || !ctrl_points) return -1;
    return guarded([&] {
        m->update(time, time_step,
                  m->input_from_result(loading, angles_of_attack, velocity, input_velocity, ctrl_points));
    });
}

int sb_csm_simulate_steady_state(sb_csm* m, const char* wind_condition_json, double ship_velocity, double loading,
                                 sb_step_result* out) {
    if (!m || !wind_condition_json) return -1;
    return guarded([&] { m->steady_state(parse_wind_condition(wind_condition_json), ship_velocity, loading, out); });
}

int sb_csm_simulate_optimal_loading(sb_csm* m, const char* wind_condition_json, double ship_velocity,
                                    long nr_loadings, sb_step_result* out, double* best_loading) {
    if (!m || !wind_condition_json) return -1;
    return guarded([&] {
        double b = m->optimal_loading(parse_wind_condition(wind_condition_json), ship_velocity, nr_loadings, out);
        if (best_loading) *best_loading = b;
    });
}

int sb_csm_freestream(sb_csm* m, const char* wind_condition_json, double ship_velocity, double time, double* out,
                      size_t n) {
    if (!m || !wind_condition_json || !out) return -1;
    if ((long)n != m->sim->N + m->sim->nr_points) return -3;
    return guarded([&] {
        m->freestream(parse_wind_condition(wind_condition_json), ship_velocity, time);
        check(hipMemcpyAsync(out, m->sim->d_fs, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, m->sim->stream),
              "download");
        check(hipStreamSynchronize(m->sim->stream), "freestream");
    });
}

// WindEnvironment queries (environment.rs:66-278)
int sb_wind_velocity(const char* environment_json, const char* wind_condition_json, int kind, const double* locations,
                     size_t n, const double* linear_velocity, double time, double* out) {
    if (!environment_json || !wind_condition_json || (n && (!locations || !out))) return -1;
    if (kind < 0 || kind > 5) {
        g_error = "unknown wind query kind";
        return -4;
    }
    return guarded([&] {
        WindEnvironmentCfg e = parse_wind_environment(environment_json);
        WindField w = wind_field(e, parse_wind_condition(wind_condition_json), 0.0, time);
        if (linear_velocity) w.linear = v3(linear_velocity[0], linear_velocity[1], linear_velocity[2]);
        for (size_t i = 0; i < n; ++i) {
            V3 p = v3(locations[3 * i], locations[3 * i + 1], locations[3 * i + 2]);
            if (kind >= 4) {  // :85-104
                double h = wind_height(w, p);
                out[i] = kind == 4 ? profile_at_height(w, h) : unsteady_parallel_at_height(w, h);
                continue;
            }
            V3 v = kind == 0   ? apparent_wind_at(w, p)            // :182-190
                   : kind == 1 ? steady_true_wind(w, p) + w.linear  // :164-180
                   : kind == 2 ? true_wind_at(w, p)                 // :106-149
                               : steady_true_wind(w, p);            // :151-162
            out[3 * i] = v.x;
            out[3 * i + 1] = v.y;
            out[3 * i + 2] = v.z;
        }
    });
}

int sb_wind_apparent_direction(const char* environment_json, const char* wind_condition_json,
                               const double* linear_velocity, double height, double* out) {
    if (!environment_json || !wind_condition_json || !linear_velocity || !out) return -1;
    return guarded([&] {
        WindEnvironmentCfg e = parse_wind_environment(environment_json);
        WindField w = wind_field(e, parse_wind_condition(wind_condition_json), 0.0, 0.0);
        w.linear = v3(linear_velocity[0], linear_velocity[1], linear_velocity[2]);
        *out = apparent_wind_direction(e, w, height);
    });
}

int sb_wind_ctrl_points_velocity(const char* environment_json, const char* wind_condition_json,
                                 const double* ctrl_points, size_t n, const double* linear_velocity, double time,
                                 const long* wing_ranges, size_t nr_wings, double* out) {
    if (!environment_json || !wind_condition_json || !linear_velocity || (n && (!ctrl_points || !out)) ||
        (nr_wings && !wing_ranges))
        return -1;
    return guarded([&] {  // environment.rs:192-223
        WindEnvironmentCfg e = parse_wind_environment(environment_json);
        WindField w = wind_field(e, parse_wind_condition(wind_condition_json), 0.0, time);
        w.linear = v3(linear_velocity[0], linear_velocity[1], linear_velocity[2]);
        std::vector<V3> p(n), v(n);
        double average_height = 0.0;
        for (size_t i = 0; i < n; ++i) {
            p[i] = v3(ctrl_points[3 * i], ctrl_points[3 * i + 1], ctrl_points[3 * i + 2]);
            v[i] = apparent_wind_at(w, p[i]);
            average_height += dot(p[i], e.up);
        }
        average_height /= (double)n;
        std::vector<std::pair<long, long>> ranges;
        for (size_t k = 0; k < nr_wings; ++k) {
            if (wing_ranges[2 * k] < 0 || wing_ranges[2 * k + 1] < wing_ranges[2 * k] ||
                wing_ranges[2 * k + 1] > (long)n)
                throw std::invalid_argument("wing index range outside the control points");
            ranges.push_back({wing_ranges[2 * k], wing_ranges[2 * k + 1]});
        }
        apply_inflow_corrections(e, apparent_wind_direction(e, w, average_height), v.data(), p.data(), ranges);
        for (size_t i = 0; i < n; ++i) {
            out[3 * i] = v[i].x;
            out[3 * i + 1] = v[i].y;
            out[3 * i + 2] = v[i].z;
        }
    });
}

int sb_wind_condition_values(const char* wind_condition_json, int which, const double* heights, size_t n, double time,
                             double* out) {
    if (!wind_condition_json || (n && (!heights || !out))) return -1;
    if (which < 0 || which > 4) {
        g_error = "unknown wind condition value";
        return -4;
    }
    return guarded([&] {  // wind_condition/mod.rs:37-61, logarithmic_model.rs:52-74
        WindField w = wind_field(WindEnvironmentCfg{}, parse_wind_condition(wind_condition_json), 0.0, time);
        for (size_t i = 0; i < n; ++i) {
            double h = heights[i];
            out[i] = which == 0   ? profile_at_height(w, h)
                     : which == 1 ? unsteady_parallel_at_height(w, h)
                     : which == 2 ? w.perpendicular
                     : which == 3 ? w.vertical
                                  : (w.profile == VP_LOGARITHMIC ? businger_dyer_unscaled(w, h) : 0.0);
        }
    });
}

}  // extern "C"
