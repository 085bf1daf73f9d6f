// Host orchestration of the dynamic lifting-line step on one GPU, and the C ABI
// declared in include/stormbird_gpu.h.
//
// Step sequence (stormbird/src/lifting_line/simulation.rs:107-203):
//   host  : line-force-model geometry update (O(N)), packed and uploaded
//   device: felt velocities -> [first step: wake initialisation] -> wake
//           shedding/convection -> panel rebuild + strength shift -> fixed
//           velocities + unit factors (frozen wake) -> fixed-point solve ->
//           wake velocities (freestream + free-wake self-induction) -> forces
//   host  : one copy of the packed results into the caller's arrays
#include "sb_host.h"

namespace sb_host {
thread_local std::string g_error;
}  // namespace sb_host

extern "C" {

const char* sb_last_error(void) { return g_error.c_str(); }

sb_sim* sb_sim_create_ex(const char* json, int device, int mode) {
    if (!json) {
        g_error = "null setup string";
        return nullptr;
    }
    if (mode != MODE_EXACT && mode != MODE_FAST && mode != MODE_FP32_KAHAN) {
        g_error = "unsupported mode";
        return nullptr;
    }
    auto sim = std::make_unique<sb_sim>();
    try {
        sim->create(json, device, mode);
    } catch (const std::exception& e) {
        g_error = e.what();
        return nullptr;
    }
    return sim.release();
}

int sb_setup_info(const char* json, long* out, size_t cap) {
    if (!json || !out) return -1;
    if (cap < 7) return -3;
    return guarded([&] { setup_counts(parse_setup(json), out); });
}

sb_sim* sb_sim_create(const char* json) { return sb_sim_create_ex(json, 0, MODE_EXACT); }

void sb_sim_destroy(sb_sim* sim) { delete sim; }

long sb_sim_nr_ctrl_points(const sb_sim* s) { return s ? s->N : -1; }
long sb_sim_nr_wings(const sb_sim* s) { return s ? s->W : -1; }
long sb_sim_nr_span_points(const sb_sim* s) { return s ? s->P : -1; }
long sb_sim_nr_wake_points(const sb_sim* s) { return s ? s->nr_points : -1; }
long sb_sim_nr_panels(const sb_sim* s) { return s ? s->nr_panels : -1; }
long sb_sim_nr_freestream_points(const sb_sim* s) { return s ? s->N + s->nr_points : -1; }

long sb_sim_get_freestream_points(const sb_sim* s, double* xyz, size_t cap) {
    if (!s || !xyz) return -1;
    if (cap == 0) return -2;
    size_t total = (size_t)(s->N + s->nr_points);
    size_t n = cap < total ? cap : total;
    int rc = guarded([&] {
        std::vector<double> tmp(3 * total);
        const LineModel& m = s->setup.lfm;
        for (int i = 0; i < s->N; ++i) {
            tmp[3 * i] = m.ctrl_points[i].x;
            tmp[3 * i + 1] = m.ctrl_points[i].y;
            tmp[3 * i + 2] = m.ctrl_points[i].z;
        }
        check(hipSetDevice(s->device), "hipSetDevice");
        check(hipMemcpy(tmp.data() + 3 * s->N, s->d_pts[s->cur], sizeof(double) * 3 * s->nr_points,
                        hipMemcpyDeviceToHost),
              "download points");
        memcpy(xyz, tmp.data(), sizeof(double) * 3 * n);
    });
    return rc != 0 ? rc : (long)n;
}

int sb_sim_do_step(sb_sim* s, double time, double dt, const double* fs, size_t n, sb_step_result* out) {
    if (!s || !fs) return -1;
    return guarded([&] { s->step(time, dt, fs, n, false, out); });
}

int sb_sim_do_step_device(sb_sim* s, double time, double dt, const double* fs, size_t n, sb_step_result* out) {
    if (!s || !fs) return -1;
    return guarded([&] { s->step(time, dt, fs, n, true, out); });
}

int sb_sim_induced_velocities(sb_sim* s, const double* pts, size_t n, double* out) {
    if (!s || ((!pts || !out) && n)) return -1;
    return guarded([&] { s->induced(pts, n, out); });
}

int sb_sim_get_previous_circulation(const sb_sim* s, double* out, size_t n) {
    if (!s || !out) return -1;
    if ((long)n != s->N) return -3;
    return guarded([&] {
        check(hipSetDevice(s->device), "hipSetDevice");
        check(hipMemcpy(out, s->d_prev, sizeof(double) * n, hipMemcpyDeviceToHost), "download");
    });
}

int sb_sim_set_previous_circulation(sb_sim* s, const double* g, size_t n) {
    if (!s || !g) return -1;
    if ((long)n != s->N) return -3;
    return guarded([&] {
        check(hipSetDevice(s->device), "hipSetDevice");
        check(hipMemcpy(s->d_prev, g, sizeof(double) * n, hipMemcpyHostToDevice), "upload");
    });
}

int sb_sim_reset(sb_sim* s) {
    if (!s) return -1;
    return guarded([&] { s->reset_state(); });
}

int sb_sim_set_solver_settings(sb_sim* s, long max_iterations, double damping, double residual_tol,
                               double strength_tol) {
    if (!s) return -1;
    if (max_iterations < 1 || !(damping > 0.0) || damping > 1.0 || residual_tol < 0.0 || strength_tol < 0.0)
        return -4;
    s->setup.solver.max_iterations = (int)max_iterations;
    s->setup.solver.damping = damping;
    s->setup.solver.residual_tol = residual_tol;
    s->setup.solver.strength_tol = strength_tol;
    return 0;
}

int sb_sim_set_translation_and_rotation_fd(sb_sim* s, double dt, const double* t, const double* r) {
    if (!s || !t || !r) return -1;
    s->setup.lfm.rbm.update_translation_fd(v3(t[0], t[1], t[2]), dt);
    s->setup.lfm.rbm.update_rotation_fd(v3(r[0], r[1], r[2]), dt);
    s->setup.lfm.update_global();
    return 0;
}

int sb_sim_set_local_wing_angles(sb_sim* s, const double* a, size_t n) {
    if (!s || !a) return -1;
    if ((long)n != s->W) return -3;
    for (size_t i = 0; i < n; ++i) s->setup.lfm.wing_angles[i] = a[i];
    s->setup.lfm.update_global();
    return 0;
}

int sb_sim_get_local_wing_angles(const sb_sim* s, double* out, size_t n) {
    if (!s || !out) return -1;
    if ((long)n != s->W) return -3;
    for (size_t i = 0; i < n; ++i) out[i] = s->setup.lfm.wing_angles[i];
    return 0;
}

int sb_sim_set_target_tiling(sb_sim* s, int span_points_per_wave) {
    if (!s) return -1;
    if (span_points_per_wave != 0 &&
        (span_points_per_wave < 1 || span_points_per_wave > 64 || 64 % span_points_per_wave != 0))
        return -4;
    return guarded([&] {
        check(hipSetDevice(s->device), "hipSetDevice");
        s->set_target_tiling(span_points_per_wave);
    });
}

int sb_sim_get_target_tiling(const sb_sim* s) { return s ? s->mc.tmap_span : -1; }

int sb_sim_get_wing_ranges(const sb_sim* s, long* ranges, size_t cap) {
    if (!s || !ranges) return -1;
    if (cap < 2 * (size_t)s->W) return -3;
    for (int w = 0; w < s->W; ++w) {
        ranges[2 * w] = s->setup.lfm.wing_start[w];
        ranges[2 * w + 1] = s->setup.lfm.wing_end[w];
    }
    return 0;
}

int sb_sim_get_rigid_body_motion(const sb_sim* s, double* motion12) {
    if (!s || !motion12) return -1;
    const RigidBody& r = s->setup.lfm.rbm;
    const V3 v[4] = {r.translation, r.rotation, r.velocity_linear, r.velocity_angular};
    for (int k = 0; k < 4; ++k) {
        motion12[3 * k] = v[k].x;
        motion12[3 * k + 1] = v[k].y;
        motion12[3 * k + 2] = v[k].z;
    }
    return 0;
}

int sb_sim_get_wake_state(const sb_sim* s, double* pts, double* vel, double* str, long* steps, int* first) {
    if (!s) return -1;
    int rc = guarded([&] {
        check(hipSetDevice(s->device), "hipSetDevice");
        if (pts)
            check(hipMemcpy(pts, s->d_pts[s->cur], sizeof(double) * 3 * s->nr_points, hipMemcpyDeviceToHost),
                  "download");
        if (vel) check(hipMemcpy(vel, s->d_vel, sizeof(double) * 3 * s->nr_points, hipMemcpyDeviceToHost), "download");
        if (str) check(hipMemcpy(str, s->d_str[s->scur], sizeof(double) * s->nr_panels, hipMemcpyDeviceToHost), "download");
    });
    if (steps) *steps = s->steps_completed;
    if (first) *first = s->first_done ? 1 : 0;
    return rc;
}

int sb_sim_set_wake_state_ex(sb_sim* s, const double* pts, size_t n_pts, const double* vel, size_t n_vel,
                             const double* str, size_t n_str, long steps, int first) {
    if (!s) return -1;
    if ((pts && n_pts != (size_t)s->nr_points) || (vel && n_vel != (size_t)s->nr_points) ||
        (str && n_str != (size_t)s->nr_panels)) {
        g_error = "wake state sizes (" + std::to_string(n_pts) + ", " + std::to_string(n_vel) + ", " +
                  std::to_string(n_str) + ") do not match the simulation's (" + std::to_string(s->nr_points) +
                  " points, " + std::to_string(s->nr_panels) + " panels)";
        return -3;
    }
    return sb_sim_set_wake_state(s, pts, vel, str, steps, first);
}

int sb_sim_set_wake_state(sb_sim* s, const double* pts, const double* vel, const double* str, long steps,
                          int first) {
    if (!s) return -1;
    return guarded([&] {
        check(hipSetDevice(s->device), "hipSetDevice");
        if (pts)
            check(hipMemcpy(s->d_pts[s->cur], pts, sizeof(double) * 3 * s->nr_points, hipMemcpyHostToDevice), "upload");
        if (vel) check(hipMemcpy(s->d_vel, vel, sizeof(double) * 3 * s->nr_points, hipMemcpyHostToDevice), "upload");
        if (str) {
            check(hipMemcpy(s->d_str[s->scur], str, sizeof(double) * s->nr_panels, hipMemcpyHostToDevice), "upload");
            int nz = 0;
            for (long k = 0; k < s->nr_panels; ++k)
                if (str[k] != 0.0) nz = (int)(k / s->N) + 1;
            s->nz_rows = nz;
        }
        s->steps_completed = steps;
        if (first && !s->first_done) {
            s->first_done = true;
            check(hipMemset(s->d_prev, 0, sizeof(double) * s->N), "memset");
        }
        if (s->nr_panels > 0) {
            Inst d = s->inst();
            check(launch_panels(d, d.pts, 0, s->stream), "panels");
        }
        check(hipStreamSynchronize(s->stream), "sync");
    });
}

int sb_sim_get_panels(const sb_sim* s, double* rec) {
    if (!s || !rec) return -1;
    return guarded([&] {
        check(hipSetDevice(s->device), "hipSetDevice");
        std::vector<PanelRec> r(s->nr_panels);
        std::vector<double> p(3 * s->nr_points);
        std::vector<double> core(s->nr_panels);
        check(hipMemcpy(r.data(), s->d_rec, sizeof(PanelRec) * s->nr_panels, hipMemcpyDeviceToHost), "download");
        check(hipMemcpy(p.data(), s->d_pts[s->cur], sizeof(double) * p.size(), hipMemcpyDeviceToHost), "download");
        check(hipMemcpy(core.data(), s->d_core, sizeof(double) * s->nr_panels, hipMemcpyDeviceToHost), "download");
        const LineModel& m = s->setup.lfm;
        for (long k = 0; k < s->nr_panels; ++k) {
            long st = k / s->N, i = k % s->N;
            long a = st * s->P + i + m.wing_of[i], b = a + s->P;
            long idx[4] = {a, a + 1, b + 1, b};
            double* o = rec + 22 * k;
            for (int q = 0; q < 4; ++q)
                for (int c = 0; c < 3; ++c) o[3 * q + c] = p[3 * idx[q] + c];
            o[12] = r[k].cx;
            o[13] = r[k].cy;
            o[14] = r[k].cz;
            o[15] = r[k].nx;
            o[16] = r[k].ny;
            o[17] = r[k].nz;
            o[18] = r[k].far2;
            o[19] = core[k];
            o[20] = r[k].area_term;
            o[21] = r[k].strength;
        }
    });
}

int sb_sim_get_frozen_wake(const sb_sim* s, double* fixed, double* factors) {
    if (!s) return -1;
    return guarded([&] {
        check(hipSetDevice(s->device), "hipSetDevice");
        if (fixed) check(hipMemcpy(fixed, s->d_fixed, sizeof(double) * 3 * s->N, hipMemcpyDeviceToHost), "download");
        if (factors) {
            long N = s->N;
            std::vector<double> F(3 * N * N);
            check(hipMemcpy(F.data(), s->d_F, sizeof(double) * F.size(), hipMemcpyDeviceToHost), "download");
            for (long j = 0; j < N; ++j)
                for (long i = 0; i < N; ++i)
                    for (int c = 0; c < 3; ++c) factors[3 * (i * N + j) + c] = F[3 * (j * N + i) + c];
        }
    });
}

int sb_sim_set_kernel_timers(sb_sim* s, int on) {
    if (!s) return -1;
    s->timers_on = on != 0;
    return 0;
}

double sb_sim_last_kernel_ms(const sb_sim* s, const char* name) {
    if (!s || !name) return -1.0;
    bool known = false;
    for (int k = 0; k < kTimers; ++k) known = known || strcmp(name, kTimerNames[k]) == 0;
    if (!known) return -2.0;
    if (!s->last_step_timed) return -3.0;
    (void)hipSetDevice(s->device);
    s->read_timers();
    for (int k = 0; k < kTimers; ++k)
        if (strcmp(name, kTimerNames[k]) == 0) return s->timer_ms[k];
    return -2.0;
}

int sb_sim_last_pair_counts(const sb_sim* s, double* c) {
    if (!s || !c) return -1;
    for (int k = 0; k < 3; ++k) c[k] = s->pairs[k];
    return 0;
}

int sb_sim_count_self_induction_pairs(sb_sim* s, double* near_total) {
    if (!s || !near_total) return -1;
    return guarded([&] { s->count_pairs(near_total); });
}

int sb_wake_induced_velocities(const double* recs, size_t n_panels, const double* targets, size_t n_targets,
                               int symmetry, int mode, double* out) {
    if ((!recs && n_panels) || (!targets && n_targets) || (!out && n_targets)) return -1;
    if (n_targets == 0) return 0;
    if (symmetry < 0 || symmetry > 3 || (mode != MODE_EXACT && mode != MODE_FAST)) return -4;
    if (n_panels > (size_t)0x7fffffff) return -4;
    PanelRec* d_rec = nullptr;
    double *d_corners = nullptr, *d_core = nullptr, *d_t = nullptr;
    hipStream_t st = nullptr;
    int rc = 0;
    try {
        std::vector<PanelRec> r(n_panels);
        std::vector<double> corners(12 * n_panels), core(n_panels);
        for (size_t k = 0; k < n_panels; ++k) {
            const double* q = recs + 22 * k;
            memcpy(&corners[12 * k], q, sizeof(double) * 12);
            r[k] = PanelRec{q[12], q[13], q[14], q[15], q[16], q[17], q[18], q[20], q[21], -1.0};
            core[k] = q[19];
        }
        check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
        d_rec = dalloc<PanelRec>(n_panels);
        d_corners = dalloc<double>(12 * n_panels);
        d_core = dalloc<double>(n_panels);
        d_t = dalloc<double>(6 * n_targets);
        check(hipMemcpyAsync(d_rec, r.data(), sizeof(PanelRec) * n_panels, hipMemcpyHostToDevice, st), "upload");
        check(hipMemcpyAsync(d_corners, corners.data(), sizeof(double) * corners.size(), hipMemcpyHostToDevice, st),
              "upload");
        check(hipMemcpyAsync(d_core, core.data(), sizeof(double) * n_panels, hipMemcpyHostToDevice, st), "upload");
        check(hipMemcpyAsync(d_t, targets, sizeof(double) * 3 * n_targets, hipMemcpyHostToDevice, st), "upload");
        Inst d{};
        d.N = (int)(n_panels ? n_panels : 1);
        d.M = 1;
        d.P = d.N + 1;
        d.symmetry = symmetry;
        d.rec = d_rec;
        d.corners = d_corners;
        d.core = d_core;
        check(launch_induced_targets(d, d_t, (long)n_targets, (long)n_panels, mode, d_t + 3 * n_targets, nullptr, st),
              "induced");
        check(hipMemcpyAsync(out, d_t + 3 * n_targets, sizeof(double) * 3 * n_targets, hipMemcpyDeviceToHost, st),
              "download");
        check(hipStreamSynchronize(st), "sync");
    } catch (const std::exception& e) {
        g_error = e.what();
        rc = -5;
    }
    if (d_rec) (void)hipFree(d_rec);
    if (d_corners) (void)hipFree(d_corners);
    if (d_core) (void)hipFree(d_core);
    if (d_t) (void)hipFree(d_t);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

int sb_debug_libm_device(int fn, const double* x, const double* y, double* out, size_t n) {
    if (!x || !out || (fn == 4 && !y)) return -1;
    if (n == 0) return 0;
    double* d = nullptr;
    int rc = 0;
    try {
        d = dalloc<double>(3 * n);
        check(hipMemcpy(d, x, sizeof(double) * n, hipMemcpyHostToDevice), "upload");
        if (y) check(hipMemcpy(d + n, y, sizeof(double) * n, hipMemcpyHostToDevice), "upload");
        check(launch_libm(fn, d, d + n, d + 2 * n, (long)n, nullptr), "libm");
        check(hipMemcpy(out, d + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost), "download");
    } catch (const std::exception& e) {
        g_error = e.what();
        rc = -5;
    }
    if (d) (void)hipFree(d);
    return rc;
}

int sb_debug_linearized_in_shared_memory(long n_strips) {
    if (n_strips < 1 || n_strips > 1024) return -1;
    return linearized_in_lds((int)n_strips) ? 1 : 0;
}

int sb_debug_gauss_device(const double* A, const double* b, size_t n, double* x, int variant) {
    if (!A || !b || !x || n < 1 || n > 1024) return -1;
    double* d = nullptr;
    int rc = 0;
    try {
        d = dalloc<double>(n * n + 2 * n + 1);
        check(hipMemcpy(d, A, sizeof(double) * n * n, hipMemcpyHostToDevice), "upload");
        check(hipMemcpy(d + n * n, b, sizeof(double) * n, hipMemcpyHostToDevice), "upload");
        int* status = (int*)(d + n * n + 2 * n);
        hipError_t e = launch_gauss_dense(d, d + n * n, (int)n, d + n * n + n, status, variant, nullptr);
        if (e == hipErrorInvalidValue) {
            rc = -1;
        } else {
            check(e, "gauss");
            int st = 0;
            check(hipMemcpy(&st, status, sizeof(int), hipMemcpyDeviceToHost), "download");
            check(hipMemcpy(x, d + n * n + n, sizeof(double) * n, hipMemcpyDeviceToHost), "download");
            rc = st;
        }
    } catch (const std::exception& e) {
        g_error = e.what();
        rc = -2;
    }
    if (d) (void)hipFree(d);
    return rc;
}

int sb_debug_gauss_blocked_device(const double* A, const double* b, size_t n, double* x) {
    if (!A || !b || !x || n < 1 || n > 1024) return -1;
    double* d = nullptr;
    int rc = 0;
    try {
        const size_t lin = n * (n + 1) + 8 * n + (n + 1) / 2 + 1;  // matrix, row data, exchange sources
        d = dalloc<double>(lin + 8 + n + 1);
        std::vector<double> cm(n * (n + 1));
        for (size_t i = 0; i < n; ++i) {
            for (size_t j = 0; j < n; ++j) cm[j * n + i] = A[i * n + j];
            cm[n * n + i] = b[i];
        }
        check(hipMemcpy(d, cm.data(), sizeof(double) * cm.size(), hipMemcpyHostToDevice), "upload");
        double* scal = d + lin;
        check(hipMemset(scal, 0, sizeof(double) * 8), "memset");
        int* status = (int*)(scal + 8 + n);
        check(launch_gauss_blocked(d, scal, (int)n, scal + 8, status, nullptr), "gauss blocked");
        int st = 0;
        check(hipMemcpy(&st, status, sizeof(int), hipMemcpyDeviceToHost), "download");
        check(hipMemcpy(x, scal + 8, sizeof(double) * n, hipMemcpyDeviceToHost), "download");
        rc = st;
    } catch (const std::exception& e) {
        g_error = e.what();
        rc = -2;
    }
    if (d) (void)hipFree(d);
    return rc;
}

int sb_debug_libm_host(int fn, const double* x, const double* y, double* out, size_t n) {
    if (!x || !out || (fn == 4 && !y)) return -1;
    for (size_t i = 0; i < n; ++i) {
        switch (fn) {
            case 0: out[i] = sbm::sin(x[i]); break;
            case 1: out[i] = sbm::cos(x[i]); break;
            case 2: out[i] = sbm::exp(x[i]); break;
            case 3: out[i] = sbm::acos(x[i]); break;
            case 4: out[i] = sbm::pow(x[i], y[i]); break;
            case 6: out[i] = sbm::atan(x[i]); break;
            default: out[i] = sbm::log(x[i]); break;
        }
    }
    return 0;
}

}  // extern "C"
