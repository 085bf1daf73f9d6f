// Host-side state of one simulation on the GPU (struct sb_sim) and the helpers
// shared by the single-instance and batched C ABI implementations.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/stormbird_gpu.h"
#include "sb_device.h"
#include "sb_model.h"

namespace sb_host {

using namespace sb;

extern thread_local std::string g_error;



struct HipError : std::runtime_error {
    int code;
    HipError(const std::string& m, int c) : std::runtime_error(m), code(c) {}
};

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e), (int)e);
}

// Device inputs handed over by pointer are written by the caller on the null
// stream (the default stream of most callers, torch's included), and the
// library's own streams (hipStreamNonBlocking) do not synchronize with it: work
// that reads such an input first waits, on the device, for everything the null
// stream holds so far.  The host does not block.
inline void wait_for_null_stream(hipStream_t st, hipEvent_t ev) {
    check(hipEventRecord(ev, nullptr), "event on the null stream");
    check(hipStreamWaitEvent(st, ev, 0), "wait for the null stream");
}

// complete_sail_model/mod.rs:215-255, control-point part: the wind field of
// condition c at `time` for a ship at `ship_velocity` (environment.rs:73-190),
// and the unsteady apparent wind at the control points of m (global positions)
// with the environment's inflow corrections at the apparent wind direction of
// the 10 m reference height.  The wake points take the returned field on the
// device (k_wind_points, kb_wind).
WindField freestream_at_ctrl_points(const WindEnvironmentCfg& e, const WindConditionCfg& c, double ship_velocity,
                                    double time, const LineModel& m, V3* out);

template <class T>
T* dalloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    check(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
    return static_cast<T*>(p);
}

// Host memory in pinned pages (hipHostMalloc) for the per-step transfers: the
// results go back by DMA without the runtime's staging copies, and the geometry and
// small freestreams are read from it by copy kernels.
template <class T>
struct PinnedAlloc {
    using value_type = T;
    PinnedAlloc() = default;
    template <class U>
    PinnedAlloc(const PinnedAlloc<U>&) {}
    T* allocate(size_t n) {
        void* p = nullptr;
        check(hipHostMalloc(&p, (n ? n : 1) * sizeof(T), hipHostMallocPortable), "hipHostMalloc");
        return static_cast<T*>(p);
    }
    void deallocate(T* p, size_t) noexcept { (void)hipHostFree(p); }
    template <class U>
    bool operator==(const PinnedAlloc<U>&) const { return true; }
    template <class U>
    bool operator!=(const PinnedAlloc<U>&) const { return false; }
};
using PinnedVec = std::vector<double, PinnedAlloc<double>>;

inline const char* kTimerNames[] = {"wake_update", "fixed", "factors", "solve", "self_induction", "forces"};
constexpr int kTimers = 6;


}  // namespace sb_host

struct sb_sim;

namespace sb_host {
// Wake export (sb_export.cpp), in the formats of wake/dynamic_wake/export.rs.
std::string rust_f64(double v);
void write_wake_obj(const sb_sim& s, const std::string& path);
void write_wake_vtk(const sb_sim& s, const std::string& path);
void write_wake_step_file(const sb_sim& s, double time, double dt);
}  // namespace sb_host

using namespace sb;
using namespace sb_host;

// Limits of this library on a setup the reference accepts (applied at create and by
// sb_setup_info, so that a setup the check accepts also builds).
inline void check_limits(const Setup& s) {
    if (s.lfm.n() > 1024) throw std::runtime_error("more than 1024 strips per simulation is not supported");
}

struct sb_sim {
    Setup setup;
    int device = 0;
    int mode = MODE_EXACT;
    hipStream_t stream = nullptr;
    int N = 0, W = 0, P = 0, M = 0;
    long nr_points = 0, nr_panels = 0;
    int nz_rows = 0;
    long steps_completed = 0;
    bool first_done = false;
    GeoLayout gl{};

    // device buffers
    double *d_pts[2] = {nullptr, nullptr}, *d_vel = nullptr, *d_str[2] = {nullptr, nullptr};
    double *d_core = nullptr, *d_geo = nullptr, *d_fs = nullptr, *d_felt = nullptr, *d_felt_span = nullptr;
    double *d_fixed = nullptr, *d_F = nullptr, *d_prev = nullptr, *d_hist = nullptr, *d_scal = nullptr;
    double *d_out = nullptr, *d_tmp = nullptr;
    double *d_lin = nullptr, *d_lin_gamma = nullptr;  // Linearized solve scratch
    double* d_segs = nullptr;  // ring-edge geometry per panel, when the free wake is self-induced
    double hs_core = 0.0;                             // quasi-steady horseshoe core, per step
    size_t tmp_cap = 0;
    PanelRec* d_rec = nullptr;
    int *d_col = nullptr, *d_wing = nullptr, *d_span_src = nullptr;
    FoilParams* d_foil = nullptr;
    McBuffers mc, tmp_mc;
    FoldBuffers fb;
    bool balanced = true;
    int cur = 0, scur = 0;

    PinnedVec h_geo, h_out;
    PinnedVec h_fs;                // host freestreams up to 1 MiB, read by the upload kernel
    bool fs_copy_pending = false;  // an upload kernel may still read h_fs (cleared once the step has ended)
    bool geo_copy_pending = false;  // the same for h_geo
    hipEvent_t ev[kTimers + 1] = {};
    hipEvent_t ev_input = nullptr;  // orders device inputs written on the null stream
    double pairs[3] = {0, 0, 0};
    mutable double timer_ms[kTimers] = {0};
    mutable bool timers_stale = false;  // the step's event intervals are read when first asked for
    bool timers_on = false;             // event records between the step's kernels (sb_sim_set_kernel_timers)
    bool last_step_timed = false;

    // An interval boundary of the per-kernel timers; no event is recorded when they
    // are off (the seven records cost ~25 us of a 0.18 ms step).
    void mark(int k) {
        if (timers_on) check(hipEventRecord(ev[k], stream), "event");
    }

    void read_timers() const {
        if (!timers_stale) return;
        static const int order[kTimers][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 4}, {4, 5}, {5, 6}};
        for (int k = 0; k < kTimers; ++k) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev[order[k][0]], ev[order[k][1]]);
            timer_ms[k] = ms;
        }
        timers_stale = false;
    }

    ~sb_sim() {
        (void)hipSetDevice(device);
        void* bufs[] = {d_pts[0], d_pts[1], d_vel, d_str[0], d_str[1], d_core, d_geo, d_fs, d_felt,
                        d_felt_span, d_fixed, d_F, d_prev, d_hist, d_scal, d_out, d_tmp, d_rec,
                        d_col, d_wing, d_span_src, d_foil, mc.partial, mc.flags, tmp_mc.partial, tmp_mc.flags,
                        fb.terms, fb.acc, d_lin, d_lin_gamma, d_segs, mc.tmap};
        for (void* p : bufs)
            if (p) (void)hipFree(p);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (ev_input) (void)hipEventDestroy(ev_input);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Self-induction targets in tiles of the structured wake grid (McBuffers::tmap): a wave
    // holds `ws` consecutive span points of 64 / ws consecutive rows, a job four such waves
    // of consecutive row groups; jobs run row band by row band.  ws = 0: flat order (the
    // default).  Measured at four sails x 40 strips x 2,000 rows from the rolled-up wake,
    // EXACT: flat 595 ms, tiles of 4 / 8 / 16 / 32 / 64 span points 637 / 692 / 647 / 704 /
    // 755 ms.  The waves of the flat order already meet the near field almost whole (the
    // ring branch runs in about as many wave slots as near pairs need), while a tile's
    // longer streamwise extent lets more panel tiles reach its bounding box and be staged.
    static constexpr int kDefaultTargetSpan = 0;
    void set_target_tiling(int ws) {
        if (ws != 0 && (ws < 1 || ws > 64 || 64 % ws != 0)) throw std::invalid_argument("span points per wave: 0 or a divisor of 64");
        const long end = setup.wake.end_index < nr_points ? setup.wake.end_index : nr_points;
        if (mc.tmap) {
            check(hipFree(mc.tmap), "hipFree");
            mc.tmap = nullptr;
        }
        mc.tmap_jobs = 0;
        mc.tmap_span = 0;
        if (ws == 0 || end <= 0 || P <= 0) return;
        const long wr = 64 / ws;
        const long rows = (end + P - 1) / P;
        const long span_bands = (P + ws - 1) / ws;
        const long row_bands = (rows + 4 * wr - 1) / (4 * wr);
        const long jobs = span_bands * row_bands;
        std::vector<int> map(jobs * 256);
        for (long rb = 0; rb < row_bands; ++rb)
            for (long sb = 0; sb < span_bands; ++sb)
                for (int l = 0; l < 256; ++l) {
                    const long row = rb * 4 * wr + (l / 64) * wr + (l % 64) / ws;
                    const long span = sb * ws + l % ws;
                    const long idx = row * P + span;
                    long v = idx;
                    if (row >= rows || span >= P || idx >= end) {  // padding: a neighbour, not stored
                        const long rr = row < rows ? row : rows - 1, ss = span < P ? span : P - 1;
                        const long c = rr * P + ss < end ? rr * P + ss : end - 1;
                        v = -(c + 1);
                    }
                    map[(rb * span_bands + sb) * 256 + l] = (int)v;
                }
        if (balanced && jobs > mc.cap_jobs) {
            if (mc.partial) check(hipFree(mc.partial), "hipFree");
            if (mc.flags) check(hipFree(mc.flags), "hipFree");
            mc.partial = nullptr;
            mc.flags = nullptr;
            mc.cap_jobs = jobs;
            mc.partial = dalloc<double>(mc.cap_jobs * 256 * 3);
            mc.flags = dalloc<long long>(mc.cap_jobs);
            check(hipMemsetAsync(mc.flags, 0xff, sizeof(long long) * mc.cap_jobs, stream), "memset");
            mc.epoch = 0;
        }
        mc.tmap = dalloc<int>(map.size());
        check(hipMemcpyAsync(mc.tmap, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice, stream), "upload");
        check(hipStreamSynchronize(stream), "sync");
        mc.tmap_jobs = jobs;
        mc.tmap_span = ws;
    }

    Inst inst() const {
        Inst d{};
        const WakeConfig& w = setup.wake;
        const SolverConfig& s = setup.solver;
        d.N = N;
        d.W = W;
        d.P = P;
        d.M = M;
        d.nz_rows = nz_rows;
        d.symmetry = setup.quasi ? setup.qs.symmetry : w.symmetry;
        d.neglect_self = w.neglect_self ? 1 : 0;
        d.first_dir = w.first_dir;
        d.steps_completed = (int)steps_completed;
        d.output_body = setup.lfm.output_body ? 1 : 0;
        d.viscous_lift_in_circulation = setup.lfm.viscous_lift_in_circulation ? 1 : 0;
        d.vc_mode = s.velocity_correction;
        d.max_iterations = s.max_iterations;
        d.end_index = w.end_index;
        d.far_ratio = w.far_ratio;
        d.first_rel = w.first_rel;
        d.last_rel = w.last_rel;
        d.damping = w.damping;
        d.rep_chord = w.rep_chord;
        d.solver_damping = s.damping;
        d.residual_tol = s.residual_tol;
        d.strength_tol = s.strength_tol;
        d.vc_ratio = s.velocity_correction_ratio;
        d.correction_factor = s.correction_factor_initial;
        d.density = setup.lfm.density;
        d.quasi = setup.quasi ? 1 : 0;
        d.solver_kind = s.kind;
        d.start_linearized = s.start_linearized ? 1 : 0;
        d.lin_disable_viscous = s.disable_viscous_corrections ? 1 : 0;
        d.lin_ivc = s.induced_velocity_correction;
        d.hs_wake_length = setup.qs.wake_length_factor * w.rep_chord;  // frozen_wake.rs:78-81
        d.hs_core = hs_core;
        d.lin = d_lin;
        d.lin_gamma = d_lin_gamma;
        d.pts = d_pts[cur];
        d.pts_next = d_pts[cur ^ 1];
        d.vel = d_vel;
        d.str = d_str[scur];
        d.str_next = d_str[scur ^ 1];
        d.core = d_core;
        d.corners = nullptr;
        d.rec = d_rec;
        d.segs = d_segs;
        d.col = d_col;
        d.wing_of = d_wing;
        d.span_src = d_span_src;
        d.foil = d_foil;
        d.geo = d_geo;
        d.gl = gl;
        d.fs = d_fs;
        d.felt = d_felt;
        d.felt_span = d_felt_span;
        d.fixed = d_fixed;
        d.F = d_F;
        d.prev_gamma = d_prev;
        d.hist = d_hist;
        d.scal = d_scal;
        d.out = d_out;
        return d;
    }

    void layout() {
        int o = 0;
        auto take = [&](int n) { int r = o; o += n; return r; };
        gl.ctrl = take(3 * N);
        gl.rel_g = take(3 * N);
        gl.dir_g = take(3 * N);
        gl.chord_g = take(3 * N);
        gl.chord_len = take(N);
        gl.local_len = take(N);
        gl.rel_local = take(3 * N);
        gl.rel_out = take(3 * N);
        gl.dir_out = take(3 * N);
        gl.ctrl_out = take(3 * N);
        gl.chord_out = take(3 * N);
        gl.is_virtual = take(N);
        gl.span_pts = take(3 * P);
        gl.end_g = take(3 * N);
        gl.chord_span = take(3 * P);
        gl.rbm = take(18);
        gl.body_axes = take(9);
        gl.total = o;
    }

    // Fixed velocities of the single-instance step: per-pair terms of up to
    // 8 Mi doubles (64 MiB) per chunk, allocated at the first step.
    void ensure_fold_buffers() {
        if (fb.terms) return;
        long ld = (N + 63) / 64 * 64;
        long need = 3L * ld * (long)N * (M > 1 ? M - 1 : 1);
        fb.cap = need < (8L << 20) ? need : (8L << 20);
        fb.terms = dalloc<double>(fb.cap);
        fb.cap_targets = N;
        fb.acc = dalloc<double>(3L * N);
    }

    // Packs the geometry into `dst` (gl.total doubles), or into h_geo and
    // uploads it to d_geo when dst is null.
    void pack_geometry(double* dst = nullptr) {
        const LineModel& m = setup.lfm;
        if (!dst) {
            if (geo_copy_pending) check(hipStreamSynchronize(stream), "previous geometry upload");
            h_geo.assign(gl.total, 0.0);
        }
        double* g = dst ? dst : h_geo.data();
        if (dst) std::fill(dst, dst + gl.total, 0.0);
        auto put = [](double* p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; };
        for (int i = 0; i < N; ++i) {
            const SpanLine& G = m.global[i];
            const SpanLine& L = m.local[i];
            const SpanLine& O = m.output_body ? L : G;
            put(g + gl.ctrl + 3 * i, m.ctrl_points[i]);
            put(g + gl.rel_g + 3 * i, G.rel());
            put(g + gl.dir_g + 3 * i, G.dir());
            put(g + gl.chord_g + 3 * i, m.chord_global[i]);
            g[gl.chord_len + i] = m.chord_len[i];
            g[gl.local_len + i] = L.length();
            put(g + gl.rel_local + 3 * i, L.rel());
            put(g + gl.rel_out + 3 * i, O.rel());
            put(g + gl.dir_out + 3 * i, O.dir());
            put(g + gl.ctrl_out + 3 * i, O.ctrl());
            put(g + gl.chord_out + 3 * i, m.output_body ? m.chord_local[i] : m.chord_global[i]);
            g[gl.is_virtual + i] = m.is_virtual[i] ? 1.0 : 0.0;
            put(g + gl.end_g + 3 * i, G.end);
        }
        for (int j = 0; j < P; ++j) {
            put(g + gl.span_pts + 3 * j, m.span_points[j]);
            put(g + gl.chord_span + 3 * j, m.chord_global_at_span[j]);
        }
        put(g + gl.rbm, m.rbm.translation);
        put(g + gl.rbm + 3, m.rbm.rotation);
        put(g + gl.rbm + 6, m.rbm.velocity_linear);
        put(g + gl.rbm + 9, m.rbm.velocity_angular);
        put(g + gl.body_axes, m.rbm.rotate(v3(1, 0, 0)));
        put(g + gl.body_axes + 3, m.rbm.rotate(v3(0, 1, 0)));
        put(g + gl.body_axes + 6, m.rbm.rotate(v3(0, 0, 1)));
        if (!dst) {  // read over the bus by a copy kernel, as the host freestream
            void* dp = nullptr;
            check(hipHostGetDevicePointer(&dp, h_geo.data(), 0), "pinned geometry");
            geo_copy_pending = true;
            check(launch_copy_doubles((const double*)dp, d_geo, (long)gl.total, stream), "upload geometry");
        }
    }

    void create(const char* json, int dev, int md) { create(parse_setup(json), dev, md); }

    void create(Setup parsed, int dev, int md) {
        setup = std::move(parsed);
        device = dev;
        mode = md;
        check(hipSetDevice(device), "hipSetDevice");
        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
        check(hipEventCreateWithFlags(&ev_input, hipEventDisableTiming), "hipEventCreate");
        const LineModel& m = setup.lfm;
        N = m.n();
        W = m.nw();
        P = (int)m.span_points.size();
        M = setup.wake.M;
        check_limits(setup);
        nr_points = setup.quasi ? 0 : (long)P * (M + 1);  // quasi-steady: no wake points
        nr_panels = (long)N * M;
        layout();
        d_pts[0] = dalloc<double>(3 * nr_points);
        d_pts[1] = dalloc<double>(3 * nr_points);
        d_vel = dalloc<double>(3 * nr_points);
        d_str[0] = dalloc<double>(nr_panels);
        d_str[1] = dalloc<double>(nr_panels);
        d_core = dalloc<double>(nr_panels);
        d_rec = dalloc<PanelRec>(nr_panels);
        // ring-edge geometry: where near pairs are many (a self-induced free wake)
        if (!setup.quasi && setup.wake.end_index > 0) d_segs = dalloc<double>(NEARD * nr_panels);
        d_geo = dalloc<double>(gl.total);
        d_fs = dalloc<double>(3 * (N + nr_points));
        d_felt = dalloc<double>(3 * N);
        d_felt_span = dalloc<double>(3 * P);
        d_fixed = dalloc<double>(3 * N);
        d_F = dalloc<double>(3L * N * N);
        d_prev = dalloc<double>(N);
        d_hist = dalloc<double>(6 * N);
        d_scal = dalloc<double>(8);
        d_out = dalloc<double>(OutLayout::total(N, W));
        d_col = dalloc<int>(N);
        d_wing = dalloc<int>(N);
        d_span_src = dalloc<int>(2 * P);
        d_foil = dalloc<FoilParams>(W);
        if (setup.solver.kind == 1 || setup.solver.start_linearized) {
            if (!linearized_in_lds(N)) d_lin = dalloc<double>((size_t)linearized_global_doubles(N));
            d_lin_gamma = dalloc<double>(N);
        }
        // SB_PAIR_SCHEDULE=plain keeps one workgroup per 256 targets for the pair
        // sums (no load balancing); the default balances over the resident workgroups.
        const char* sched = getenv("SB_PAIR_SCHEDULE");
        balanced = !(sched && strcmp(sched, "plain") == 0);
        mc.cap_jobs = balanced ? (nr_points + 255) / 256 : 0;
        if (mc.cap_jobs > 0) {
            mc.partial = dalloc<double>(mc.cap_jobs * 256 * 3);
            mc.flags = dalloc<long long>(mc.cap_jobs);
            check(hipMemsetAsync(mc.flags, 0xff, sizeof(long long) * mc.cap_jobs, stream), "memset");
        }
        if (!setup.quasi && setup.wake.end_index > 0) set_target_tiling(kDefaultTargetSpan);
        std::vector<int> col(N), span_src(2 * P);
        for (int i = 0; i < N; ++i) col[i] = i + m.wing_of[i];
        int j = 0;
        for (int w = 0; w < W; ++w) {  // value_mapping.rs:87-137 without extrapolation
            span_src[2 * j] = span_src[2 * j + 1] = m.wing_start[w];
            ++j;
            for (int i = m.wing_start[w]; i < m.wing_end[w]; ++i, ++j) {
                span_src[2 * j] = i;
                span_src[2 * j + 1] = (i == m.wing_end[w] - 1) ? i : i + 1;
            }
        }
        check(hipMemcpyAsync(d_col, col.data(), sizeof(int) * N, hipMemcpyHostToDevice, stream), "upload");
        check(hipMemcpyAsync(d_wing, m.wing_of.data(), sizeof(int) * N, hipMemcpyHostToDevice, stream), "upload");
        check(hipMemcpyAsync(d_span_src, span_src.data(), sizeof(int) * 2 * P, hipMemcpyHostToDevice, stream), "upload");
        check(hipMemcpyAsync(d_foil, m.foils.data(), sizeof(FoilParams) * W, hipMemcpyHostToDevice, stream), "upload");
        if (nr_panels > 0)
            check(hipMemcpyAsync(d_core, setup.wake.core.data(), sizeof(double) * nr_panels, hipMemcpyHostToDevice,
                                 stream),
                  "upload");
        h_out.assign(OutLayout::total(N, W), 0.0);
        reset_state();
    }

    // The state of a newly built simulation: straight chord-based wake, zero
    // strengths, no completed step (the line-force model's motion is kept).
    void reset_state() {
        check(hipSetDevice(device), "hipSetDevice");
        const LineModel& m = setup.lfm;
        first_done = false;
        steps_completed = 0;
        nz_rows = 0;
        cur = 0;
        scur = 0;
        check(hipMemsetAsync(d_prev, 0, sizeof(double) * N, stream), "memset");
        check(hipMemsetAsync(d_hist, 0, sizeof(double) * 6 * N, stream), "memset");
        check(hipMemsetAsync(d_scal, 0, sizeof(double) * 8, stream), "memset");
        check(hipMemsetAsync(d_str[1], 0, sizeof(double) * nr_panels, stream), "memset");
        check(hipMemsetAsync(d_fixed, 0, sizeof(double) * 3 * N, stream), "memset");
        pack_geometry();
        if (setup.quasi) {  // no shed wake
            check(hipStreamSynchronize(stream), "reset");
            return;
        }
        // DynamicWakeBuilder::build -> initialize_based_on_chord_length (initialization.rs:46-71)
        V3 sum = v3(0, 0, 0);
        for (const V3& c : m.chord_global_at_span) sum = sum + c;
        V3 avg = sum / (double)m.chord_global_at_span.size();
        double wl = setup.wake.initial_relative_wake_length * len(avg);
        V3 vel = (wl / (double)M) * avg;
        double v[3] = {vel.x, vel.y, vel.z};
        Inst d = inst();
        check(launch_init_wake(d, 1.0, 0, v, stream), "init wake");
        check(launch_panels(d, d.pts, 0, stream), "panels");
        check(hipStreamSynchronize(stream), "reset");
    }

    void upload_freestream(const double* fs, size_t n, bool on_device) {
        size_t bytes = sizeof(double) * 3 * n;
        if (on_device) wait_for_null_stream(stream, ev_input);
        // Host freestreams up to 1 MiB are copied into pinned memory and read over the
        // bus by a copy kernel on the step's stream: for the 101 KB of a single wing sail
        // this measured faster than the runtime's copy of the pageable array, and a DMA
        // copy from pinned memory measured ~0.2 ms slower.  Larger arrays take the
        // runtime's copy.  The pinned buffer is reused once the previous step has ended.
        if (!on_device && bytes <= (1u << 20)) {
            if (fs_copy_pending) check(hipStreamSynchronize(stream), "previous step");
            h_fs.resize(3 * n);
            std::memcpy(h_fs.data(), fs, bytes);
            void* dp = nullptr;
            check(hipHostGetDevicePointer(&dp, h_fs.data(), 0), "pinned freestream");
            fs_copy_pending = true;
            check(launch_copy3((const double*)dp, d_fs, (long)n, stream), "upload freestream");
            return;
        }
        check(hipMemcpyAsync(d_fs, fs, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream),
              "upload freestream");
    }

    // fs == nullptr: the freestream is already in d_fs (written by the caller's
    // kernels on this stream).
    void step(double time, double dt, const double* fs, size_t n, bool fs_on_device, sb_step_result* out) {
        check(hipSetDevice(device), "hipSetDevice");
        if ((long)n != N + nr_points) throw std::invalid_argument("freestream length mismatch");
        setup.lfm.update_global();
        pack_geometry();
        if (fs) upload_freestream(fs, n, fs_on_device);
        if (setup.quasi) update_horseshoe_core();
        Inst d = inst();
        bool first = !first_done;
        mark(0);
        check(launch_prepare(d, first ? 1 : 0, stream), "prepare");
        if (setup.quasi) {
            quasi_steady_step(d, time, dt, out);
            return;
        }
        if (first) {
            check(launch_init_wake(d, dt, 1, nullptr, stream), "init wake");
            first_done = true;
            nz_rows = 0;
        }
        // update_before_solving (update_data.rs:28-42)
        check(launch_wake_step(d, dt, stream), "wake step");
        cur ^= 1;
        d = inst();
        check(launch_panels(d, d.pts, 1, stream), "panels");
        scur ^= 1;
        nz_rows = nz_rows == 0 ? 0 : (nz_rows + 1 < M ? nz_rows + 1 : M);
        d = inst();
        mark(1);
        // FrozenWake::update_from_full_wake (frozen_wake.rs:189-261)
        ensure_fold_buffers();
        check(launch_fixed(d, mode, &fb, stream), "fixed velocities");
        mark(2);
        check(launch_factors(d, stream), "factors");
        mark(3);
        // Solver::solve (SimpleIterative or Linearized), then update_wing_strength
        launch_solver(d);
        nz_rows = nz_rows < 1 ? 1 : nz_rows;
        d = inst();
        // update_velocity_at_points (update_data.rs:125-141)
        check(launch_velocity_copy(d, stream), "velocity copy");
        mark(4);
        long end = d.end_index < nr_points ? d.end_index : nr_points;
        pairs[0] = 0;
        if (end > 0 && steps_completed > 2) {
            check(launch_self_induction(d, mode, &mc, stream), "self induction");
            pairs[0] = (double)end * (double)N * (double)nz_rows;
        }
        steps_completed += 1;
        pairs[1] = (double)N * (double)N * (double)(nz_rows > 1 ? nz_rows - 1 : 0);
        pairs[2] = (double)N * (double)N;
        mark(5);
        check(launch_forces(d, dt, stream), "forces");
        mark(6);
        finish_step(time, out);
        if (setup.wake.write_files) write_wake_step_file(*this, time, dt);  // simulation.rs:181-186
    }

    // frozen_wake.rs:84-99: Relative core = factor * mean global span-line length
    void update_horseshoe_core() {
        const QuasiConfig& q = setup.qs;
        if (q.core_kind == 0) {
            double sum = 0.0;
            for (int i = 0; i < N; ++i) sum += setup.lfm.global[i].length();
            hs_core = q.core_value * (sum / (double)N);
        } else {
            hs_core = q.core_kind == 1 ? q.core_value : 0.0;
        }
    }

    void launch_solver(const Inst& d) {
        if (setup.solver.kind == 1) {
            check(launch_linearized(d, 0, stream), "linearized solve");
            return;
        }
        if (setup.solver.start_linearized) check(launch_linearized(d, 1, stream), "linearized start");
        check(launch_solve(d, stream), "solve");
    }

    // simulation.rs:107-203 with WakeData::QuasiSteady: horseshoe factors from
    // the felt span-point velocities, solve, forces (no wake to update).
    void quasi_steady_step(const Inst& d, double time, double dt, sb_step_result* out) {
        first_done = true;
        mark(1);
        mark(2);
        check(launch_horseshoe_factors(d, stream), "horseshoe factors");
        mark(3);
        launch_solver(d);
        mark(4);
        mark(5);
        check(launch_forces(d, dt, stream), "forces");
        mark(6);
        steps_completed += 1;
        pairs[0] = 0;
        pairs[1] = 0;
        pairs[2] = (double)N * (double)N;
        finish_step(time, out);
    }

    void finish_step(double time, sb_step_result* out) {
        check(hipMemcpyAsync(h_out.data(), d_out, sizeof(double) * h_out.size(), hipMemcpyDeviceToHost, stream),
              "download");
        check(hipStreamSynchronize(stream), "step");
        fs_copy_pending = geo_copy_pending = false;
        timers_stale = timers_on;
        last_step_timed = timers_on;
        const double singular = h_out[OutLayout::singular(N, W)];  // the Linearized solve's flag
        if (singular != 0.0) {  // linalg.rs:77-81, unwrapped at linearized.rs:139
            long row = (long)singular - 1;
            check(hipMemsetAsync(d_scal + 4, 0, sizeof(double), stream), "memset");
            check(hipStreamSynchronize(stream), "step");
            throw std::runtime_error("Linearized solver: matrix is singular or nearly singular at row " +
                                     std::to_string(row));
        }
        write_result(time, out);
    }

    void write_result(double time, sb_step_result* o) const {
        if (!o) return;
        const double* h = h_out.data();
        o->time = time;
        o->iterations = (long)h[OutLayout::iters(N, W)];
        o->residual = h[OutLayout::residual(N, W)];
        auto cp = [](double* dst, const double* src, size_t n) {
            if (dst) memcpy(dst, src, sizeof(double) * n);
        };
        cp(o->circulation_strength, h + OutLayout::gamma(N), N);
        cp(o->angles_of_attack, h + OutLayout::alpha(N), N);
        cp(o->ctrl_points_velocity, h + OutLayout::u(N), 3 * N);
        cp(o->solver_input_velocity, h + OutLayout::felt(N), 3 * N);
        cp(o->ctrl_points, h + OutLayout::ctrl(N), 3 * N);
        cp(o->sectional_forces, h + OutLayout::sect(N), 18 * N);
        cp(o->integrated_forces, h + OutLayout::integ_f(N), 18 * W);
        cp(o->integrated_moments, h + OutLayout::integ_m(N, W), 18 * W);
    }

    // Near-field and total pair counts of the free-wake self-induction at the
    // current state (the sum update_velocity_at_points would evaluate).
    void count_pairs(double* near_total) {
        check(hipSetDevice(device), "hipSetDevice");
        long end = setup.wake.end_index < nr_points ? setup.wake.end_index : nr_points;
        long blocks = 0;
        double* buf = tmp(2 * ((end + 255) / 256) + 2);
        Inst d = inst();
        check(launch_count_pairs(d, d.pts, end, (long)N * nz_rows, buf, &blocks, stream), "count pairs");
        std::vector<double> h(2 * blocks);
        if (blocks)
            check(hipMemcpyAsync(h.data(), buf, sizeof(double) * 2 * blocks, hipMemcpyDeviceToHost, stream), "download");
        check(hipStreamSynchronize(stream), "count pairs");
        double a = 0.0, b = 0.0;
        for (long q = 0; q < blocks; ++q) {
            a += h[2 * q];
            b += h[2 * q + 1];
        }
        near_total[0] = a;
        near_total[1] = b;
    }

    double* tmp(size_t n) {
        if (n > tmp_cap) {
            if (d_tmp) (void)hipFree(d_tmp);
            d_tmp = dalloc<double>(n);
            tmp_cap = n;
        }
        return d_tmp;
    }

    void induced(const double* pts, size_t n, double* out) {
        check(hipSetDevice(device), "hipSetDevice");
        if (setup.quasi)  // simulation.rs:207-216
            throw std::runtime_error("Induced velocities not implemented for pure quasi-steady simulations yet");
        if (n == 0) return;
        double* buf = tmp(6 * n);
        check(hipMemcpyAsync(buf, pts, sizeof(double) * 3 * n, hipMemcpyHostToDevice, stream), "upload points");
        Inst d = inst();
        McBuffers* m = (long)((n + 255) / 256) <= mc.cap_jobs ? &mc : nullptr;
        if (!m && n > 2048 && balanced) {  // larger than the wake: scratch sized for n
            if (!tmp_mc.partial || tmp_mc.cap_jobs < (long)((n + 255) / 256)) {
                if (tmp_mc.partial) (void)hipFree(tmp_mc.partial);
                if (tmp_mc.flags) (void)hipFree(tmp_mc.flags);
                tmp_mc.cap_jobs = (long)((n + 255) / 256);
                tmp_mc.partial = dalloc<double>(tmp_mc.cap_jobs * 256 * 3);
                tmp_mc.flags = dalloc<long long>(tmp_mc.cap_jobs);
                check(hipMemsetAsync(tmp_mc.flags, 0xff, sizeof(long long) * tmp_mc.cap_jobs, stream), "memset");
                tmp_mc.epoch = 0;
            }
            m = &tmp_mc;
        }
        check(launch_induced_targets(d, buf, (long)n, (long)N * nz_rows, mode, buf + 3 * n, m, stream), "induced");
        check(hipMemcpyAsync(out, buf + 3 * n, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, stream), "download");
        check(hipStreamSynchronize(stream), "induced");
    }
};

namespace sb_host {

// Sizes of a parsed setup: strips, wings, span points, wake rows, freestream
// points, wake panels, 1 for quasi-steady.
inline void setup_counts(const Setup& s, long* out) {
    check_limits(s);
    const long N = s.lfm.n(), W = s.lfm.nw(), P = (long)s.lfm.span_points.size(), M = s.wake.M;
    const long points = s.quasi ? 0 : P * (M + 1);
    out[0] = N;
    out[1] = W;
    out[2] = P;
    out[3] = M;
    out[4] = N + points;
    out[5] = N * M;
    out[6] = s.quasi ? 1 : 0;
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const HipError& e) {
        g_error = e.what();
        return -5;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return -3;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -4;
    }
}

}  // namespace sb_host
