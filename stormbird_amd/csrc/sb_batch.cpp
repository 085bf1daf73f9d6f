// Many simulations stepped together on one GPU: the batched form of
// Simulation::do_step (stormbird/src/lifting_line/simulation.rs:107-203).
//
// Each instance keeps its own wake buffers (an sb_sim created from its own
// setup string); the batch adds one slab per per-step transfer (geometry,
// freestream, packed results), the instance descriptors of the four stages of
// a step, and the job lists of the pair sums.  A step is one host update of
// every instance's line-force model, three uploads, fourteen launches over all
// instances and one download.  From the second step on, the launches and the
// transfers between pinned staging buffers and the device are replayed from a
// captured HIP graph, re-captured only when the launch shapes change (the
// free-wake self-induction starting, or a different job count).
//
// Every instance computes exactly what its own single-instance step computes,
// bit for bit: the kernel bodies are shared, and the batched pair sums keep
// each target's ascending panel order.
//
// A batch of quasi-steady instances (no shed wake) runs its own short step:
// prepare, horseshoe factors, one-workgroup Linearized solve (and/or the
// fixed-point solve), forces.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <exception>
#include <fstream>
#include <numeric>
#include <sstream>
#include <thread>

#include "sb_host.h"
#include "sb_json.h"
#include "sb_pool.h"

namespace {

template <class T>
T* halloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    check(hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
    return static_cast<T*>(p);
}

constexpr int kStages = 4;  // before the wake step, after it, after the strength shift, after the solve

using sb_pool::parallel_for;  // per-instance host work, which touches only that instance

}  // namespace

struct sb_batch {
    std::vector<std::unique_ptr<sb_sim>> sims;
    int device = 0, mode = MODE_EXACT, n = 0;
    hipStream_t stream = nullptr;
    std::vector<long> geo_off, fs_off, out_off;  // doubles
    long geo_total = 0, fs_total = 0, out_total = 0;
    double *d_geo = nullptr, *d_fs = nullptr, *d_out = nullptr;
    double *h_geo = nullptr, *h_out = nullptr;
    Inst *h_inst = nullptr, *d_inst = nullptr;  // [kStages][n]
    long* d_off = nullptr;                      // [BK_KINDS][n + 1]
    int2 *h_jobs = nullptr, *d_jobs = nullptr;  // fixed jobs then target jobs
    long cap_fixed_jobs = 0, cap_target_jobs = 0;
    unsigned* d_counters = nullptr;
    BatchDev bd;
    bool first_done = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long graph_sig[3] = {-1, -1, -1};
    hipEvent_t ev[2] = {};
    hipEvent_t ev_input = nullptr;  // orders device inputs written on the null stream
    double last_ms = 0.0;
    int use_graph = 1;
    // learned warm start
    sb_mlp* mlp = nullptr;
    int warm_mode = 0;          // 0 off, 1 first step only, 2 every step, 3 every step judged
                                // (plus 4: the prediction is the change from the previous circulation)
    bool record_features = false;
    long n_strips = 0;
    std::vector<long> h_strip_off;  // [n + 1]
    long* d_strip_off = nullptr;
    float *d_feat = nullptr, *d_scale = nullptr, *d_pred = nullptr;
    unsigned *d_fallbacks = nullptr, *h_fallbacks = nullptr;
    bool last_applied = false;
    // quasi-steady batches
    bool quasi = false;
    int lin_variant = 0;
    size_t lin_shm = 0;
    double *d_singular = nullptr, *h_singular = nullptr;  // [n]
    // work of the last dynamic step, for the tuning key: pairs of the fixed-velocity and
    // self-induction sums as the kernels evaluate them
    double fixed_work = 0.0, target_work = 0.0;
    // wind steps: one wind environment, one wind condition per instance
    bool has_wind = false;
    WindEnvironmentCfg wind_env;
    std::vector<WindConditionCfg> wind_cond;
    long* d_wind_blk = nullptr;  // [n + 1] 256-point blocks of each instance's freestream
    long wind_blocks = 0;
    WindJob *h_wind = nullptr, *d_wind = nullptr;
    double *h_wind_ctrl = nullptr, *d_wind_ctrl = nullptr;  // control-point freestream [n_strips][3]

    ~sb_batch() {
        (void)hipSetDevice(device);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        void* dev[] = {d_geo, d_fs, d_out, d_inst, d_off, d_jobs, d_counters, d_strip_off, d_feat, d_scale, d_pred,
                       d_fallbacks, d_singular, d_wind_blk, d_wind, d_wind_ctrl};
        for (void* p : dev)
            if (p) (void)hipFree(p);
        void* host[] = {h_geo, h_out, h_inst, h_jobs, h_fallbacks, h_singular, h_wind, h_wind_ctrl};
        for (void* p : host)
            if (p) (void)hipHostFree(p);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (ev_input) (void)hipEventDestroy(ev_input);
        if (stream) (void)hipStreamDestroy(stream);
        sims.clear();
    }

    void create(const char* const* jsons, long count, int dev, int md) {
        if (count <= 0) throw std::invalid_argument("empty batch");
        device = dev;
        mode = md;
        n = (int)count;
        check(hipSetDevice(device), "hipSetDevice");
        for (long i = 0; i < count; ++i) {
            if (!jsons[i]) throw std::invalid_argument("null setup string in batch");
            auto s = std::make_unique<sb_sim>();
            s->create(jsons[i], dev, md);
            if (s->N > 1024) throw std::invalid_argument("batched instances support at most 1024 strips");
            if (!s->setup.quasi && (s->setup.solver.kind != 0 || s->setup.solver.start_linearized))
                throw std::invalid_argument(
                    "batched dynamic instances use the SimpleIterative solver without "
                    "start_with_linearized_solution");
            if (i > 0 && s->setup.quasi != sims[0]->setup.quasi)
                throw std::invalid_argument("a batch holds either quasi-steady or dynamic instances, not both");
            sims.push_back(std::move(s));
        }
        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
        check(hipEventCreateWithFlags(&ev_input, hipEventDisableTiming), "hipEventCreate");
        quasi = sims[0]->setup.quasi;
        bool any_lin = false;
        for (auto& s : sims) any_lin = any_lin || s->setup.solver.kind == 1 || s->setup.solver.start_linearized;
        if (any_lin) {
            std::vector<int> ns(n);
            int mx = 1;
            for (int i = 0; i < n; ++i) mx = std::max(mx, ns[i] = sims[i]->N);
            lin_variant = batch_linearized_variant(mx);
            lin_shm = lin_variant ? batch_linearized_shared(ns.data(), n, lin_variant) : 0;
            if (!lin_variant || !lin_shm)
                throw std::invalid_argument("batched Linearized solves need every instance's system in one "
                                            "workgroup's shared memory (" + std::to_string(mx) + " strips is too many)");
        }
        d_singular = dalloc<double>(n);
        h_singular = halloc<double>(n);
        geo_off.resize(n + 1);
        fs_off.resize(n + 1);
        out_off.resize(n + 1);
        geo_off[0] = fs_off[0] = out_off[0] = 0;
        int max_n = 1, max_w = 1, sym = 0;
        for (int i = 0; i < n; ++i) {
            const sb_sim& s = *sims[i];
            geo_off[i + 1] = geo_off[i] + s.gl.total;
            fs_off[i + 1] = fs_off[i] + 3L * (s.N + s.nr_points);
            out_off[i + 1] = out_off[i] + OutLayout::total(s.N, s.W);
            max_n = std::max(max_n, s.N);
            max_w = std::max(max_w, s.W);
            sym |= s.setup.wake.symmetry != 0 ? 1 : 0;
        }
        geo_total = geo_off[n];
        fs_total = fs_off[n];
        out_total = out_off[n];
        d_geo = dalloc<double>(geo_total);
        d_fs = dalloc<double>(fs_total);
        d_out = dalloc<double>(out_total);
        h_geo = halloc<double>(geo_total);
        h_out = halloc<double>(out_total);
        h_inst = halloc<Inst>((size_t)kStages * n);
        d_inst = dalloc<Inst>((size_t)kStages * n);

        // flat block offsets per kind (256-thread blocks)
        std::vector<long> off((size_t)BK_KINDS * (n + 1), 0);
        for (int k = 0; k < BK_KINDS; ++k) {
            long* o = off.data() + (long)k * (n + 1);
            for (int i = 0; i < n; ++i) {
                const sb_sim& s = *sims[i];
                long items = 0;
                switch (k) {
                    case BK_POINTS: items = (long)s.P * (s.M + 1); break;
                    case BK_SPAN: items = s.P; break;
                    case BK_PANELS: items = (long)s.N * s.M; break;
                    case BK_FACTORS: items = (long)s.N * s.N; break;
                    default: items = 3L * s.P * (s.M + 1); break;
                }
                o[i + 1] = o[i] + (items + 255) / 256;
            }
            bd.blocks[k] = o[n];
        }
        d_off = dalloc<long>(off.size());
        check(hipMemcpyAsync(d_off, off.data(), sizeof(long) * off.size(), hipMemcpyHostToDevice, stream), "upload");

        cap_fixed_jobs = 0;
        cap_target_jobs = 0;
        for (auto& s : sims) {
            cap_fixed_jobs += (s->N + 63) / 64;
            long np = (long)s->P * (s->M + 1);
            long end = std::min<long>(s->setup.wake.end_index, np);
            cap_target_jobs += (end + 255) / 256;
        }
        h_jobs = halloc<int2>(cap_fixed_jobs + cap_target_jobs);
        d_jobs = dalloc<int2>(cap_fixed_jobs + cap_target_jobs);
        d_counters = dalloc<unsigned>(2);
        check(hipMemsetAsync(d_counters, 0, 2 * sizeof(unsigned), stream), "memset");

        bd.n = n;
        bd.off = d_off;
        bd.max_threads = std::min(1024, ((max_n + 63) / 64) * 64);
        bd.solve_shm = solve_shared_bytes(max_n, max_w);
        bd.counters = d_counters;
        bd.sym = sym;
        bd.stage_near = 1;
        for (auto& s : sims)
            if (std::min<long>(s->setup.wake.end_index, (long)s->P * (s->M + 1)) > 0 && !s->d_segs) bd.stage_near = 0;
        bd.fixed_jobs = d_jobs;
        bd.target_jobs = d_jobs + cap_fixed_jobs;
        set_tuning(bd.fixed_per, bd.targets_tile);
        const char* g = getenv("SB_BATCH_GRAPH");
        use_graph = !(g && strcmp(g, "0") == 0);
        // per-strip arrays of the warm start
        std::vector<long> so(n + 1, 0);
        for (int i = 0; i < n; ++i) so[i + 1] = so[i] + sims[i]->N;
        n_strips = so[n];
        h_strip_off = so;
        d_strip_off = dalloc<long>(n + 1);
        check(hipMemcpyAsync(d_strip_off, so.data(), sizeof(long) * (n + 1), hipMemcpyHostToDevice, stream), "upload");
        d_feat = dalloc<float>(n_strips * MLP_FEATURES);
        d_scale = dalloc<float>(n_strips);
        d_pred = dalloc<float>(n_strips);
        d_fallbacks = dalloc<unsigned>(2);
        h_fallbacks = halloc<unsigned>(2);
        h_fallbacks[0] = h_fallbacks[1] = 0;
        bd.strip_off = d_strip_off;
        bd.n_strips = n_strips;
        bd.feat = d_feat;
        bd.scale = d_scale;
        bd.pred = d_pred;
        bd.fallbacks = d_fallbacks;
        check(hipStreamSynchronize(stream), "batch create");
    }

    // complete_sail_model/mod.rs:215-255 for every instance: parsed once, evaluated
    // each wind step.
    void set_wind(const char* environment_json, const char* const* conditions_json) {
        if (quasi) throw std::invalid_argument("wind steps run on dynamic batches");
        if (!conditions_json) throw std::invalid_argument("null wind condition list");
        WindEnvironmentCfg e = parse_wind_environment(environment_json ? environment_json : "{}");
        std::vector<WindConditionCfg> c;
        for (int i = 0; i < n; ++i) {
            if (!conditions_json[i]) throw std::invalid_argument("null wind condition string");
            c.push_back(parse_wind_condition(conditions_json[i]));
        }
        check(hipSetDevice(device), "hipSetDevice");
        if (!d_wind) {
            std::vector<long> blk(n + 1, 0);
            for (int i = 0; i < n; ++i) blk[i + 1] = blk[i] + (fs_off[i + 1] / 3 - fs_off[i] / 3 + 255) / 256;
            wind_blocks = blk[n];
            d_wind_blk = dalloc<long>(n + 1);
            check(hipMemcpyAsync(d_wind_blk, blk.data(), sizeof(long) * (n + 1), hipMemcpyHostToDevice, stream),
                  "upload");
            h_wind = halloc<WindJob>(n);
            d_wind = dalloc<WindJob>(n);
            h_wind_ctrl = halloc<double>(3 * n_strips);
            d_wind_ctrl = dalloc<double>(3 * n_strips);
            check(hipStreamSynchronize(stream), "wind setup");
        }
        wind_env = std::move(e);
        wind_cond = std::move(c);
        has_wind = true;
    }

    void reset() {
        check(hipSetDevice(device), "hipSetDevice");
        for (auto& s : sims) s->reset_state();
        first_done = false;
    }

    // Descriptor of instance i at the current host state, with the batch slabs.
    Inst stage_inst(int i, double dt) const {
        const sb_sim& s = *sims[i];
        Inst d = s.inst();
        d.geo = d_geo + geo_off[i];
        d.fs = d_fs + fs_off[i];
        d.out = d_out + out_off[i];
        d.dt = dt;
        return d;
    }

    // ship_velocities != nullptr: a wind step (fs unused), the freestream of every
    // instance from its wind condition at times[i] (set_wind).
    void step(const double* times, const double* dts, const double* fs, size_t n_fs, bool fs_on_device,
              sb_step_result* results, const double* ship_velocities = nullptr) {
        check(hipSetDevice(device), "hipSetDevice");
        const bool wind = ship_velocities != nullptr;
        if (wind && (!has_wind || !times))
            throw std::invalid_argument(!has_wind ? "no wind conditions set" : "a wind step needs the times");
        if (!wind && (long)n_fs != fs_total / 3) throw std::invalid_argument("freestream length mismatch");
        if (!dts) throw std::invalid_argument("null time steps");
        for (int i = 0; i < n; ++i) {
            if (sims[i]->first_done != first_done || sims[i]->steps_completed != sims[0]->steps_completed)
                throw std::invalid_argument("batched instances must be stepped together");
        }
        const bool first = !first_done;
        // host: line-force-model update and packed geometry of every instance
        parallel_for(n, [&](int i) {
            sims[i]->setup.lfm.update_global();
            sims[i]->pack_geometry(h_geo + geo_off[i]);
            if (wind) {
                const sb_sim& s = *sims[i];
                WindJob& j = h_wind[i];
                j.w = freestream_at_ctrl_points(wind_env, wind_cond[i], ship_velocities[i], times[i], s.setup.lfm,
                                                reinterpret_cast<V3*>(h_wind_ctrl + 3 * h_strip_off[i]));
                j.ctrl = d_wind_ctrl + 3 * h_strip_off[i];
                j.dst = d_fs + fs_off[i];
                j.n_ctrl = s.N;
                j.n_pts = (fs_off[i + 1] - fs_off[i]) / 3 - s.N;
            }
        });
        if (quasi) {
            if (wind) throw std::invalid_argument("wind steps run on dynamic batches");
            step_quasi(first, times, dts, fs, fs_on_device, results);
            return;
        }
        // stage descriptors (the host state transitions of sb_sim::step)
        for (int i = 0; i < n; ++i) {
            sb_sim& s = *sims[i];
            h_inst[0 * n + i] = stage_inst(i, dts[i]);
            if (first) s.nz_rows = 0;
            s.cur ^= 1;
            h_inst[1 * n + i] = stage_inst(i, dts[i]);
            s.scur ^= 1;
            s.nz_rows = s.nz_rows == 0 ? 0 : (s.nz_rows + 1 < s.M ? s.nz_rows + 1 : s.M);
            h_inst[2 * n + i] = stage_inst(i, dts[i]);
            s.nz_rows = s.nz_rows < 1 ? 1 : s.nz_rows;
            h_inst[3 * n + i] = stage_inst(i, dts[i]);
        }
        // pair-sum jobs: fixed velocities longest first; self-induction where it runs
        std::vector<std::pair<long, int2>> fj;
        fj.reserve(cap_fixed_jobs);
        long nt = 0;
        for (int i = 0; i < n; ++i) {
            const Inst& d = h_inst[2 * n + i];
            long len = (long)d.N * std::max(d.nz_rows - 1, 0);
            for (int tb = 0; tb < (d.N + 63) / 64; ++tb) fj.push_back({len, int2{i, tb}});
        }
        std::stable_sort(fj.begin(), fj.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
        for (size_t k = 0; k < fj.size(); ++k) h_jobs[k] = fj[k].second;
        for (int i = 0; i < n; ++i) {
            const sb_sim& s = *sims[i];
            const Inst& d = h_inst[3 * n + i];
            long np = (long)d.P * (d.M + 1);
            long end = std::min<long>(d.end_index, np);
            if (end > 0 && s.steps_completed > 2 && d.nz_rows > 0)
                for (long tb = 0; tb < (end + 255) / 256; ++tb) h_jobs[cap_fixed_jobs + nt++] = int2{i, (int)tb};
        }
        bd.n_fixed_jobs = (int)fj.size();
        bd.n_target_jobs = (int)nt;
        fixed_work = 0.0;
        for (const auto& f : fj) fixed_work += 64.0 * (double)f.first;
        target_work = 0.0;
        for (long k = 0; k < nt; ++k) {
            const Inst& d = h_inst[3 * n + h_jobs[cap_fixed_jobs + k].x];
            target_work += 256.0 * (double)d.N * (double)d.nz_rows;
        }

        // freestream into the slab (outside the graph: the caller's pointer may change)
        if (wind) {
            // the wake points where they are before this step (stage 0)
            for (int i = 0; i < n; ++i) h_wind[i].pts = h_inst[i].pts;
            check(hipMemcpyAsync(d_wind_ctrl, h_wind_ctrl, sizeof(double) * 3 * n_strips, hipMemcpyHostToDevice,
                                 stream),
                  "upload control-point wind");
            check(hipMemcpyAsync(d_wind, h_wind, sizeof(WindJob) * n, hipMemcpyHostToDevice, stream), "upload wind");
            check(launch_batch_wind(d_wind, d_wind_blk, n, wind_blocks, stream), "wind at every point");
        } else {
            if (fs_on_device) wait_for_null_stream(stream, ev_input);
            check(hipMemcpyAsync(d_fs, fs, sizeof(double) * fs_total,
                                 fs_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream),
                  "upload freestream");
        }
        check(hipEventRecord(ev[0], stream), "event");
        const int when = warm_mode & 3;
        const bool apply = mlp && (when >= 2 || (when == 1 && first));
        const bool feats = apply || record_features;
        last_applied = apply;
        long sig[3] = {first ? 1 : 0, bd.n_fixed_jobs, (bd.n_target_jobs * 8 + (apply ? warm_mode : 0)) * 2 + (feats ? 1 : 0)};
        if (first || !use_graph) {
            enqueue(first, apply, feats);
        } else {
            if (!exec || sig[0] != graph_sig[0] || sig[1] != graph_sig[1] || sig[2] != graph_sig[2]) {
                if (exec) (void)hipGraphExecDestroy(exec);
                if (graph) (void)hipGraphDestroy(graph);
                exec = nullptr;
                graph = nullptr;
                check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
                try {
                    enqueue(false, apply, feats);
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(stream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                check(hipStreamEndCapture(stream, &graph), "end capture");
                check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
                std::copy(sig, sig + 3, graph_sig);
            }
            check(hipGraphLaunch(exec, stream), "graph launch");
        }
        check(hipEventRecord(ev[1], stream), "event");
        check(hipStreamSynchronize(stream), "batch step");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
        last_ms = ms;
        if (first) first_done = true;
        parallel_for(n, [&](int i) {
            sb_sim& s = *sims[i];
            s.first_done = true;
            s.steps_completed += 1;
            std::copy(h_out + out_off[i], h_out + out_off[i + 1], s.h_out.begin());
            if (results) s.write_result(times ? times[i] : 0.0, results + i);
        });
        for (int i = 0; i < n; ++i)
            if (sims[i]->setup.wake.write_files) write_wake_step_file(*sims[i], times ? times[i] : 0.0, dts[i]);
    }

    // simulation.rs:107-203 with WakeData::QuasiSteady for every instance.
    void step_quasi(bool first, const double* times, const double* dts, const double* fs, bool fs_on_device,
                    sb_step_result* results) {
        if (mlp) throw std::invalid_argument("the learned warm start runs on dynamic batches");
        bool any_lin_start = false, any_lin_solver = false, any_iter = false;
        parallel_for(n, [&](int i) {
            sims[i]->update_horseshoe_core();
            h_inst[i] = stage_inst(i, dts[i]);
        });
        for (int i = 0; i < n; ++i) {
            sb_sim& s = *sims[i];
            any_lin_solver = any_lin_solver || s.setup.solver.kind == 1;
            any_lin_start = any_lin_start || (s.setup.solver.kind == 0 && s.setup.solver.start_linearized);
            any_iter = any_iter || s.setup.solver.kind == 0;
        }
        if (any_lin_solver && any_iter)
            throw std::invalid_argument("a quasi-steady batch uses one solver (Linearized or SimpleIterative)");
        if (fs_on_device) wait_for_null_stream(stream, ev_input);
        check(hipMemcpyAsync(d_fs, fs, sizeof(double) * fs_total,
                             fs_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream),
              "upload freestream");
        check(hipEventRecord(ev[0], stream), "event");
        check(hipMemcpyAsync(d_geo, h_geo, sizeof(double) * geo_total, hipMemcpyHostToDevice, stream), "upload geometry");
        check(hipMemcpyAsync(d_inst, h_inst, sizeof(Inst) * n, hipMemcpyHostToDevice, stream), "upload");
        BatchDev b = bd;
        b.inst = d_inst;
        check(launch_batch_prepare(b, first ? 1 : 0, stream), "prepare");
        check(launch_batch_horseshoe_factors(b, stream), "horseshoe factors");
        if (any_lin_solver) {
            check(launch_batch_linearized(b, lin_variant, lin_shm, 0, stream), "linearized solve");
        } else {
            if (any_lin_start) check(launch_batch_linearized(b, lin_variant, lin_shm, 1, stream), "linearized start");
            check(launch_batch_solve(b, stream), "solve");
        }
        check(launch_batch_forces(b, stream), "forces");
        if (any_lin_solver || any_lin_start) {
            check(launch_batch_take_singular(b, d_singular, stream), "singular flags");
            check(hipMemcpyAsync(h_singular, d_singular, sizeof(double) * n, hipMemcpyDeviceToHost, stream), "download");
        }
        check(hipMemcpyAsync(h_out, d_out, sizeof(double) * out_total, hipMemcpyDeviceToHost, stream), "download");
        check(hipEventRecord(ev[1], stream), "event");
        check(hipStreamSynchronize(stream), "batch step");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
        last_ms = ms;
        if (first) first_done = true;
        for (int i = 0; i < n; ++i) {
            sb_sim& s = *sims[i];
            s.first_done = true;
            s.steps_completed += 1;
        }
        if (any_lin_solver || any_lin_start)
            for (int i = 0; i < n; ++i)
                if (h_singular[i] != 0.0)  // linalg.rs:77-81, unwrapped at linearized.rs:139
                    throw std::runtime_error("instance " + std::to_string(i) +
                                             ": Linearized solver: matrix is singular or nearly singular at row " +
                                             std::to_string((long)h_singular[i] - 1));
        parallel_for(n, [&](int i) {
            sb_sim& s = *sims[i];
            std::copy(h_out + out_off[i], h_out + out_off[i + 1], s.h_out.begin());
            if (results) s.write_result(times ? times[i] : 0.0, results + i);
        });
    }

    // The launches of one step (captured into the graph after the first step).
    // Kernel variants (results do not depend on them: every target's sum keeps its
    // ascending panel order); the captured step graph is rebuilt at the next step.
    void set_tuning(int fixed_per, int targets_tile) {
        if (fixed_per < 1 || fixed_per > 3) throw std::invalid_argument("fixed_per must be 1, 2 or 3");
        if (targets_tile != 64 && targets_tile != 128 && targets_tile != 256)
            throw std::invalid_argument("targets_tile must be 64, 128 or 256");
        check(hipSetDevice(device), "hipSetDevice");
        int fg = batch_resident_groups(0, mode, bd.sym, fixed_per, 0);
        int tg = batch_resident_groups(1, mode, bd.sym, targets_tile, bd.stage_near);
        if (fg <= 0 || tg <= 0) throw HipError("occupancy query failed", -1);
        bd.fixed_per = fixed_per;
        bd.targets_tile = targets_tile;
        bd.fixed_grid = fg;
        bd.target_grid = tg;
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_sig[0] = graph_sig[1] = graph_sig[2] = -1;
    }

    // Key of the tuning cache: the device, the arithmetic, and the shape class of the
    // last step's work (powers of two of the instance count and of both pair sums, the
    // largest strip count in units of 64).
    std::string tuning_key() const {
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, device), "device properties");
        int max_n = 0;
        for (auto& s : sims) max_n = std::max(max_n, s->N);
        auto lg = [](double x) { return x >= 1.0 ? (int)std::floor(std::log2(x)) : -1; };
        std::ostringstream k;
        k << prop.gcnArchName << "|" << prop.name << "|" << (mode == MODE_FAST ? "fast" : "exact") << "|sym"
          << bd.sym << "|n" << lg((double)n) << "|f" << lg(fixed_work) << "|t" << lg(target_work) << "|N"
          << (max_n + 63) / 64;
        return k.str();
    }

    // Times every variant of the batched pair kernels on the batch's own last-step state
    // (outputs saved and restored, so the state is unchanged) and applies the fastest.
    // timings: kb_fixed for 1, 2, 3 panels per wave, then kb_targets for 64, 128, 256
    // records per tile (ms, -1 where a kernel had no work).
    void autotune(double* timings, int reps, long max_jobs) {
        if (quasi) throw std::invalid_argument("the autotuner covers dynamic batches");
        if (!first_done) throw std::invalid_argument("take a step before tuning: the jobs come from the last step");
        check(hipSetDevice(device), "hipSetDevice");
        long nf = 0, nv = 0;
        for (auto& s : sims) {
            nf += 3L * s->N;
            nv += 3L * s->nr_points;
        }
        const bool targets = bd.n_target_jobs > 0;
        double* save = dalloc<double>(nf + (targets ? nv : 0));
        auto copy_all = [&](bool to_save) {
            long o = 0;
            for (auto& s : sims) {
                size_t bytes = sizeof(double) * 3 * s->N;
                check(hipMemcpyAsync(to_save ? save + o : s->d_fixed, to_save ? s->d_fixed : save + o, bytes,
                                     hipMemcpyDeviceToDevice, stream),
                      "tuning copy");
                o += 3L * s->N;
            }
            if (!targets) return;
            for (auto& s : sims) {
                size_t bytes = sizeof(double) * 3 * s->nr_points;
                check(hipMemcpyAsync(to_save ? save + o : s->d_vel, to_save ? s->d_vel : save + o, bytes,
                                     hipMemcpyDeviceToDevice, stream),
                      "tuning copy");
                o += 3L * s->nr_points;
            }
        };
        const int prev_per = bd.fixed_per, prev_tile = bd.targets_tile;
        int best_per = prev_per, best_tile = prev_tile;
        try {
            copy_all(true);
            auto time_launch = [&](auto&& launch) {
                double best = 1e300;
                for (int r = 0; r <= reps; ++r) {  // the first launch loads the code object
                    check(hipEventRecord(ev[0], stream), "event");
                    launch();
                    check(hipEventRecord(ev[1], stream), "event");
                    check(hipStreamSynchronize(stream), "tuning");
                    float ms = 0.f;
                    check(hipEventElapsedTime(&ms, ev[0], ev[1]), "event time");
                    if (r > 0) best = std::min(best, (double)ms);
                }
                return best;
            };
            const int pers[3] = {1, 2, 3}, tiles[3] = {64, 128, 256};
            double tbest = 1e300;
            for (int v = 0; v < 3; ++v) {
                double t = -1.0;
                if (bd.n_fixed_jobs > 0) {
                    set_tuning(pers[v], prev_tile);
                    BatchDev b = bd;
                    b.inst = d_inst + 2L * n;
                    if (max_jobs > 0 && b.n_fixed_jobs > max_jobs) b.n_fixed_jobs = (int)max_jobs;
                    t = time_launch([&] { check(launch_batch_fixed(b, mode, stream), "fixed velocities"); });
                    if (t < tbest) {
                        tbest = t;
                        best_per = pers[v];
                    }
                }
                if (timings) timings[v] = t;
            }
            tbest = 1e300;
            for (int v = 0; v < 3; ++v) {
                double t = -1.0;
                if (targets) {
                    set_tuning(best_per, tiles[v]);
                    BatchDev b = bd;
                    b.inst = d_inst + 3L * n;
                    if (max_jobs > 0 && b.n_target_jobs > max_jobs) b.n_target_jobs = (int)max_jobs;
                    t = time_launch([&] { check(launch_batch_self_induction(b, mode, stream), "self induction"); });
                    if (t < tbest) {
                        tbest = t;
                        best_tile = tiles[v];
                    }
                }
                if (timings) timings[3 + v] = t;
            }
            copy_all(false);
            check(hipStreamSynchronize(stream), "tuning restore");
        } catch (...) {
            (void)hipStreamSynchronize(stream);
            (void)hipFree(save);
            set_tuning(prev_per, prev_tile);
            throw;
        }
        check(hipFree(save), "hipFree");
        set_tuning(best_per, best_tile);
    }

    void enqueue(bool first, bool apply, bool feats) {
        check(hipMemcpyAsync(d_geo, h_geo, sizeof(double) * geo_total, hipMemcpyHostToDevice, stream), "upload geometry");
        check(hipMemcpyAsync(d_inst, h_inst, sizeof(Inst) * kStages * n, hipMemcpyHostToDevice, stream), "upload");
        check(hipMemcpyAsync(d_jobs, h_jobs, sizeof(int2) * (cap_fixed_jobs + cap_target_jobs), hipMemcpyHostToDevice,
                             stream),
              "upload jobs");
        BatchDev b = bd;
        b.inst = d_inst;
        check(launch_batch_prepare(b, first ? 1 : 0, stream), "prepare");
        const bool judged = apply && (warm_mode & 3) == 3;
        if (feats) check(launch_batch_warm_start(b, mlp, apply ? (judged ? 2 : 1) : 0, stream), "warm start");
        if (first) check(launch_batch_init_wake(b, stream), "init wake");
        check(launch_batch_wake_step(b, stream), "wake step");
        b.inst = d_inst + 1L * n;
        check(launch_batch_panels(b, 1, stream), "panels");
        b.inst = d_inst + 2L * n;
        check(launch_batch_fixed(b, mode, stream), "fixed velocities");
        check(launch_batch_factors(b, stream), "factors");
        if (judged) check(launch_batch_warm_judge(b, stream), "warm judge");
        if (apply)
            check(hipMemcpyAsync(h_fallbacks, d_fallbacks, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, stream),
                  "download");
        check(launch_batch_solve(b, stream), "solve");
        b.inst = d_inst + 3L * n;
        check(launch_batch_velocity_copy(b, stream), "velocity copy");
        check(launch_batch_self_induction(b, mode, stream), "self induction");
        check(launch_batch_forces(b, stream), "forces");
        check(hipMemcpyAsync(h_out, d_out, sizeof(double) * out_total, hipMemcpyDeviceToHost, stream), "download");
    }
};

namespace {

// Tuning cache file: {"entries": [{"key": ..., "fixed_per": p, "targets_tile": t,
// "ms": [six timings]}, ...]}.  Unreadable or malformed files count as empty.
struct TuneEntry {
    std::string key;
    int fixed_per = 3, targets_tile = 256;
    double ms[6] = {-1, -1, -1, -1, -1, -1};
};

std::vector<TuneEntry> read_cache(const char* path) {
    std::vector<TuneEntry> out;
    std::ifstream f(path);
    if (!f) return out;
    std::stringstream ss;
    ss << f.rdbuf();
    std::string text = ss.str();
    try {
        auto root = json::parse(text.c_str());
        const json::Node* es = root->find("entries");
        if (!es || es->t != json::Node::Arr) return out;
        for (auto& e : es->items) {
            const json::Node* k = e->find("key");
            const json::Node* p = e->find("fixed_per");
            const json::Node* t = e->find("targets_tile");
            if (!k || k->t != json::Node::Str || !p || p->t != json::Node::Num || !t || t->t != json::Node::Num)
                continue;
            TuneEntry te;
            te.key = k->s;
            te.fixed_per = (int)p->n;
            te.targets_tile = (int)t->n;
            if (const json::Node* ms = e->find("ms"))
                if (ms->t == json::Node::Arr)
                    for (size_t i = 0; i < ms->items.size() && i < 6; ++i)
                        if (ms->items[i]->t == json::Node::Num) te.ms[i] = ms->items[i]->n;
            out.push_back(te);
        }
    } catch (const std::exception&) {
        out.clear();
    }
    return out;
}

std::string escape_json(const std::string& s) {
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        if ((unsigned char)c < 0x20) continue;
        o += c;
    }
    return o;
}

void write_cache(const char* path, const std::vector<TuneEntry>& es) {
    std::ostringstream o;
    o << "{\"entries\": [";
    for (size_t i = 0; i < es.size(); ++i) {
        const TuneEntry& e = es[i];
        o << (i ? ",\n  " : "\n  ") << "{\"key\": \"" << escape_json(e.key) << "\", \"fixed_per\": " << e.fixed_per
          << ", \"targets_tile\": " << e.targets_tile << ", \"ms\": [";
        for (int k = 0; k < 6; ++k) {
            char buf[32];
            snprintf(buf, sizeof buf, "%.4f", e.ms[k]);
            o << (k ? ", " : "") << buf;
        }
        o << "]}";
    }
    o << "\n]}\n";
    std::string tmp = std::string(path) + ".tmp";
    {
        std::ofstream f(tmp);
        if (!f) throw std::runtime_error(std::string("cannot write the tuning cache ") + tmp);
        f << o.str();
    }
    if (std::rename(tmp.c_str(), path) != 0) throw std::runtime_error(std::string("cannot replace ") + path);
}

}  // namespace

extern "C" {

int sb_batch_set_tuning(sb_batch* b, int fixed_per, int targets_tile) {
    if (!b) return -1;
    return guarded([&] { b->set_tuning(fixed_per, targets_tile); });
}

int sb_batch_get_tuning(const sb_batch* b, int* fixed_per, int* targets_tile) {
    if (!b || !fixed_per || !targets_tile) return -1;
    *fixed_per = b->bd.fixed_per;
    *targets_tile = b->bd.targets_tile;
    return 0;
}

int sb_batch_autotune(sb_batch* b, const char* cache_path, int flags, int reps, long max_jobs, double* timings,
                      size_t cap) {
    if (!b) return -1;
    if (timings && cap < 6) return -3;
    if (reps < 1 || max_jobs < 0) return -4;
    int cached = 0;
    int rc = guarded([&] {
        std::string key = b->tuning_key();
        std::vector<TuneEntry> es = cache_path ? read_cache(cache_path) : std::vector<TuneEntry>{};
        if ((flags & 1) && cache_path) {
            for (const TuneEntry& e : es)
                if (e.key == key) {
                    b->set_tuning(e.fixed_per, e.targets_tile);
                    if (timings) std::copy(e.ms, e.ms + 6, timings);
                    cached = 1;
                    return;
                }
        }
        TuneEntry te;
        te.key = key;
        b->autotune(te.ms, reps, max_jobs);
        te.fixed_per = b->bd.fixed_per;
        te.targets_tile = b->bd.targets_tile;
        if (timings) std::copy(te.ms, te.ms + 6, timings);
        if ((flags & 2) && cache_path) {
            es.erase(std::remove_if(es.begin(), es.end(), [&](const TuneEntry& e) { return e.key == key; }),
                     es.end());
            es.push_back(te);
            write_cache(cache_path, es);
        }
    });
    return rc != 0 ? rc : cached;
}

sb_batch* sb_batch_create(const char* const* setup_jsons, long n, int device, int mode) {
    if (!setup_jsons) {
        g_error = "null setup array";
        return nullptr;
    }
    if (mode != MODE_EXACT && mode != MODE_FAST) {
        g_error = "unsupported mode";
        return nullptr;
    }
    auto b = std::make_unique<sb_batch>();
    int rc = guarded([&] { b->create(setup_jsons, n, device, mode); });
    return rc == 0 ? b.release() : nullptr;
}

void sb_batch_destroy(sb_batch* b) { delete b; }

long sb_batch_size(const sb_batch* b) { return b ? b->n : -1; }

sb_sim* sb_batch_instance(sb_batch* b, long i) {
    if (!b || i < 0 || i >= b->n) return nullptr;
    return b->sims[i].get();
}

long sb_batch_freestream_offsets(const sb_batch* b, long* offsets, size_t cap) {
    if (!b || !offsets) return -1;
    if (cap < (size_t)b->n + 1) return -2;
    for (int i = 0; i <= b->n; ++i) offsets[i] = b->fs_off[i] / 3;
    return b->n + 1;
}

int sb_batch_do_step(sb_batch* b, const double* times, const double* dts, const double* fs, size_t n_points,
                     sb_step_result* results) {
    if (!b || !fs) return -1;
    return guarded([&] { b->step(times, dts, fs, n_points, false, results); });
}

int sb_batch_do_step_device(sb_batch* b, const double* times, const double* dts, const double* fs, size_t n_points,
                            sb_step_result* results) {
    if (!b || !fs) return -1;
    return guarded([&] { b->step(times, dts, fs, n_points, true, results); });
}

int sb_batch_set_wind(sb_batch* b, const char* environment_json, const char* const* wind_conditions_json) {
    if (!b || !wind_conditions_json) return -1;
    return guarded([&] { b->set_wind(environment_json, wind_conditions_json); });
}

int sb_batch_do_step_wind(sb_batch* b, const double* times, const double* dts, const double* ship_velocities,
                          sb_step_result* results) {
    if (!b || !times || !dts || !ship_velocities) return -1;
    return guarded([&] { b->step(times, dts, nullptr, 0, false, results, ship_velocities); });
}

int sb_batch_get_freestream(const sb_batch* b, double* out, size_t n_points) {
    if (!b || !out) return -1;
    if (n_points != (size_t)(b->fs_total / 3)) return -3;
    return guarded([&] {
        check(hipSetDevice(b->device), "hipSetDevice");
        check(hipMemcpyAsync(out, b->d_fs, sizeof(double) * b->fs_total, hipMemcpyDeviceToHost, b->stream),
              "download freestream");
        check(hipStreamSynchronize(b->stream), "get freestream");
    });
}

double sb_batch_last_step_ms(const sb_batch* b) { return b ? b->last_ms : -1.0; }

int sb_batch_reset(sb_batch* b) {
    if (!b) return -1;
    return guarded([&] { b->reset(); });
}

int sb_batch_set_warm_start(sb_batch* b, sb_mlp* m, int mode) {
    if (!b) return -1;
    if (mode < 0 || mode > 7 || mode == 4 || (mode > 0 && !m)) return -4;
    if (b->quasi && mode > 0) {
        g_error = "the learned warm start runs on dynamic batches";
        return -4;
    }
    return guarded([&] {
        b->mlp = mode > 0 ? m : nullptr;
        b->warm_mode = mode;
        b->bd.warm_residual = (mode & 4) ? 1 : 0;
    });
}

int sb_batch_set_record_features(sb_batch* b, int on) {
    if (!b) return -1;
    if (b->quasi && on) {
        g_error = "features are recorded on dynamic batches";
        return -4;
    }
    b->record_features = on != 0;
    return 0;
}

long sb_batch_nr_strips(const sb_batch* b) { return b ? b->n_strips : -1; }

int sb_batch_get_features(const sb_batch* b, float* features, float* scale, size_t cap_strips) {
    if (!b) return -1;
    if (cap_strips < (size_t)b->n_strips) return -3;
    return guarded([&] {
        check(hipSetDevice(b->device), "hipSetDevice");
        if (features)
            check(hipMemcpy(features, b->d_feat, sizeof(float) * MLP_FEATURES * b->n_strips, hipMemcpyDeviceToHost),
                  "download");
        if (scale) check(hipMemcpy(scale, b->d_scale, sizeof(float) * b->n_strips, hipMemcpyDeviceToHost), "download");
    });
}

long sb_batch_last_fallbacks(const sb_batch* b) {
    if (!b) return -1;
    return b->last_applied ? (long)b->h_fallbacks[0] : 0;
}

long sb_batch_last_warm_taken(const sb_batch* b) {
    if (!b) return -1;
    if (!b->last_applied) return 0;
    return (b->warm_mode & 3) == 3 ? (long)b->h_fallbacks[1] : (long)b->n - (long)b->h_fallbacks[0];
}

}  // extern "C"
