"""Benchmark of the dynamic lifting-line step on one MI355X.

Workload: four 30 m x 10 m wing sails 25 m apart, 40 strips each (N = 160),
free wake of 2,000 panel rows (320,000 panels, 328,164 wake points), every
wake point receiving the induced velocity of every panel (~1.05e11
Biot-Savart pairs per step).  The timed steps start from a synthesized
developed wake (all 2,000 rows shed, seeded perturbation of a convected
sheet); the freestream at all points is resident in device memory.

A "step" is one Simulation::do_step: wake shedding and convection, panel
rebuild, frozen-wake velocities and factors, the damped fixed-point solve,
free-wake self-induction on all wake points, forces.

Prints one JSON line (metric/unit from the headline: time-steps/s, with
Biot-Savart pairs/s and the FP64 roofline fraction of the self-induction
kernel), plus the CPU baseline: the f64 C++ restatement of the reference
(OpenMP over targets, as the reference's rayon map) timed on this host on a
bounded sample of the same step.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-steps/s + Biot-Savart pairs/s (% FP64 peak), batched+warm vs plain"
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector/matrix dense peak
FLOP_FAR, FLOP_NEAR = 38.0, 334.0  # algorithmic flops per far (doublet) / near (4 segments) pair


def synthesize_developed_state(sim, case, seed=1):
    """All rows shed: convected sheet + seeded roll-up noise, strengths from the first solves."""
    import stormbird_amd.setups as S

    fs = S.uniform_freestream(sim.nr_freestream_points(), case["velocity"])
    dt = case["time_step"]
    for k in range(3):
        sim.do_step(time=k * dt, time_step=dt, freestream_velocity=fs)
    st = sim.get_wake_state()
    n = sim.nr_ctrl_points
    P = sim.nr_span_points
    M = sim.nr_panels // n
    rng = np.random.default_rng(seed)
    pts = st["points"].reshape(M + 1, P, 3).copy()
    rows = np.arange(M + 1, dtype=np.float64)[:, None, None]
    pts[2:] += rng.normal(0.0, 1.0, pts[2:].shape) * 0.05 * np.sqrt(rows[2:])
    gamma0 = st["strengths"][:n]
    gam = np.tile(gamma0, (M, 1)) * (1.0 + 0.02 * rng.standard_normal((M, n)))
    vel = np.tile(np.asarray(case["velocity"], dtype=np.float64), (pts.shape[0] * pts.shape[1], 1))
    prev = sim.previous_circulation_strength
    return pts.reshape(-1, 3), vel, gam.ravel(), prev, fs


def time_gpu(sim, fs_dev_ptr, n_points, dt, steps, warmup, t0_step):
    import torch

    sim.set_kernel_timers(True)  # seven event records per step, ~25 us of a 590 ms step
    for k in range(warmup):
        sim.do_step_device(time=(t0_step + k) * dt, time_step=dt, freestream_ptr=fs_dev_ptr, n_points=n_points)
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for k in range(steps):
        sim.do_step_device(time=(t0_step + warmup + k) * dt, time_step=dt, freestream_ptr=fs_dev_ptr,
                           n_points=n_points)
        kms.append(sim.last_kernel_ms("self_induction"))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0), kms


def cpu_baseline(js, state, n_panels_total, pairs_per_step, budget_s=12.0):
    """Times the oracle's self-induction sum on a strided sample of the wake points."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py

    orc = oracle_py.Simulation(js, "glibc")
    pts, vel, gam, prev = state
    orc.set_wake_state(pts, vel, gam, 2000)
    orc.set_previous_circulation(prev)
    cal = pts[:: max(1, pts.shape[0] // 256)][:256]
    t0 = time.perf_counter()
    orc.induced_velocities(cal)
    t_cal = time.perf_counter() - t0
    rate = cal.shape[0] * n_panels_total / max(t_cal, 1e-9)
    n_sample = int(min(pts.shape[0], max(256, budget_s * rate / n_panels_total)))
    sample = pts[:: max(1, pts.shape[0] // n_sample)][:n_sample]
    t0 = time.perf_counter()
    orc.induced_velocities(sample)
    t = time.perf_counter() - t0
    pairs_s = sample.shape[0] * n_panels_total / t
    threads = oracle_py.set_threads(0)
    # one thread on a small sample of the same sum
    one = pts[:: max(1, pts.shape[0] // 64)][:64]
    oracle_py.set_threads(1)
    t0 = time.perf_counter()
    orc.induced_velocities(one)
    t1 = time.perf_counter() - t0
    oracle_py.set_threads(threads)
    pairs_1 = one.shape[0] * n_panels_total / t1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": pairs_s / pairs_per_step,
        "unit": "time-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"free-wake self-induction sum at {sample.shape[0]} of {pts.shape[0]} wake points "
                   f"x {n_panels_total} panels ({t:.1f} s, {pairs_s:.3e} pairs/s), scaled to the "
                   f"{pairs_per_step:.3e} pairs of one step; C++ f64 restatement of the reference, "
                   f"OpenMP over targets"),
        "pairs_per_s": pairs_s,
        "one_thread": {"pairs_per_s": pairs_1, "time_steps_per_s": pairs_1 / pairs_per_step,
                       "sample": f"{one.shape[0]} wake points x {n_panels_total} panels ({t1:.2f} s)"},
        "cpu_model": model,
        "host_threads_visible": os.cpu_count(),
    }


def develop_wake(sim, steps_completed, rng):
    """Fills every wake row of `sim` with shed strengths (row-0 circulation x (1 + 2 % noise))
    and perturbs the free rows, keeping the points and velocities it has."""
    st = sim.get_wake_state()
    n, P = sim.nr_ctrl_points, sim.nr_span_points
    M = sim.nr_panels // n
    pts = st["points"].reshape(M + 1, P, 3).copy()
    rows = np.arange(M + 1, dtype=np.float64)[:, None, None]
    pts[2:] += rng.normal(0.0, 1.0, pts[2:].shape) * 0.01 * np.sqrt(rows[2:])
    g0 = st["strengths"][:n]
    gam = np.tile(g0, (M, 1)) * (1.0 + 0.02 * rng.standard_normal((M, n)))
    sim.set_wake_state(pts.reshape(-1, 3), st["velocities"], gam.ravel(), steps_completed)


def batched_polar_sweep(n_inst=512, seed=2, steps=3, plain_steps=1, mode=0, mlp=None):
    """512 polar-sweep instances (seeded: 1-6 sails, 16-64 strips, chords 5-15 m, AR 2-6,
    200-3000 wake rows, AWA 20-180 deg, AWS 6-18 m/s) from a developed wake: one batched
    step (sb_batch, graph-replayed) against the plain path (each instance stepped alone),
    both in the same arithmetic mode (0 EXACT, 1 FAST).  With a trained warm-start network,
    the same developed steps are then timed with the judged warm start (per instance the
    solver's first evaluation keeps the better of the previous circulation and the
    network's guess)."""
    import torch

    import stormbird_amd as sb
    from stormbird_amd import setups as S

    rng = np.random.default_rng(seed)
    made = [S.polar_sweep_instance(rng) for _ in range(n_inst)]
    setups = [m[0] for m in made]
    cases = [m[1] for m in made]
    t_create = time.perf_counter()
    batch = sb.Batch(setups, mode=mode)
    t_create = time.perf_counter() - t_create
    dts = np.array([c["time_step"] for c in cases])
    fs_all = np.concatenate([S.uniform_freestream(s.nr_freestream_points(), c["velocity"])
                             for s, c in zip(batch.instances, cases)])
    fs_dev = torch.from_numpy(fs_all).to("cuda")
    n_pts = fs_all.shape[0]
    offs = batch.freestream_offsets
    for k in range(3):
        batch.do_step_device(times=k * dts, time_steps=dts, freestream_ptr=fs_dev.data_ptr(), n_points=n_pts,
                             want_result=False)
    drng = np.random.default_rng(seed + 100)
    for s in batch.instances:
        develop_wake(s, 3, drng)
    n_ctrl = np.array([s.nr_ctrl_points for s in batch.instances], dtype=np.float64)
    rows = np.array([s.nr_panels // s.nr_ctrl_points for s in batch.instances], dtype=np.float64)
    pairs = float(np.sum(n_ctrl * n_ctrl * (rows - 1) + n_ctrl * n_ctrl))
    torch.cuda.synchronize()
    # plain: every instance stepped alone
    k0 = 3
    t0 = time.perf_counter()
    for k in range(plain_steps):
        for i, s in enumerate(batch.instances):
            s.do_step_device(time=(k0 + k) * dts[i], time_step=dts[i],
                             freestream_ptr=fs_dev.data_ptr() + 24 * int(offs[i]),
                             n_points=int(offs[i + 1] - offs[i]), want_result=False)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    k0 += plain_steps
    # batched: one step on the developed state, the autotuner on it, then the first call
    # with the chosen variants captures the step graph
    batch.do_step_device(times=k0 * dts, time_steps=dts, freestream_ptr=fs_dev.data_ptr(), n_points=n_pts,
                         want_result=False)
    k0 += 1
    _, tuning, tune_ms = batch.autotune(None)
    batch.do_step_device(times=k0 * dts, time_steps=dts, freestream_ptr=fs_dev.data_ptr(), n_points=n_pts,
                         want_result=False)
    k0 += 1
    torch.cuda.synchronize()
    dev_ms, iters = [], []
    t0 = time.perf_counter()
    for k in range(steps):
        res = batch.do_step_device(times=(k0 + k) * dts, time_steps=dts, freestream_ptr=fs_dev.data_ptr(),
                                   n_points=n_pts, want_result=True)
        dev_ms.append(batch.last_step_ms())
        iters.append(float(np.mean(res.iterations)))
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    k0 += steps
    plain_rate = n_inst * plain_steps / t_plain
    batch_rate = n_inst * steps / t_batch
    warm = None
    if mlp is not None:
        batch.set_warm_start(mlp, 3)
        batch.do_step_device(times=k0 * dts, time_steps=dts, freestream_ptr=fs_dev.data_ptr(), n_points=n_pts,
                             want_result=False)  # captures the warm-start step graph
        k0 += 1
        torch.cuda.synchronize()
        wms, witers, taken = [], [], 0
        t0 = time.perf_counter()
        for k in range(steps):
            res = batch.do_step_device(times=(k0 + k) * dts, time_steps=dts, freestream_ptr=fs_dev.data_ptr(),
                                       n_points=n_pts, want_result=True)
            wms.append(batch.last_step_ms())
            witers.append(float(np.mean(res.iterations)))
            taken += batch.last_warm_taken()
        torch.cuda.synchronize()
        t_warm = time.perf_counter() - t0
        warm_rate = n_inst * steps / t_warm
        warm = {"mode": "judged (per instance the better of the previous circulation and the guess)",
                "batched_warm_instance_steps_per_s": warm_rate, "batched_warm_ms_per_step": 1000.0 * t_warm / steps,
                "batched_warm_device_ms_per_step": float(np.mean(wms)),
                "iterations_mean_warm": float(np.mean(witers)), "iterations_mean_without": float(np.mean(iters)),
                "guess_taken_fraction": taken / (n_inst * steps), "batched_warm_vs_plain": warm_rate / plain_rate}
        batch.set_warm_start(None, 0)
    return {
        "workload": (f"{n_inst} polar-sweep instances (seed {seed}): 1-6 sails, 16-64 strips per sail, chords "
                     "5-15 m, AR 2-6, 200-3000 wake rows, AWA 20-180 deg, AWS 6-18 m/s, convected wakes, "
                     "developed-wake steady state"),
        "n_instances": n_inst,
        "n_ctrl_points_total": int(n_ctrl.sum()),
        "biot_savart_pairs_per_batched_step": pairs,
        "batched_instance_steps_per_s": batch_rate,
        "batched_ms_per_step": 1000.0 * t_batch / steps,
        "batched_device_ms_per_step": float(np.mean(dev_ms)),
        "batched_pairs_per_s": pairs * steps / t_batch,
        "plain_instance_steps_per_s": plain_rate,
        "plain_ms_per_sweep_step": 1000.0 * t_plain / plain_steps,
        "batched_vs_plain": batch_rate / plain_rate,
        "steps": steps,
        "plain_steps": plain_steps,
        "create_s": t_create,
        "mode": {0: "exact", 1: "fast"}[mode],
        "iterations_mean": float(np.mean(iters)),
        "autotuner": {"fixed_per": tuning[0], "targets_tile": tuning[1],
                      "fixed_ms_per_panels_per_wave_1_2_3": [float(x) for x in tune_ms[:3]]},
        "warm_start": warm,
    }


def _perturbed(cases, rng):
    """Per instance: the freestream rotated by up to +-6 deg and scaled by 0.8-1.2."""
    out = []
    for c in cases:
        u = np.asarray(c["velocity"], dtype=np.float64)
        a = rng.uniform(-6.0, 6.0) * math.pi / 180.0
        k = rng.uniform(0.8, 1.2)
        out.append(k * np.array([u[0] * math.cos(a) - u[1] * math.sin(a), u[0] * math.sin(a) + u[1] * math.cos(a), 0.0]))
    return out


def polar_training_set(n_solves=200_000, seed=30, n_geometries=4096, per_batch=512):
    """Training rows of the learned warm start: first-step circulation solves of fresh
    polar-sweep geometries (n_geometries drawn from the polar-sweep distribution, in
    batches of per_batch; seeds apart from the held-out instances), each under
    ceil(n_solves / n_geometries) perturbed inflows, solved tight (max |Gamma_est - Gamma|
    < 1e-9 |u| c / 2).  The first step sheds one wake row, so the solve does not depend on
    the wake length: the training batches carry 8 rows.  Returns (features [rows, 64],
    targets Gamma / scale [rows], solves, seconds)."""
    import stormbird_amd as sb
    from stormbird_amd import setups as S

    n_batches = max(1, n_geometries // per_batch)
    conditions = -(-n_solves // (n_batches * per_batch))
    t0 = time.perf_counter()
    X, T = [], []
    solves = 0
    for b in range(n_batches):
        grng = np.random.default_rng([seed, b])
        made = [S.polar_sweep_instance(grng, max_panels=8, min_panels=8) for _ in range(per_batch)]
        cases = [m[1] for m in made]
        batch = sb.Batch([m[0] for m in made], mode=sb.MODE_EXACT)
        dts = np.array([c["time_step"] for c in cases])
        scale0 = np.array([0.5 * math.sqrt(sum(x * x for x in c["velocity"])) * 10.0 for c in cases])
        for s, t in zip(batch.instances, 1e-9 * scale0):
            s.set_solver_settings(max_iterations=5000, damping=0.05, residual_tolerance=0.0, strength_tolerance=t)
        batch.set_record_features(True)
        npts = [s.nr_freestream_points() for s in batch.instances]
        crng = np.random.default_rng([seed, b, 1])
        for _ in range(conditions):
            vel = _perturbed(cases, crng)
            batch.reset()
            fs = np.repeat(np.asarray(vel), npts, axis=0)
            res = batch.do_step(times=None, time_steps=dts, freestream_velocity=fs)
            x, sc = batch.get_features()
            g = np.concatenate([r.force_input.circulation_strength for r in res])
            X.append(x)
            T.append((g / sc).astype(np.float32))
            solves += per_batch
        del batch
    return np.concatenate(X), np.concatenate(T), solves, time.perf_counter() - t0


def warm_start_polar(n_inst=512, seed=3, train_solves=200_000, train_steps=3000, minibatch=65536):
    """Learned warm start on the polar-sweep distribution: the network is trained on the GPU
    (bf16 matrix-core forward/backward, Adam) on `train_solves` first-step solves of fresh
    geometries (polar_training_set); the held-out 512 instances (seed 3) are then solved
    under perturbed inflow to a 1e-6 relative tolerance (max |Gamma_est - Gamma| < 1e-6
    max |Gamma*|) from zero (cold) and from the network's guess (warm), batched, and cold one
    instance at a time."""
    import torch

    import stormbird_amd as sb
    from stormbird_amd import setups as S

    rng = np.random.default_rng(seed)
    made = [S.polar_sweep_instance(rng) for _ in range(n_inst)]
    setups = [m[0] for m in made]
    cases = [m[1] for m in made]
    batch = sb.Batch(setups, mode=sb.MODE_EXACT)
    dts = np.array([c["time_step"] for c in cases])
    scale0 = np.array([0.5 * math.sqrt(sum(x * x for x in c["velocity"])) * 10.0 for c in cases])

    def solve(vel):
        batch.reset()
        fs = np.concatenate([S.uniform_freestream(s.nr_freestream_points(), v) for s, v in zip(batch.instances, vel)])
        return batch.do_step(times=None, time_steps=dts, freestream_velocity=fs)

    def settings(tol):
        for s, t in zip(batch.instances, tol):
            s.set_solver_settings(max_iterations=5000, damping=0.05, residual_tolerance=0.0, strength_tolerance=t)

    X, T, n_solves, t_data = polar_training_set(train_solves)
    # training on the GPU
    t0 = time.perf_counter()
    mlp = sb.Mlp(seed=seed)
    xd = torch.from_numpy(X).to("cuda")
    td = torch.from_numpy(T).to("cuda")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    nb = min(minibatch, X.shape[0])
    loss = None
    for it in range(train_steps):
        idx = torch.randint(0, X.shape[0], (nb,), device="cuda", generator=gen)
        xb = xd.index_select(0, idx).contiguous()
        tb = td.index_select(0, idx).contiguous()
        lr = 1e-3 if it < 0.7 * train_steps else 2e-4
        loss = mlp.train_step(xb.data_ptr(), tb.data_ptr(), nb, lr)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    # held-out conditions
    vel = _perturbed(cases, np.random.default_rng(seed + 1000))
    settings(1e-11 * scale0)
    ref = solve(vel)
    gref = [r.force_input.circulation_strength.copy() for r in ref]
    tol = np.array([1e-6 * np.max(np.abs(g)) for g in gref])
    settings(tol)
    batch.set_warm_start(None, 0)
    batch.set_record_features(True)
    cold = solve(vel)
    ms_cold = batch.last_step_ms()
    x, sc = batch.get_features()
    batch.set_warm_start(mlp, 1)
    warm = solve(vel)
    ms_warm = batch.last_step_ms()
    fallbacks = batch.last_fallbacks()
    # quality of the guess itself
    pd = torch.zeros(x.shape[0], dtype=torch.float32, device="cuda")
    xh = torch.from_numpy(x).to("cuda")
    mlp.forward(xh.data_ptr(), x.shape[0], pd.data_ptr())
    guess = pd.cpu().numpy().astype(np.float64) * sc
    off = np.concatenate([[0], np.cumsum([s.nr_ctrl_points for s in batch.instances])])
    guess_err = np.array([np.max(np.abs(guess[off[i]:off[i + 1]] - gref[i])) / np.max(np.abs(gref[i]))
                          for i in range(n_inst)])
    it_cold = np.array([r.iterations for r in cold], dtype=np.float64)
    it_warm = np.array([r.iterations for r in warm], dtype=np.float64)
    dev_c = max(float(np.max(np.abs(c.force_input.circulation_strength - g)) / np.max(np.abs(g)))
                for c, g in zip(cold, gref))
    dev_w = max(float(np.max(np.abs(w.force_input.circulation_strength - g)) / np.max(np.abs(g)))
                for w, g in zip(warm, gref))
    # plain: each instance's cold first-step solve alone, same settings
    batch.set_warm_start(None, 0)
    batch.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, s in enumerate(batch.instances):
        fs = S.uniform_freestream(s.nr_freestream_points(), vel[i])
        s.do_step(time=0.0, time_step=dts[i], freestream_velocity=fs)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    result = {
        "workload": (f"{n_inst} polar-sweep instances (seed {seed}), first-step circulation solves under "
                     "perturbed inflow (+-6 deg, x0.8-1.2), damping 0.05, stop at max |Gamma_est - Gamma| < "
                     "1e-6 max |Gamma*| (residual test off); network trained on first-step solves of 4096 fresh "
                     "geometries of the same distribution"),
        "train_solves": int(n_solves),
        "train_geometries": 4096,
        "train_rows": int(X.shape[0]),
        "train_steps": train_steps,
        "minibatch": nb,
        "final_train_mse": loss,
        "data_s": t_data,
        "train_s": t_train,
        "guess_rel_error_median": float(np.median(guess_err)),
        "guess_rel_error_p90": float(np.percentile(guess_err, 90)),
        "iterations_cold_mean": float(it_cold.mean()),
        "iterations_warm_mean": float(it_warm.mean()),
        "iterations_cold_max": float(it_cold.max()),
        "iterations_warm_max": float(it_warm.max()),
        "fallback_rate": fallbacks / n_inst,
        "max_rel_dev_cold_vs_converged": dev_c,
        "max_rel_dev_warm_vs_converged": dev_w,
        "batched_cold_ms": ms_cold,
        "batched_warm_ms": ms_warm,
        "plain_cold_ms": 1000.0 * t_plain,
        "solves_per_s_batched_warm": n_inst / (ms_warm * 1e-3),
        "solves_per_s_plain_cold": n_inst / t_plain,
        "batched_warm_vs_plain": (n_inst / (ms_warm * 1e-3)) / (n_inst / t_plain),
    }
    return result, mlp


def route_simulation(n_ships=64, seed=5, steps=1, mlp=None):
    """Route simulation: 64 ships x 4 wing sails x 48 strips, free wake of 1,500 rows
    (self-induction on every wake point), in 2 wind conditions (power-law true wind, ship
    speed, von Karman gusts along and across the wind: 15 %, 200 m), each ship with a
    turbulence realization of its own, roll 10 deg at 12 s.  Every ship starts from the
    rolled-up wake of its condition after 1,500 steps from the start (points, velocities and
    strengths stored in stormbird_amd/data/route_developed_wakes_1500rows.npz, made by
    `bench.py --only route-rollup`); the timed steps run from there with the default kernel
    variants, and are extrapolated to 10,000 steps.  With a trained network, one more step
    runs with the judged per-step warm start."""
    import torch

    from stormbird_amd.route import DEVELOPED_ROUTE, RouteSimulation

    t0 = time.perf_counter()
    route = RouteSimulation(n_ships=n_ships, seed=seed)
    route.load_developed(DEVELOPED_ROUTE)
    t_setup = time.perf_counter() - t0
    print(f"route: {n_ships} ships loaded in {t_setup:.1f} s", file=sys.stderr, flush=True)
    pairs = 0.0
    for s in route.batch.instances:
        n, P = s.nr_ctrl_points, s.nr_span_points
        M = s.nr_panels // n
        pairs += float(P * (M + 1)) * float(n * M) + float(n * n * (M - 1)) + float(n * n)
    near = []
    for c in range(len(route.conditions)):
        nn, tot = route.batch.instances[c].count_self_induction_pairs()
        near.append(nn / tot)
    torch.cuda.synchronize()
    dev_ms, iters = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        res = route.step(want_result=True)
        dev_ms.append(route.batch.last_step_ms())
        iters.append(np.mean([r.iterations for r in res]))
        print(f"route: step {route.k}, {dev_ms[-1]:.0f} ms on the device", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    out = {
        "workload": (f"{n_ships} ships (seed {seed}) x 4 wing sails 30x10 m x 48 strips, free wake of 1500 rows, "
                     "all wake points self-induced, 2 wind conditions (power-law true wind 8-14 m/s, ship "
                     "4-7 m/s), von Karman gusts along and across the wind 15 % / 200 m (1000 linspace "
                     "frequencies, 0.0005-0.5 Hz), roll 10 deg at 12 s, rolled-up wakes of a 1,500-step run "
                     "from the start"),
        "n_ships": n_ships,
        "biot_savart_pairs_per_batched_step": pairs,
        "near_field_pair_fraction_per_condition": near,
        "batched_s_per_step": t,
        "batched_device_ms_per_step": float(np.mean(dev_ms)),
        "ship_steps_per_s": n_ships / t,
        "pairs_per_s": pairs / t,
        "extrapolated_10000_steps_h": 10000 * t / 3600.0,
        "iterations_mean": float(np.mean(iters)),
        "setup_s": t_setup,
        "freestream": ("WindEnvironment semantics evaluated by the batch: control points on the host (gusts, "
                       "profile, ship speed), every wake point on the device, inside the timed step"),
        "kernel_variants": dict(zip(("fixed_per", "targets_tile"), route.batch.get_tuning())),
    }
    if mlp is not None:
        route.batch.set_warm_start(mlp, 3, residual=True)
        res = route.step(want_result=True)
        out["judged_warm_start"] = {"iterations_mean": float(np.mean([r.iterations for r in res])),
                                    "taken_fraction": route.batch.last_warm_taken() / n_ships,
                                    "fallbacks": route.batch.last_fallbacks(),
                                    "device_ms_per_step": route.batch.last_step_ms()}
    return out


def route_warm_start(n_ships=64, train_seed=11, eval_seed=5, nr_panels=60, warmup=40, record=120, eval_steps=40,
                     train_steps=1500, minibatch=65536):
    """The judged per-step warm start of the route, trained on route steps: the features
    include the solver's default initial guess (the previous step's circulation), and the
    network predicts the change of the circulation from it over this step (gusts, roll), in
    units of the scale, so that a zero prediction is the default guess.  Training data: `record` steps of a route of other wind conditions and
    turbulence (seed train_seed, 8 conditions) with short wakes (`nr_panels` rows), each
    solved tight (max |Gamma_est - Gamma| < 1e-9 |u| c / 2 or 400 iterations).  Evaluation:
    the route's own conditions (seed eval_seed) with short wakes and the default solver
    settings, `eval_steps` steps without and then with the judged warm start, from the same
    state."""
    import torch

    import stormbird_amd as sb
    from stormbird_amd.route import RouteSimulation

    t0 = time.perf_counter()
    tr = RouteSimulation(n_ships=n_ships, n_conditions=8, seed=train_seed, nr_panels=nr_panels)
    for s in tr.batch.instances:
        s.set_solver_settings(max_iterations=400, damping=0.05, residual_tolerance=0.0, strength_tolerance=1e-9 * 60.0)
    for _ in range(warmup):
        tr.step(want_result=False)
    tr.batch.set_record_features(True)
    X, T = [], []
    for _ in range(record):
        res = tr.step(want_result=True)
        x, sc = tr.batch.get_features()
        g = np.concatenate([r.force_input.circulation_strength for r in res])
        X.append(x)
        # the change from the previous circulation (feature 53: the strip's own, over its scale)
        T.append((g / sc - x[:, 53]).astype(np.float32))
    X, T = np.concatenate(X), np.concatenate(T)
    del tr
    t_data = time.perf_counter() - t0
    print(f"route warm start: {X.shape[0]} training rows in {t_data:.1f} s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    mlp = sb.Mlp(seed=train_seed)
    xd = torch.from_numpy(X).to("cuda")
    td = torch.from_numpy(T).to("cuda")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(train_seed)
    nb = min(minibatch, X.shape[0])
    loss = None
    for it in range(train_steps):
        idx = torch.randint(0, X.shape[0], (nb,), device="cuda", generator=gen)
        xb = xd.index_select(0, idx).contiguous()
        tb = td.index_select(0, idx).contiguous()
        loss = mlp.train_step(xb.data_ptr(), tb.data_ptr(), nb, 1e-3 if it < 0.7 * train_steps else 2e-4)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    print(f"route warm start: trained in {t_train:.1f} s (mse {loss:.3e})", file=sys.stderr, flush=True)
    # evaluation: the same steps from the same state, without and with the judged warm start
    ev = RouteSimulation(n_ships=n_ships, seed=eval_seed, nr_panels=nr_panels)
    for _ in range(warmup):
        ev.step(want_result=False)
    saved = [(s.get_wake_state(), s.previous_circulation_strength.copy()) for s in ev.batch.instances]
    k0 = ev.k

    def run(warm, max_iterations=40):
        for s in ev.batch.instances:
            s.set_solver_settings(max_iterations=max_iterations, damping=0.05, residual_tolerance=1e-4,
                                  strength_tolerance=1e-6)
        ev.batch.set_warm_start(mlp if warm else None, 3 if warm else 0, residual=True)
        for s, (st, prev) in zip(ev.batch.instances, saved):
            s.set_wake_state(st["points"], st["velocities"], st["strengths"], st["steps_completed"])
            s.previous_circulation_strength = prev
        ev.set_motion(k0 - 2)
        ev.set_motion(k0 - 1)
        ev.k = k0
        its, resid, taken = [], [], 0
        for _ in range(eval_steps):
            r = ev.step(want_result=True)
            its.append(np.asarray(r.iterations, dtype=np.float64))
            resid.append(np.asarray(r.residual, dtype=np.float64))
            taken += ev.batch.last_warm_taken() if warm else 0
        return np.concatenate(its), np.concatenate(resid), taken

    it_cold, res_cold, _ = run(False)
    it_warm, res_warm, taken = run(True)
    # the same steps with the solver allowed to converge (the default tolerances, at most 1000
    # iterations): what the guess saves in iterations
    itc_cold, _, _ = run(False, 1000)
    itc_warm, _, taken_c = run(True, 1000)
    return {
        "workload": (f"route steps with short wakes ({nr_panels} rows), {n_ships} ships; trained on {record} steps of "
                     f"another route (seed {train_seed}, 8 wind conditions), evaluated on {eval_steps} steps of the "
                     f"benchmark route's conditions (seed {eval_seed}) with the default solver settings "
                     "(at most 40 iterations, residual 1e-4)"),
        "train_rows": int(X.shape[0]), "train_steps": train_steps, "final_train_mse": loss,
        "data_s": t_data, "train_s": t_train,
        "iterations_mean_without": float(it_cold.mean()), "iterations_mean_judged": float(it_warm.mean()),
        "residual_mean_without": float(res_cold.mean()), "residual_mean_judged": float(res_warm.mean()),
        "guess_taken_fraction": taken / float(n_ships * eval_steps),
        "converging_solver": {"max_iterations": 1000, "iterations_mean_without": float(itc_cold.mean()),
                              "iterations_mean_judged": float(itc_warm.mean()),
                              "iterations_max_without": float(itc_cold.max()),
                              "iterations_max_judged": float(itc_warm.max()),
                              "guess_taken_fraction": taken_c / float(n_ships * eval_steps)},
    }, mlp


def route_rollup(steps=1500, n_conditions=2, seed=5, dump=None, resume=None, budget_s=None):
    """The route's ships 0..n_conditions-1 (one per wind condition) run from the start for
    `steps` steps in FAST arithmetic (within 1e-10 of EXACT per step), so that their wakes
    roll up; `dump` stores them (RouteSimulation.store_developed).  With `budget_s`, the run
    stops after that many seconds, and `resume` continues from what it stored.  Progress on
    stderr every 100 steps or 60 s."""
    import stormbird_amd as sb
    from stormbird_amd.route import RouteSimulation

    route = RouteSimulation(n_ships=n_conditions, n_conditions=n_conditions, seed=seed, mode=sb.MODE_FAST)
    if resume:
        route.load_developed(resume)
    t0 = last = time.perf_counter()
    res = None
    while route.k < steps:
        k = route.k
        res = route.step(want_result=True)
        now = time.perf_counter()
        if (k + 1) % 100 == 0 or k == steps - 1 or now - last > 60.0:
            last = now
            print(f"route rollup: step {k + 1}, {now - t0:.1f} s, iterations "
                  f"{[int(x) for x in res.iterations]}", file=sys.stderr, flush=True)
        if budget_s is not None and time.perf_counter() - t0 > budget_s:
            break
    total = time.perf_counter() - t0
    done = route.k >= steps
    near = []
    for c in range(n_conditions):
        nn, tot = route.batch.instances[c].count_self_induction_pairs()
        near.append(nn / tot)
    if dump:
        route.store_developed(dump)
    return {"steps_completed": route.k, "done": done, "total_s": total,
            "near_field_pair_fraction_per_condition": near}


def mlp_kernels(rows=65536, steps=20, infer_rows=73468):
    """The warm-start network's GEMMs alone: `steps` Adam steps on a minibatch of `rows`
    random rows, and `steps` forward passes over `infer_rows` rows (the strips of the
    512-instance polar sweep), timed with HIP events around the stream's work (the library
    synchronizes each call).  Algorithmic flops per row: forward 2 (64*256 + 3*256*256 +
    256); a training step adds the weight gradients of every layer and the input gradients
    of layers 1-3, 2 (64*256 + 3*256*256) + 2 (3*256*256)."""
    import torch

    import stormbird_amd as sb

    rng = np.random.default_rng(0)
    X = torch.from_numpy(rng.standard_normal((max(rows, infer_rows), 64)).astype(np.float32)).to("cuda")
    T = torch.from_numpy(rng.standard_normal(rows).astype(np.float32)).to("cuda")
    Y = torch.zeros(infer_rows, dtype=torch.float32, device="cuda")
    m = sb.Mlp(seed=1)
    fwd = 2.0 * (64 * 256 + 3 * 256 * 256 + 256)
    train = fwd + 2.0 * (64 * 256 + 3 * 256 * 256) + 2.0 * (3 * 256 * 256)
    for _ in range(2):
        m.train_step(X.data_ptr(), T.data_ptr(), rows, 1e-4)
        m.forward(X.data_ptr(), infer_rows, Y.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        m.train_step(X.data_ptr(), T.data_ptr(), rows, 1e-4)
    t_train = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        m.forward(X.data_ptr(), infer_rows, Y.data_ptr())
    t_fwd = (time.perf_counter() - t0) / steps
    return {"train_rows": rows, "train_step_ms": 1e3 * t_train,
            "train_tflops_host_timed": train * rows / t_train / 1e12,
            "forward_rows": infer_rows, "forward_ms": 1e3 * t_fwd,
            "forward_tflops_host_timed": fwd * infer_rows / t_fwd / 1e12,
            "note": "host-timed whole calls (launch and synchronization included); per-kernel times in the "
                    "rocprofv3 summary of `bench.py --only mlp`"}


def single_wing_sail(steps=200):
    """The plain single-instance workload: one 30 x 10 m wing sail, 20 cosine-spaced strips,
    frozen-shape wake of 200 rows (4,000 panels), 10 m/s at 10 deg, 200 steps from the start.
    Timed host to host (the reference API: freestream and results in host memory), with the
    freestream resident on the device, and on the CPU restatement (one thread: this path is
    serial in the reference apart from its per-target loops, which are short here).  At
    7.96e4 fixed-velocity pairs per step the step is latency-bound: a chain of small
    kernels and the device-side fixed-point loop, not the pair rate."""
    import torch

    import stormbird_amd as sb
    from stormbird_amd import setups as S

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py

    js = S.single_wing_sail(nr_panels=200)
    case = S.single_wing_sail_case()
    dt = case["time_step"]
    out = {"workload": "single 30x10 m wing sail, 20 cosine strips, frozen-shape wake of 200 rows "
                       "(4,000 panels), 10 m/s at 10 deg, 200 steps from the start",
           "steps": steps}
    for label in ("host_to_host", "device_resident"):
        sim = sb.Simulation(js, mode=sb.MODE_EXACT)
        fs = S.uniform_freestream(sim.nr_freestream_points(), case["velocity"])
        fs_dev = torch.from_numpy(np.ascontiguousarray(fs)).to("cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            if label == "host_to_host":
                sim.do_step(time=k * dt, time_step=dt, freestream_velocity=fs)
            else:
                sim.do_step_device(time=k * dt, time_step=dt, freestream_ptr=fs_dev.data_ptr(),
                                   n_points=fs.shape[0], want_result=(k == steps - 1))
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        out[label] = {"time_steps_per_s": steps / t, "ms_per_step": 1000.0 * t / steps}
    threads = oracle_py.set_threads(0)
    oracle_py.set_threads(1)
    orc = oracle_py.Simulation(js, "glibc")
    fs = S.uniform_freestream(orc.nr_freestream_points(), case["velocity"])
    t0 = time.perf_counter()
    for k in range(steps):
        orc.do_step(time=k * dt, time_step=dt, freestream_velocity=fs)
    t = time.perf_counter() - t0
    oracle_py.set_threads(threads)
    out["cpu_restatement_one_thread"] = {"time_steps_per_s": steps / t, "ms_per_step": 1000.0 * t / steps}
    out["gpu_host_to_host_vs_cpu"] = out["host_to_host"]["time_steps_per_s"] / out["cpu_restatement_one_thread"][
        "time_steps_per_s"]
    out["bound"] = "latency (launch chain and the device-side fixed-point loop; 7.96e4 pairs per step)"
    return out


def quasi_steady(steps=50):
    """Quasi-steady path (horseshoe frozen wake + Linearized solve, the reference's default
    settings): the reference's speed-test setup (3 wing sails x 40 strips,
    wing_sails_speed_test.rs) and a 3 x 340-strip case whose 1020 x 1020 system sits in
    global memory; per-step GPU time beside the CPU restatement (one thread)."""
    import stormbird_amd as sb
    from stormbird_amd import setups as S

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py

    cases = [("3 wing sails x 40 strips (N = 120), Linearized", S.wing_sails_quasi_steady(), [1.2, 0.0, 0.0], steps),
             ("3 wing sails x 340 strips (N = 1020), Linearized", S.wing_sails_quasi_steady(nr_strips=340),
              [1.2, 0.0, 0.0], max(3, steps // 10)),
             ("3 wing sails x 40 strips (N = 120), quasi-steady SimpleIterative from the Linearized start",
              S.wing_sails_quasi_steady(solver={"start_with_linearized_solution": True}), [1.2, 0.0, 0.0], steps)]
    out = []
    for name, js, u, n in cases:
        g = sb.Simulation(js)
        fs = S.uniform_freestream(g.nr_freestream_points(), u)
        for _ in range(2):
            g.do_step(time=0.0, time_step=1.0, freestream_velocity=fs)
        t0 = time.perf_counter()
        for _ in range(n):
            r = g.do_step(time=0.0, time_step=1.0, freestream_velocity=fs)
        t_gpu = (time.perf_counter() - t0) / n
        # the kernel split in further steps with the interval timers on (their event
        # records would add ~25 us to each host-to-host step above)
        g.set_kernel_timers(True)
        dev = []
        for _ in range(n):
            g.do_step(time=0.0, time_step=1.0, freestream_velocity=fs)
            dev.append(sum(g.last_kernel_ms(k) for k in ("factors", "solve", "forces")))
        o = oracle_py.Simulation(js, "glibc")
        o.do_step(time=0.0, time_step=1.0, freestream_velocity=fs)
        n_cpu = max(1, min(n, 5)) if g.nr_ctrl_points > 500 else n
        t0 = time.perf_counter()
        for _ in range(n_cpu):
            o.do_step(time=0.0, time_step=1.0, freestream_velocity=fs)
        t_cpu = (time.perf_counter() - t0) / n_cpu
        out.append({"workload": name, "n_ctrl_points": g.nr_ctrl_points, "iterations": r.iterations,
                    "gpu_ms_per_step": 1e3 * t_gpu,
                    "gpu_device_ms_per_step": float(np.mean(dev)),
                    "kernel_ms": {k: g.last_kernel_ms(k) for k in ("factors", "solve", "forces")},
                    "linearized_matrix_in_shared_memory": bool(
                        sb.lib().sb_debug_linearized_in_shared_memory(g.nr_ctrl_points) == 1),
                    "cpu_ms_per_step": 1e3 * t_cpu, "cpu": "oracle restatement, 1 thread",
                    "gpu_vs_cpu": t_cpu / t_gpu})
        del g, o
    # batched polar: many quasi-steady configurations, one workgroup per instance
    n_inst, seed = 1024, 4
    rng = np.random.default_rng(seed)
    setups, pcases = zip(*[S.quasi_steady_polar_instance(rng) for _ in range(n_inst)])
    batch = sb.Batch(list(setups))
    fs_all = np.concatenate([S.uniform_freestream(inst.nr_freestream_points(), c["velocity"])
                             for inst, c in zip(batch.instances, pcases)])
    dts = np.array([c["time_step"] for c in pcases])
    for k in range(2):
        batch.do_step(times=k * dts, time_steps=dts, freestream_velocity=fs_all)
    reps = 5
    dev = []
    t0 = time.perf_counter()
    for k in range(reps):
        batch.do_step(times=k * dts, time_steps=dts, freestream_velocity=fs_all)
        dev.append(batch.last_step_ms())
    t_b = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for k in range(reps):
        batch.do_step(times=k * dts, time_steps=dts, freestream_velocity=fs_all, want_result=False)
    t_nr = (time.perf_counter() - t0) / reps
    sample = 64
    alone = [sb.Simulation(js) for js in setups[:sample]]
    parts = [S.uniform_freestream(a.nr_freestream_points(), c["velocity"]) for a, c in zip(alone, pcases)]
    for a, f, dt in zip(alone, parts, dts):
        a.do_step(time=0.0, time_step=dt, freestream_velocity=f)
    t0 = time.perf_counter()
    for a, f, dt in zip(alone, parts, dts):
        a.do_step(time=0.0, time_step=dt, freestream_velocity=f)
    t_plain = (time.perf_counter() - t0) / sample
    orcs = [oracle_py.Simulation(js, "glibc") for js in setups[:sample]]
    t0 = time.perf_counter()
    for o, f, dt in zip(orcs, parts, dts):
        o.do_step(time=0.0, time_step=dt, freestream_velocity=f)
    t_cpu = (time.perf_counter() - t0) / sample
    n_tot = int(sum(inst.nr_ctrl_points for inst in batch.instances))
    polar = {"workload": (f"{n_inst} quasi-steady polar instances (seed {seed}): 1-4 sails, 8-32 strips per sail, "
                          "chords 5-15 m, AR 2-6, AWA 20-180 deg, AWS 6-18 m/s, default settings (horseshoe wake, "
                          "Linearized solve)"),
             "n_instances": n_inst, "n_ctrl_points_total": n_tot,
             "batched_ms_per_step": 1e3 * t_b, "batched_device_ms_per_step": float(np.mean(dev)),
             "batched_ms_per_step_results_left_on_device": 1e3 * t_nr,
             "batched_instance_steps_per_s": n_inst / t_b,
             "plain_gpu_ms_per_instance_step": 1e3 * t_plain, "plain_gpu_instance_steps_per_s": 1.0 / t_plain,
             "cpu_ms_per_instance_step": 1e3 * t_cpu, "cpu_instance_steps_per_s": 1.0 / t_cpu,
             "cpu": f"oracle restatement, 1 thread, first {sample} instances",
             "batched_vs_plain_gpu": t_plain * n_inst / t_b, "batched_vs_cpu_1_thread": t_cpu * n_inst / t_b}
    del batch, alone, orcs
    return {"cases": out, "batched_polar": polar,
            "step_times": "host-to-host do_step (freestream upload and result download included); "
                          "device time is the sum of the step's kernel intervals"}


def complete_sail_model_section(n_directions=36, n_loadings=5, steps=3):
    """CompleteSailModel (complete_sail_model/mod.rs): a polar of optimal-loading steady
    states for the reference's speed-test sails (quasi-steady, 3 x 40 strips), GPU beside the
    CPU restatement on one thread; and dynamic steps of the four-sail free wake (2,000 rows)
    driven by a power-law wind with gusts, the wake-point freestream evaluated on the device,
    beside the plain step with a device-resident freestream."""
    import torch

    import stormbird_amd as sb
    from stormbird_amd import setups as S

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_csm

    gust = {"frequencies": [0.05, 0.12, 0.31], "amplitudes": [0.6, 0.3, 0.15], "phases": [0.2, 1.9, 4.0]}
    conditions = [{"direction_coming_from": math.radians(-180.0 + 360.0 * i / n_directions),
                   "velocity_variation": {"PowerModel": {"reference_velocity": 9.0}}}
                  for i in range(n_directions)]
    js = S.complete_sail_model(S.wing_sails_quasi_steady(), 3)
    g = sb.CompleteSailModel(js)
    g.simulate_condition_optimal_controller_loading(wind_condition=sb.WindCondition(conditions[0]), ship_velocity=6.0,
                                                    nr_loadings_to_test=n_loadings)
    t0 = time.perf_counter()
    for c in conditions:
        g.simulate_condition_optimal_controller_loading(wind_condition=sb.WindCondition(c), ship_velocity=6.0,
                                                        nr_loadings_to_test=n_loadings)
    t_gpu = (time.perf_counter() - t0) / len(conditions)
    o = oracle_csm.CompleteSailModel(js, "glibc")
    sample = conditions[::6]
    t0 = time.perf_counter()
    for c in sample:
        o.optimal_loading(c, 6.0, n_loadings)
    t_cpu = (time.perf_counter() - t0) / len(sample)
    polar = {"workload": (f"3 wing sails x 40 strips (quasi-steady, Linearized), {n_directions} wind directions, "
                          f"power-law wind 9 m/s at 10 m, ship 6 m/s, optimal loading of {n_loadings}"),
             "gpu_ms_per_condition": 1e3 * t_gpu, "cpu_ms_per_condition": 1e3 * t_cpu,
             "cpu": f"restatement (Python wind and controller, C++ lifting line), 1 thread, {len(sample)} conditions",
             "gpu_vs_cpu": t_cpu / t_gpu}
    del g, o

    js4, case = S.four_sail_free_wake()
    dt = case["time_step"]
    wc = {"direction_coming_from": math.radians(case["awa_deg"]),
          "velocity_variation": {"PowerModel": {"reference_velocity": 12.0}},
          "parallel_gust": gust, "perpendicular_gust": gust}
    m = sb.CompleteSailModel(S.complete_sail_model(js4, 4))
    sim = m.simulation
    pts, vel, gam, prev, _ = S.developed_four_sail_state(case)
    fs = S.uniform_freestream(sim.nr_freestream_points(), case["velocity"])
    sim.set_wake_state(pts, vel, gam, 2000)
    sim.previous_circulation_strength = prev
    w = sb.WindCondition(wc)
    m.do_step(time=2000 * dt, time_step=dt, wind_condition=w, ship_velocity=0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        m.do_step(time=(2001 + k) * dt, time_step=dt, wind_condition=w, ship_velocity=0.0)
    torch.cuda.synchronize()
    t_csm = (time.perf_counter() - t0) / steps
    n_points = sim.nr_freestream_points()
    fs_dev = torch.from_numpy(np.ascontiguousarray(fs)).to("cuda")
    e, _ = time_gpu(sim, fs_dev.data_ptr(), n_points, dt, steps, 1, 2010)
    t_plain = e / steps
    dynamic = {"workload": ("four 30x10 m wing sails, 40 strips each, free wake of 2000 rows (328164 wake "
                            "points), power-law wind 12 m/s at 10 m with gusts along and across, from the rolled-up "
                            "wake of a 2,000-step run"),
               "csm_ms_per_step": 1e3 * t_csm, "plain_device_freestream_ms_per_step": 1e3 * t_plain,
               "wind_evaluation_overhead_ms": 1e3 * (t_csm - t_plain),
               "freestream": "control points on the host (inflow corrections, controller), wake points on the device"}
    del m, sim
    return {"optimal_loading_polar": polar, "dynamic_steps": dynamic}


def full_run(steps=2000, nr_strips=40, dump=None):
    """The four-sail free wake run from the start for `steps` steps (the wake fills one row
    per step, so the self-induction work grows with the square of the step number); progress
    on stderr every 100 steps."""
    import torch

    import stormbird_amd as sb
    from stormbird_amd import setups as S

    js, case = S.four_sail_free_wake(nr_panels=steps, nr_strips=nr_strips)
    dt = case["time_step"]
    sim = sb.Simulation(js, mode=sb.MODE_EXACT)
    sim.set_kernel_timers(True)
    fs = S.uniform_freestream(sim.nr_freestream_points(), case["velocity"])
    fs_dev = torch.from_numpy(np.ascontiguousarray(fs)).to("cuda")
    n_pts = sim.nr_freestream_points()
    marks = {}
    t0 = time.perf_counter()
    res = None
    for k in range(steps):
        want = (k + 1) % 100 == 0 or k == steps - 1
        res = sim.do_step_device(time=k * dt, time_step=dt, freestream_ptr=fs_dev.data_ptr(), n_points=n_pts,
                                 want_result=want)
        if want:
            torch.cuda.synchronize()
            marks[k + 1] = time.perf_counter() - t0
            print(f"full run: step {k + 1}, {marks[k + 1]:.1f} s, iterations {res.iterations}", file=sys.stderr,
                  flush=True)
    total = time.perf_counter() - t0
    gam = res.force_input.circulation_strength
    st = sim.get_wake_state()
    near, pairs = sim.count_self_induction_pairs()
    kernel_ms = {k: sim.last_kernel_ms(k) for k in ("wake_update", "fixed", "factors", "solve", "self_induction",
                                                     "forces")}
    if dump:
        np.savez(dump, points=st["points"], velocities=st["velocities"], strengths=st["strengths"],
                 previous_circulation=sim.previous_circulation_strength, steps_completed=st["steps_completed"])
    last = steps - 100 if steps > 100 else 0
    return {
        "workload": (f"four 30x10 m wing sails, {nr_strips} strips each, free wake run from the start for "
                     f"{steps} steps (every wake point self-induced), EXACT"),
        "steps": steps, "total_s": total, "mean_time_steps_per_s": steps / total,
        "last_100_time_steps_per_s": (steps - last) / (total - marks.get(last, 0.0)) if last else None,
        "seconds_at": {str(k): v for k, v in marks.items() if k % 500 == 0 or k == steps},
        "final": {"iterations": res.iterations, "residual": res.residual,
                  "circulation_max_abs": float(np.max(np.abs(gam))),
                  "all_finite": bool(np.all(np.isfinite(gam)) and np.all(np.isfinite(st["points"]))),
                  "wake_extent_m": float(np.max(np.linalg.norm(st["points"] - st["points"][0], axis=1))),
                  "integrated_force": [float(x) for x in res.integrated_forces_sum()]},
        "last_step": {"self_induction_pairs": {"near": near, "far": pairs - near}, "kernel_ms": kernel_ms},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--panels", type=int, default=2000, help="wake panel rows (default 2000)")
    ap.add_argument("--strips", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fast", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-instances", type=int, default=512)
    ap.add_argument("--no-warm", action="store_true")
    ap.add_argument("--no-route", action="store_true")
    ap.add_argument("--route-ships", type=int, default=64)
    ap.add_argument("--train-solves", type=int, default=200_000, help="first-step solves in the warm start's training set")
    ap.add_argument("--no-quasi", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-csm", action="store_true")
    ap.add_argument("--resume", default=None, help="with --only route-rollup: continue from this checkpoint")
    ap.add_argument("--budget-s", type=float, default=None,
                    help="with --only route-rollup: stop after this many seconds and write a checkpoint")
    ap.add_argument("--dump-state", default=None,
                    help="with --only full or route-rollup: write the final wake state (.npz)")
    ap.add_argument("--only", choices=("batched", "warm", "route", "route-rollup", "route-warm", "quasi", "csm",
                                       "full", "single", "mlp"),
                    default=None,
                    help="run one extra section alone and print its JSON (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU layout (target points "
                                   "sharded, panels replicated, one all-gather of wake points per step) "
                                   "is described in DESIGN.md but not built", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    import stormbird_amd as sb
    from stormbird_amd import setups as S

    if args.only:
        if args.only == "batched":
            out = {"exact": batched_polar_sweep(n_inst=args.batch_instances, steps=args.steps),
                   "fast": batched_polar_sweep(n_inst=args.batch_instances, steps=args.steps, mode=1)}
        elif args.only == "warm":
            out, _ = warm_start_polar(n_inst=args.batch_instances, train_solves=args.train_solves)
        elif args.only == "quasi":
            out = quasi_steady()
        elif args.only == "single":
            out = single_wing_sail()
        elif args.only == "csm":
            out = complete_sail_model_section()
        elif args.only == "full":
            out = full_run(steps=args.panels, nr_strips=args.strips, dump=args.dump_state)
        elif args.only == "route-warm":
            out, _ = route_warm_start()
        elif args.only == "mlp":
            out = mlp_kernels()
        elif args.only == "route-rollup":
            out = route_rollup(steps=args.panels, dump=args.dump_state, resume=args.resume,
                               budget_s=args.budget_s)
        else:
            out = route_simulation(n_ships=args.route_ships, steps=args.steps)
        print(json.dumps({"section": args.only, **out}))
        return

    js, case = S.four_sail_free_wake(nr_panels=args.panels, nr_strips=args.strips)
    dt = case["time_step"]
    sim = sb.Simulation(js, mode=sb.MODE_EXACT)
    if args.panels == 2000 and args.strips == 40:
        pts, vel, gam, prev, _ = S.developed_four_sail_state(case)
        fs = S.uniform_freestream(sim.nr_freestream_points(), case["velocity"])
        state = "rolled-up wake of a 2,000-step run from the start (stored state)"
    else:
        pts, vel, gam, prev, fs = synthesize_developed_state(sim, case)
        state = "synthesized developed wake (convected sheet with seeded noise)"
    sim.set_wake_state(pts, vel, gam, 2000)
    sim.previous_circulation_strength = prev
    n_points = sim.nr_freestream_points()
    fs_dev = torch.from_numpy(np.ascontiguousarray(fs)).to("cuda")
    near, total = sim.count_self_induction_pairs()

    elapsed, kms = time_gpu(sim, fs_dev.data_ptr(), n_points, dt, args.steps, args.warmup, 2000)
    ms_per_step = 1000.0 * elapsed / args.steps
    steps_per_s = args.steps / elapsed
    pc = sim.last_pair_counts()
    pairs_per_step = float(pc[0] + pc[1] + pc[2])
    k_ms = float(np.mean(kms))
    flops = (total - near) * FLOP_FAR + near * FLOP_NEAR
    achieved = flops / (k_ms * 1e-3) / 1e12

    fast = fp32 = None
    if not args.no_fast:
        # FAST and FP32-Kahan: one step from the developed state against the EXACT step from
        # the same state (EXACT equals the f64 restatement bit for bit), then each mode's rate.
        refs = {}
        for md in (sb.MODE_EXACT, sb.MODE_FAST, sb.MODE_FP32_KAHAN):
            s_ = sb.Simulation(js, mode=md)
            s_.set_wake_state(pts, vel, gam, 2000)
            s_.previous_circulation_strength = prev
            r_ = s_.do_step(time=2000 * dt, time_step=dt, freestream_velocity=fs)
            refs[md] = (s_, r_.as_arrays(), s_.get_wake_state()["velocities"])
        _, a64, v64 = refs.pop(sb.MODE_EXACT)

        def rel(a, b):
            return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))

        def mode_line(md):
            s_, a_, v_ = refs.pop(md)
            e_, k_ = time_gpu(s_, fs_dev.data_ptr(), n_points, dt, args.steps, args.warmup, 2001)
            km = float(np.mean(k_))
            return {"time_steps_per_s": args.steps / e_, "ms_per_step": 1000.0 * e_ / args.steps,
                    "self_induction_ms": km, "self_induction_tflops": flops / (km * 1e-3) / 1e12,
                    "max_rel_dev_vs_f64": {"circulation": rel(a_["circulation_strength"], a64["circulation_strength"]),
                                           "integrated_forces": rel(a_["integrated_forces"], a64["integrated_forces"]),
                                           "wake_velocities": rel(v_, v64)},
                    "deviation_measured": "one step from the developed state, against the EXACT (f64) step"}

        fast = mode_line(sb.MODE_FAST)
        fp32 = mode_line(sb.MODE_FP32_KAHAN)
        del refs

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "self_induction_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(js, (pts, vel, gam, prev), float(sim.nr_panels), pairs_per_step)

    line = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"four 30x10 m wing sails 25 m apart, {args.strips} strips each, free wake of "
                         f"{args.panels} panel rows ({sim.nr_panels} panels, {sim.nr_wake_points} wake points), "
                         f"all wake points self-induced, steady state from the {state}"),
            "mode": "exact (bit-identical to the f64 restatement)",
            "n_ctrl_points": sim.nr_ctrl_points,
            "n_panels": sim.nr_panels,
            "n_wake_points": sim.nr_wake_points,
        },
        "biot_savart_pairs_per_step": pairs_per_step,
        "biot_savart_pairs_per_s": pairs_per_step * steps_per_s,
        "self_induction_pairs": {"near": near, "far": total - near},
        "self_induction_kernel_ms": k_ms,
        "roofline": {
            # compute-bound on the FP64 vector ALUs (a pairwise sum, not a matrix product);
            # their peak on MI355X is 78.6 TFLOP/s with a fused multiply-add counted as two
            "bound": "valu",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "compute": "fp64 vector ALU (pairwise Biot-Savart sum, not a matrix product)",
        },
        "kernel_ms": {k: sim.last_kernel_ms(k) for k in
                      ("wake_update", "fixed", "factors", "solve", "self_induction", "forces")},
        "fast_mode": fast,
        "fp32_kahan_mode": fp32,
        "cpu_baseline": cpu,
    }
    mlp = None
    if not args.no_single:
        line["single_wing_sail"] = single_wing_sail()
    if not args.no_quasi:
        line["quasi_steady"] = quasi_steady()
    if not args.no_csm:
        line["complete_sail_model"] = complete_sail_model_section()
    del sim
    if not args.no_warm:
        import gc

        gc.collect()
        line["warm_start"], mlp = warm_start_polar(n_inst=args.batch_instances, train_solves=args.train_solves)
    if not args.no_batch:
        import gc

        gc.collect()
        line["batched"] = batched_polar_sweep(n_inst=args.batch_instances, mlp=mlp)
        gc.collect()
        line["batched_fast"] = batched_polar_sweep(n_inst=args.batch_instances, mode=1, mlp=mlp)
    if not args.no_route:
        import gc

        gc.collect()
        route_warm, route_mlp = route_warm_start(n_ships=args.route_ships)
        line["route"] = route_simulation(n_ships=args.route_ships, mlp=route_mlp)
        line["route"]["warm_start_short_wakes"] = route_warm
    print(json.dumps(line))


if __name__ == "__main__":
    main()
