"""GPU: many instances stepped together (sb_batch_*) against each instance stepped alone.

Bar: bit-identical results and wake states per instance and per step, in
EXACT and FAST mode, over uneven instances (strips, wings, wake rows, time
steps, free and convected wakes, symmetry), across the first step (wake
initialisation), the start of the free-wake self-induction and several
replays of the captured step graph.  One instance is also held to the
oracle directly.
"""
import numpy as np
import pytest

import stormbird_amd as sb
from stormbird_amd import setups as S

from test_gpu_parity import compare_results, same

pytestmark = pytest.mark.gpu


def mixed_setups(seed=7, n=10):
    rng = np.random.default_rng(seed)
    setups, cases = [], []
    for _ in range(n):
        js, case = S.polar_sweep_instance(rng, max_panels=40, min_panels=6)
        setups.append(js)
        cases.append(case)
    # a free-wake instance and a symmetric one (self-induction and mirror paths)
    js, case = S.polar_sweep_instance(rng, max_panels=30, min_panels=20)
    d = __import__("json").loads(js)
    d["simulation_settings"]["Dynamic"]["wake"]["ratio_of_wake_affected_by_induced_velocities"] = 1.0
    setups.append(__import__("json").dumps(d))
    cases.append(case)
    js, case = S.polar_sweep_instance(rng, max_panels=25, min_panels=10)
    d = __import__("json").loads(js)
    d["simulation_settings"]["Dynamic"]["wake"]["symmetry_condition"] = "Z"
    setups.append(__import__("json").dumps(d))
    cases.append(case)
    return setups, cases


def run_pair(mode, steps):
    setups, cases = mixed_setups()
    batch = sb.Batch(setups, mode=mode)
    alone = [sb.Simulation(js, mode=mode) for js in setups]
    fs_parts = [S.uniform_freestream(s.nr_freestream_points(), c["velocity"]) for s, c in zip(alone, cases)]
    fs_all = np.concatenate(fs_parts)
    assert fs_all.shape[0] == batch.nr_freestream_points
    dts = np.array([c["time_step"] for c in cases])
    for k in range(steps):
        print(f"batch step {k} mode {mode}", flush=True)
        rb = batch.do_step(times=k * dts, time_steps=dts, freestream_velocity=fs_all)
        for i, sim in enumerate(alone):
            ra = sim.do_step(time=k * dts[i], time_step=dts[i], freestream_velocity=fs_parts[i])
            compare_results(rb[i], ra, exact=True, where=f"step {k} instance {i}")
        if k in (0, 3, steps - 1):
            for i, sim in enumerate(alone):
                wb, wa = batch.instances[i].get_wake_state(), sim.get_wake_state()
                for key in ("points", "velocities", "strengths"):
                    assert same(wb[key], wa[key]), f"step {k} instance {i}: wake {key}"
    return batch


def test_batch_equals_alone_exact():
    batch = run_pair(sb.MODE_EXACT, 9)
    assert batch.last_step_ms() > 0.0


def test_batch_equals_alone_fast():
    run_pair(sb.MODE_FAST, 6)


def test_batch_instance_matches_oracle(oracle):
    rng = np.random.default_rng(11)
    setups, cases = [], []
    for _ in range(3):
        js, case = S.polar_sweep_instance(rng, max_panels=20, min_panels=8)
        setups.append(js)
        cases.append(case)
    batch = sb.Batch(setups)
    orc = oracle.Simulation(setups[1], variant="pm")
    parts = [S.uniform_freestream(s.nr_freestream_points(), c["velocity"]) for s, c in zip(batch.instances, cases)]
    dts = np.array([c["time_step"] for c in cases])
    for k in range(8):
        rb = batch.do_step(times=k * dts, time_steps=dts, freestream_velocity=np.concatenate(parts))
        ro = orc.do_step(time=k * dts[1], time_step=dts[1], freestream_velocity=parts[1])
        compare_results(rb[1], ro, exact=True, where=f"step {k}")


def test_batch_rejects_bad_input():
    js, _ = S.polar_sweep_instance(np.random.default_rng(1), max_panels=10, min_panels=5)
    with pytest.raises(ValueError):
        sb.Batch([js, "{not json"])
    batch = sb.Batch([js, js])
    with pytest.raises(RuntimeError):
        batch.do_step(times=None, time_steps=0.1, freestream_velocity=np.zeros((3, 3)))


def test_route_simulation_batch_equals_alone(oracle):
    """Rolling ships in gusty wind with free wakes: the batch's freestream is the reference's
    WindEnvironment at every point (the oracle's restatement, bit for bit: power-law profile,
    gusts along and across the wind, ship speed), and the batched steps equal each ship
    stepped alone with that freestream."""
    import oracle_csm

    from stormbird_amd.route import RouteSimulation, wind_environment

    route = RouteSimulation(n_ships=3, n_conditions=2, seed=5, nr_strips=8, nr_panels=12, nr_freqs=50)
    alone = [sb.Simulation(js, mode=sb.MODE_EXACT) for js in route.setups]
    m = oracle_csm.Libm("pm")
    env = oracle_csm.WindEnvironment(wind_environment(), m)
    conds = [oracle_csm.WindCondition(wc, m) for wc in route.wind_conditions]
    off = route.batch.freestream_offsets
    for k in range(7):
        print(f"route step {k}", flush=True)
        pts = []
        for i, sim in enumerate(alone):
            dt = float(route.dt[i])
            sim.set_translation_and_rotation_with_finite_difference_for_the_velocity(
                dt, [0.0, 0.0, 0.0], [route.roll(route.k * dt), 0.0, 0.0])
            pts.append(sim.get_freestream_velocity_points())
        t = route.k * route.dt
        rb = route.step()
        fs_all = route.batch.get_freestream()
        for i, sim in enumerate(alone):
            fs = fs_all[off[i]:off[i + 1]]
            linear = oracle_csm.scale(env.zero, float(route.ship_speed[i]))
            want = np.array([env.unsteady_apparent(conds[i], tuple(p), linear, float(t[i])) for p in pts[i]])
            assert same(fs, want), f"step {k} ship {i}: freestream"
            ra = sim.do_step(time=float(t[i]), time_step=float(route.dt[i]), freestream_velocity=fs)
            compare_results(rb[i], ra, exact=True, where=f"step {k} ship {i}")
    # a judged warm start (untrained network) changes only the initial guess
    net = sb.Mlp(seed=1)
    route.batch.set_warm_start(net, 3)
    res = route.step()
    assert route.batch.last_warm_taken() + route.batch.last_fallbacks() <= 3
    for r in res:
        assert np.all(np.isfinite(r.force_input.circulation_strength))
    with pytest.raises(ValueError):
        route.batch.set_wind(wind_environment(), route.wind_conditions[:2])


@pytest.mark.parametrize("solver", [None, {"start_with_linearized_solution": True,
                                           "max_iterations_per_time_step": 400}])
def test_quasi_steady_batch_matches_single_instances(oracle, solver):
    """A batch of quasi-steady instances (horseshoe wake; Linearized, or SimpleIterative from
    the Linearized start) equals every instance stepped alone, bit for bit, over several steps;
    two instances are held to the oracle too."""
    rng = np.random.default_rng(11)
    setups, cases = zip(*[S.quasi_steady_polar_instance(rng, solver=solver) for _ in range(12)])
    batch = sb.Batch(list(setups))
    alone = [sb.Simulation(js) for js in setups]
    orcs = {0: oracle.Simulation(setups[0], "pm"), 5: oracle.Simulation(setups[5], "pm")}
    assert batch.nr_freestream_points == sum(s.nr_ctrl_points for s in alone)
    dts = np.array([c["time_step"] for c in cases])
    for k in range(3):
        print(f"quasi batch step {k}", flush=True)
        parts = [S.uniform_freestream(s.nr_freestream_points(), np.array(c["velocity"]) * (1.0 + 0.05 * k))
                 for s, c in zip(alone, cases)]
        rb = batch.do_step(times=k * dts, time_steps=dts, freestream_velocity=np.concatenate(parts))
        for i, sim in enumerate(alone):
            ra = sim.do_step(time=k * dts[i], time_step=dts[i], freestream_velocity=parts[i])
            compare_results(rb[i], ra, exact=True, where=f"step {k} instance {i}")
            if i in orcs:
                ro = orcs[i].do_step(time=k * dts[i], time_step=dts[i], freestream_velocity=parts[i])
                compare_results(rb[i], ro, exact=True, where=f"oracle step {k} instance {i}")


def test_batch_rejects_mixed_and_warm_quasi():
    rng = np.random.default_rng(3)
    q, _ = S.quasi_steady_polar_instance(rng)
    d, _ = S.polar_sweep_instance(rng, max_panels=10, min_panels=5)
    with pytest.raises(ValueError, match="either quasi-steady or dynamic"):
        sb.Batch([q, d])
    b = sb.Batch([q, q])
    m = sb.Mlp(seed=1)
    with pytest.raises(RuntimeError, match="dynamic batches"):
        b.set_warm_start(m, 1)


@pytest.mark.parametrize("symmetric", [False, True])
def test_autotuner_and_every_variant_keep_results(tmp_path, symmetric):
    """The autotuner times the kernel variants on the batch's own state without changing
    it; every variant (panels per wave of the fixed-velocity kernel, LDS tile length of the
    self-induction kernel) gives the same bits; the choice is cached per device and shape
    class in a JSON file and reused by a batch of the same shape."""
    import json

    setups, cases = mixed_setups()
    if not symmetric:
        setups, cases = setups[:-1], cases[:-1]
    ref, tuned = sb.Batch(setups), sb.Batch(setups)
    fs_all = np.concatenate([S.uniform_freestream(s.nr_freestream_points(), c["velocity"])
                             for s, c in zip(ref.instances, cases)])
    dts = np.array([c["time_step"] for c in cases])

    def step_both(k):
        print(f"tuning step {k}", flush=True)
        ra = ref.do_step(times=k * dts, time_steps=dts, freestream_velocity=fs_all)
        rb = tuned.do_step(times=k * dts, time_steps=dts, freestream_velocity=fs_all)
        for i in range(len(setups)):
            compare_results(rb[i], ra[i], exact=True, where=f"step {k} instance {i}")

    for k in range(5):  # the free wake's self-induction runs from the fourth step on
        step_both(k)
    cache = str(tmp_path / "tuning.json")
    cached, choice, t = tuned.autotune(cache)
    assert not cached and choice[0] in (1, 2, 3) and choice[1] in (64, 128, 256)
    assert np.all(t > 0.0), t  # both kernels had work
    for i in range(len(setups)):  # the state is unchanged
        wa, wb = ref.instances[i].get_wake_state(), tuned.instances[i].get_wake_state()
        for key in ("points", "velocities", "strengths"):
            assert same(wa[key], wb[key]), f"instance {i}: wake {key} after tuning"
    k = 5
    for variant in (choice, (1, 64), (2, 128), (3, 256)):
        tuned.set_tuning(*variant)
        assert tuned.get_tuning() == variant
        step_both(k)
        k += 1
    entries = json.load(open(cache))["entries"]
    assert len(entries) == 1 and entries[0]["fixed_per"] == choice[0] and entries[0]["targets_tile"] == choice[1]
    other = sb.Batch(setups)
    for j in range(5):
        other.do_step(times=j * dts, time_steps=dts, freestream_velocity=fs_all, want_result=False)
    hit, choice2, t2 = other.autotune(cache)
    assert hit and choice2 == choice and np.allclose(t2, t, atol=1e-4)
    with pytest.raises(RuntimeError):
        other.set_tuning(4, 256)
    fresh = sb.Batch(setups)
    with pytest.raises(RuntimeError, match="take a step"):
        fresh.autotune(None)


def test_residual_warm_start_with_a_zero_network_is_the_default_guess():
    """In residual mode the network predicts the change from the previous circulation: a
    network whose every parameter is zero predicts no change, and the steps then equal the
    steps without a warm start bit for bit."""
    from stormbird_amd.route import RouteSimulation

    a = RouteSimulation(n_ships=3, n_conditions=2, seed=6, nr_strips=8, nr_panels=12, nr_freqs=50)
    b = RouteSimulation(n_ships=3, n_conditions=2, seed=6, nr_strips=8, nr_panels=12, nr_freqs=50)
    net = sb.Mlp(seed=2)
    net.set_params(np.zeros(net.nr_params, dtype=np.float32))
    b.batch.set_warm_start(net, 2, residual=True)
    for k in range(6):
        print(f"residual warm start step {k}", flush=True)
        ra, rb = a.step(), b.step()
        for i in range(3):
            compare_results(rb[i], ra[i], exact=True, where=f"step {k} ship {i}")
    assert b.batch.last_fallbacks() == 0
