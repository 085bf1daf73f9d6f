"""GPU: the batched workloads at their full sizes, one step each, bit for bit.

* The 512-instance polar sweep (seed 2: 1-6 sails, 16-64 strips per sail, 200-3000 wake
  rows), developed as the benchmark develops it: every instance of the batched step
  equals the same instance stepped alone from the same state (results and wake state),
  and a seeded sample of 8 instances, among them the one with the most strips and the
  one with the longest wake, equals the oracle; in FAST arithmetic the same 8 stay within
  1e-10 of it.
* The 64-ship route (4 sails x 48 strips, 1,500-row free wake, gusts, roll) from the
  stored rolled-up wakes: the freestream the batch evaluates at the control points and at
  sampled wake points equals the oracle's WindEnvironment, a sample of ships equals the
  same ships stepped alone (results and wake state), and sampled wake points of each
  equal freestream + the oracle's induced velocity at the GPU's post-step wake.
"""
import numpy as np
import pytest

import stormbird_amd as sb
from stormbird_amd import setups as S

from test_gpu_parity import FIELDS, compare_results, rel_dev, same

pytestmark = pytest.mark.gpu


def wake_equal(a, b, where):
    for key in ("points", "velocities", "strengths"):
        assert same(a[key], b[key]), f"{where}: wake {key}"
    assert a["steps_completed"] == b["steps_completed"], where


@pytest.mark.timeout(900)
def test_polar_sweep_512_instances_at_size(oracle):
    import torch

    import bench

    rng = np.random.default_rng(2)
    made = [S.polar_sweep_instance(rng) for _ in range(512)]
    setups = [m[0] for m in made]
    cases = [m[1] for m in made]
    print("polar sweep: building 512 instances", flush=True)
    batch = sb.Batch(setups, mode=sb.MODE_EXACT)
    dts = np.array([c["time_step"] for c in cases])
    parts = [S.uniform_freestream(s.nr_freestream_points(), c["velocity"]) for s, c in zip(batch.instances, cases)]
    fs_dev = torch.from_numpy(np.concatenate(parts)).to("cuda")
    n_pts = int(fs_dev.shape[0])
    for k in range(3):
        batch.do_step_device(times=k * dts, time_steps=dts, freestream_ptr=fs_dev.data_ptr(), n_points=n_pts,
                             want_result=False)
    drng = np.random.default_rng(102)
    for s in batch.instances:
        bench.develop_wake(s, 3, drng)
    before = [(s.get_wake_state(), s.previous_circulation_strength.copy()) for s in batch.instances]
    print("polar sweep: batched step from the developed state", flush=True)
    k0 = 3
    rb = batch.do_step_device(times=k0 * dts, time_steps=dts, freestream_ptr=fs_dev.data_ptr(), n_points=n_pts,
                              want_result=True)
    n_ctrl = np.array([s.nr_ctrl_points for s in batch.instances])
    rows = np.array([s.nr_panels // s.nr_ctrl_points for s in batch.instances])
    for i, js in enumerate(setups):
        if i % 64 == 0:
            print(f"polar sweep: instance {i} alone (N {n_ctrl[i]}, M {rows[i]})", flush=True)
        st, prev = before[i]
        alone = sb.Simulation(js, mode=sb.MODE_EXACT)
        alone.set_wake_state(st["points"], st["velocities"], st["strengths"], st["steps_completed"])
        alone.previous_circulation_strength = prev
        ra = alone.do_step(time=k0 * dts[i], time_step=dts[i], freestream_velocity=parts[i])
        compare_results(rb[i], ra, exact=True, where=f"instance {i}")
        wake_equal(batch.instances[i].get_wake_state(), alone.get_wake_state(), f"instance {i}")
        del alone
    pick = {int(np.argmax(n_ctrl)), int(np.argmax(rows))}
    srng = np.random.default_rng(21)
    while len(pick) < 8:
        pick.add(int(srng.integers(0, 512)))
    for i in sorted(pick):
        print(f"polar sweep: instance {i} against the oracle (N {n_ctrl[i]}, M {rows[i]})", flush=True)
        st, prev = before[i]
        orc = oracle.Simulation(setups[i], "pm")
        orc.set_wake_state(st["points"], st["velocities"], st["strengths"], st["steps_completed"])
        orc.set_previous_circulation(prev)
        ro = orc.do_step(time=k0 * dts[i], time_step=dts[i], freestream_velocity=parts[i])
        compare_results(rb[i], ro, exact=True, where=f"instance {i} vs oracle")
        wake_equal(batch.instances[i].get_wake_state(), orc.get_wake_state(), f"instance {i} vs oracle")
        # FAST arithmetic (the batched FAST step equals FAST alone bit for bit): within 1e-10
        # of the oracle from the same state
        fast = sb.Simulation(setups[i], mode=sb.MODE_FAST)
        fast.set_wake_state(st["points"], st["velocities"], st["strengths"], st["steps_completed"])
        fast.previous_circulation_strength = prev
        rf = fast.do_step(time=k0 * dts[i], time_step=dts[i], freestream_velocity=parts[i])
        af, ao = rf.as_arrays(), ro.as_arrays()
        for f in FIELDS:
            assert rel_dev(af[f], ao[f]) <= 1e-10, f"instance {i} FAST {f}: {rel_dev(af[f], ao[f]):.3e}"
        assert abs(rf.iterations - ro.iterations) <= 1  # a 1e-14 difference may cross a stopping test
        del fast


@pytest.mark.timeout(900)
def test_route_64_ships_at_size(oracle):
    import oracle_csm

    from stormbird_amd.route import RouteSimulation, wind_environment

    print("route: building 64 ships, loading the rolled-up wakes", flush=True)
    route = RouteSimulation(n_ships=64, seed=5)
    route.load_developed()
    route.step(want_result=False)  # the first step from the stored wakes
    sample = [0, 21, 42, 63]
    before = {i: (route.batch.instances[i].get_wake_state(), route.batch.instances[i].previous_circulation_strength.copy())
              for i in sample}
    k = route.k
    print("route: batched step of 64 ships", flush=True)
    res = route.step(want_result=True)
    fs_all = route.batch.get_freestream()
    off = route.batch.freestream_offsets
    m = oracle_csm.Libm("pm")
    env = oracle_csm.WindEnvironment(wind_environment(), m)
    for i in sample:
        print(f"route: ship {i} alone", flush=True)
        dt = float(route.dt[i])
        fs = fs_all[off[i]:off[i + 1]]
        alone = sb.Simulation(route.setups[i], mode=sb.MODE_EXACT)
        # the motion state of step k: rotation at (k - 1) dt, then at k dt (angular velocity by difference)
        for kk in (k - 1, k):
            alone.set_translation_and_rotation_with_finite_difference_for_the_velocity(
                dt, [0, 0, 0], [route.roll(kk * dt), 0, 0])
        st, prev = before[i]
        alone.set_wake_state(st["points"], st["velocities"], st["strengths"], st["steps_completed"])
        alone.previous_circulation_strength = prev
        # the wind at the control points and at sampled wake points, as the reference evaluates it
        pts = alone.get_freestream_velocity_points()
        n = alone.nr_ctrl_points
        idx = np.concatenate([np.arange(n), n + np.random.default_rng(i).choice(pts.shape[0] - n, 600, replace=False)])
        c = oracle_csm.WindCondition(route.wind_conditions[i], m)
        linear = oracle_csm.scale(env.zero, float(route.ship_speed[i]))
        want = np.array([env.unsteady_apparent(c, tuple(pts[q]), linear, k * dt) for q in idx])
        assert same(fs[idx], want), f"ship {i}: freestream"
        ra = alone.do_step(time=k * dt, time_step=dt, freestream_velocity=fs)
        compare_results(res[i], ra, exact=True, where=f"ship {i}")
        w = route.batch.instances[i].get_wake_state()
        wake_equal(w, alone.get_wake_state(), f"ship {i}")
        del alone
        print(f"route: ship {i} sampled wake points against the oracle", flush=True)
        orc = oracle.Simulation(route.setups[i], "pm")
        orc.set_wake_state(w["points"], w["velocities"], w["strengths"], w["steps_completed"])
        wi = np.random.default_rng(100 + i).choice(w["points"].shape[0], 600, replace=False)
        u = orc.induced_velocities(w["points"][wi])
        assert same(w["velocities"][wi], fs[n:][wi] + u), f"ship {i}: sampled wake velocities"
        assert np.max(np.abs(u)) > 0.0
