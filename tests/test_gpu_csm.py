"""GPU parity of the complete sail model (complete_sail_model/mod.rs:38-330) against the
oracle built with the portable libm, bit for bit: the freestream of a step (wind profile,
gusts, ship speed and inflow corrections at the control points; the wake points on the
device), steps with the controller applied from the wind condition and from a result,
steady-state conditions, the optimal-loading search, do_multiple_steps with its wake
restart, and the reference's C interface (interfaces/cstormbird).
"""
import ctypes
import json
import os

import numpy as np
import pytest

import stormbird_amd as sb
from stormbird_amd import CompleteSailModel, WindCondition

from test_gpu_parity import compare_results

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "setups")
QUASI = "csm_wing_sail_quasi_steady"
DYNAMIC = "csm_three_sails_dynamic_inflow"


def setup(name):
    return open(os.path.join(GOLDEN, name + ".json")).read()


def conditions():
    gust = {"frequencies": [0.05, 0.11, 0.3], "amplitudes": [0.8, 0.4, 0.2], "phases": [0.3, 2.0, 4.1]}
    return [
        {"direction_coming_from": 0.6, "velocity_variation": {"PowerModel": {"reference_velocity": 9.0}},
         "parallel_gust": gust, "perpendicular_gust": gust, "vertical_gust": dict(gust, amplitudes=[0.1, 0.0, 0.05])},
        {"direction_coming_from": -1.9, "velocity_variation": {"LogarithmicModel": {
            "friction_velocity": 0.45, "surface_roughness": 0.0005, "obukhov_length": -60.0}}},
        {"direction_coming_from": 2.8, "velocity_variation": {"Constant": 7.5}},
    ]


@pytest.fixture(scope="module")
def oc(oracle):
    import oracle_csm

    return oracle_csm


def pair(oc, name, mode=sb.MODE_EXACT):
    js = setup(name)
    return CompleteSailModel(js, mode=mode), oc.CompleteSailModel(js, "pm")


def same_angles(g, o):
    assert g.local_wing_angles() == [float(x) for x in o.sim.line_model()["wing_angles"]]


@pytest.mark.parametrize("name", [QUASI, DYNAMIC])
def test_freestream_matches_oracle(oc, name):
    g, o = pair(oc, name)
    for k, wc in enumerate(conditions()):
        got = g.freestream_velocity(wind_condition=WindCondition(wc), ship_velocity=4.0 + k, time=2.5 * k)
        want = o.freestream(wc, 4.0 + k, 2.5 * k)
        assert got.shape == want.shape and np.array_equal(got, want), k
        # after steps the wake points have moved; the device evaluates them where they are
        for step in range(2):
            rg = g.do_step(time=0.5 * step, time_step=0.5, wind_condition=WindCondition(wc), ship_velocity=4.0)
            ro = o.do_step(0.5 * step, 0.5, wc, 4.0)
            compare_results(rg, ro, where=f"{name} condition {k} step {step}")
        got = g.freestream_velocity(wind_condition=WindCondition(wc), ship_velocity=3.0, time=7.0)
        assert np.array_equal(got, o.freestream(wc, 3.0, 7.0))


def test_steps_with_controller_match_oracle(oc):
    g, o = pair(oc, DYNAMIC)
    wc = conditions()[0]
    dt = 0.5
    for k in range(6):
        t = k * dt
        loading = 0.6 + 0.05 * k
        g.apply_controller_based_on_wind_condition(time=t, time_step=dt, wind_condition=WindCondition(wc),
                                                   ship_velocity=5.0, controller_loading=loading)
        o.apply_controller_wind(t, dt, wc, 5.0, loading)
        same_angles(g, o)
        rg = g.do_step(time=t, time_step=dt, wind_condition=WindCondition(wc), ship_velocity=5.0)
        ro = o.do_step(t, dt, wc, 5.0)
        compare_results(rg, ro, where=f"step {k}")
        g.apply_controller_based_on_simulation_result(time=t, time_step=dt, loading=loading, simulation_result=rg)
        o.apply_controller_result(t, dt, loading, ro)
        same_angles(g, o)
    assert any(abs(a) > 0.01 for a in g.local_wing_angles())


def test_controller_on_a_moving_ship_matches_oracle(oc):
    """Translation and roll/pitch/yaw of the ship: rotated chord vectors in the apparent wind
    direction and the rotational velocity removed from the measured velocity
    (results/simulation.rs:118-149, environment.rs:297-327)."""
    g, o = pair(oc, DYNAMIC)
    wc = conditions()[1]
    dt = 0.5
    for k in range(5):
        t = k * dt
        tr = [0.3 * k, 0.1 * k, 0.0]
        rot = [0.02 * k, -0.01 * k, 0.05 * k]
        g.simulation.set_translation_and_rotation_with_finite_difference_for_the_velocity(dt, tr, rot)
        o.sim.set_translation_and_rotation_with_finite_difference_for_the_velocity(dt, tr, rot)
        g.apply_controller_based_on_wind_condition(time=t, time_step=dt, wind_condition=WindCondition(wc),
                                                   ship_velocity=4.0, controller_loading=0.9)
        o.apply_controller_wind(t, dt, wc, 4.0, 0.9)
        same_angles(g, o)
        rg = g.do_step(time=t, time_step=dt, wind_condition=WindCondition(wc), ship_velocity=4.0)
        ro = o.do_step(t, dt, wc, 4.0)
        compare_results(rg, ro, where=f"moving step {k}")
        g.apply_controller_based_on_simulation_result(time=t, time_step=dt, loading=0.9, simulation_result=rg)
        o.apply_controller_result(t, dt, 0.9, ro)
        same_angles(g, o)


@pytest.mark.parametrize("name", [QUASI, DYNAMIC])
def test_steady_state_and_optimal_loading_match_oracle(oc, name):
    g, o = pair(oc, name)
    for k, wc in enumerate(conditions()):
        rg = g.simulate_steady_state_condition(wind_condition=WindCondition(wc), ship_velocity=6.0,
                                               controller_loading=0.7)
        ro = o.steady_state(wc, 6.0, 0.7)
        compare_results(rg, ro, where=f"{name} steady {k}")
        same_angles(g, o)
        simple = g.simulate_steady_state_condition_simple_output(wind_condition=WindCondition(wc), ship_velocity=6.0,
                                                                 controller_loading=0.7)
        ro = o.steady_state(wc, 6.0, 0.7)
        for w, s in enumerate(simple):
            assert np.array_equal(s.force, ro.integrated_forces[w].total)
            assert np.array_equal(s.moment, ro.integrated_moments[w].total)
            assert s.input_power == 0.0
    wc = conditions()[0]
    rg = g.simulate_condition_optimal_controller_loading(wind_condition=WindCondition(wc), ship_velocity=6.0,
                                                         nr_loadings_to_test=4)
    ro, best = o.optimal_loading(wc, 6.0, 4)
    assert g.last_optimal_loading == best
    compare_results(rg, ro, where=f"{name} optimal loading")
    same_angles(g, o)


def test_do_multiple_steps_restarts_the_wake(oc):
    g, o = pair(oc, DYNAMIC)
    wc = conditions()[1]
    for run in range(2):
        rg = g.do_multiple_steps(end_time=1.6, time_step=0.4, wind_condition=WindCondition(wc), ship_velocity=3.0)
        ro = o.do_multiple_steps(1.6, 0.4, wc, 3.0)
        assert len(rg) == len(ro) == 4
        for k, (a, b) in enumerate(zip(rg, ro)):
            compare_results(a, b, where=f"run {run} step {k}")


def test_fast_mode_close_to_exact(oc):
    exact = CompleteSailModel(setup(DYNAMIC))
    fast = CompleteSailModel(setup(DYNAMIC), mode=sb.MODE_FAST)
    wc = WindCondition(conditions()[0])
    for k in range(3):
        a = exact.do_step(time=0.5 * k, time_step=0.5, wind_condition=wc, ship_velocity=5.0)
        b = fast.do_step(time=0.5 * k, time_step=0.5, wind_condition=wc, ship_velocity=5.0)
        fa, fb = a.integrated_forces_sum(), b.integrated_forces_sum()
        assert np.max(np.abs(fa - fb)) <= 1e-9 * np.max(np.abs(fa))


class CWind(ctypes.Structure):
    _fields_ = [("velocity", ctypes.c_double), ("direction_coming_from", ctypes.c_double)]


class CSail(ctypes.Structure):
    _fields_ = [("force", ctypes.c_double * 3), ("moment", ctypes.c_double * 3), ("input_power", ctypes.c_double)]


def test_cstormbird_interface(oc):
    """interfaces/cstormbird/src/lifting_line/complete_sail_model.rs:17-109 with its WindCondition
    conversion (wind.rs:14-33: power profile, reference height 10, power factor 1/9)."""
    L = sb.lib()
    L.complete_sail_model_new.restype = ctypes.c_void_p
    L.complete_sail_model_new.argtypes = [ctypes.c_char_p]
    L.complete_sail_model_get_number_of_sails.restype = ctypes.c_int32
    L.complete_sail_model_get_number_of_sails.argtypes = [ctypes.c_void_p]
    L.complete_sail_model_drop.argtypes = [ctypes.c_void_p]
    L.complete_sail_model_simulate_steady_state_condition.restype = ctypes.c_int32
    L.complete_sail_model_simulate_steady_state_condition.argtypes = [
        ctypes.c_void_p, CWind, ctypes.c_double, ctypes.c_double, ctypes.POINTER(CSail), ctypes.c_size_t]
    assert L.complete_sail_model_new(None) is None
    assert L.complete_sail_model_new(b"{}") is None
    assert L.complete_sail_model_get_number_of_sails(None) == -1
    L.complete_sail_model_drop(None)
    js = setup(DYNAMIC)
    h = L.complete_sail_model_new(js.encode())
    assert h
    try:
        assert L.complete_sail_model_get_number_of_sails(h) == 3
        out = (CSail * 3)()
        wind = CWind(8.5, 0.9)
        assert L.complete_sail_model_simulate_steady_state_condition(None, wind, 5.0, 0.8, out, 3) == -1
        assert L.complete_sail_model_simulate_steady_state_condition(h, wind, 5.0, 0.8, None, 3) == -1
        assert L.complete_sail_model_simulate_steady_state_condition(h, wind, 5.0, 0.8, out, 0) == -2
        o = oc.CompleteSailModel(js, "pm")
        wc = {"direction_coming_from": 0.9, "velocity_variation": {"PowerModel": {
            "reference_velocity": 8.5, "reference_height": 10.0, "power_factor": 1.0 / 9.0}}}
        for length in (3, 2):
            n = L.complete_sail_model_simulate_steady_state_condition(h, wind, 5.0, 0.8, out, length)
            assert n == length
            ro = o.steady_state(wc, 5.0, 0.8)
            for w in range(n):
                assert list(out[w].force) == ro.integrated_forces[w].total.tolist()
                assert list(out[w].moment) == ro.integrated_moments[w].total.tolist()
                assert out[w].input_power == 0.0
    finally:
        L.complete_sail_model_drop(h)


def test_controller_errors_where_the_reference_panics():
    base = json.loads(setup(QUASI))
    wc = WindCondition(conditions()[2])
    zero = json.loads(json.dumps(base))
    zero["controller"]["time_steps_between_updates"] = 0
    m = CompleteSailModel(json.dumps(zero))
    with pytest.raises(RuntimeError, match="time_steps_between_updates"):
        m.apply_controller_based_on_wind_condition(time=0.0, time_step=1.0, wind_condition=wc, ship_velocity=1.0,
                                                   controller_loading=1.0)
    extra = json.loads(json.dumps(base))
    extra["controller"]["set_points"] *= 2
    m = CompleteSailModel(json.dumps(extra))
    with pytest.raises(RuntimeError, match="set points"):
        m.simulate_steady_state_condition(wind_condition=wc, ship_velocity=1.0, controller_loading=1.0)
    late = json.loads(json.dumps(base))
    late["controller"]["start_time"] = 5.0
    m = CompleteSailModel(json.dumps(late))
    m.apply_controller_based_on_wind_condition(time=1.0, time_step=1.0, wind_condition=wc, ship_velocity=1.0,
                                               controller_loading=1.0)
    assert m.local_wing_angles() == [0.0]  # before start_time the controller does not act
