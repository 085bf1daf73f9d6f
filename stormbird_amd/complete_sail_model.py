"""The complete sail model and its wind inputs, with pystormbird's Python interface.

``WindCondition`` mirrors interfaces/pystormbird/src/wind/wind_condition.rs:15-232,
``WindEnvironment`` interfaces/pystormbird/src/wind/environment.rs:18-190 and
``CompleteSailModel`` interfaces/pystormbird/src/lifting_line/complete_sail_model.rs:17-165:
same constructors, keyword-only arguments and return values.  A wind condition
is held as the JSON of Stormbird's ``WindCondition`` (wind/wind_condition/mod.rs:18-29)
and every evaluation runs in libstormbird_amd.so; the model's lifting-line
steps run on the GPU.  Where the reference panics (an unwrap on bad JSON, an
index out of range), these raise.
"""
from __future__ import annotations

import copy
import ctypes
import json

import numpy as np

from . import MODE_EXACT, Simulation, _err, _f64, _p, lib
from .results import SimulationResult, StepBuffers

__all__ = ["WindCondition", "WindEnvironment", "CompleteSailModel", "csm_setup_info"]

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)
_registered = False


def _L():
    global _registered
    L = lib()
    if not _registered:
        vp, c = ctypes.c_void_p, ctypes.c_char_p
        d, sz = ctypes.c_double, ctypes.c_size_t
        L.sb_csm_create_ex.restype = vp
        L.sb_csm_create_ex.argtypes = [c, ctypes.c_int, ctypes.c_int]
        L.sb_csm_destroy.argtypes = [vp]
        L.sb_csm_simulation.restype = vp
        L.sb_csm_simulation.argtypes = [vp]
        L.sb_csm_do_step.argtypes = [vp, d, d, c, d, vp]
        L.sb_csm_restart.argtypes = [vp]
        L.sb_csm_apply_controller_wind.argtypes = [vp, d, d, c, d, d]
        L.sb_csm_apply_controller_result.argtypes = [vp, d, d, d, _dp, _dp, _dp, _dp]
        L.sb_csm_simulate_steady_state.argtypes = [vp, c, d, d, vp]
        L.sb_csm_simulate_optimal_loading.argtypes = [vp, c, d, ctypes.c_long, vp, _dp]
        L.sb_csm_freestream.argtypes = [vp, c, d, d, _dp, sz]
        L.sb_wind_velocity.argtypes = [c, c, ctypes.c_int, _dp, sz, _dp, d, _dp]
        L.sb_wind_apparent_direction.argtypes = [c, c, _dp, d, _dp]
        L.sb_wind_ctrl_points_velocity.argtypes = [c, c, _dp, sz, _dp, d, _lp, sz, _dp]
        L.sb_wind_condition_values.argtypes = [c, ctypes.c_int, _dp, sz, d, _dp]
        L.sb_sim_get_local_wing_angles.argtypes = [vp, _dp, sz]
        L.sb_csm_setup_info.argtypes = [c, _lp, sz]
        _registered = True
    return L


def csm_setup_info(setup_string: str) -> dict:
    """Parses a CompleteSailModelBuilder string without a GPU (sb_csm_setup_info);
    ValueError on a bad setup."""
    out = (ctypes.c_long * 9)()
    rc = _L().sb_csm_setup_info(setup_string.encode(), out, 9)
    if rc != 0:
        raise ValueError(_L().sb_last_error().decode())
    keys = ("nr_ctrl_points", "nr_wings", "nr_span_points", "nr_wake_rows", "nr_freestream_points",
            "nr_panels", "quasi_steady", "nr_set_points", "inflow_corrections")
    return dict(zip(keys, list(out)))


def _vec3(v) -> np.ndarray:
    a = _f64(v)
    if a.shape != (3,):
        raise ValueError("expected three components")
    return a


class WindCondition:
    """Direction the wind comes from, a velocity profile over height and optional
    gust spectra (wind/wind_condition/mod.rs:18-29)."""

    def __init__(self, data: dict):
        self.data = data

    # ---- constructors (pystormbird wind_condition.rs:23-123) -------------------
    @classmethod
    def new_constant(cls, *, direction_coming_from: float, velocity: float) -> "WindCondition":
        return cls({"direction_coming_from": float(direction_coming_from),
                    "velocity_variation": {"Constant": float(velocity)}})

    @classmethod
    def new_power_model(cls, *, direction_coming_from: float, reference_velocity: float,
                        reference_height: float = 10.0, power_factor: float = 0.11111) -> "WindCondition":
        return cls({"direction_coming_from": float(direction_coming_from),
                    "velocity_variation": {"PowerModel": {"reference_velocity": float(reference_velocity),
                                                          "reference_height": float(reference_height),
                                                          "power_factor": float(power_factor)}}})

    @classmethod
    def new_logarithmic_model(cls, *, direction_coming_from: float, friction_velocity: float,
                              surface_roughness: float, obukhov_length: float = 0.0) -> "WindCondition":
        model = {"friction_velocity": float(friction_velocity), "surface_roughness": float(surface_roughness),
                 "obukhov_length": None if obukhov_length == 0.0 else float(obukhov_length),
                 "von_karman_constant": 0.41, "stable_coefficient": 6.0, "unstable_coefficient": 19.3}
        return cls({"direction_coming_from": float(direction_coming_from),
                    "velocity_variation": {"LogarithmicModel": model}})

    @classmethod
    def from_json_string(cls, s: str) -> "WindCondition":
        return cls(json.loads(s))

    def to_json_string(self) -> str:
        return json.dumps(self.data)

    def _json(self) -> bytes:
        return self.to_json_string().encode()

    def _set_gust(self, key: str, gust_string: str):
        gust = json.loads(gust_string)
        for k in ("frequencies", "amplitudes", "phases"):
            if k not in gust:
                raise ValueError(f"DiscretizedSpectrum: missing field `{k}`")
        self.data[key] = gust

    def set_parallel_gust_from_json_string(self, gust_string: str):
        self._set_gust("parallel_gust", gust_string)

    def set_perpendicular_gust_from_json_string(self, gust_string: str):
        self._set_gust("perpendicular_gust", gust_string)

    def set_vertical_gust_from_json_string(self, gust_string: str):
        self._set_gust("vertical_gust", gust_string)

    # ---- evaluation (wind_condition/mod.rs:37-61) -----------------------------
    def _values(self, which: int, heights, time: float = 0.0) -> np.ndarray:
        h = np.atleast_1d(_f64(heights))
        out = np.zeros(h.shape[0])
        rc = _L().sb_wind_condition_values(self._json(), which, _p(h), h.shape[0], float(time), _p(out))
        if rc != 0:
            _err("WindCondition", rc)
        return out

    def steady_true_wind_velocity_at_height(self, height: float) -> float:
        return float(self._values(0, height)[0])

    def unsteady_parallel_true_wind_velocity_at_height(self, height: float, time: float) -> float:
        return float(self._values(1, height, time)[0])

    def unsteady_perpendicular_true_wind_velocity(self, time: float) -> float:
        return float(self._values(2, 0.0, time)[0])

    def unsteady_vertical_true_wind_velocity(self, time: float) -> float:
        return float(self._values(3, 0.0, time)[0])

    def businger_dyer_unscaled_correction(self, height: float) -> float:
        return float(self._values(4, height)[0])

    # ---- LogarithmicModel fields (pystormbird wind_condition.rs:152-231) ------
    def _log(self):
        return self.data["velocity_variation"].get("LogarithmicModel")

    def _get(self, key, default):
        m = self._log()
        if m is None:
            return 0.0
        v = m.get(key, default)
        return 0.0 if v is None else float(v)

    def _set(self, key, value):
        m = self._log()
        if m is not None:
            m[key] = float(value)

    @property
    def von_karman_constant(self) -> float:
        return self._get("von_karman_constant", 0.41)

    @property
    def stable_coefficient(self) -> float:
        return self._get("stable_coefficient", 6.0)

    @property
    def unstable_coefficient(self) -> float:
        return self._get("unstable_coefficient", 19.3)

    @property
    def friction_velocity(self) -> float:
        return self._get("friction_velocity", 0.0)

    @friction_velocity.setter
    def friction_velocity(self, value: float):
        self._set("friction_velocity", value)

    @property
    def surface_roughness(self) -> float:
        return self._get("surface_roughness", 0.0)

    @surface_roughness.setter
    def surface_roughness(self, value: float):
        self._set("surface_roughness", value)

    @property
    def obukhov_length(self) -> float:
        return self._get("obukhov_length", None)

    @obukhov_length.setter
    def obukhov_length(self, value: float):
        self._set("obukhov_length", value)

    def copy(self) -> "WindCondition":
        return WindCondition(copy.deepcopy(self.data))


class WindEnvironment:
    """Directions, water-plane height and inflow corrections (wind/environment.rs:21-34)."""

    def __init__(self, setup_string: str):
        self.setup_string = setup_string
        self._b = setup_string.encode()
        self._check()

    def _check(self):
        z = np.zeros(3)
        constant = WindCondition.new_constant(direction_coming_from=0.0, velocity=0.0)
        rc = _L().sb_wind_velocity(self._b, constant._json(), 0, _p(z), 0, None, 0.0, _p(z))
        if rc != 0:
            raise ValueError(_L().sb_last_error().decode())

    def _query(self, kind, condition, locations, linear_velocity=None, time=0.0):
        loc = _f64(locations, (-1, 3))
        n = loc.shape[0]
        out = np.zeros((n, 3) if kind < 4 else n)
        lin = None if linear_velocity is None else _p(_vec3(linear_velocity))
        rc = _L().sb_wind_velocity(self._b, condition._json(), kind, _p(loc), n, lin, float(time), _p(out))
        if rc != 0:
            _err("WindEnvironment", rc)
        return out

    def steady_true_wind_velocity_at_location(self, *, condition, location) -> float:
        return float(self._query(4, condition, location)[0])

    def steady_true_wind_velocity_vector_at_location(self, *, condition, location):
        return self._query(3, condition, location)[0].tolist()

    def unsteady_true_wind_velocity_vector_at_location(self, *, condition, location, time):
        return self._query(2, condition, location, None, time)[0].tolist()

    def steady_apparent_wind_velocity_vector_at_location(self, *, condition, location, linear_velocity):
        return self._query(1, condition, location, linear_velocity)[0].tolist()

    def unsteady_apparent_wind_velocity_vector_at_location(self, *, condition, location, linear_velocity, time):
        return self._query(0, condition, location, linear_velocity, time)[0].tolist()

    def unsteady_apparent_wind_velocity_vectors_at_locations(self, *, condition, locations, linear_velocity,
                                                             time) -> np.ndarray:
        """Array form of unsteady_apparent_wind_velocity_vector_at_location."""
        return self._query(0, condition, locations, linear_velocity, time)

    def apparent_wind_direction_from_condition_and_linear_velocity(self, *, condition, linear_velocity,
                                                                   height: float = 10.0) -> float:
        out = np.zeros(1)
        rc = _L().sb_wind_apparent_direction(self._b, condition._json(), _p(_vec3(linear_velocity)), float(height),
                                             _p(out))
        if rc != 0:
            _err("apparent_wind_direction", rc)
        return float(out[0])

    def apparent_wind_velocity_vectors_at_ctrl_points_with_corrections_applied(
            self, *, condition, ctrl_points, linear_velocity, time, wing_indices):
        pts = _f64(ctrl_points, (-1, 3))
        ranges = np.ascontiguousarray(np.asarray(wing_indices, dtype=np.int64).reshape(-1, 2), dtype=np.int64)
        out = np.zeros_like(pts)
        rc = _L().sb_wind_ctrl_points_velocity(self._b, condition._json(), _p(pts), pts.shape[0],
                                               _p(_vec3(linear_velocity)), float(time),
                                               ranges.ctypes.data_as(_lp), ranges.shape[0], _p(out))
        if rc != 0:
            _err("apparent_wind_velocity_vectors_at_ctrl_points_with_corrections_applied", rc)
        return out.tolist()


class CompleteSailModel:
    """Lifting-line simulation + wind environment + controller
    (stormbird/src/lifting_line/complete_sail_model/mod.rs:38-330) on the GPU."""

    def __init__(self, setup_string: str, *, device: int = 0, mode: int = MODE_EXACT):
        L = _L()
        self._h = L.sb_csm_create_ex(setup_string.encode(), device, mode)
        if not self._h:
            raise ValueError(L.sb_last_error().decode())
        self.mode = mode
        self.simulation = Simulation._borrowed(L.sb_csm_simulation(self._h), mode, self)
        self._env = json.loads(setup_string)["wind_environment"]
        self._buf = StepBuffers(self.simulation.nr_ctrl_points, self.simulation.nr_wings)

    def __del__(self):
        if getattr(self, "_h", None):
            _L().sb_csm_destroy(self._h)
        self._h = None

    def get_number_of_sails(self) -> int:
        return self.simulation.nr_wings

    def _step_out(self):
        return ctypes.byref(self._buf.c)

    def do_step(self, *, time: float, time_step: float, wind_condition: WindCondition,
                ship_velocity: float) -> SimulationResult:
        rc = _L().sb_csm_do_step(self._h, float(time), float(time_step), wind_condition._json(),
                                 float(ship_velocity), self._step_out())
        if rc != 0:
            _err("do_step", rc)
        return self._buf.snapshot()

    def do_multiple_steps(self, *, end_time: float, time_step: float, wind_condition: WindCondition,
                          ship_velocity: float):
        """complete_sail_model/mod.rs:145-172: the wake starts afresh, then steps at
        0, dt, 2 dt, ... while the time is below end_time."""
        _L().sb_csm_restart(self._h)
        results = []
        t = 0.0
        while t < end_time:
            results.append(self.do_step(time=t, time_step=time_step, wind_condition=wind_condition,
                                        ship_velocity=ship_velocity))
            t += time_step
        return results

    def apply_controller_based_on_wind_condition(self, *, time: float, time_step: float,
                                                 wind_condition: WindCondition, ship_velocity: float,
                                                 controller_loading: float):
        rc = _L().sb_csm_apply_controller_wind(self._h, float(time), float(time_step), wind_condition._json(),
                                               float(ship_velocity), float(controller_loading))
        if rc != 0:
            _err("apply_controller_based_on_wind_condition", rc)

    def apply_controller_based_on_simulation_result(self, *, time: float, time_step: float, loading: float,
                                                    simulation_result: SimulationResult):
        r = simulation_result
        aoa = _f64(r.force_input.angles_of_attack)
        vel = _f64(r.force_input.velocity)
        inp = _f64(r.solver_input_ctrl_points_velocity)
        pts = _f64(r.ctrl_points)
        if aoa.shape[0] != self.simulation.nr_ctrl_points:
            raise ValueError("simulation result of another model")
        rc = _L().sb_csm_apply_controller_result(self._h, float(time), float(time_step), float(loading),
                                                 _p(aoa), _p(vel), _p(inp), _p(pts))
        if rc != 0:
            _err("apply_controller_based_on_simulation_result", rc)

    def simulate_steady_state_condition(self, *, wind_condition: WindCondition, ship_velocity: float,
                                        controller_loading: float) -> SimulationResult:
        rc = _L().sb_csm_simulate_steady_state(self._h, wind_condition._json(), float(ship_velocity),
                                               float(controller_loading), self._step_out())
        if rc != 0:
            _err("simulate_steady_state_condition", rc)
        return self._buf.snapshot()

    def simulate_steady_state_condition_simple_output(self, *, wind_condition: WindCondition,
                                                      ship_velocity: float, controller_loading: float):
        return self.simulate_steady_state_condition(wind_condition=wind_condition, ship_velocity=ship_velocity,
                                                    controller_loading=controller_loading).as_simplified()

    def simulate_condition_optimal_controller_loading(self, *, wind_condition: WindCondition,
                                                      ship_velocity: float, nr_loadings_to_test: int):
        best = np.zeros(1)
        rc = _L().sb_csm_simulate_optimal_loading(self._h, wind_condition._json(), float(ship_velocity),
                                                  int(nr_loadings_to_test), self._step_out(), _p(best))
        if rc != 0:
            _err("simulate_condition_optimal_controller_loading", rc)
        self.last_optimal_loading = float(best[0])
        return self._buf.snapshot()

    def freestream_velocity(self, *, wind_condition: WindCondition, ship_velocity: float, time: float):
        """Freestream at the control points and wake points (mod.rs:213-255)."""
        n = self.simulation.nr_freestream_points()
        out = np.zeros((n, 3))
        rc = _L().sb_csm_freestream(self._h, wind_condition._json(), float(ship_velocity), float(time), _p(out), n)
        if rc != 0:
            _err("freestream_velocity", rc)
        return out

    def section_models_internal_state(self):
        """Foil sections carry no internal state (line_force_model/data_access.rs:101-117)."""
        return [0.0] * self.simulation.nr_wings

    def local_wing_angles(self):
        w = self.simulation.nr_wings
        out = np.zeros(w)
        rc = _L().sb_sim_get_local_wing_angles(self.simulation._h, _p(out), w)
        if rc != 0:
            _err("local_wing_angles", rc)
        return out.tolist()

    def set_local_wing_angles(self, local_wing_angles):
        self.simulation.set_local_wing_angles(local_wing_angles)

    def set_section_models_internal_state(self, internal_state):
        """No effect on Foil sections (line_force_model/data_update.rs:145-158)."""
        del internal_state

    def get_wind_environment(self) -> WindEnvironment:
        return WindEnvironment(json.dumps(self._env))
