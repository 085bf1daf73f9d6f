"""stormbird_amd — Stormbird's dynamic lifting-line step on an AMD Instinct MI355X.

``Simulation`` mirrors the Python interface of Stormbird's ``Simulation``
(interfaces/pystormbird/src/lifting_line/simulation.rs:16-165):
construct from a SimulationBuilder JSON string, ``get_freestream_velocity_points``,
``do_step(time=, time_step=, freestream_velocity=)`` returning a
``SimulationResult``, ``induced_velocities(points)``, rigid-body and wing-angle
setters, and the ``previous_circulation_strength`` warm-start hook.

All work runs in libstormbird_amd.so (HIP kernels for gfx950) behind the C ABI
of include/stormbird_gpu.h.  There is no CPU fallback: if the library is not
built or no GPU is present, constructing a Simulation raises.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from .results import FIELDS, BatchResults, StepBuffers, StepResultC, SimulationResult  # noqa: F401

__all__ = ["Simulation", "SimulationResult", "MODE_EXACT", "MODE_FAST", "library_path", "lib",
           "wake_induced_velocities", "CompleteSailModel", "WindCondition", "WindEnvironment"]

MODE_EXACT = 0
MODE_FAST = 1
MODE_FP32_KAHAN = 2

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libstormbird_amd.so")
_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def lib():
    """Loads libstormbird_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise RuntimeError(f"{library_path} is missing: build it with `python -m stormbird_amd.build`")
        L = ctypes.CDLL(library_path)
        vp = ctypes.c_void_p
        L.sb_last_error.restype = ctypes.c_char_p
        L.sb_sim_create.restype = vp
        L.sb_sim_create.argtypes = [ctypes.c_char_p]
        L.sb_sim_create_ex.restype = vp
        L.sb_sim_create_ex.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
        L.sb_sim_destroy.argtypes = [vp]
        for f in ("nr_ctrl_points", "nr_wings", "nr_span_points", "nr_wake_points", "nr_panels",
                  "nr_freestream_points"):
            fn = getattr(L, "sb_sim_" + f)
            fn.restype = ctypes.c_long
            fn.argtypes = [vp]
        L.sb_sim_get_freestream_points.restype = ctypes.c_long
        L.sb_sim_get_freestream_points.argtypes = [vp, _dp, ctypes.c_size_t]
        step_args = [vp, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_size_t, ctypes.POINTER(StepResultC)]
        L.sb_sim_do_step.argtypes = step_args
        L.sb_sim_do_step_device.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                            ctypes.c_size_t, ctypes.POINTER(StepResultC)]
        L.sb_sim_induced_velocities.argtypes = [vp, _dp, ctypes.c_size_t, _dp]
        L.sb_sim_get_previous_circulation.argtypes = [vp, _dp, ctypes.c_size_t]
        L.sb_sim_set_previous_circulation.argtypes = [vp, _dp, ctypes.c_size_t]
        L.sb_sim_set_translation_and_rotation_fd.argtypes = [vp, ctypes.c_double, _dp, _dp]
        L.sb_sim_set_local_wing_angles.argtypes = [vp, _dp, ctypes.c_size_t]
        L.sb_sim_get_wake_state.argtypes = [vp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_long),
                                            ctypes.POINTER(ctypes.c_int)]
        L.sb_sim_set_wake_state.argtypes = [vp, _dp, _dp, _dp, ctypes.c_long, ctypes.c_int]
        L.sb_sim_set_wake_state_ex.argtypes = [vp, _dp, ctypes.c_size_t, _dp, ctypes.c_size_t, _dp,
                                               ctypes.c_size_t, ctypes.c_long, ctypes.c_int]
        L.sb_sim_get_panels.argtypes = [vp, _dp]
        L.sb_sim_get_frozen_wake.argtypes = [vp, _dp, _dp]
        L.sb_sim_set_kernel_timers.restype = ctypes.c_int
        L.sb_sim_set_kernel_timers.argtypes = [vp, ctypes.c_int]
        L.sb_sim_last_kernel_ms.restype = ctypes.c_double
        L.sb_sim_last_kernel_ms.argtypes = [vp, ctypes.c_char_p]
        L.sb_sim_last_pair_counts.argtypes = [vp, _dp]
        L.sb_sim_count_self_induction_pairs.argtypes = [vp, _dp]
        L.sb_sim_set_target_tiling.argtypes = [vp, ctypes.c_int]
        L.sb_sim_get_target_tiling.argtypes = [vp]
        L.sb_wake_induced_velocities.argtypes = [_dp, ctypes.c_size_t, _dp, ctypes.c_size_t, ctypes.c_int,
                                                 ctypes.c_int, _dp]
        L.sb_debug_libm_device.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_size_t]
        L.sb_frozen_matvec.argtypes = [_dp, _dp, _dp, ctypes.c_size_t, _dp]
        L.sb_wake_induced_velocities_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                                         ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                                         ctypes.c_void_p, ctypes.c_void_p]
        L.sb_noalloc_max_panels.restype = ctypes.c_long
        L.sb_debug_linearized_in_shared_memory.argtypes = [ctypes.c_long]
        L.sb_setup_info.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_long), ctypes.c_size_t]
        L.sb_debug_libm_host.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_size_t]
        L.sb_batch_create.restype = vp
        L.sb_batch_create.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.c_long, ctypes.c_int, ctypes.c_int]
        L.sb_batch_destroy.argtypes = [vp]
        L.sb_batch_size.restype = ctypes.c_long
        L.sb_batch_size.argtypes = [vp]
        L.sb_batch_instance.restype = vp
        L.sb_batch_instance.argtypes = [vp, ctypes.c_long]
        L.sb_batch_freestream_offsets.restype = ctypes.c_long
        L.sb_batch_freestream_offsets.argtypes = [vp, ctypes.POINTER(ctypes.c_long), ctypes.c_size_t]
        L.sb_batch_do_step.argtypes = [vp, _dp, _dp, _dp, ctypes.c_size_t, vp]
        L.sb_batch_do_step_device.argtypes = [vp, _dp, _dp, ctypes.c_void_p, ctypes.c_size_t, vp]
        L.sb_batch_set_wind.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]
        L.sb_batch_do_step_wind.argtypes = [vp, _dp, _dp, _dp, vp]
        L.sb_batch_get_freestream.argtypes = [vp, _dp, ctypes.c_size_t]
        L.sb_batch_last_step_ms.restype = ctypes.c_double
        L.sb_batch_last_step_ms.argtypes = [vp]
        L.sb_batch_reset.argtypes = [vp]
        L.sb_sim_write_wake_obj.argtypes = [vp, ctypes.c_char_p]
        L.sb_sim_write_wake_vtk.argtypes = [vp, ctypes.c_char_p]
        L.sb_sim_get_panel_point_indices.argtypes = [vp, ctypes.POINTER(ctypes.c_long), ctypes.c_size_t]
        L.sb_debug_format_f64.argtypes = [ctypes.c_double, ctypes.c_char_p, ctypes.c_size_t]
        L.sb_sim_reset.argtypes = [vp]
        L.sb_sim_set_solver_settings.argtypes = [vp, ctypes.c_long, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        _fp = ctypes.POINTER(ctypes.c_float)
        L.sb_mlp_create.restype = vp
        L.sb_mlp_create.argtypes = [ctypes.c_int, ctypes.c_ulonglong]
        L.sb_mlp_destroy.argtypes = [vp]
        L.sb_mlp_nr_params.restype = ctypes.c_long
        L.sb_mlp_nr_params.argtypes = [vp]
        L.sb_mlp_get_params.argtypes = [vp, _fp, ctypes.c_size_t]
        L.sb_mlp_set_params.argtypes = [vp, _fp, ctypes.c_size_t]
        L.sb_mlp_forward.argtypes = [vp, ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p]
        L.sb_mlp_gradients.argtypes = [vp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, _fp, _fp]
        L.sb_mlp_train_step.argtypes = [vp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_float, _fp]
        L.sb_batch_set_warm_start.argtypes = [vp, vp, ctypes.c_int]
        L.sb_batch_set_record_features.argtypes = [vp, ctypes.c_int]
        L.sb_batch_nr_strips.restype = ctypes.c_long
        L.sb_batch_nr_strips.argtypes = [vp]
        L.sb_batch_get_features.argtypes = [vp, _fp, _fp, ctypes.c_size_t]
        L.sb_batch_last_fallbacks.restype = ctypes.c_long
        L.sb_batch_last_fallbacks.argtypes = [vp]
        L.sb_batch_last_warm_taken.restype = ctypes.c_long
        L.sb_batch_last_warm_taken.argtypes = [vp]
        L.sb_batch_set_tuning.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.sb_batch_get_tuning.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.sb_batch_autotune.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_long, _dp,
                                        ctypes.c_size_t]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


def _err(what: str, rc: int):
    raise RuntimeError(f"{what} failed ({rc}): {lib().sb_last_error().decode()}")


class Simulation:
    """GPU lifting-line simulation with Stormbird's Simulation interface."""

    def __init__(self, setup_string: str, *, device: int = 0, mode: int = MODE_EXACT):
        L = lib()
        self._h = L.sb_sim_create_ex(setup_string.encode(), device, mode)
        if not self._h:
            raise ValueError(L.sb_last_error().decode())
        self.mode = mode
        self.nr_ctrl_points = L.sb_sim_nr_ctrl_points(self._h)
        self.nr_wings = L.sb_sim_nr_wings(self._h)
        self.nr_span_points = L.sb_sim_nr_span_points(self._h)
        self.nr_wake_points = L.sb_sim_nr_wake_points(self._h)
        self.nr_panels = L.sb_sim_nr_panels(self._h)
        self._buf = StepBuffers(self.nr_ctrl_points, self.nr_wings)

    @classmethod
    def _borrowed(cls, handle, mode: int, owner):
        """View of an instance owned by a Batch (not destroyed by this object)."""
        self = cls.__new__(cls)
        L = lib()
        self._h = handle
        self._owner = owner
        self.mode = mode
        self.nr_ctrl_points = L.sb_sim_nr_ctrl_points(handle)
        self.nr_wings = L.sb_sim_nr_wings(handle)
        self.nr_span_points = L.sb_sim_nr_span_points(handle)
        self.nr_wake_points = L.sb_sim_nr_wake_points(handle)
        self.nr_panels = L.sb_sim_nr_panels(handle)
        self._buf = StepBuffers(self.nr_ctrl_points, self.nr_wings)
        return self

    def __del__(self):
        if getattr(self, "_h", None) and getattr(self, "_owner", None) is None:
            lib().sb_sim_destroy(self._h)
        self._h = None

    def nr_freestream_points(self) -> int:
        return self.nr_ctrl_points + self.nr_wake_points

    def get_freestream_velocity_points(self) -> np.ndarray:
        n = self.nr_freestream_points()
        out = np.zeros((n, 3))
        rc = lib().sb_sim_get_freestream_points(self._h, _p(out), n)
        if rc < 0:
            _err("get_freestream_velocity_points", rc)
        return out

    def do_step(self, *, time: float, time_step: float, freestream_velocity):
        fs = _f64(freestream_velocity, (-1, 3))
        rc = lib().sb_sim_do_step(self._h, time, time_step, _p(fs), fs.shape[0], ctypes.byref(self._buf.c))
        if rc != 0:
            _err("do_step", rc)
        return self._buf.snapshot()

    def do_step_device(self, *, time: float, time_step: float, freestream_ptr: int, n_points: int,
                       want_result: bool = True):
        """do_step with the freestream already in device memory (e.g. a torch tensor's data_ptr())."""
        out = ctypes.byref(self._buf.c) if want_result else None
        rc = lib().sb_sim_do_step_device(self._h, time, time_step, ctypes.c_void_p(freestream_ptr), n_points, out)
        if rc != 0:
            _err("do_step_device", rc)
        return self._buf.snapshot() if want_result else None

    def induced_velocities(self, points) -> np.ndarray:
        p = _f64(points, (-1, 3))
        out = np.zeros_like(p)
        rc = lib().sb_sim_induced_velocities(self._h, _p(p), p.shape[0], _p(out))
        if rc != 0:
            _err("induced_velocities", rc)
        return out

    @property
    def previous_circulation_strength(self) -> np.ndarray:
        out = np.zeros(self.nr_ctrl_points)
        rc = lib().sb_sim_get_previous_circulation(self._h, _p(out), out.size)
        if rc != 0:
            _err("get_previous_circulation", rc)
        return out

    @previous_circulation_strength.setter
    def previous_circulation_strength(self, g):
        g = _f64(g)
        rc = lib().sb_sim_set_previous_circulation(self._h, _p(g), g.size)
        if rc != 0:
            _err("set_previous_circulation", rc)

    def get_previous_circulation(self) -> np.ndarray:
        return self.previous_circulation_strength

    def set_previous_circulation(self, g):
        self.previous_circulation_strength = g

    def reset_previous_circulation_strength(self):
        self.previous_circulation_strength = np.zeros(self.nr_ctrl_points)

    def set_translation_and_rotation_with_finite_difference_for_the_velocity(self, time_step, translation,
                                                                             rotation):
        t, r = _f64(translation), _f64(rotation)
        if t.shape != (3,) or r.shape != (3,):
            raise ValueError(f"translation and rotation must be 3-vectors, got shapes {t.shape} and {r.shape}")
        rc = lib().sb_sim_set_translation_and_rotation_fd(self._h, time_step, _p(t), _p(r))
        if rc != 0:
            _err("set_translation_and_rotation", rc)

    def set_local_wing_angles(self, angles):
        a = _f64(angles)
        rc = lib().sb_sim_set_local_wing_angles(self._h, _p(a), a.size)
        if rc != 0:
            raise ValueError("The number of wing angles must match the number of wings in the simulation.")

    def get_wake_state(self):
        pts = np.zeros((self.nr_wake_points, 3))
        vel = np.zeros((self.nr_wake_points, 3))
        st = np.zeros(self.nr_panels)
        steps = ctypes.c_long()
        first = ctypes.c_int()
        rc = lib().sb_sim_get_wake_state(self._h, _p(pts), _p(vel), _p(st), ctypes.byref(steps), ctypes.byref(first))
        if rc != 0:
            _err("get_wake_state", rc)
        return {"points": pts, "velocities": vel, "strengths": st, "steps_completed": steps.value,
                "first_step_done": bool(first.value)}

    def set_wake_state(self, points, velocities, strengths, steps_completed, first_step_done=True):
        """points, velocities: nr_wake_points x 3 values; strengths: nr_panels values (any shape
        holding exactly that many; ValueError otherwise)."""
        arrays = []
        for name, a, shape in (("points", points, (self.nr_wake_points, 3)),
                               ("velocities", velocities, (self.nr_wake_points, 3)),
                               ("strengths", strengths, (self.nr_panels,))):
            a = _f64(a)
            if a.size != int(np.prod(shape)):
                raise ValueError(f"set_wake_state: {name} has shape {a.shape}, the simulation needs {shape}")
            arrays.append(a.reshape(shape))
        pts, vel, st = arrays
        rc = lib().sb_sim_set_wake_state_ex(self._h, _p(pts), pts.shape[0], _p(vel), vel.shape[0], _p(st), st.size,
                                            int(steps_completed), 1 if first_step_done else 0)
        if rc != 0:
            _err("set_wake_state", rc)

    @property
    def target_tiling(self) -> int:
        """Span points per wave of the self-induction's target tiles (0: flat order)."""
        return lib().sb_sim_get_target_tiling(self._h)

    @target_tiling.setter
    def target_tiling(self, span_points_per_wave: int):
        rc = lib().sb_sim_set_target_tiling(self._h, int(span_points_per_wave))
        if rc != 0:
            _err("set_target_tiling", rc)

    def get_panels(self) -> np.ndarray:
        out = np.zeros((self.nr_panels, 22))
        rc = lib().sb_sim_get_panels(self._h, _p(out))
        if rc != 0:
            _err("get_panels", rc)
        return out

    def get_frozen_wake(self):
        n = self.nr_ctrl_points
        fixed = np.zeros((n, 3))
        fac = np.zeros((n, n, 3))
        rc = lib().sb_sim_get_frozen_wake(self._h, _p(fixed), _p(fac))
        if rc != 0:
            _err("get_frozen_wake", rc)
        return fixed, fac

    def write_wake_to_obj_file(self, path: str):
        """DynamicWake::write_wake_to_obj_file (wake/dynamic_wake/export.rs:30-64)."""
        rc = lib().sb_sim_write_wake_obj(self._h, str(path).encode())
        if rc != 0:
            _err("write_wake_to_obj_file", rc)

    def write_wake_to_vtk_file(self, path: str):
        """DynamicWake::write_wake_to_vtk_file (wake/dynamic_wake/export.rs:66-159)."""
        rc = lib().sb_sim_write_wake_vtk(self._h, str(path).encode())
        if rc != 0:
            _err("write_wake_to_vtk_file", rc)

    def panel_point_indices(self) -> np.ndarray:
        out = np.zeros((self.nr_panels, 4), dtype=np.int64)
        rc = lib().sb_sim_get_panel_point_indices(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), out.size)
        if rc != 0:
            _err("panel_point_indices", rc)
        return out

    def export_to_plotly_mesh(self) -> dict:
        """DynamicWake::export_to_plotly_mesh (export.rs:165-210): points x, y, z; two triangles
        (0, 1, 2) and (0, 2, 3) per panel as i, j, k; the panel strength per triangle."""
        st = self.get_wake_state()
        p = st["points"]
        ix = self.panel_point_indices().astype(np.float64)
        tri_i = np.stack([ix[:, 0], ix[:, 0]], axis=1).reshape(-1)
        tri_j = np.stack([ix[:, 1], ix[:, 2]], axis=1).reshape(-1)
        tri_k = np.stack([ix[:, 2], ix[:, 3]], axis=1).reshape(-1)
        strength = np.repeat(st["strengths"], 2)
        return {"x": p[:, 0].tolist(), "y": p[:, 1].tolist(), "z": p[:, 2].tolist(), "i": tri_i.tolist(),
                "j": tri_j.tolist(), "k": tri_k.tolist(), "strength": strength.tolist()}

    def reset(self):
        """State of a newly built simulation (straight initial wake, no completed step)."""
        rc = lib().sb_sim_reset(self._h)
        if rc != 0:
            _err("reset", rc)

    def set_solver_settings(self, *, max_iterations: int, damping: float, residual_tolerance: float,
                            strength_tolerance: float):
        """SimpleIterative's public fields (max_iterations_per_time_step, damping_factor,
        residual_tolerance_absolute, strength_difference_tolerance), from the next step on."""
        rc = lib().sb_sim_set_solver_settings(self._h, int(max_iterations), float(damping),
                                              float(residual_tolerance), float(strength_tolerance))
        if rc != 0:
            _err("set_solver_settings", rc)

    def set_kernel_timers(self, on: bool = True):
        """Per-kernel interval timers of later steps on or off (off by default: the
        event records cost about 25 us per step)."""
        rc = lib().sb_sim_set_kernel_timers(self._h, 1 if on else 0)
        if rc != 0:
            _err("set_kernel_timers", rc)

    def last_kernel_ms(self, name: str) -> float:
        """Milliseconds of one kernel interval of the last step ("wake_update", "fixed",
        "factors", "solve", "forces", "self_induction"); the last step must have run with
        the timers on."""
        v = lib().sb_sim_last_kernel_ms(self._h, name.encode())
        if v == -2.0:
            raise ValueError(f"unknown kernel interval {name!r}")
        if v == -3.0:
            raise RuntimeError("the last step ran without kernel timers (set_kernel_timers(True))")
        return v

    def count_self_induction_pairs(self):
        """(near, total) pairs with nonzero strength of the self-induction sum at the current state."""
        c = np.zeros(2)
        rc = lib().sb_sim_count_self_induction_pairs(self._h, _p(c))
        if rc != 0:
            _err("count_self_induction_pairs", rc)
        return float(c[0]), float(c[1])

    def last_pair_counts(self) -> np.ndarray:
        c = np.zeros(3)
        lib().sb_sim_last_pair_counts(self._h, _p(c))
        return c


def wake_induced_velocities(panel_records, targets, symmetry: int = 0, mode: int = MODE_EXACT) -> np.ndarray:
    """Induced velocity at `targets` from panel records [n][22] (kernel-level entry)."""
    r = _f64(panel_records, (-1, 22))
    t = _f64(targets, (-1, 3))
    out = np.zeros_like(t)
    rc = lib().sb_wake_induced_velocities(_p(r), r.shape[0], _p(t), t.shape[0], symmetry, mode, _p(out))
    if rc != 0:
        _err("wake_induced_velocities", rc)
    return out


def setup_info(setup_string: str) -> dict:
    """Parses a setup string without a GPU (sb_setup_info); ValueError on a bad setup."""
    out = (ctypes.c_long * 7)()
    rc = lib().sb_setup_info(setup_string.encode(), out, 7)
    if rc != 0:
        raise ValueError(lib().sb_last_error().decode())
    keys = ("nr_ctrl_points", "nr_wings", "nr_span_points", "nr_wake_rows", "nr_freestream_points",
            "nr_panels", "quasi_steady")
    return dict(zip(keys, list(out)))


def frozen_matvec(fixed, factors, strengths) -> np.ndarray:
    """FrozenWake::update_induced_velocities_at_control_points (frozen_wake.rs:268-301) on the GPU:
    fixed [n][3] + factors [n][n][3] (row-major) times strengths [n]."""
    g = _f64(strengths)
    n = g.size
    fx = _f64(fixed)
    fa = _f64(factors)
    if fx.size != 3 * n or fa.size != 3 * n * n:
        raise ValueError(f"frozen_matvec: fixed {fx.shape} and factors {fa.shape} do not fit {n} strengths")
    out = np.zeros((n, 3))
    rc = lib().sb_frozen_matvec(_p(fx), _p(fa), _p(g), n, _p(out))
    if rc != 0:
        _err("frozen_matvec", rc)
    return out


def libm_device(fn: int, x, y=None) -> np.ndarray:
    x = _f64(x)
    y = _f64(y) if y is not None else x
    out = np.zeros_like(x)
    rc = lib().sb_debug_libm_device(fn, _p(x), _p(y), _p(out), x.size)
    if rc != 0:
        _err("libm_device", rc)
    return out


def libm_host(fn: int, x, y=None) -> np.ndarray:
    x = _f64(x)
    y = _f64(y) if y is not None else x
    out = np.zeros_like(x)
    lib().sb_debug_libm_host(fn, _p(x), _p(y), _p(out), x.size)
    return out


class Batch:
    """Many simulations stepped together on one GPU (``sb_batch_*``).

    ``instances[i]`` is a ``Simulation`` view of instance i: queries and state
    setters apply to it; advance it only through ``Batch.do_step``.  Each
    instance's results equal stepping it alone (same mode) bit for bit.
    """

    def __init__(self, setup_strings, *, device: int = 0, mode: int = MODE_EXACT):
        L = lib()
        enc = [s.encode() for s in setup_strings]
        n = len(enc)
        arr = (ctypes.c_char_p * n)(*enc)
        self._h = L.sb_batch_create(arr, n, device, mode)
        if not self._h:
            raise ValueError(L.sb_last_error().decode())
        self.mode = mode
        self.n = n
        self.instances = [Simulation._borrowed(L.sb_batch_instance(self._h, i), mode, self) for i in range(n)]
        offs = (ctypes.c_long * (n + 1))()
        if L.sb_batch_freestream_offsets(self._h, offs, n + 1) != n + 1:
            _err("sb_batch_freestream_offsets", -1)
        self.freestream_offsets = np.array(list(offs), dtype=np.int64)
        # batch-wide output arrays; every instance's buffers are views into them
        ns = [s.nr_ctrl_points for s in self.instances]
        ws = [s.nr_wings for s in self.instances]
        self._strip_off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
        self._wing_off = np.concatenate([[0], np.cumsum(ws)]).astype(np.int64)
        self._out = {}
        for name, per_strip, tail in FIELDS:
            per = int(np.prod(tail)) if tail else 1
            total = int(self._strip_off[-1] if per_strip else self._wing_off[-1])
            self._out[name] = np.zeros(total * per)
        self._res = (StepResultC * n)()
        for i, sim in enumerate(self.instances):
            storage = {}
            for name, per_strip, tail in FIELDS:
                per = int(np.prod(tail)) if tail else 1
                off = self._strip_off if per_strip else self._wing_off
                storage[name] = self._out[name][off[i] * per:off[i + 1] * per]
            sim._buf = StepBuffers(ns[i], ws[i], storage)
            self._res[i] = sim._buf.c

    def __del__(self):
        if getattr(self, "_h", None):
            for sim in getattr(self, "instances", []):
                sim._h = None
            lib().sb_batch_destroy(self._h)
            self._h = None

    @property
    def nr_freestream_points(self) -> int:
        return int(self.freestream_offsets[-1])

    def get_freestream_velocity_points(self) -> np.ndarray:
        return np.concatenate([s.get_freestream_velocity_points() for s in self.instances])

    def _arrays(self, times, time_steps):
        dts = _f64(np.broadcast_to(np.asarray(time_steps, dtype=np.float64), (self.n,)).copy())
        ts = _f64(np.broadcast_to(np.asarray(0.0 if times is None else times, dtype=np.float64), (self.n,)).copy())
        return ts, dts

    # sb_step_result as a numpy record: time, iterations, residual, then eight pointers
    _RES_DTYPE = np.dtype([("time", "<f8"), ("iterations", "<i8"), ("residual", "<f8")] +
                          [(f, "<u8") for f, _ in StepResultC._fields_[3:]])

    def _collect(self):
        arr = np.frombuffer(self._res, dtype=self._RES_DTYPE)
        return BatchResults(arr["time"].copy(), arr["iterations"].copy(), arr["residual"].copy(),
                            {k: v.copy() for k, v in self._out.items()}, self._strip_off, self._wing_off)

    def do_step(self, *, times, time_steps, freestream_velocity, want_result: bool = True):
        fs = _f64(freestream_velocity, (-1, 3))
        ts, dts = self._arrays(times, time_steps)
        res = ctypes.cast(self._res, ctypes.c_void_p) if want_result else None
        rc = lib().sb_batch_do_step(self._h, _p(ts), _p(dts), _p(fs), fs.shape[0], res)
        if rc != 0:
            _err("sb_batch_do_step", rc)
        return self._collect() if want_result else None

    def do_step_device(self, *, times, time_steps, freestream_ptr: int, n_points: int, want_result: bool = True):
        ts, dts = self._arrays(times, time_steps)
        res = ctypes.cast(self._res, ctypes.c_void_p) if want_result else None
        rc = lib().sb_batch_do_step_device(self._h, _p(ts), _p(dts), ctypes.c_void_p(freestream_ptr), n_points, res)
        if rc != 0:
            _err("sb_batch_do_step_device", rc)
        return self._collect() if want_result else None

    def set_wind(self, environment, conditions):
        """Wind of every instance: a WindEnvironment (JSON string or dict; None for the
        defaults) and one WindCondition per instance (JSON strings or dicts), as the
        reference's serde forms (wind/environment.rs:21-34, wind_condition/mod.rs:18-39)."""
        if len(conditions) != len(self.instances):
            raise ValueError(f"{len(conditions)} wind conditions for {len(self.instances)} instances")
        env = environment if isinstance(environment, (str, type(None))) else json.dumps(environment)
        enc = [(c if isinstance(c, str) else json.dumps(c)).encode() for c in conditions]
        arr = (ctypes.c_char_p * len(enc))(*enc)
        rc = lib().sb_batch_set_wind(self._h, env.encode() if env else None, arr)
        if rc != 0:
            _err("sb_batch_set_wind", rc)

    def do_step_wind(self, *, times, time_steps, ship_velocities, want_result: bool = True):
        """One step of every instance in its wind (set_wind) at times[i], the ship moving
        at ship_velocities[i]: CompleteSailModel::freestream_velocity then the step
        (complete_sail_model/mod.rs:174-255), without a controller."""
        if times is None:
            raise ValueError("a wind step needs the times")
        ts, dts = self._arrays(times, time_steps)
        sv = _f64(ship_velocities, (-1,))
        if sv.shape[0] != len(self.instances):
            raise ValueError(f"{sv.shape[0]} ship velocities for {len(self.instances)} instances")
        res = ctypes.cast(self._res, ctypes.c_void_p) if want_result else None
        rc = lib().sb_batch_do_step_wind(self._h, _p(ts), _p(dts), _p(sv), res)
        if rc != 0:
            _err("sb_batch_do_step_wind", rc)
        return self._collect() if want_result else None

    def get_freestream(self) -> np.ndarray:
        """The freestream of the last step [points, 3] (split by freestream_offsets)."""
        out = np.zeros((int(self.freestream_offsets[-1]), 3))
        rc = lib().sb_batch_get_freestream(self._h, _p(out), out.shape[0])
        if rc != 0:
            _err("sb_batch_get_freestream", rc)
        return out

    def last_step_ms(self) -> float:
        return lib().sb_batch_last_step_ms(self._h)

    def reset(self):
        rc = lib().sb_batch_reset(self._h)
        if rc != 0:
            _err("sb_batch_reset", rc)

    def set_tuning(self, fixed_per: int, targets_tile: int):
        """Kernel variants of the batched pair sums: panels per compute wave per block of the
        fixed-velocity kernel (1, 2, 3) and panel records per LDS tile of the self-induction
        kernel (64, 128, 256).  Results do not depend on them."""
        rc = lib().sb_batch_set_tuning(self._h, int(fixed_per), int(targets_tile))
        if rc != 0:
            _err("sb_batch_set_tuning", rc)

    def get_tuning(self):
        p, t = ctypes.c_int(), ctypes.c_int()
        lib().sb_batch_get_tuning(self._h, ctypes.byref(p), ctypes.byref(t))
        return p.value, t.value

    def autotune(self, cache_path=None, use_cache=True, write_cache=True, reps=3, max_jobs=0):
        """Times every variant on the batch's last-step state (take a step first; the state is
        left unchanged) and applies the fastest.  With a cache file, a stored choice for the
        same device and shape class is applied without timing; `max_jobs` > 0 times only the first
        jobs of each kernel's queue (a sample, for very long steps).  Returns
        (cached, (fixed_per, targets_tile), timings_ms [6])."""
        t = np.zeros(6)
        flags = (1 if use_cache else 0) | (2 if write_cache else 0)
        path = cache_path.encode() if cache_path else None
        rc = lib().sb_batch_autotune(self._h, path, flags, int(reps), int(max_jobs), _p(t), 6)
        if rc < 0:
            _err("sb_batch_autotune", rc)
        return bool(rc == 1), self.get_tuning(), t

    def set_warm_start(self, mlp, mode: int = 1, residual: bool = False):
        """mode 0 off, 1 first step, 2 every step, 3 every step judged by the solver's first
        evaluation (keeps a reference to the network); residual: the network predicts the
        change from the previous circulation, in units of the scale."""
        self._mlp = mlp if mode else None
        m = int(mode) | (4 if residual and mode else 0)
        rc = lib().sb_batch_set_warm_start(self._h, mlp._h if mlp is not None else None, m)
        if rc != 0:
            _err("sb_batch_set_warm_start", rc)

    def set_record_features(self, on: bool = True):
        lib().sb_batch_set_record_features(self._h, 1 if on else 0)

    @property
    def nr_strips(self) -> int:
        return lib().sb_batch_nr_strips(self._h)

    def get_features(self):
        """(features [strips, 64], scale [strips]) of the last step."""
        n = self.nr_strips
        x = np.zeros((n, 64), dtype=np.float32)
        sc = np.zeros(n, dtype=np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        rc = lib().sb_batch_get_features(self._h, x.ctypes.data_as(fp), sc.ctypes.data_as(fp), n)
        if rc != 0:
            _err("sb_batch_get_features", rc)
        return x, sc

    def last_fallbacks(self) -> int:
        return lib().sb_batch_last_fallbacks(self._h)

    def last_warm_taken(self) -> int:
        return lib().sb_batch_last_warm_taken(self._h)


class Mlp:
    """Warm-start network (sb_mlp_*): 64 -> 256 x 4 (GELU) -> 1, bf16 matrix-core GEMMs,
    fp32 master weights.  Inputs and targets are device pointers (e.g. torch tensors)."""

    def __init__(self, *, device: int = 0, seed: int = 1):
        L = lib()
        self._h = L.sb_mlp_create(device, seed)
        if not self._h:
            raise RuntimeError(L.sb_last_error().decode())
        self.nr_params = L.sb_mlp_nr_params(self._h)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.sb_mlp_destroy(h)

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.nr_params, dtype=np.float32)
        rc = lib().sb_mlp_get_params(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), out.size)
        if rc != 0:
            _err("sb_mlp_get_params", rc)
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        rc = lib().sb_mlp_set_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), p.size)
        if rc != 0:
            _err("sb_mlp_set_params", rc)

    def forward(self, x_ptr: int, n: int, y_ptr: int):
        rc = lib().sb_mlp_forward(self._h, ctypes.c_void_p(x_ptr), int(n), ctypes.c_void_p(y_ptr))
        if rc != 0:
            _err("sb_mlp_forward", rc)

    def gradients(self, x_ptr: int, t_ptr: int, n: int):
        g = np.zeros(self.nr_params, dtype=np.float32)
        loss = ctypes.c_float(0.0)
        fp = ctypes.POINTER(ctypes.c_float)
        rc = lib().sb_mlp_gradients(self._h, ctypes.c_void_p(x_ptr), ctypes.c_void_p(t_ptr), int(n),
                                    g.ctypes.data_as(fp), ctypes.byref(loss))
        if rc != 0:
            _err("sb_mlp_gradients", rc)
        return g, loss.value

    def train_step(self, x_ptr: int, t_ptr: int, n: int, lr: float) -> float:
        loss = ctypes.c_float(0.0)
        rc = lib().sb_mlp_train_step(self._h, ctypes.c_void_p(x_ptr), ctypes.c_void_p(t_ptr), int(n), float(lr),
                                     ctypes.byref(loss))
        if rc != 0:
            _err("sb_mlp_train_step", rc)
        return loss.value


def format_f64(v: float) -> str:
    """The library's Rust-`{}` formatting of an f64 (test hook)."""
    buf = ctypes.create_string_buffer(512)
    n = lib().sb_debug_format_f64(float(v), buf, 512)
    if n < 0:
        _err("sb_debug_format_f64", n)
    return buf.value.decode()


from .complete_sail_model import CompleteSailModel, WindCondition, WindEnvironment  # noqa: E402
