"""ORACLE — test infrastructure only.

ctypes wrapper of oracle/liboracle.so, the CPU f64 restatement of Stormbird's
dynamic lifting-line path.  Imported only by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.
Its Simulation class exposes the same methods as stormbird_amd.Simulation so
that parity tests drive both the same way.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from stormbird_amd.results import StepBuffers, StepResultC

_HERE = os.path.dirname(os.path.abspath(__file__))
# "glibc": elementary functions from the host libm (what the Rust build uses on Linux);
# "pm": the portable libm shared with the device code (bit-for-bit trajectory comparison).
_LIB_PATHS = {"glibc": os.path.join(_HERE, "liboracle.so"), "pm": os.path.join(_HERE, "liboracle_pm.so")}
_libs = {}
_dp = ctypes.POINTER(ctypes.c_double)


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(variant: str = "glibc"):
    if variant not in _libs:
        path = _LIB_PATHS[variant]
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_sim_create.restype = vp
        L.orc_sim_create.argtypes = [ctypes.c_char_p]
        L.orc_sim_destroy.argtypes = [vp]
        L.orc_sim_clone.restype = vp
        L.orc_sim_clone.argtypes = [vp]
        for f in ("nr_ctrl_points", "nr_wings", "nr_span_points", "nr_wake_points", "nr_panels",
                  "nr_freestream_points"):
            fn = getattr(L, "orc_sim_" + f)
            fn.restype = ctypes.c_long
            fn.argtypes = [vp]
        L.orc_sim_get_freestream_points.restype = ctypes.c_long
        L.orc_sim_get_freestream_points.argtypes = [vp, _dp, ctypes.c_size_t]
        L.orc_sim_do_step.restype = ctypes.c_int
        L.orc_sim_do_step.argtypes = [vp, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_size_t,
                                      ctypes.POINTER(StepResultC)]
        L.orc_sim_induced_velocities.argtypes = [vp, _dp, ctypes.c_size_t, _dp]
        L.orc_sim_get_previous_circulation.argtypes = [vp, _dp, ctypes.c_size_t]
        L.orc_sim_set_previous_circulation.argtypes = [vp, _dp, ctypes.c_size_t]
        L.orc_sim_set_translation_and_rotation_fd.argtypes = [vp, ctypes.c_double, _dp, _dp]
        L.orc_sim_set_local_wing_angles.argtypes = [vp, _dp, ctypes.c_size_t]
        L.orc_sim_get_wake_state.argtypes = [vp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_long),
                                             ctypes.POINTER(ctypes.c_int)]
        L.orc_sim_set_wake_state.argtypes = [vp, _dp, _dp, _dp, ctypes.c_long, ctypes.c_int]
        L.orc_sim_update_panels.argtypes = [vp]
        L.orc_sim_get_panels.argtypes = [vp, _dp]
        L.orc_sim_get_frozen_wake.argtypes = [vp, _dp, _dp]
        L.orc_segment_unit_velocity.argtypes = [_dp, _dp, _dp, ctypes.c_double, _dp]
        L.orc_panel_unit_velocity.argtypes = [_dp, ctypes.c_double, ctypes.c_double, _dp,
                                              ctypes.c_int, _dp]
        L.orc_foil_cl.restype = ctypes.c_double
        L.orc_foil_cl.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_int]
        L.orc_sigmoid.restype = ctypes.c_double
        L.orc_sigmoid.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_libm_variant.restype = ctypes.c_int
        L.orc_libm_eval.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_size_t]
        L.orc_solve_gaussian_elimination.argtypes = [_dp, _dp, ctypes.c_size_t, _dp]
        L.orc_horseshoe_unit_velocity.argtypes = [_dp, _dp, _dp, _dp, _dp, ctypes.c_double, _dp]
        L.orc_sim_get_line_model.argtypes = [vp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_long), _dp, _dp]
        L.orc_sim_restart.argtypes = [vp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_rotate.argtypes = [_dp, _dp, ctypes.c_int, _dp]
        L.orc_rotate_around_axis.argtypes = [_dp, ctypes.c_double, _dp, _dp]
        L.orc_in_rotated_coordinate_system.argtypes = [_dp, _dp, ctypes.c_int, _dp]
        L.orc_moment_in_new_coordinate_system.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, _dp]
        L.orc_rigid_body_transform_point.argtypes = [_dp, ctypes.c_int, _dp, _dp]
        L.orc_rigid_body_velocity_at_point.argtypes = [_dp, ctypes.c_int, _dp, _dp]
        L.orc_rigid_body_update_fd.argtypes = [_dp, _dp, _dp, ctypes.c_double]
        L.orc_frozen_matvec.argtypes = [_dp, _dp, _dp, ctypes.c_size_t, _dp]
        L.orc_corrected_velocity.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, ctypes.c_size_t, _dp]
        L.orc_set_threads.restype = ctypes.c_int
        _libs[variant] = L
    return _libs[variant]


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class Simulation:
    """CPU oracle simulation; same interface as stormbird_amd.Simulation."""

    def __init__(self, setup_string: str, variant: str = "glibc", _handle=None):
        self._v = variant
        L = lib(variant)
        self._h = _handle if _handle is not None else L.orc_sim_create(setup_string.encode())
        if not self._h:
            raise ValueError(L.orc_last_error().decode())
        self.nr_ctrl_points = L.orc_sim_nr_ctrl_points(self._h)
        self.nr_wings = L.orc_sim_nr_wings(self._h)
        self.nr_span_points = L.orc_sim_nr_span_points(self._h)
        self.nr_wake_points = L.orc_sim_nr_wake_points(self._h)
        self.nr_panels = L.orc_sim_nr_panels(self._h)
        self._buf = StepBuffers(self.nr_ctrl_points, self.nr_wings)

    def __del__(self):
        if getattr(self, "_h", None):
            lib(self._v).orc_sim_destroy(self._h)
            self._h = None

    def clone(self) -> "Simulation":
        return Simulation("", self._v, _handle=lib(self._v).orc_sim_clone(self._h))

    def nr_freestream_points(self) -> int:
        return self.nr_ctrl_points + self.nr_wake_points

    def get_freestream_velocity_points(self) -> np.ndarray:
        n = self.nr_freestream_points()
        out = np.zeros((n, 3))
        lib(self._v).orc_sim_get_freestream_points(self._h, _p(out), n)
        return out

    def do_step(self, *, time: float, time_step: float, freestream_velocity):
        fs = _f64(freestream_velocity, (-1, 3))
        rc = lib(self._v).orc_sim_do_step(self._h, time, time_step, _p(fs), fs.shape[0],
                                   ctypes.byref(self._buf.c))
        if rc != 0:
            raise RuntimeError(f"orc_sim_do_step failed ({rc}): {lib(self._v).orc_last_error().decode()}")
        return self._buf.snapshot()

    def induced_velocities(self, points) -> np.ndarray:
        p = _f64(points, (-1, 3))
        out = np.zeros_like(p)
        lib(self._v).orc_sim_induced_velocities(self._h, _p(p), p.shape[0], _p(out))
        return out

    def get_previous_circulation(self) -> np.ndarray:
        out = np.zeros(self.nr_ctrl_points)
        lib(self._v).orc_sim_get_previous_circulation(self._h, _p(out), out.size)
        return out

    def set_previous_circulation(self, g):
        g = _f64(g)
        if lib(self._v).orc_sim_set_previous_circulation(self._h, _p(g), g.size) != 0:
            raise ValueError("wrong length")

    def set_translation_and_rotation_with_finite_difference_for_the_velocity(self, time_step,
                                                                             translation, rotation):
        t, r = _f64(translation), _f64(rotation)
        lib(self._v).orc_sim_set_translation_and_rotation_fd(self._h, time_step, _p(t), _p(r))

    def set_local_wing_angles(self, angles):
        a = _f64(angles)
        if lib(self._v).orc_sim_set_local_wing_angles(self._h, _p(a), a.size) != 0:
            raise ValueError("wrong number of wing angles")

    def line_model(self) -> dict:
        """Global span lines, chord vectors, wing ranges, wing angles and rigid-body state."""
        n, w = self.nr_ctrl_points, self.nr_wings
        d = {"span_lines": np.zeros((n, 2, 3)), "chord_global": np.zeros((n, 3)),
             "chord_local_not_rotated": np.zeros((n, 3)), "wing_ranges": np.zeros((w, 2), dtype=np.int64),
             "wing_angles": np.zeros(w), "rigid_body": np.zeros((4, 3))}
        lib(self._v).orc_sim_get_line_model(self._h, _p(d["span_lines"]), _p(d["chord_global"]),
                                            _p(d["chord_local_not_rotated"]),
                                            d["wing_ranges"].ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
                                            _p(d["wing_angles"]), _p(d["rigid_body"]))
        return d

    def restart(self):
        lib(self._v).orc_sim_restart(self._h)

    def get_wake_state(self):
        pts = np.zeros((self.nr_wake_points, 3))
        vel = np.zeros((self.nr_wake_points, 3))
        st = np.zeros(self.nr_panels)
        steps = ctypes.c_long()
        first = ctypes.c_int()
        lib(self._v).orc_sim_get_wake_state(self._h, _p(pts), _p(vel), _p(st), ctypes.byref(steps),
                                     ctypes.byref(first))
        return {"points": pts, "velocities": vel, "strengths": st, "steps_completed": steps.value,
                "first_step_done": bool(first.value)}

    def set_wake_state(self, points, velocities, strengths, steps_completed, first_step_done=True):
        pts, vel, st = _f64(points), _f64(velocities), _f64(strengths)
        lib(self._v).orc_sim_set_wake_state(self._h, _p(pts), _p(vel), _p(st), int(steps_completed),
                                     1 if first_step_done else 0)

    def get_panels(self) -> np.ndarray:
        out = np.zeros((self.nr_panels, 22))
        lib(self._v).orc_sim_get_panels(self._h, _p(out))
        return out

    def get_frozen_wake(self):
        n = self.nr_ctrl_points
        fixed = np.zeros((n, 3))
        fac = np.zeros((n, n, 3))
        lib(self._v).orc_sim_get_frozen_wake(self._h, _p(fixed), _p(fac))
        return fixed, fac


def segment_unit_velocity(p0, p1, x, core):
    a, b, c = _f64(p0), _f64(p1), _f64(x)
    out = np.zeros(3)
    lib().orc_segment_unit_velocity(_p(a), _p(b), _p(c), core, _p(out))
    return out


def panel_unit_velocity(corners, x, far_ratio=5.0, core=0.0, mode=0):
    """mode 0: near/far switch, 1: vortex ring, 2: point doublet."""
    c, xx = _f64(corners, (12,)), _f64(x)
    out = np.zeros(3)
    lib().orc_panel_unit_velocity(_p(c), far_ratio, core, _p(xx), mode, _p(out))
    return out


def foil_value(foil_json: str, alpha: float, which: int = 0) -> float:
    """which: 0 CL, 1 CL pre-stall with drop-off, 2 CD, 3 amount of stall, 4 post-stall CL."""
    return lib().orc_foil_cl(foil_json.encode(), alpha, which)


def sigmoid(x, x0, r):
    return lib().orc_sigmoid(x, x0, r)


def libm_eval(fn: int, x, y=None, variant: str = "glibc") -> np.ndarray:
    """fn: 0 sin, 1 cos, 2 exp, 3 acos, 4 pow(x, y), 5 log, 6 atan."""
    x = _f64(x)
    y = _f64(y) if y is not None else x
    out = np.zeros_like(x)
    lib(variant).orc_libm_eval(fn, _p(x), _p(y), _p(out), x.size)
    return out


def solve_gaussian_elimination(A, b):
    """stormath Matrix::solve_gaussian_elimination (linalg.rs:41-104); ValueError when singular."""
    A = _f64(A)
    b = _f64(b)
    n = b.size
    x = np.zeros(n)
    if lib().orc_solve_gaussian_elimination(_p(A), _p(b), n, _p(x)) != 0:
        raise ValueError(lib().orc_last_error().decode())
    return x


def horseshoe_unit_velocity(start, end, wake_start, wake_end, x, core=0.0):
    """HorseshoeVortex::induced_velocity_with_unit_strength (horseshoe_vortex.rs:23-103)."""
    out = np.zeros(3)
    lib().orc_horseshoe_unit_velocity(_p(_f64(start)), _p(_f64(end)), _p(_f64(wake_start)), _p(_f64(wake_end)),
                                      _p(_f64(x)), float(core), _p(out))
    return out


def set_threads(n: int, variant: str = "glibc") -> int:
    """OpenMP threads of the oracle's per-target loops (n > 0 sets it); returns the count."""
    return lib(variant).orc_set_threads(int(n))


# ---- stormath transformations and rigid-body motion (rotation_type 0 XYZ, 1 YawPitchRoll) ----
_ROT = {"XYZ": 0, "YawPitchRoll": 1}


def _vec_call(fn, *args):
    out = np.zeros(3)
    fn(*args, _p(out))
    return out


def rotate(v, rotation, rotation_type="XYZ"):
    """SpatialVector::rotate (transformations.rs:43-48)."""
    return _vec_call(lib().orc_rotate, _p(_f64(v)), _p(_f64(rotation)), _ROT[rotation_type])


def rotate_around_axis(v, angle, axis):
    """SpatialVector::rotate_around_axis, Rodrigues' formula (transformations.rs:29-41)."""
    return _vec_call(lib().orc_rotate_around_axis, _p(_f64(v)), float(angle), _p(_f64(axis)))


def in_rotated_coordinate_system(v, rotation, rotation_type="XYZ"):
    """SpatialVector::in_rotated_coordinate_system (transformations.rs:145-155)."""
    return _vec_call(lib().orc_in_rotated_coordinate_system, _p(_f64(v)), _p(_f64(rotation)),
                     _ROT[rotation_type])


def moment_in_new_coordinate_system(moment, rotation, translation, force, rotation_type="XYZ"):
    """SpatialVector::moment_in_new_coordinate_system (transformations.rs:157-180)."""
    return _vec_call(lib().orc_moment_in_new_coordinate_system, _p(_f64(moment)), _p(_f64(rotation)),
                     _p(_f64(translation)), _p(_f64(force)), _ROT[rotation_type])


def _motion12(translation, rotation, velocity_linear, velocity_angular):
    return np.concatenate([_f64(translation), _f64(rotation), _f64(velocity_linear), _f64(velocity_angular)])


def rigid_body_transform_point(motion, point, rotation_type="XYZ"):
    """RigidBodyMotion::transform_point (rigid_body_motion.rs:32-34); motion = (t, r, v, w)."""
    m = _motion12(*motion)
    return _vec_call(lib().orc_rigid_body_transform_point, _p(m), _ROT[rotation_type], _p(_f64(point)))


def rigid_body_velocity_at_point(motion, point, rotation_type="XYZ"):
    """RigidBodyMotion::velocity_at_point (rigid_body_motion.rs:57-65); motion = (t, r, v, w)."""
    m = _motion12(*motion)
    return _vec_call(lib().orc_rigid_body_velocity_at_point, _p(m), _ROT[rotation_type], _p(_f64(point)))


def rigid_body_update_fd(motion, translation, rotation, time_step):
    """update_{translation,rotation}_with_velocity_using_finite_difference (rigid_body_motion.rs:67-103).
    Returns the new (translation, rotation, velocity_linear, velocity_angular)."""
    m = _motion12(*motion)
    lib().orc_rigid_body_update_fd(_p(m), _p(_f64(translation)), _p(_f64(rotation)), float(time_step))
    return m[0:3].copy(), m[3:6].copy(), m[6:9].copy(), m[9:12].copy()


def corrected_velocity(kind, freestream, induced, ratio=0.0):
    """VelocityCorrections::corrected_velocity (velocity_corrections.rs:28-100); kind
    "MaxInducedVelocityMagnitudeRatio", "FixedMagnitudeEqualToFreestream" or "NoCorrection"."""
    k = {"NoCorrection": 0, "MaxInducedVelocityMagnitudeRatio": 1, "FixedMagnitudeEqualToFreestream": 2}[kind]
    fs, ui = _f64(freestream, (-1, 3)), _f64(induced, (-1, 3))
    out = np.zeros_like(fs)
    lib().orc_corrected_velocity(k, float(ratio), _p(fs), _p(ui), fs.shape[0], _p(out))
    return out


def frozen_matvec(fixed, factors, strengths, variant: str = "glibc"):
    """FrozenWake::update_induced_velocities_at_control_points (frozen_wake.rs:268-301)."""
    g = _f64(strengths)
    out = np.zeros((g.size, 3))
    lib(variant).orc_frozen_matvec(_p(_f64(fixed)), _p(_f64(factors)), _p(g), g.size, _p(out))
    return out
