"""ORACLE — test infrastructure only.

Python restatement of Stormbird's wind model, controller and CompleteSailModel
(stormbird 0.9.0): wind/environment.rs, wind/wind_condition/*, wind/inflow_corrections.rs,
controller/*.rs and lifting_line/complete_sail_model/mod.rs.  The lifting-line
steps themselves run in the C oracle (oracle_py.Simulation).  Imported only by
tests/ as the checker.

Every function follows the Rust source it cites, in its evaluation order, one
scalar operation at a time on Python floats (IEEE binary64, no fused
multiply-add).  Elementary functions come from a `Libm`: the host libm
(`math`, what the Rust build calls on Linux) or the portable libm that the
device code shares (`orc_libm_eval` of liboracle_pm.so), for bit-for-bit
comparison with the GPU library.

Parity pinning (tests/test_oracle_wind.py): the velocity profiles against
vectors that the reference's own Python models wrote (tests/golden/wind_profiles.json),
the wind directions against the reference's test_true_wind_velocity_vectors,
closed forms of gusts, rotations and interpolation, and the controller's
set-point logic against hand-derived cases.
"""
from __future__ import annotations

import json
import math

import numpy as np

import oracle_py

PI = math.pi
TAU = 2.0 * math.pi
MIN_POSITIVE = 2.2250738585072014e-308  # f64::MIN_POSITIVE
U64 = 1 << 64


class Libm:
    """Elementary functions: "glibc" (host libm) or "pm" (portable, shared with the device)."""

    def __init__(self, variant: str = "glibc"):
        self.variant = variant

    def _pm(self, fn, x, y=0.0):
        return float(oracle_py.libm_eval(fn, np.array([x]), np.array([y]), variant="pm")[0])

    def sin(self, x):
        return math.sin(x) if self.variant == "glibc" else self._pm(0, x)

    def cos(self, x):
        return math.cos(x) if self.variant == "glibc" else self._pm(1, x)

    def acos(self, x):
        return math.acos(x) if self.variant == "glibc" else self._pm(3, x)

    def pow(self, x, y):
        return math.pow(x, y) if self.variant == "glibc" else self._pm(4, x, y)

    def log(self, x):
        return math.log(x) if self.variant == "glibc" else self._pm(5, x)

    def atan(self, x):
        return math.atan(x) if self.variant == "glibc" else self._pm(6, x)


# ---- stormath SpatialVector (spatial_vector/vector_math.rs, transformations.rs) ----------
def add(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def scale(a, s):  # SpatialVector * Float and Float * SpatialVector
    return (a[0] * s, a[1] * s, a[2] * s)


def dot(a, b):  # vector_math.rs:11-13
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def cross(a, b):  # vector_math.rs:55-61
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def length(a):  # vector_math.rs:18-26
    return math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])


def normalize(a):  # vector_math.rs:30-45
    n = length(a)
    if n > 0.0:
        return (a[0] / n, a[1] / n, a[2] / n)
    return (0.0, 0.0, 0.0)


def absolute_angle_between(a, b, m: Libm):  # vector_math.rs:65-80
    la = a[0] * a[0] + a[1] * a[1] + a[2] * a[2]
    lb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2]
    if la == 0.0 or lb == 0.0:
        return 0.0
    c = dot(a, b) / math.sqrt(la * lb)
    c = min(max(c, -1.0), 1.0)
    return m.acos(c)


def signed_angle_between(a, b, axis, m: Libm):  # vector_math.rs:85-95
    triple = dot(a, cross(b, axis))
    ang = absolute_angle_between(a, b, m)
    return ang if triple > 0.0 else -ang


def rotate_around_axis(v, angle, axis, m: Libm):  # transformations.rs:29-40
    k = normalize(axis)
    c, s = m.cos(angle), m.sin(angle)
    t1 = scale(v, c)
    t2 = scale(cross(k, v), s)
    t3 = scale(scale(k, dot(k, v)), 1.0 - c)
    return add(add(t1, t2), t3)


def rotate(v, rotation, ypr: bool, m: Libm):  # transformations.rs:42-111
    if ypr:
        yaw = (0.0, 0.0, 1.0)
        pitch = rotate_around_axis((0.0, 1.0, 0.0), rotation[2], yaw, m)
        roll = rotate_around_axis(rotate_around_axis((1.0, 0.0, 0.0), rotation[2], yaw, m), rotation[1], pitch, m)
        r = rotate_around_axis(v, rotation[2], yaw, m)
        r = rotate_around_axis(r, rotation[1], pitch, m)
        return rotate_around_axis(r, rotation[0], roll, m)
    x, y, z = v
    if rotation[0] != 0.0:
        c, s = m.cos(rotation[0]), m.sin(rotation[0])
        y, z = y * c - z * s, y * s + z * c
    if rotation[1] != 0.0:
        c, s = m.cos(rotation[1]), m.sin(rotation[1])
        x, z = x * c + z * s, -x * s + z * c
    if rotation[2] != 0.0:
        c, s = m.cos(rotation[2]), m.sin(rotation[2])
        x, y = x * c - y * s, x * s + y * c
    return (x, y, z)


def vector(node, default):
    """serde form of SpatialVector: {"x", "y", "z"} or [x, y, z] (serde_implementations.rs)."""
    if node is None:
        return default
    if isinstance(node, dict):
        return (float(node["x"]), float(node["y"]), float(node["z"]))
    return (float(node[0]), float(node[1]), float(node[2]))


# ---- stormath interpolation.rs:17-100 ------------------------------------------------------
def _two_points(x, xd, yd, lerp):
    if x <= xd[0]:
        return yd[0]
    if x >= xd[1]:
        return yd[1]
    return lerp(yd[0], yd[1], (x - xd[0]) / (xd[1] - xd[0]))


def linear_interpolation(x, xd, yd, vector_values=False):
    if len(xd) != len(yd):
        raise ValueError("x_data and y_data must have the same length")
    if vector_values:
        def lerp(a, b, t):
            return add(a, scale(sub(b, a), t))
    else:
        def lerp(a, b, t):
            return a + (b - a) * t
    n = len(xd)
    if n == 0:
        raise ValueError("x_data is empty")
    if n == 1:
        return yd[0]
    if n == 2:
        return _two_points(x, xd, yd, lerp)
    lo, hi = 0, n - 1
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if xd[mid] > x:
            hi = mid
        else:
            lo = mid
    if lo == n - 1:
        return yd[lo]
    return _two_points(x, xd[lo:lo + 2], yd[lo:lo + 2], lerp)


# ---- WindCondition (wind_condition/mod.rs, velocity_variation/*, discretized_spectrum.rs) ----
class WindCondition:
    def __init__(self, d, m: Libm):
        if isinstance(d, str):
            d = json.loads(d)
        self.m = m
        self.direction_coming_from = float(d["direction_coming_from"])
        (kind, p), = d["velocity_variation"].items()
        self.kind = kind
        if kind == "Constant":
            self.value = float(p)
        elif kind == "PowerModel":  # power_model.rs:10-23
            self.ref_velocity = float(p["reference_velocity"])
            self.ref_height = float(p.get("reference_height", 10.0))
            self.power_factor = float(p.get("power_factor", 1.0 / 9.0))
        elif kind == "LogarithmicModel":  # logarithmic_model.rs:10-28
            self.u_star = float(p["friction_velocity"])
            self.z0 = float(p["surface_roughness"])
            L = p.get("obukhov_length")
            self.obukhov = None if L is None else float(L)
            self.kappa = float(p.get("von_karman_constant", 0.41))
            self.stable = float(p.get("stable_coefficient", 6.0))
            self.unstable = float(p.get("unstable_coefficient", 19.3))
        else:
            raise ValueError(kind)
        self.gusts = [d.get(k) for k in ("parallel_gust", "perpendicular_gust", "vertical_gust")]

    def businger_dyer_unscaled(self, h):  # logarithmic_model.rs:52-74
        if self.obukhov is None:
            return 0.0
        zeta = h / self.obukhov
        if self.obukhov > 0.0:
            return -self.stable * zeta
        m = self.m
        x = m.pow(1.0 - self.unstable * zeta, 0.25)
        first = 2.0 * m.log((1.0 + x) / 2.0)
        second = m.log((1.0 + x * x) / 2.0)
        third = -2.0 * m.atan(x) + PI / 2.0
        return first + second + third

    def steady_at_height(self, h):  # velocity_variation/mod.rs:22-28
        if self.kind == "Constant":
            return self.value
        if self.kind == "PowerModel":  # power_model.rs:28-40
            f = self.m.pow(h / self.ref_height, self.power_factor) if self.power_factor > 0.0 else 1.0
            return self.ref_velocity * f
        scale_factor = self.u_star / self.kappa  # logarithmic_model.rs:26-50, 76-80
        neutral = 0.0 if (self.z0 <= 0.0 or h <= 0.0) else scale_factor * self.m.log(h / self.z0)
        return neutral - self.businger_dyer_unscaled(h) * scale_factor

    def gust(self, k, t):  # discretized_spectrum.rs:37-45
        g = self.gusts[k]
        if g is None:
            return 0.0
        out = 0.0
        for f, a, ph in zip(g["frequencies"], g["amplitudes"], g["phases"]):
            out += float(a) * self.m.sin(TAU * float(f) * t + float(ph))
        return out

    def unsteady_parallel_at_height(self, h, t):  # wind_condition/mod.rs:41-48
        u = self.steady_at_height(h)
        if self.gusts[0] is not None:
            u += self.gust(0, t)
        return u


# ---- WindEnvironment (environment.rs:21-360, inflow_corrections.rs) ----------------------
class WindEnvironment:
    def __init__(self, d, m: Libm):
        if isinstance(d, str):
            d = json.loads(d)
        self.m = m
        self.up = vector(d.get("up_direction"), (0.0, 0.0, 1.0))
        self.axis = vector(d.get("wind_rotation_axis"), (0.0, 0.0, 1.0))
        self.zero = vector(d.get("zero_direction_vector"), (1.0, 0.0, 0.0))
        self.water_plane_height = float(d.get("water_plane_height", 0.0))
        self.inflow = d.get("inflow_corrections")

    def height(self, loc):  # :66-71
        return max(dot(loc, self.up) - self.water_plane_height, 0.0)

    def direction_vector(self, c: WindCondition):  # :73-79
        return rotate_around_axis(self.zero, c.direction_coming_from, self.axis, self.m)

    def steady_true(self, c, loc):  # :151-162
        return scale(self.direction_vector(c), c.steady_at_height(self.height(loc)))

    def unsteady_true(self, c, loc, t):  # :106-149
        u = c.unsteady_parallel_at_height(self.height(loc), t)
        par = self.direction_vector(c)
        perp = c.gust(1, t)
        perp_dir = cross(self.up, par)
        vert = c.gust(2, t)
        return add(add(scale(par, u), scale(perp_dir, perp)), scale(self.up, vert))

    def unsteady_apparent(self, c, loc, linear, t):  # :182-190
        return add(self.unsteady_true(c, loc, t), linear)

    def steady_apparent(self, c, loc, linear):  # :164-180
        return add(self.steady_true(c, loc), linear)

    def apparent_direction(self, c, linear, height):  # :258-278
        apparent = add(self.steady_true(c, scale(self.up, height)), linear)
        return signed_angle_between(self.zero, apparent, self.axis, self.m)

    def _correct_single(self, one, wing, h, v):  # inflow_corrections.rs:21-45
        lo, hi = _range(one["wing_indices"][wing])
        hs = [float(x) for x in one["height_values"][lo:hi]]
        mags = [float(x) for x in one["magnitude_corrections"][lo:hi]]
        angs = [float(x) for x in one["angle_corrections"][lo:hi]]
        if hi > len(one["height_values"]) or hi > len(one["magnitude_corrections"]) or hi > len(one["angle_corrections"]):
            raise ValueError("range outside the table")
        mag = linear_interpolation(h, hs, mags)
        ang = linear_interpolation(h, hs, angs)
        return scale(rotate_around_axis(v, ang, self.up, self.m), mag)

    def correct(self, wing, direction, h, v):  # inflow_corrections.rs:55-85
        dirs = [float(x) for x in self.inflow["apparent_wind_directions"]]
        per = [self._correct_single(self.inflow["corrections"][i], wing, h, v) for i in range(len(dirs))]
        return linear_interpolation(direction, dirs, per, vector_values=True)

    def apply_inflow_corrections(self, direction, fs, ctrl_points, wing_ranges):  # :225-254
        if self.inflow is None:
            return
        heights = [dot(p, self.up) for p in ctrl_points]
        for w, (lo, hi) in enumerate(wing_ranges):
            for i in range(lo, hi):
                fs[i] = self.correct(w, direction, heights[i], fs[i])

    def ctrl_points_with_corrections(self, c, ctrl_points, linear, t, wing_ranges):  # :192-223
        v = []
        avg = 0.0
        for p in ctrl_points:
            v.append(self.unsteady_apparent(c, p, linear, t))
            avg += dot(p, self.up)
        avg /= float(len(ctrl_points))
        self.apply_inflow_corrections(self.apparent_direction(c, linear, avg), v, ctrl_points, wing_ranges)
        return v


def _range(r):
    if isinstance(r, dict):
        return int(r["start"]), int(r["end"])
    return int(r[0]), int(r[1])


# ---- controller (controller/*.rs) ------------------------------------------------------------
def measure_float_values(values, wing_ranges, settings):  # measurements.rs:43-77
    kind = settings.get("measurement_type", "Mean")
    start = int(settings.get("start_index", 0))
    end_offset = int(settings.get("end_offset", 0))
    out = []
    for lo, hi in wing_ranges:
        w = values[lo:hi]
        n = len(w)
        end = (n - end_offset - 1) % U64  # usize arithmetic
        end = min(max(end, start + 1), (n - 1) % U64)
        if start > end or end > n:
            raise ValueError("measurement range outside the wing")
        x = w[start:end]
        if len(x) == 0:
            raise ValueError("empty measurement range")
        r = x[0]
        if kind == "Mean":  # stormath statistics.rs:11-24
            for v in x[1:]:
                r = r + v
            r = r / float(len(x))
        elif kind == "Max":  # :43-53
            for v in x[1:]:
                if v > r:
                    r = v
        else:  # :56-66
            for v in x[1:]:
                if v < r:
                    r = v
        out.append(r)
    return out


def signum(x):  # f64::signum
    return math.copysign(1.0, x)


def limit_value(old, raw, max_change):  # set_points.rs:28-46
    diff = raw - old
    return old + max_change * signum(diff) if abs(diff) > max_change else raw


def wrap(a):  # set_points.rs:188-200
    while a > PI:
        a -= TAU
    while a < -PI:
        a += TAU
    return a


def new_output(sp, inp, dt):  # set_points.rs:70-186
    dirs = [float(x) for x in sp["apparent_wind_directions_data"]]
    aoa_data = sp.get("angle_of_attack_data")
    angle = 0.0
    if aoa_data is not None:
        set_point = inp["loading"] * linear_interpolation(inp["direction"], dirs, [float(x) for x in aoa_data])
        if sp.get("use_effective_angle_of_attack", False):
            err = wrap(inp["aoa"] - set_point)
            angle = wrap(inp["wing_angle"] + err)
        else:
            angle = inp["direction"] - set_point
    if sp.get("max_local_wing_angle_change_rate") is not None:
        angle = limit_value(inp["wing_angle"], angle, float(sp["max_local_wing_angle_change_rate"]) * dt)
    state = 0.0
    st_data = sp.get("section_model_internal_state_data")
    if st_data is not None:
        raw = inp["loading"] * linear_interpolation(inp["direction"], dirs, [float(x) for x in st_data])
        state = raw
        ist = sp.get("internal_state_type", "Generic")
        if isinstance(ist, dict) and "SpinRatio" in ist:  # :12-25
            conv = ist["SpinRatio"]
            rps = raw * inp["velocity"] / (PI * float(conv["diameter"]))
            state = float(conv["max_rps"]) * signum(rps) if abs(rps) > float(conv["max_rps"]) else rps
    if sp.get("max_internal_section_state_change_rate") is not None:
        state = limit_value(inp["state"], state, float(sp["max_internal_section_state_change_rate"]) * dt)
    return angle, state


# ---- CompleteSailModel (complete_sail_model/mod.rs) -------------------------------------------
class CompleteSailModel:
    def __init__(self, setup_string: str, variant: str = "glibc"):
        d = json.loads(setup_string)
        self.m = Libm(variant)
        self.sim = oracle_py.Simulation(json.dumps(d["lifting_line_simulation"]), variant)
        self.env = WindEnvironment(d["wind_environment"], self.m)
        c = d["controller"]
        self.set_points = c["set_points"]
        fm = c.get("flow_measurement_settings") or {}
        self.meas = {k: fm.get(k) or {} for k in ("angle_of_attack", "wind_direction", "wind_velocity")}
        self.steps_between_updates = int(c.get("time_steps_between_updates", 1))
        self.start_time = float(c.get("start_time", 0.0))
        self.use_input_velocity = bool(c.get("use_input_velocity_for_apparent_wind_direction", False))

    def condition(self, wc) -> WindCondition:
        return wc if isinstance(wc, WindCondition) else WindCondition(wc, self.m)

    def _lm(self):
        lm = self.sim.line_model()
        ranges = [(int(a), int(b)) for a, b in lm["wing_ranges"]]
        return lm, ranges

    def freestream(self, wc, ship_velocity, time):  # mod.rs:213-255
        c = self.condition(wc)
        pts = [tuple(p) for p in self.sim.get_freestream_velocity_points()]
        linear = scale(self.env.zero, ship_velocity)
        fs = [self.env.unsteady_apparent(c, p, linear, time) for p in pts]
        direction = self.env.apparent_direction(c, linear, 10.0)
        n = self.sim.nr_ctrl_points
        _, ranges = self._lm()
        self.env.apply_inflow_corrections(direction, fs, pts[:n], ranges)
        return np.array(fs, dtype=np.float64).reshape(-1, 3)

    def do_step(self, time, time_step, wc, ship_velocity):  # mod.rs:174-191
        fs = self.freestream(wc, ship_velocity, time)
        return self.sim.do_step(time=time, time_step=time_step, freestream_velocity=fs)

    def do_multiple_steps(self, end_time, time_step, wc, ship_velocity):  # mod.rs:145-172
        self.sim.restart()
        out, t = [], 0.0
        while t < end_time:
            out.append(self.do_step(t, time_step, wc, ship_velocity))
            t += time_step
        return out

    def _strip_direction(self, lm, ranges, i, v):  # environment.rs:297-327
        w = next(k for k, (lo, hi) in enumerate(ranges) if lo <= i < hi)
        first = ranges[w][0]
        s = lm["span_lines"][first]
        axis = normalize(sub(tuple(s[1]), tuple(s[0])))
        rb = lm["rigid_body"]
        # the built model always rotates XYZ: LineForceModelBuilder does not pass rotation_type on
        chord = rotate(tuple(lm["chord_local_not_rotated"][first]), tuple(rb[1]), False, self.m)
        return signed_angle_between(chord, v, axis, self.m)

    def _inputs(self, loading, lm, ranges, aoa, speed, direction):  # input.rs:33-157
        a = measure_float_values(aoa, ranges, self.meas["angle_of_attack"])
        u = measure_float_values(speed, ranges, self.meas["wind_velocity"])
        d = measure_float_values(direction, ranges, self.meas["wind_direction"])
        return [{"loading": loading, "wing_angle": float(lm["wing_angles"][w]), "state": 0.0, "aoa": a[w],
                 "velocity": u[w], "direction": d[w]} for w in range(len(ranges))]

    def input_from_velocity(self, loading, v):  # input.rs:91-157
        lm, ranges = self._lm()
        speed = [length(x) for x in v]
        aoa = []
        for i, x in enumerate(v):  # force_calculations.rs:79-107
            s = lm["span_lines"][i]
            if length(x) > MIN_POSITIVE:
                aoa.append(signed_angle_between(tuple(lm["chord_global"][i]), x,
                                                normalize(sub(tuple(s[1]), tuple(s[0]))), self.m))
            else:
                aoa.append(0.0)
        direction = [self._strip_direction(lm, ranges, i, x) for i, x in enumerate(v)]
        return self._inputs(loading, lm, ranges, aoa, speed, direction)

    def input_from_result(self, loading, result):  # input.rs:33-88, results/simulation.rs:118-149
        lm, ranges = self._lm()
        rb = lm["rigid_body"]
        va = tuple(rb[3])
        trans = tuple(rb[0])
        vel = [tuple(x) for x in result.force_input.velocity]
        rel = [tuple(x) for x in (result.solver_input_ctrl_points_velocity if self.use_input_velocity else
                                  result.force_input.velocity)]
        pts = [tuple(x) for x in result.ctrl_points]
        speed = [length(x) for x in vel]
        direction = [self._strip_direction(lm, ranges, i, add(rel[i], cross(va, sub(pts[i], trans))))
                     for i in range(len(pts))]
        aoa = [float(x) for x in result.force_input.angles_of_attack]
        return self._inputs(loading, lm, ranges, aoa, speed, direction)

    def update(self, time, dt, inputs):  # controller/mod.rs:25-60, data_update.rs:160-177
        if self.steps_between_updates == 0:
            raise ZeroDivisionError("time_steps_between_updates is 0")
        if not time >= self.start_time:
            return
        outs = []
        for k, sp in enumerate(self.set_points):
            if k >= len(inputs):
                raise IndexError("more set points than sails")
            outs.append(new_output(sp, inputs[k], dt))
        lm, _ = self._lm()
        angles = [float(x) for x in lm["wing_angles"]]
        for k, (angle, _state) in enumerate(outs):
            angles[k] = angle
        self.sim.set_local_wing_angles(angles)

    def apply_controller_wind(self, time, dt, wc, ship_velocity, loading):  # mod.rs:193-211, 257-288
        fs = self.freestream(wc, ship_velocity, time)
        n = self.sim.nr_ctrl_points
        self.update(time, dt, self.input_from_velocity(loading, [tuple(x) for x in fs[:n]]))

    def apply_controller_result(self, time, dt, loading, result):  # mod.rs:290-318
        self.update(time, dt, self.input_from_result(loading, result))

    def steady_state(self, wc, ship_velocity, loading):  # mod.rs:99-122
        self.apply_controller_wind(0.0, 1.0, wc, ship_velocity, loading)
        return self.do_step(0.0, 1.0, wc, ship_velocity)

    def optimal_loading(self, wc, ship_velocity, n):  # mod.rs:58-97
        dx = (1.0 - 0.1) / float(n - 1) if n > 1 else math.inf
        x, results, best, best_i, loadings = 0.1, [], -math.inf, 0, []
        for i in range(n):
            r = self.steady_state(wc, ship_velocity, x)
            s = 0.0  # results/simulation.rs:61-69, x component
            for w in range(self.sim.nr_wings):
                s = s + float(r.integrated_forces[w].total[0])
            effective = -s * ship_velocity - 0.0
            if effective > best:
                best, best_i = effective, i
            results.append(r)
            loadings.append(x)
            x += dx
        return results[best_i], loadings[best_i]
