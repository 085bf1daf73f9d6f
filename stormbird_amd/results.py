"""Result structures returned by ``Simulation.do_step``.

Mirrors the Python-visible shape of Stormbird's ``SimulationResult``
(stormbird/src/common_utils/results/simulation.rs:27-40 and the pystormbird
stubs interfaces/pystormbird/pystormbird/__init__.pyi:4-70): sectional
forces input, sectional forces, per-wing integrated forces and moments,
iterations and residual.

The C ABI fills caller-owned arrays through ``StepResultC`` (same layout as
``sb_step_result`` in include/stormbird_gpu.h).
"""
from __future__ import annotations

import ctypes

import numpy as np

FORCE_KINDS = ("circulatory", "viscous_lift", "sectional_drag", "added_mass", "gyroscopic", "total")

_dp = ctypes.POINTER(ctypes.c_double)


class StepResultC(ctypes.Structure):
    _fields_ = [
        ("time", ctypes.c_double),
        ("iterations", ctypes.c_long),
        ("residual", ctypes.c_double),
        ("circulation_strength", _dp),
        ("angles_of_attack", _dp),
        ("ctrl_points_velocity", _dp),
        ("solver_input_velocity", _dp),
        ("ctrl_points", _dp),
        ("sectional_forces", _dp),
        ("integrated_forces", _dp),
        ("integrated_moments", _dp),
    ]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


# Output fields of one step: name, per-strip (True) or per-wing, shape after the count.
FIELDS = (("circulation_strength", True, ()), ("angles_of_attack", True, ()),
          ("ctrl_points_velocity", True, (3,)), ("solver_input_velocity", True, (3,)),
          ("ctrl_points", True, (3,)), ("sectional_forces", True, (6, 3)),
          ("integrated_forces", False, (6, 3)), ("integrated_moments", False, (6, 3)))


def _field_shape(per_strip, tail, n, w):
    count = n if per_strip else w
    if len(tail) == 2:  # [6][count][3]
        return (6, count, 3)
    return (count,) + tail


class StepBuffers:
    """Caller-owned output arrays for one step (N strips, W wings).  `storage` maps each
    field to a flat array to use (views into batch-wide arrays); default: own arrays."""

    def __init__(self, n: int, w: int, storage=None):
        self.n, self.w = n, w
        for name, per_strip, tail in FIELDS:
            shape = _field_shape(per_strip, tail, n, w)
            if storage is None:
                arr = np.zeros(shape)
            else:
                arr = storage[name].reshape(shape)
            setattr(self, name, arr)
        self.c = StepResultC()
        self.c.circulation_strength = _ptr(self.circulation_strength)
        self.c.angles_of_attack = _ptr(self.angles_of_attack)
        self.c.ctrl_points_velocity = _ptr(self.ctrl_points_velocity)
        self.c.solver_input_velocity = _ptr(self.solver_input_velocity)
        self.c.ctrl_points = _ptr(self.ctrl_points)
        self.c.sectional_forces = _ptr(self.sectional_forces)
        self.c.integrated_forces = _ptr(self.integrated_forces)
        self.c.integrated_moments = _ptr(self.integrated_moments)

    def snapshot(self) -> "SimulationResult":
        return SimulationResult(
            time=self.c.time,
            iterations=int(self.c.iterations),
            residual=float(self.c.residual),
            circulation_strength=self.circulation_strength.copy(),
            angles_of_attack=self.angles_of_attack.copy(),
            velocity=self.ctrl_points_velocity.copy(),
            solver_input_velocity=self.solver_input_velocity.copy(),
            ctrl_points=self.ctrl_points.copy(),
            sectional=self.sectional_forces.copy(),
            integrated_forces=self.integrated_forces.copy(),
            integrated_moments=self.integrated_moments.copy(),
        )


class _Fields:
    def __init__(self, arr: np.ndarray):
        for k, name in enumerate(FORCE_KINDS):
            setattr(self, name, arr[k])


class SectionalForcesInput:
    def __init__(self, circulation_strength, velocity, angles_of_attack):
        self.circulation_strength = circulation_strength
        self.velocity = velocity
        self.angles_of_attack = angles_of_attack


class SingleSailResult:
    """results/simplfied.rs:9-15: total force and moment of one sail and its input power."""

    def __init__(self, force, moment, input_power):
        self.force = force
        self.moment = moment
        self.input_power = input_power


class SimulationResult:
    """One step's result (same field names as Stormbird's SimulationResult).  The nested
    views (force_input, sectional_forces, per-wing integrated forces) are built on first use."""

    def __init__(self, *, time, iterations, residual, circulation_strength, angles_of_attack,
                 velocity, solver_input_velocity, ctrl_points, sectional, integrated_forces,
                 integrated_moments):
        self.time = time
        self.iterations = iterations
        self.residual = residual
        self.ctrl_points = ctrl_points
        self.solver_input_ctrl_points_velocity = solver_input_velocity
        self._gamma, self._vel, self._alpha, self._sect = circulation_strength, velocity, angles_of_attack, sectional
        self._if = integrated_forces
        self._im = integrated_moments
        self._cache = {}

    def _lazy(self, key, make):
        v = self._cache.get(key)
        if v is None:
            v = self._cache[key] = make()
        return v

    @property
    def force_input(self):
        return self._lazy("fi", lambda: SectionalForcesInput(self._gamma, self._vel, self._alpha))

    @property
    def sectional_forces(self):
        return self._lazy("sf", lambda: _Fields(self._sect))

    @property
    def integrated_forces(self):
        return self._lazy("if", lambda: [_Fields(self._if[:, w, :]) for w in range(self._if.shape[1])])

    @property
    def integrated_moments(self):
        return self._lazy("im", lambda: [_Fields(self._im[:, w, :]) for w in range(self._im.shape[1])])

    @property
    def input_power(self):
        """Per-wing input power: zero, as for every sail whose input power model is
        NoPower (line_force_model/input_power.rs:76-98), the only model this path takes."""
        return [0.0] * self._if.shape[1]

    def input_power_sum(self) -> float:
        """results/simulation.rs:81-89."""
        return 0.0

    def as_simplified(self):
        """Per-wing total force and moment and input power (results/simulation.rs:150-166)."""
        return [SingleSailResult(self._if[5, w].copy(), self._im[5, w].copy(), 0.0)
                for w in range(self._if.shape[1])]

    def integrated_forces_sum(self) -> np.ndarray:
        """Sum of the per-wing total forces (results/simulation.rs:61-69)."""
        s = np.zeros(3)
        for w in range(self._if.shape[1]):
            s = s + self._if[5, w]
        return s

    def integrated_moments_sum(self) -> np.ndarray:
        s = np.zeros(3)
        for w in range(self._im.shape[1]):
            s = s + self._im[5, w]
        return s

    def as_arrays(self) -> dict:
        """Every numeric output as a flat dict (used by parity checks)."""
        return {
            "circulation_strength": self._gamma,
            "angles_of_attack": self._alpha,
            "velocity": self._vel,
            "sectional_forces": self._sect,
            "integrated_forces": self._if,
            "integrated_moments": self._im,
        }


class BatchResults:
    """The results of one batched step: a sequence of per-instance SimulationResult built on
    access from one copy of the batch-wide output arrays, plus whole-batch arrays."""

    def __init__(self, time, iterations, residual, arrays, strip_off, wing_off):
        self.time = time                # [n]
        self.iterations = iterations    # [n]
        self.residual = residual        # [n]
        self._a = arrays                # field -> flat array over all instances
        self._so, self._wo = strip_off, wing_off
        self._items = {}

    def __len__(self):
        return len(self.iterations)

    def field(self, name: str, i: int) -> np.ndarray:
        per_strip, tail = next((p, t) for f, p, t in FIELDS if f == name)
        n, w = self._so[i + 1] - self._so[i], self._wo[i + 1] - self._wo[i]
        shape = _field_shape(per_strip, tail, n, w)
        size = int(np.prod(shape))
        start = (self._so[i] if per_strip else self._wo[i]) * size // (n if per_strip else w)
        return self._a[name][start:start + size].reshape(shape)

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[k] for k in range(*i.indices(len(self)))]
        if i < 0:
            i += len(self)
        if not 0 <= i < len(self):
            raise IndexError(i)
        r = self._items.get(i)
        if r is None:
            f = lambda name: self.field(name, i)  # noqa: E731
            r = self._items[i] = SimulationResult(
                time=float(self.time[i]), iterations=int(self.iterations[i]), residual=float(self.residual[i]),
                circulation_strength=f("circulation_strength"), angles_of_attack=f("angles_of_attack"),
                velocity=f("ctrl_points_velocity"), solver_input_velocity=f("solver_input_velocity"),
                ctrl_points=f("ctrl_points"), sectional=f("sectional_forces"),
                integrated_forces=f("integrated_forces"), integrated_moments=f("integrated_moments"))
        return r

    def __iter__(self):
        for i in range(len(self)):
            yield self[i]
