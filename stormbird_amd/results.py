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


class StepBuffers:
    """Caller-owned output arrays for one step (N strips, W wings)."""

    def __init__(self, n: int, w: int):
        self.n, self.w = n, w
        self.circulation_strength = np.zeros(n)
        self.angles_of_attack = np.zeros(n)
        self.ctrl_points_velocity = np.zeros((n, 3))
        self.solver_input_velocity = np.zeros((n, 3))
        self.ctrl_points = np.zeros((n, 3))
        self.sectional_forces = np.zeros((6, n, 3))
        self.integrated_forces = np.zeros((6, w, 3))
        self.integrated_moments = np.zeros((6, w, 3))
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


class SimulationResult:
    """One step's result (same field names as Stormbird's SimulationResult)."""

    def __init__(self, *, time, iterations, residual, circulation_strength, angles_of_attack,
                 velocity, solver_input_velocity, ctrl_points, sectional, integrated_forces,
                 integrated_moments):
        self.time = time
        self.iterations = iterations
        self.residual = residual
        self.ctrl_points = ctrl_points
        self.solver_input_ctrl_points_velocity = solver_input_velocity
        self.force_input = SectionalForcesInput(circulation_strength, velocity, angles_of_attack)
        self.sectional_forces = _Fields(sectional)
        self.integrated_forces = [_Fields(integrated_forces[:, w, :]) for w in range(integrated_forces.shape[1])]
        self.integrated_moments = [_Fields(integrated_moments[:, w, :]) for w in range(integrated_moments.shape[1])]
        self._if = integrated_forces
        self._im = integrated_moments
        self.input_power = [0.0] * integrated_forces.shape[1]

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
            "circulation_strength": self.force_input.circulation_strength,
            "angles_of_attack": self.force_input.angles_of_attack,
            "velocity": self.force_input.velocity,
            "sectional_forces": np.stack([getattr(self.sectional_forces, k) for k in FORCE_KINDS]),
            "integrated_forces": self._if,
            "integrated_moments": self._im,
        }
