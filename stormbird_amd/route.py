"""Time-domain route simulation of many ships on the batched GPU path.

Each ship carries four 30 m x 10 m wing sails 25 m apart with a free wake; the
apparent wind has a seeded angle and speed per ship, a spatially uniform gust
on its magnitude, and the ship rolls sinusoidally.  Every step:

- the roll angle is set through the finite-difference motion setter
  (line_force_model/data_update.rs:188-205), which also gives the motion
  velocity felt by the sails;
- the freestream at every control and wake point of a ship is its apparent
  wind vector scaled by (speed + gust(t)) / speed, written into one device
  array for all ships;
- all ships advance together through ``Batch.do_step_device``.

Gusts follow the reference's ``DiscretizedSpectrum``
(wind/wind_condition/discretized_spectrum.rs:13-47): a sum of harmonics
a_k sin(2 pi f_k t + phi_k), here with amplitudes from the longitudinal von
Karman spectrum S(f) = 4 sigma^2 (L/U) / (1 + 70.8 (f L / U)^2)^(5/6),
sigma = intensity * U, on log-spaced frequency bins, and seeded phases.
"""
from __future__ import annotations

import json
import math

import numpy as np

from . import setups as S

DEG = math.pi / 180.0


def von_karman_spectrum(mean_speed: float, intensity: float = 0.15, length_scale: float = 200.0, n: int = 64,
                        f_min: float = 0.002, f_max: float = 0.5, rng=None) -> dict:
    """Discretized longitudinal von Karman gust spectrum (frequencies [Hz], amplitudes [m/s], phases [rad])."""
    rng = rng if rng is not None else np.random.default_rng(0)
    edges = np.geomspace(f_min, f_max, n + 1)
    f = np.sqrt(edges[:-1] * edges[1:])
    df = np.diff(edges)
    sigma = intensity * mean_speed
    lu = length_scale / mean_speed
    s = 4.0 * sigma * sigma * lu / (1.0 + 70.8 * (f * lu) ** 2) ** (5.0 / 6.0)
    return {"frequencies": f.tolist(), "amplitudes": np.sqrt(2.0 * s * df).tolist(),
            "phases": rng.uniform(0.0, 2.0 * math.pi, n).tolist()}


def gust_at(spectrum: dict, t: float) -> float:
    """DiscretizedSpectrum::value_at_time: sum of a_k sin(2 pi f_k t + phi_k)."""
    f = np.asarray(spectrum["frequencies"])
    a = np.asarray(spectrum["amplitudes"])
    p = np.asarray(spectrum["phases"])
    return float(np.sum(a * np.sin(2.0 * math.pi * f * t + p)))


def ship_setup(awa_deg: float, nr_strips: int = 48, nr_panels: int = 1500, ratio: float = 1.0):
    """Four 30 m x 10 m wing sails 25 m apart along x, chords 10 deg from the apparent wind."""
    awa = awa_deg * DEG
    u_dir = [-math.cos(awa), -math.sin(awa), 0.0]
    chord = S.rotate_around_axis([10.0 * u_dir[0], 10.0 * u_dir[1], 0.0], -10.0 * DEG, [0.0, 0.0, 1.0])
    wings = []
    for x in (-37.5, -12.5, 12.5, 37.5):
        wings.append({
            "section_points": [S._v(x, 0, 0), S._v(x, 0, 30.0)],
            "chord_vectors": [S._v(*chord), S._v(*chord)],
            "section_model": S.foil(**S.SAIL_FOIL),
        })
    setup = {
        "line_force_model": {"wing_builders": wings, "nr_sections": nr_strips},
        "simulation_settings": S.dynamic_settings(nr_panels=nr_panels, core={"Absolute": 1.0}, ratio=ratio,
                                                  first_rel=0.25),
    }
    return json.dumps(setup), u_dir


class RouteSimulation:
    """``n_ships`` ships stepped together (see module docstring)."""

    def __init__(self, n_ships: int = 64, seed: int = 5, nr_strips: int = 48, nr_panels: int = 1500,
                 roll_amplitude_deg: float = 10.0, roll_period: float = 12.0, intensity: float = 0.15,
                 length_scale: float = 200.0, mode: int = 0, device: int = 0):
        import torch

        from . import Batch

        rng = np.random.default_rng(seed)
        self.n = n_ships
        self.awa = rng.uniform(30.0, 150.0, n_ships)
        self.aws = rng.uniform(8.0, 16.0, n_ships)
        setups, self.dirs = [], []
        for a in self.awa:
            js, d = ship_setup(a, nr_strips, nr_panels)
            setups.append(js)
            self.dirs.append(np.asarray(d))
        self.spectra = [von_karman_spectrum(u, intensity, length_scale, rng=rng) for u in self.aws]
        self.dt = 0.25 * 10.0 / self.aws
        self.roll_amp = roll_amplitude_deg * DEG
        self.roll_period = roll_period
        self.batch = Batch(setups, mode=mode, device=device)
        self.offsets = self.batch.freestream_offsets
        self.fs = torch.zeros((int(self.offsets[-1]), 3), dtype=torch.float64, device=f"cuda:{device}")
        self.k = 0

    def velocity(self, i: int, t: float) -> np.ndarray:
        return self.dirs[i] * (self.aws[i] + gust_at(self.spectra[i], t))

    def roll(self, t: float) -> float:
        return self.roll_amp * math.sin(2.0 * math.pi * t / self.roll_period)

    def prepare_step(self):
        import torch

        times = self.k * self.dt
        for i, sim in enumerate(self.batch.instances):
            t = float(times[i])
            sim.set_translation_and_rotation_with_finite_difference_for_the_velocity(
                float(self.dt[i]), [0.0, 0.0, 0.0], [self.roll(t), 0.0, 0.0])
            v = torch.from_numpy(self.velocity(i, t)).to(self.fs.device)
            self.fs[int(self.offsets[i]):int(self.offsets[i + 1])] = v
        return times

    def step(self, want_result: bool = True):
        times = self.prepare_step()
        res = self.batch.do_step_device(times=times, time_steps=self.dt, freestream_ptr=self.fs.data_ptr(),
                                        n_points=int(self.offsets[-1]), want_result=want_result)
        self.k += 1
        return res

    def freestream_host(self, i: int) -> np.ndarray:
        """The freestream of ship i at the current step (as written by prepare_step)."""
        return self.fs[int(self.offsets[i]):int(self.offsets[i + 1])].cpu().numpy()
