"""Time-domain route simulation of many ships on the batched GPU path.

Each ship carries four 30 m x 10 m wing sails 25 m apart along its deck (x), the
deck 10 m above the water, with a free wake.  The wind follows the reference's
WindEnvironment / WindCondition semantics (wind/environment.rs:60-190,
wind_condition/mod.rs:18-85), evaluated by the batch itself
(``Batch.set_wind`` / ``Batch.do_step_wind``, as CompleteSailModel::freestream_velocity,
complete_sail_model/mod.rs:215-255): a true wind with a power-law profile (1/9,
reference height 10 m), the ship's own speed along the environment's zero direction,
and two gust signals, one along the true wind and one across it.  Ships sail in
``n_conditions`` wind conditions (true wind direction and speed, ship speed), each
with a turbulence realization of its own.

Gusts: the von Karman longitudinal and lateral spectra (intensity 15 % of the true
wind speed, length scale 200 m) discretized as the reference's
``GustSpectrum.discretize`` (interfaces/stormbird_setup/.../gust_spectrum.py:17-37):
``nr_freqs`` frequencies on a linspace from ``min_freq`` to ``max_freq``,
amplitude sqrt(2 S(f_k) df) with df the linspace step, phases uniform in [0, 2 pi)
(seeded here); the signal is DiscretizedSpectrum::value_at_time
(wind_condition/discretized_spectrum.rs:37-45).

Roll: every step the roll angle is set through the finite-difference motion setter
(line_force_model/data_update.rs:188-205), which also gives the motion velocity felt by
the sails.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import setups as S

DEG = math.pi / 180.0
DECK_HEIGHT = 10.0  # the sails' foot above the water plane
CHORD = 10.0
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEVELOPED_ROUTE = os.path.join(DATA, "route_developed_wakes_1500rows.npz")


class VonKarman:
    """von Karman turbulence spectrum (one-sided, in Hz) of the gust along the mean wind
    (``lateral=False``: S(f) = 4 s^2 (L/U) / (1 + 70.8 x^2)^(5/6)) or across it
    (``lateral=True``: S(f) = 4 s^2 (L/U) (1 + 755.2 x^2) / (1 + 283.2 x^2)^(11/6)),
    x = f L / U, s = intensity U.  Both integrate to s^2 over 0 < f < infinity."""

    def __init__(self, mean_velocity: float, intensity: float = 0.15, length_scale: float = 200.0,
                 lateral: bool = False):
        self.mean_velocity = float(mean_velocity)
        self.standard_deviation = float(intensity) * self.mean_velocity
        self.length_scale = float(length_scale)
        self.lateral = bool(lateral)

    def spectrum_value(self, frequency):
        lu = self.length_scale / self.mean_velocity
        x = np.asarray(frequency, dtype=np.float64) * lu
        s2 = self.standard_deviation ** 2
        if self.lateral:
            return 4.0 * s2 * lu * (1.0 + 755.2 * x * x) / (1.0 + 283.2 * x * x) ** (11.0 / 6.0)
        return 4.0 * s2 * lu / (1.0 + 70.8 * x * x) ** (5.0 / 6.0)

    def discretize(self, min_freq: float, max_freq: float, nr_freqs: int, rng: np.random.Generator) -> dict:
        """DiscretizedSpectrum fields, as GustSpectrum.discretize (gust_spectrum.py:17-37)."""
        frequencies = np.linspace(min_freq, max_freq, nr_freqs)
        delta_freq = frequencies[1] - frequencies[0]
        amplitudes = np.sqrt(2.0 * self.spectrum_value(frequencies) * delta_freq)
        phases = rng.uniform(0.0, 2.0 * math.pi, nr_freqs)
        return {"frequencies": frequencies.tolist(), "amplitudes": amplitudes.tolist(), "phases": phases.tolist()}


def gust_at(spectrum: dict, t: float) -> float:
    """DiscretizedSpectrum::value_at_time: sum of a_k sin(2 pi f_k t + phi_k)."""
    f = np.asarray(spectrum["frequencies"])
    a = np.asarray(spectrum["amplitudes"])
    p = np.asarray(spectrum["phases"])
    return float(np.sum(a * np.sin(2.0 * math.pi * f * t + p)))


def apparent_wind(direction_coming_from: float, true_wind_speed: float, ship_speed: float) -> np.ndarray:
    """Steady apparent wind at the 10 m reference height: the true wind along the zero
    direction (+x) rotated by direction_coming_from about z, plus the ship's speed along the
    zero direction (environment.rs:73-79, 151-180; complete_sail_model/mod.rs:223)."""
    c, s = math.cos(direction_coming_from), math.sin(direction_coming_from)
    return np.array([true_wind_speed * c + ship_speed, true_wind_speed * s, 0.0])


def ship_setup(apparent_direction: float, nr_strips: int = 48, nr_panels: int = 1500, ratio: float = 1.0):
    """Four 30 m x 10 m wing sails 25 m apart along x; the chords 10 deg from the
    apparent wind direction (radians, the angle of the apparent wind vector from +x)."""
    u_dir = [math.cos(apparent_direction), math.sin(apparent_direction), 0.0]
    chord = S.rotate_around_axis([CHORD * u_dir[0], CHORD * u_dir[1], 0.0], -10.0 * DEG, [0.0, 0.0, 1.0])
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


def wind_environment() -> dict:
    """WindEnvironment (environment.rs:21-34): default directions, water plane 10 m below the deck."""
    return {"water_plane_height": -DECK_HEIGHT}


def route_conditions(n_conditions: int, seed: int):
    """Seeded wind conditions: true wind coming from 50-150 deg off the zero direction,
    8-14 m/s at 10 m, ship speed 4-7 m/s."""
    rng = np.random.default_rng([seed, 0])
    out = []
    for _ in range(n_conditions):
        out.append({"direction_coming_from": float(rng.uniform(50.0, 150.0)) * DEG,
                    "true_wind_speed": float(rng.uniform(8.0, 14.0)),
                    "ship_speed": float(rng.uniform(4.0, 7.0))})
    return out


class RouteSimulation:
    """``n_ships`` ships stepped together: ship i sails in condition i % n_conditions with
    turbulence realization i (see the module docstring)."""

    def __init__(self, n_ships: int = 64, n_conditions: int = 2, seed: int = 5, nr_strips: int = 48,
                 nr_panels: int = 1500, roll_amplitude_deg: float = 10.0, roll_period: float = 12.0,
                 intensity: float = 0.15, length_scale: float = 200.0, min_freq: float = 0.0005,
                 max_freq: float = 0.5, nr_freqs: int = 1000, mode: int = 0, device: int = 0):
        from . import Batch

        self.n = n_ships
        self.conditions = route_conditions(n_conditions, seed)
        self.condition_of = np.arange(n_ships) % n_conditions
        setups, self.wind_conditions = [], []
        self.ship_speed = np.empty(n_ships)
        self.dt = np.empty(n_ships)
        self.apparent_direction = np.empty(n_ships)
        for i in range(n_ships):
            c = self.conditions[self.condition_of[i]]
            a = apparent_wind(c["direction_coming_from"], c["true_wind_speed"], c["ship_speed"])
            self.apparent_direction[i] = math.atan2(a[1], a[0])
            setups.append(ship_setup(self.apparent_direction[i], nr_strips, nr_panels)[0])
            self.ship_speed[i] = c["ship_speed"]
            self.dt[i] = 0.25 * CHORD / float(np.linalg.norm(a))
            grng = np.random.default_rng([seed, 1, i])
            u = c["true_wind_speed"]
            self.wind_conditions.append({
                "direction_coming_from": c["direction_coming_from"],
                "velocity_variation": {"PowerModel": {"reference_velocity": u}},
                "parallel_gust": VonKarman(u, intensity, length_scale).discretize(min_freq, max_freq, nr_freqs, grng),
                "perpendicular_gust": VonKarman(u, intensity, length_scale, lateral=True).discretize(
                    min_freq, max_freq, nr_freqs, grng),
            })
        self.setups = setups
        self.roll_amp = roll_amplitude_deg * DEG
        self.roll_period = roll_period
        self.batch = Batch(setups, mode=mode, device=device)
        self.batch.set_wind(wind_environment(), self.wind_conditions)
        self.k = 0

    def roll(self, t: float) -> float:
        return self.roll_amp * math.sin(2.0 * math.pi * t / self.roll_period)

    def set_motion(self, k: int):
        """The roll of step k on every ship (time k dt)."""
        for i, sim in enumerate(self.batch.instances):
            sim.set_translation_and_rotation_with_finite_difference_for_the_velocity(
                float(self.dt[i]), [0.0, 0.0, 0.0], [self.roll(k * float(self.dt[i])), 0.0, 0.0])

    def step(self, want_result: bool = True):
        self.set_motion(self.k)
        res = self.batch.do_step_wind(times=self.k * self.dt, time_steps=self.dt, ship_velocities=self.ship_speed,
                                      want_result=want_result)
        self.k += 1
        return res

    def load_developed(self, path: str = DEVELOPED_ROUTE):
        """Every ship takes the stored wake of its condition (store_developed), and the route
        continues from that step: the roll of the two steps before it is replayed so that the
        motion setter's finite difference sees the same angles."""
        d = np.load(path)
        k0 = int(d["steps_completed"])
        if int(d["n_conditions"]) != len(self.conditions):
            raise ValueError("the stored wakes are of another set of conditions")
        while self.k < 3:  # the batch's own first steps (its first-step state), then the stored wakes
            self.step(want_result=False)
        for i, sim in enumerate(self.batch.instances):
            c = int(self.condition_of[i])
            sim.set_wake_state(d[f"points_{c}"].astype(np.float64), d[f"velocities_{c}"].astype(np.float64),
                               d[f"strengths_{c}"].astype(np.float64), k0)
            sim.previous_circulation_strength = d[f"previous_circulation_{c}"]
        self.set_motion(k0 - 2)
        self.set_motion(k0 - 1)
        self.k = k0

    def store_developed(self, path: str = DEVELOPED_ROUTE):
        """The wake of the first n_conditions ships (one per condition): points, velocities
        and strengths in float32 (0.25 mm at the wake's extent), the previous circulation in
        float64."""
        n_c = len(self.conditions)
        if self.n < n_c:
            raise ValueError("store_developed needs one ship per condition")
        out = {"n_conditions": n_c, "steps_completed": self.k}
        for c in range(n_c):
            st = self.batch.instances[c].get_wake_state()
            out[f"points_{c}"] = st["points"].astype(np.float32)
            out[f"velocities_{c}"] = st["velocities"].astype(np.float32)
            out[f"strengths_{c}"] = st["strengths"].astype(np.float32)
            out[f"previous_circulation_{c}"] = self.batch.instances[c].previous_circulation_strength.copy()
        np.savez_compressed(path, **out)
