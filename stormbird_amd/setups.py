"""Setup-string builders for the lifting-line workloads used in tests and benches.

Every function returns a JSON string in the schema Stormbird's
``SimulationBuilder`` deserialises (stormbird/src/lifting_line/simulation_builder.rs:34-66,
dynamic_wake/builder.rs:58-119, line_force_model/builder/mod.rs:23-46,
builder/single_wing.rs:35-47, section_models/foil.rs:43-117).

Workloads:
  * ``rectangular_wing``   — the reference's own test wing (lifting_line/tests/test_setup.rs:12-66).
  * ``single_wing_sail``   — one 30 m x 10 m wing sail, 20 cosine-spaced strips, frozen-shape
                              wake of 200 panel rows (4,000 panels), 10 m/s at 10 deg.
  * ``four_sail_free_wake``— four 30 m x 10 m sails 25 m apart, 40 strips each, 2,000 panel rows
                              (320,000 panels), free wake (all wake points get induced velocity).
  * ``polar_sweep_instance``— one seeded instance of the 512-instance batched polar sweep.
"""
from __future__ import annotations

import json
import os
import math

import numpy as np


def _v(x, y, z):
    return {"x": float(x), "y": float(y), "z": float(z)}


def rotate_around_axis(v, angle, axis):
    """Rodrigues rotation in the reference's operation order (transformations.rs:29-40)."""
    n = math.sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2])
    a = [axis[0] / n, axis[1] / n, axis[2] / n] if n > 0.0 else [0.0, 0.0, 0.0]
    c, s = math.cos(angle), math.sin(angle)
    t1 = [v[0] * c, v[1] * c, v[2] * c]
    cr = [a[1] * v[2] - a[2] * v[1], a[2] * v[0] - a[0] * v[2], a[0] * v[1] - a[1] * v[0]]
    t2 = [cr[0] * s, cr[1] * s, cr[2] * s]
    d = a[0] * v[0] + a[1] * v[1] + a[2] * v[2]
    t3 = [(a[i] * d) * (1.0 - c) for i in range(3)]
    return [t1[i] + t2[i] + t3[i] for i in range(3)]


def foil(**kw):
    return {"Foil": dict(kw)}


def dynamic_settings(nr_panels=100, core=None, ratio=0.0, first_rel=None, last_rel=None,
                     symmetry=None, neglect_self=False, solver=None, extra_wake=None):
    wake = {"nr_panels_per_line_element": int(nr_panels)}
    if core is not None:
        wake["viscous_core_length"] = core
    if ratio:
        wake["ratio_of_wake_affected_by_induced_velocities"] = float(ratio)
    if first_rel is not None:
        wake["first_panel_relative_length"] = float(first_rel)
    if last_rel is not None:
        wake["last_panel_relative_length"] = float(last_rel)
    if symmetry is not None:
        wake["symmetry_condition"] = symmetry
    if neglect_self:
        wake["neglect_self_induced_velocities"] = True
    if extra_wake:
        wake.update(extra_wake)
    d = {"wake": wake}
    if solver is not None:
        d["solver"] = {"SimpleIterative": dict(solver)} if "Linearized" not in solver else dict(solver)
    return {"Dynamic": d}


def quasi_steady_settings(solver=None, wake_length_factor=None, core=None, symmetry=None):
    """QuasiSteadySettings (simulation_builder.rs:21-29, wake/settings.rs:33-58).

    ``solver`` is None (the default Linearized solver), {"Linearized": {...}} or the fields of a
    QuasiSteadySimpleIterativeBuilder (simple_iterative.rs:14-45)."""
    d = {}
    if solver is not None:
        d["solver"] = dict(solver) if "Linearized" in solver else {"SimpleIterative": dict(solver)}
    wake = {}
    if wake_length_factor is not None:
        wake["wake_length_factor"] = float(wake_length_factor)
    if core is not None:
        wake["viscous_core_length"] = core
    if symmetry is not None:
        wake["symmetry_condition"] = symmetry
    if wake:
        d["wake"] = wake
    return {"QuasiSteady": d}


def with_settings(setup_json: str, settings) -> str:
    """The same line force model with other simulation settings (None drops the key, which the
    reference reads as the default QuasiSteady settings, simulation_builder.rs:51-64)."""
    d = json.loads(setup_json)
    if settings is None:
        d.pop("simulation_settings", None)
    else:
        d["simulation_settings"] = settings
    return json.dumps(d)


def rectangular_wing(aspect_ratio=5.0, cl_zero_angle=0.5, angle_of_attack=0.0, nr_strips=20,
                     negative_span_orientation=False, nr_panels=100, **dyn):
    """The reference's RectangularWing test setup (lifting_line/tests/test_setup.rs:37-66)."""
    axis = [0.0, 0.0, -1.0] if negative_span_orientation else [0.0, 0.0, 1.0]
    chord = rotate_around_axis([1.0, 0.0, 0.0], -angle_of_attack, axis)
    last_z = -aspect_ratio if negative_span_orientation else aspect_ratio
    wing = {
        "section_points": [_v(0, 0, 0), _v(0, 0, last_z)],
        "chord_vectors": [_v(*chord), _v(*chord)],
        "section_model": foil(cl_zero_angle=cl_zero_angle,
                              mean_positive_stall_angle=45.0 * (math.pi / 180.0),
                              mean_negative_stall_angle=45.0 * (math.pi / 180.0)),
        "non_zero_circulation_at_ends": [False, False],
    }
    setup = {
        "line_force_model": {"wing_builders": [wing], "nr_sections": int(nr_strips), "density": 13.2},
        "simulation_settings": dynamic_settings(nr_panels=nr_panels, **dyn),
    }
    return json.dumps(setup)


def wing_sails_quasi_steady(nr_strips=40, solver=None):
    """Three 10 m wing sails, chord 1 m, at (x, y) = (-5, 5), (0, -5), (5, 5), z from 10 m, with
    default quasi-steady Linearized settings: the reference's speed test
    (lifting_line/tests/wing_sails_speed_test.rs:25-78); freestream 1.2 m/s along x."""
    wings = []
    for x, y in ((-5.0, 5.0), (0.0, -5.0), (5.0, 5.0)):
        wings.append({
            "section_points": [_v(x, y, 10.0), _v(x, y, 20.0)],
            "chord_vectors": [_v(1, 0, 0), _v(1, 0, 0)],
            "section_model": foil(cl_zero_angle=0.5),
            "non_zero_circulation_at_ends": [False, False],
        })
    return json.dumps({"line_force_model": {"wing_builders": wings, "nr_sections": int(nr_strips)},
                       "simulation_settings": quasi_steady_settings(solver=solver)})


DEG = math.pi / 180.0

# Section model of the wing-sail workloads: CL = 2*pi*sin(alpha) blended to zero around 15 deg
# (logistic blend, equivalent to a tanh blend), CD = 0.01 + 0.05 alpha^2.
SAIL_FOIL = dict(cl_initial_slope=2.0 * math.pi, mean_positive_stall_angle=15.0 * DEG,
                 mean_negative_stall_angle=15.0 * DEG, cd_min=0.01, cd_second_order_factor=0.05)


def single_wing_sail(nr_panels=200, nr_strips=20):
    """One wing sail: span 30 m (+z), chord 10 m (+x), cosine-spaced strips.

    With 21 cosine-spaced section points and nr_sections = 12 the unchanged WingBuilder
    (builder/single_wing.rs:83-99) makes exactly one strip per segment -> 20 strips.
    """
    zs = [15.0 * (1.0 - math.cos(k * math.pi / nr_strips)) for k in range(nr_strips + 1)]
    wing = {
        "section_points": [_v(0, 0, z) for z in zs],
        "chord_vectors": [_v(10, 0, 0) for _ in zs],
        "section_model": foil(**SAIL_FOIL),
    }
    setup = {
        "line_force_model": {"wing_builders": [wing], "nr_sections": 12},
        "simulation_settings": dynamic_settings(nr_panels=nr_panels, core={"Absolute": 1.0},
                                                first_rel=0.25),
    }
    return json.dumps(setup)


def single_wing_sail_case():
    """Freestream, time step and step count of the single-wing-sail workload."""
    u = [10.0 * math.cos(10.0 * DEG), 10.0 * math.sin(10.0 * DEG), 0.0]
    return {"velocity": u, "time_step": 0.25, "steps": 200}


def four_sail_free_wake(nr_panels=2000, nr_strips=40, seed=1, ratio=1.0, alpha_deg=10.0):
    """Four 30 m x 10 m wing sails at x = -37.5, -12.5, 12.5, 37.5 m; free wake.

    Apparent wind 12 m/s from a seeded angle AWA ~ U[30, 150] deg; each sail's chord is set
    alpha_deg from the apparent-wind direction.  Returns (setup_json, case dict).
    """
    rng = np.random.default_rng(seed)
    awa = float(rng.uniform(30.0, 150.0)) * DEG
    aws = 12.0
    u = [-aws * math.cos(awa), -aws * math.sin(awa), 0.0]
    udir = [u[0] / aws, u[1] / aws, 0.0]
    chord = rotate_around_axis([10.0 * udir[0], 10.0 * udir[1], 0.0], -alpha_deg * DEG, [0.0, 0.0, 1.0])
    wings = []
    for x in (-37.5, -12.5, 12.5, 37.5):
        wings.append({
            "section_points": [_v(x, 0, 0), _v(x, 0, 30)],
            "chord_vectors": [_v(*chord), _v(*chord)],
            "section_model": foil(**SAIL_FOIL),
        })
    setup = {
        "line_force_model": {"wing_builders": wings, "nr_sections": int(nr_strips)},
        "simulation_settings": dynamic_settings(nr_panels=nr_panels, core={"Absolute": 1.0},
                                                ratio=ratio, first_rel=0.25),
    }
    case = {"velocity": u, "time_step": 0.25 * 10.0 / aws, "steps": int(nr_panels), "awa_deg": awa / DEG}
    return json.dumps(setup), case


def complete_sail_model(setup_json: str, n_sails: int, environment=None, max_angle_deg=15.0) -> str:
    """CompleteSailModelBuilder JSON (complete_sail_model/builder.rs:14-20) around a
    SimulationBuilder string: the default wind environment (or `environment`) and, per sail,
    the default single-element wing-sail set points of the reference's setup package
    (stormbird_setup/controller.py: angle of attack +-max_angle_deg beyond 15 deg of apparent
    wind, zero within 10 deg) with its measurement defaults (start_index 1, end_offset 1)."""
    dirs = [math.radians(a) for a in (-180.0, -15.0, -10.0, 10.0, 15.0, 180.0)]
    a = math.radians(max_angle_deg)
    set_point = {"apparent_wind_directions_data": dirs, "angle_of_attack_data": [-a, -a, 0.0, 0.0, a, a]}
    measure = {"measurement_type": "Mean", "start_index": 1, "end_offset": 1}
    return json.dumps({
        "lifting_line_simulation": json.loads(setup_json),
        "wind_environment": environment if environment is not None else {},
        "controller": {"set_points": [set_point] * int(n_sails),
                       "flow_measurement_settings": {"angle_of_attack": measure, "wind_direction": measure,
                                                     "wind_velocity": measure}},
    })


DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEVELOPED_FOUR_SAIL = os.path.join(DATA, "four_sail_developed_wake_2000rows.npz")


def compact_state(src: str, dst: str = DEVELOPED_FOUR_SAIL) -> None:
    """Stores a dumped wake state (bench.py --only full --dump-state src) compactly: wake
    points and strengths in float32 (0.25 mm at the wake's 5 km extent), the previous
    circulation in float64; the velocities are not kept."""
    d = np.load(src)
    np.savez_compressed(dst, points=d["points"].astype(np.float32), strengths=d["strengths"].astype(np.float32),
                        previous_circulation=d["previous_circulation"], steps_completed=d["steps_completed"])


def developed_four_sail_state(case):
    """The wake of four_sail_free_wake() (defaults: 40 strips, 2,000 rows) after 2,000 steps run
    from the start on the GPU in EXACT arithmetic: rolled-up tip and root vortices, 17 % of
    the self-induction pairs in the near field.  Returns (points, velocities, strengths,
    previous circulation, steps completed); the velocities are the freestream (the first step
    convects with them, later steps with the induced field)."""
    d = np.load(DEVELOPED_FOUR_SAIL)
    pts = d["points"].astype(np.float64)
    vel = np.tile(np.asarray(case["velocity"], dtype=np.float64), (pts.shape[0], 1))
    return (pts, vel, d["strengths"].astype(np.float64), np.array(d["previous_circulation"], dtype=np.float64),
            int(d["steps_completed"]))


def polar_sweep_instance(rng: np.random.Generator, max_panels=3000, min_panels=200):
    """One instance of the batched polar sweep.

    W ~ U{1..6} sails, strips/sail ~ U{16..64}, chord ~ U[5,15] m, AR ~ U[2,6],
    wake rows M ~ U{min_panels..max_panels}, AWA ~ U[20,180] deg, AWS ~ U[6,18] m/s.
    Sails stand 2.5 chords apart along x; chords sit 8 deg from the wind.
    """
    w = int(rng.integers(1, 7))
    strips = int(rng.integers(16, 65))
    c = float(rng.uniform(5.0, 15.0))
    ar = float(rng.uniform(2.0, 6.0))
    m = int(rng.integers(min_panels, max_panels + 1))
    awa = float(rng.uniform(20.0, 180.0)) * DEG
    aws = float(rng.uniform(6.0, 18.0))
    u = [-aws * math.cos(awa), -aws * math.sin(awa), 0.0]
    chord = rotate_around_axis([c * u[0] / aws, c * u[1] / aws, 0.0], -8.0 * DEG, [0.0, 0.0, 1.0])
    span = ar * c
    wings = []
    for k in range(w):
        x = (k - 0.5 * (w - 1)) * 2.5 * c
        wings.append({
            "section_points": [_v(x, 0, 0), _v(x, 0, span)],
            "chord_vectors": [_v(*chord), _v(*chord)],
            "section_model": foil(**SAIL_FOIL),
        })
    setup = {
        "line_force_model": {"wing_builders": wings, "nr_sections": strips},
        "simulation_settings": dynamic_settings(nr_panels=m, core={"Absolute": 0.1 * c}),
    }
    case = {"velocity": u, "time_step": 0.25 * c / aws, "steps": m, "wings": w, "strips": strips}
    return json.dumps(setup), case


def quasi_steady_polar_instance(rng: np.random.Generator, solver=None):
    """One instance of the quasi-steady polar sweep (default QuasiSteady settings: horseshoe
    wake, Linearized solve).

    W ~ U{1..4} sails, strips/sail ~ U{8..32} (N <= 128), chord ~ U[5,15] m, AR ~ U[2,6],
    AWA ~ U[20,180] deg, AWS ~ U[6,18] m/s; sails 2.5 chords apart along x, chords 8 deg
    from the wind (the sail foil of the wing-sail workloads).
    """
    w = int(rng.integers(1, 5))
    strips = int(rng.integers(8, 33))
    c = float(rng.uniform(5.0, 15.0))
    ar = float(rng.uniform(2.0, 6.0))
    awa = float(rng.uniform(20.0, 180.0)) * DEG
    aws = float(rng.uniform(6.0, 18.0))
    u = [-aws * math.cos(awa), -aws * math.sin(awa), 0.0]
    chord = rotate_around_axis([c * u[0] / aws, c * u[1] / aws, 0.0], -8.0 * DEG, [0.0, 0.0, 1.0])
    span = ar * c
    wings = []
    for k in range(w):
        x = (k - 0.5 * (w - 1)) * 2.5 * c
        wings.append({
            "section_points": [_v(x, 0, 0), _v(x, 0, span)],
            "chord_vectors": [_v(*chord), _v(*chord)],
            "section_model": foil(**SAIL_FOIL),
        })
    setup = {
        "line_force_model": {"wing_builders": wings, "nr_sections": strips},
        "simulation_settings": quasi_steady_settings(solver=solver),
    }
    case = {"velocity": u, "time_step": 0.25 * c / aws, "wings": w, "strips": strips}
    return json.dumps(setup), case


def uniform_freestream(n_points: int, velocity) -> np.ndarray:
    out = np.empty((n_points, 3))
    out[:] = np.asarray(velocity, dtype=np.float64)
    return out
