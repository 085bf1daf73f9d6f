"""GPU: the setup strings written by the reference's own setup package (tests/golden/setups/)
run unchanged and equal the oracle bit for bit, moving wings included; omitting every field
whose value is the Rust default changes nothing."""
import glob
import json
import math
import os

import numpy as np
import pytest

import stormbird_amd as sb
from stormbird_amd import setups as S

from test_gpu_parity import compare_results

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "setups")
# SimulationBuilder strings; csm_*.json hold CompleteSailModelBuilder strings
FIXTURES = sorted(p for p in glob.glob(os.path.join(GOLDEN, "*.json")) if not os.path.basename(p).startswith("csm_"))


def heave(t):
    return [0.0, 0.1 * math.sin(2.0 * t), 0.0], [0.0, 0.0, 0.0]


def run(oracle, text, steps, u=(8.0, 0.5, 0.0), dt=0.0125, motion=heave):
    g, o = sb.Simulation(text), oracle.Simulation(text, "pm")
    assert g.nr_freestream_points() == o.nr_freestream_points()
    out = []
    for k in range(steps):
        t = (k + 1) * dt
        if motion:
            tr, rot = motion(t)
            for s in (g, o):
                s.set_translation_and_rotation_with_finite_difference_for_the_velocity(dt, tr, rot)
        fs = S.uniform_freestream(g.nr_freestream_points(), list(u))
        rg = g.do_step(time=t, time_step=dt, freestream_velocity=fs)
        ro = o.do_step(time=t, time_step=dt, freestream_velocity=fs)
        print(f"{os.path.basename(text[:0]) or 'setup'} step {k}", flush=True)
        compare_results(rg, ro, where=f"step {k}")
        out.append(rg.as_arrays())
    return out


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_reference_setup_package_output_runs(oracle, path):
    text = open(path).read()
    steps = 3 if "quasi" in path else 10
    run(oracle, text, steps)


def strip_rust_defaults(d):
    """Deletes every field whose value equals the Rust default (builder/mod.rs:27-45,
    single_wing.rs:39-46, dynamic_wake/builder.rs:59-120, simple_iterative.rs:60-80)."""
    lfm = d["line_force_model"]
    for k, v in (("density", 1.225), ("circulation_correction", "None"), ("output_coordinate_system", "Global"),
                 ("force_calculation_settings", {"include_viscous_lift_in_the_circulation": False}),
                 ("rotation", {"x": 0.0, "y": 0.0, "z": 0.0}), ("translation", {"x": 0.0, "y": 0.0, "z": 0.0})):
        if lfm.get(k) == v:
            del lfm[k]
    if all(a == 0.0 for a in lfm.get("local_wing_angles", [])):
        lfm.pop("local_wing_angles", None)
    for wb in lfm["wing_builders"]:
        for k, v in (("line_segment_is_virtual", None), ("non_zero_circulation_at_ends", [False, False]),
                     ("nr_sections", None), ("input_power_model", "NoPower")):
            if k in wb and wb[k] == v:
                del wb[k]
    dyn = d["simulation_settings"]["Dynamic"]
    w = dyn["wake"]
    for k, v in (("nr_panels_per_line_element", 100), ("viscous_core_length", {"Relative": 0.1}),
                 ("viscous_core_length_evolution", "Constant"), ("symmetry_condition", "NoSymmetry"),
                 ("first_panel_relative_length", 0.75), ("last_panel_relative_length", 0.25),
                 ("first_wake_points_direction", "Chord"), ("ratio_of_wake_affected_by_induced_velocities", 0.0),
                 ("shape_damping_factor", 0.0), ("write_wake_data_to_file", False), ("wake_files_folder_path", "")):
        if w.get(k) == v:
            del w[k]
    s = dyn["solver"]["SimpleIterative"]
    for k, v in (("max_iterations_per_time_step", 40), ("damping_factor", 0.05), ("residual_tolerance_absolute", 1e-4),
                 ("strength_difference_tolerance", 1e-6), ("velocity_corrections", "NoCorrection"),
                 ("start_with_linearized_solution", False), ("correction_factor_initial_solution", 1.0)):
        if s.get(k) == v:
            del s[k]
    return d


def test_rust_defaults_equal_absent_fields(oracle):
    text = open(os.path.join(GOLDEN, "heaving_wing_dynamic.json")).read()
    full = json.loads(text)
    short = strip_rust_defaults(json.loads(text))
    assert len(json.dumps(short)) < 0.6 * len(json.dumps(full))
    a = run(oracle, json.dumps(full), 6)
    b = run(oracle, json.dumps(short), 6)
    for ra, rb in zip(a, b):
        for k in ra:
            assert np.array_equal(ra[k], rb[k]), k
