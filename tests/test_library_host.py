"""Host-side checks of libstormbird_amd.so that need no GPU.

* the library loads and exports every function include/stormbird_gpu.h declares;
* the portable elementary functions (used by host and device code) are within
  1 ulp of the host libm (2 for pow) and bit-identical between the library's
  host code and the portable-libm oracle build;
* malformed and unsupported setups are rejected before any GPU call.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import stormbird_amd as sb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    for h in ("stormbird_gpu.h", "cstormbird_compat.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b((?:sb|complete_sail_model)_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_function():
    names = header_functions()
    assert len(names) >= 20
    assert "complete_sail_model_simulate_steady_state_condition" in names
    lib = ctypes.CDLL(sb.library_path)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def ulps(a, b):
    ia, ib = a.view(np.int64), b.view(np.int64)
    ia = np.where(ia < 0, np.int64(-2 ** 63) - ia, ia)
    ib = np.where(ib < 0, np.int64(-2 ** 63) - ib, ib)
    return np.abs(ia - ib)


CASES = {
    "sin": (0, lambda r, n: (r.uniform(-8, 8, n), None), 1),
    "cos": (1, lambda r, n: (r.uniform(-8, 8, n), None), 1),
    "exp": (2, lambda r, n: (r.uniform(-60, 60, n), None), 1),
    "acos": (3, lambda r, n: (r.uniform(-1, 1, n), None), 1),
    "acos_near_one": (3, lambda r, n: (1 - 10 ** r.uniform(-14, -0.3, n), None), 1),
    "pow": (4, lambda r, n: (r.uniform(0, 1, n), r.uniform(0.2, 4, n)), 2),
    "log": (5, lambda r, n: (np.exp(r.uniform(-60, 60, n)), None), 1),
    "atan": (6, lambda r, n: (np.sinh(r.uniform(-8, 8, n)), None), 2),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_portable_libm_accuracy_and_host_identity(oracle, name):
    fn, gen, max_ulp = CASES[name]
    rng = np.random.default_rng(hash(name) % 2 ** 32)
    x, y = gen(rng, 300_000)
    host = sb.libm_host(fn, x, y)
    glibc = oracle.libm_eval(fn, x, y, variant="glibc")
    pm = oracle.libm_eval(fn, x, y, variant="pm")
    assert np.array_equal(host, pm)
    assert int(ulps(host, glibc).max()) <= max_ulp
    assert np.mean(ulps(host, glibc) == 0) > 0.7


def test_libm_special_values():
    x = np.array([0.0, -0.0, 1e-300, 1.0, -1.0])
    assert np.array_equal(sb.libm_host(0, x[:2]), x[:2])  # sin(+-0) = +-0
    assert np.all(sb.libm_host(1, x[:2]) == 1.0)
    assert sb.libm_host(3, np.array([1.0]))[0] == 0.0
    assert abs(sb.libm_host(3, np.array([-1.0]))[0] - np.pi) < 1e-15
    assert sb.libm_host(2, np.array([0.0]))[0] == 1.0
    assert np.isnan(sb.libm_host(3, np.array([1.5]))[0])
    a = sb.libm_host(6, np.array([0.0, -0.0, np.inf, -np.inf, 1.0]))
    assert a[0] == 0.0 and np.signbit(a[1]) and a[2] == np.pi / 2 and a[3] == -np.pi / 2
    assert a[4] == np.pi / 4


@pytest.mark.parametrize("bad,msg", [
    ("not json", "invalid JSON"),
    ('{"line_force_model": 3}', "expected an object"),
    ('{"line_force_model": {"wing_builders": [], "nr_sections": 4}}', "wing_builders"),
])
def test_malformed_setups_rejected_without_gpu(bad, msg):
    with pytest.raises(ValueError, match=msg):
        sb.Simulation(bad)


def test_unsupported_settings_rejected(oracle):
    import json

    import stormbird_amd.setups as S

    d = json.loads(S.single_wing_sail())
    d["simulation_settings"] = {"Steady": {}}
    with pytest.raises(ValueError, match="unknown SimulationSettings"):
        sb.Simulation(json.dumps(d))
    d["simulation_settings"] = {"QuasiSteady": {"solver": {"QuasiNewton": {}}}}
    with pytest.raises(ValueError, match="unknown Solver"):
        sb.Simulation(json.dumps(d))
    d["simulation_settings"] = {"QuasiSteady": {"solver": {"SimpleIterative": {
        "use_raw_circulation_during_iterations": True}}}}
    with pytest.raises(ValueError, match="unknown field"):
        sb.Simulation(json.dumps(d))
    d = json.loads(S.single_wing_sail())
    d["simulation_settings"]["Dynamic"]["wake"]["bogus"] = 1
    with pytest.raises(ValueError, match="unknown field"):
        sb.Simulation(json.dumps(d))


def test_single_wing_sail_discretisation(oracle):
    """21 cosine-spaced section points with nr_sections = 12 give one strip per segment."""
    import stormbird_amd.setups as S

    sim = oracle.Simulation(S.single_wing_sail())
    assert sim.nr_ctrl_points == 20 and sim.nr_panels == 4000
    ctrl = sim.get_freestream_velocity_points()[:20]
    z = [15.0 * (1 - np.cos(k * np.pi / 20)) for k in range(21)]
    assert np.allclose(ctrl[:, 2], 0.5 * (np.array(z[:-1]) + np.array(z[1:])), atol=1e-12)


C_PROGRAM = r"""
#include <stdio.h>
#include "stormbird_gpu.h"
#include "cstormbird_compat.h"

/* the reference's C interface, as a cstormbird user calls it */
int run(const char* setup) {
    CompleteSailModel* m = complete_sail_model_new(setup);
    if (!m) return -1;
    SingleSailResult out[8];
    WindCondition wind = {8.0, 0.6};
    int32_t n = complete_sail_model_simulate_steady_state_condition(m, wind, 6.0, 0.8, out, 8);
    int32_t sails = complete_sail_model_get_number_of_sails(m);
    complete_sail_model_drop(m);
    return n == sails ? 0 : 1;
}

int main(int argc, char** argv) {
    long info[9];
    /* GPU-free entry points run here; the steady state needs a GPU and is only linked */
    if (argc > 1) return run(argv[1]);
    int rc = sb_csm_setup_info("{}", info, 9);
    printf("%d %s\n", rc, sb_last_error());
    double out[1];
    rc = sb_wind_condition_values(
        "{\"direction_coming_from\": 0.0, \"velocity_variation\": {\"Constant\": 7.5}}", 0, out, 0, 0.0, out);
    double h = 10.0;
    rc |= sb_wind_condition_values(
        "{\"direction_coming_from\": 0.0, \"velocity_variation\": {\"Constant\": 7.5}}", 0, &h, 1, 0.0, out);
    printf("%d %.3f\n", rc, out[0]);
    return 0;
}
"""


def test_c_headers_compile_and_link(tmp_path):
    """A C program written against cstormbird's interface and the sb_* ABI compiles with gcc
    (C99) and links against the library; its GPU-free calls run."""
    import subprocess

    src = tmp_path / "prog.c"
    src.write_text(C_PROGRAM)
    exe = tmp_path / "prog"
    libdir = os.path.dirname(sb.library_path)
    cmd = ["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           "-L", libdir, "-lstormbird_amd", f"-Wl,-rpath,{libdir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    first, second = r.stdout.strip().splitlines()
    assert first.startswith("-4 ") and "lifting_line_simulation" in first  # missing field named
    assert second == "0 7.500"


def test_stored_developed_state_loads():
    """The rolled-up four-sail wake the bench starts from: shapes of the default setup
    (40 strips x 4 sails, 2,000 rows), finite values, a wake convected downstream."""
    from stormbird_amd import setups as S

    js, case = S.four_sail_free_wake()
    info = sb.setup_info(js)
    pts, vel, gam, prev, steps = S.developed_four_sail_state(case)
    assert pts.shape == (info["nr_freestream_points"] - info["nr_ctrl_points"], 3)
    assert gam.shape == (info["nr_panels"],) and prev.shape == (info["nr_ctrl_points"],)
    assert vel.shape == pts.shape and steps == 2000
    assert np.all(np.isfinite(pts)) and np.all(np.isfinite(gam)) and np.all(np.isfinite(prev))
    u = np.asarray(case["velocity"])
    downstream = (pts - pts[:164].mean(axis=0)) @ (u / np.linalg.norm(u))
    assert downstream.max() > 4000.0  # 2,000 steps x 0.208 s x 12 m/s = 5 km of wake


def test_setup_check_applies_the_create_time_limits():
    """sb_setup_info rejects what sb_sim_create would reject: more than 1024 strips."""
    import stormbird_amd.setups as S

    ok = sb.setup_info(S.rectangular_wing(nr_strips=1024, nr_panels=2))
    assert ok["nr_ctrl_points"] == 1024
    with pytest.raises(ValueError, match="1024"):
        sb.setup_info(S.rectangular_wing(nr_strips=1025, nr_panels=2))
