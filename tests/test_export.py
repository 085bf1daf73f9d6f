"""Wake export in the reference's formats (wake/dynamic_wake/export.rs:16-210).

CPU: the library's number formatting against Rust's `{}` output for f64 (values
from Rust's documented behaviour: shortest round-trip digits in positional
notation, "-0", "NaN", "inf") and against an independent Python restatement on
many random values.  GPU: OBJ and VTK files written by the library after a few
steps equal, byte for byte, the texts restated here from the oracle's wake (same
bits as the GPU wake), and the per-step files of `write_wake_data_to_file`.
"""
import math
import os
from decimal import Decimal

import numpy as np
import pytest

import stormbird_amd as sb
from stormbird_amd import setups as S


def rust_display(v: float) -> str:
    """Rust `{}` for f64, restated from Python's shortest repr digits."""
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    neg = math.copysign(1.0, v) < 0
    d = Decimal(repr(abs(float(v))))
    t = format(d, "f")
    if "." in t:
        t = t.rstrip("0").rstrip(".")
    return ("-" if neg else "") + t


@pytest.mark.parametrize("value,text", [
    (1.0, "1"), (0.1, "0.1"), (1e21, "1000000000000000000000"), (1e-7, "0.0000001"), (-0.0, "-0"),
    (123.456, "123.456"), (100.0, "100"), (-2.5, "-2.5"), (float("nan"), "NaN"), (float("inf"), "inf"),
    (float("-inf"), "-inf"), (1.7976931348623157e308, "17976931348623157" + "0" * 292),
    (5e-324, "0." + "0" * 323 + "5"),
])
def test_number_format_known_values(value, text):
    assert sb.format_f64(value) == text


def test_number_format_matches_restatement():
    rng = np.random.default_rng(4)
    vals = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-30, 30, 2000),
                           rng.uniform(-100, 100, 2000), np.round(rng.uniform(-1e6, 1e6, 500), 3)])
    for v in vals:
        assert sb.format_f64(float(v)) == rust_display(float(v)), repr(v)


def expected_obj(points, idx):
    lines = ["o wake"]
    lines += [f"v {rust_display(p[0])} {rust_display(p[1])} {rust_display(p[2])}" for p in points]
    lines += [f"f {a + 1} {b + 1} {c + 1} {d + 1}" for a, b, c, d in idx]
    return "\n".join(lines) + "\n"


def expected_vtk(points, idx, strength, core):
    o = ['<?xml version="1.0"?>', '<VTKFile type="PolyData" version="0.1" byte_order="LittleEndian">',
         "\t<PolyData>",
         f'\t\t<Piece NumberOfPoints="{len(points)}" NumberOfVerts="0" NumberOfLines="0" NumberOfStrips="0" '
         f'NumberOfPolys="{len(strength)}">',
         "\t\t\t<Points>", '\t\t\t\t<DataArray type="Float32" NumberOfComponents="3" format="ascii">']
    o += [f"\t\t\t\t\t{rust_display(p[0])} {rust_display(p[1])} {rust_display(p[2])}" for p in points]
    o += ["\t\t\t\t</DataArray>", "\t\t\t</Points>", "\t\t\t<Polys>",
          '\t\t\t\t<DataArray type="Int32" Name="connectivity" format="ascii">']
    o += [f"\t\t\t\t\t{a} {b} {c} {d}" for a, b, c, d in idx]
    o += ["\t\t\t\t</DataArray>", '\t\t\t\t<DataArray type="Int32" Name="offsets" format="ascii">']
    o += [f"\t\t\t\t\t{(k + 1) * 4}" for k in range(len(strength))]
    o += ["\t\t\t\t</DataArray>", "\t\t\t</Polys>", '\t\t\t<CellData Scalars="strength">',
          '\t\t\t\t<DataArray type="Float32" Name="strength" format="ascii">']
    o += [f"\t\t\t\t\t{rust_display(g)}" for g in strength]
    o += ["\t\t\t\t</DataArray>", '\t\t\t\t<DataArray type="Float32" Name="viscous_core_length" format="ascii">']
    o += [f"\t\t\t\t\t{rust_display(c)}" for c in core]
    o += ["\t\t\t\t</DataArray>", "\t\t\t</CellData>", "\t\t</Piece>", "\t</PolyData>", "</VTKFile>"]
    return "\n".join(o) + "\n"


def reference_indices(n_strips_per_wing, M):
    """panel_point_indices (dynamic_wake/mod.rs:82-91) for wings of the given strip counts."""
    N = sum(n_strips_per_wing)
    P = N + len(n_strips_per_wing)
    wing = np.concatenate([[w] * n for w, n in enumerate(n_strips_per_wing)])
    out = []
    for k in range(N * M):
        s, i = divmod(k, N)
        a = s * P + i + wing[i]
        out.append((a, a + 1, a + P + 1, a + P))
    return out


@pytest.mark.gpu
def test_wake_files_match_reference_format(oracle, tmp_path):
    js, case = S.four_sail_free_wake(nr_panels=8, nr_strips=5)
    d = __import__("json").loads(js)
    d["simulation_settings"]["Dynamic"]["wake"]["write_wake_data_to_file"] = True
    d["simulation_settings"]["Dynamic"]["wake"]["wake_files_folder_path"] = str(tmp_path)
    js = __import__("json").dumps(d)
    g = sb.Simulation(js)
    o = oracle.Simulation(js, variant="pm")
    fs = S.uniform_freestream(g.nr_freestream_points(), case["velocity"])
    dt = case["time_step"]
    for k in range(5):
        g.do_step(time=k * dt, time_step=dt, freestream_velocity=fs)
        o.do_step(time=k * dt, time_step=dt, freestream_velocity=fs)
    st = o.get_wake_state()
    core = o.get_panels()[:, 19]
    idx = reference_indices([5, 5, 5, 5], 8)
    assert g.panel_point_indices().tolist() == [list(t) for t in idx]
    g.write_wake_to_obj_file(tmp_path / "w.obj")
    g.write_wake_to_vtk_file(tmp_path / "w.vtp")
    assert (tmp_path / "w.obj").read_text() == expected_obj(st["points"], idx)
    vtk = expected_vtk(st["points"], idx, st["strengths"], core)
    assert (tmp_path / "w.vtp").read_text() == vtk
    # per-step files: wake_{(time / time_step) as usize}.vtp, the last equal to the direct export
    names = sorted(os.listdir(tmp_path))
    for k in range(5):
        assert f"wake_{int(k * dt / dt)}.vtp" in names
    assert (tmp_path / f"wake_{int(4 * dt / dt)}.vtp").read_text() == vtk
    mesh = g.export_to_plotly_mesh()
    assert len(mesh["i"]) == 2 * len(idx) and mesh["j"][1] == float(idx[0][2]) and mesh["k"][1] == float(idx[0][3])
