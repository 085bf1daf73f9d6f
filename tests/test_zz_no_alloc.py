"""sb_wake_induced_velocities_noalloc: the allocation-free standalone Biot-Savart sum
against the allocating entry point sb_wake_induced_velocities, bit for bit, on two
calls in a row (the second smaller), through the C ABI."""
import ctypes

import numpy as np
import pytest

import stormbird_amd as sb
from stormbird_amd import setups as S

pytestmark = pytest.mark.gpu


def _records():
    js, case = S.four_sail_free_wake(nr_panels=60, nr_strips=20)
    sim = sb.Simulation(js)
    fs = S.uniform_freestream(sim.nr_freestream_points(), case["velocity"])
    for k in range(6):
        sim.do_step(time=k * case["time_step"], time_step=case["time_step"], freestream_velocity=fs)
    return sim.get_panels()


def test_noalloc_matches_allocating_entry():
    import torch

    rec = _records()
    assert rec.shape[0] == 4800 and np.count_nonzero(rec[:, 21]) > 0
    rng = np.random.default_rng(17)
    targets = rng.uniform(-60, 60, (6000, 3))
    L = sb.lib()
    stream = torch.cuda.current_stream()
    d_rec = torch.from_numpy(np.ascontiguousarray(rec)).to("cuda")
    d_t = torch.from_numpy(targets).to("cuda")
    d_out = torch.full((6000, 3), float("nan"), dtype=torch.float64, device="cuda")
    for symmetry, mode in ((0, sb.MODE_EXACT), (3, sb.MODE_FAST)):
        # first call: 4800 panels, 6000 targets (the 256-target block kernel); second call:
        # 1000 panels, 500 targets (the few-target kernel), reusing the scratch of the first
        for n_p, n_t in ((4800, 6000), (1000, 500)):
            print(f"noalloc: symmetry {symmetry} mode {mode} panels {n_p} targets {n_t}", flush=True)
            rc = L.sb_wake_induced_velocities_noalloc(d_rec.data_ptr(), n_p, d_t.data_ptr(), n_t, symmetry, mode,
                                                      d_out.data_ptr(), ctypes.c_void_p(stream.cuda_stream))
            assert rc == 0, L.sb_last_error().decode()
            torch.cuda.synchronize()
            got = d_out[:n_t].cpu().numpy()
            want = sb.wake_induced_velocities(rec[:n_p], targets[:n_t], symmetry=symmetry, mode=mode)
            assert np.array_equal(got, want)
            assert np.all(np.isfinite(got)) and np.max(np.abs(got)) > 0.0


def test_noalloc_limits_and_empty_inputs():
    import torch

    L = sb.lib()
    cap = L.sb_noalloc_max_panels()
    assert cap > 100_000
    d = torch.zeros(22 * 8, dtype=torch.float64, device="cuda")
    t = torch.zeros(3, dtype=torch.float64, device="cuda")
    out = torch.full((1, 3), 7.0, dtype=torch.float64, device="cuda")
    # beyond the scratch: refused before any device work
    assert L.sb_wake_induced_velocities_noalloc(d.data_ptr(), cap + 1, t.data_ptr(), 1, 0, 0, out.data_ptr(),
                                                None) == -4
    assert "limit" in L.sb_last_error().decode()
    # no panels: zero velocity
    assert L.sb_wake_induced_velocities_noalloc(d.data_ptr(), 0, t.data_ptr(), 1, 0, 0, out.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), np.zeros((1, 3)))
    # no targets: nothing to do; null records with panels: -1; FP32-Kahan mode: -4
    assert L.sb_wake_induced_velocities_noalloc(d.data_ptr(), 8, None, 0, 0, 0, None, None) == 0
    assert L.sb_wake_induced_velocities_noalloc(None, 8, t.data_ptr(), 1, 0, 0, out.data_ptr(), None) == -1
    assert L.sb_wake_induced_velocities_noalloc(d.data_ptr(), 8, t.data_ptr(), 1, 0, 2, out.data_ptr(), None) == -4
