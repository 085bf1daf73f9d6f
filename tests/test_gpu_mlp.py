"""GPU: the warm-start network's bf16 matrix-core kernels and what the warm start gains.

- forward outputs and every parameter gradient against the numpy float64
  restatement with the same bf16 rounding points (tests/mlp_reference.py);
  the same comparison flags a restatement with a planted mistake (a transposed
  first-layer weight);
- training with the GPU's own forward/backward/Adam reduces the loss;
- on a small polar sweep, a network trained on first-step solves lowers the
  fixed-point iterations to a 1e-6 relative tolerance, and the solve still
  ends at the same circulation;
- a first-step solve does not depend on the wake length (the training set's
  instances carry 8 wake rows).
"""
import math

import numpy as np
import pytest

import stormbird_amd as sb
from stormbird_amd import setups as S

import mlp_reference as R
from test_gpu_parity import compare_results

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to("cuda")


def test_mlp_forward_and_gradients_match_restatement():
    rng = np.random.default_rng(5)
    n = 777
    X = rng.standard_normal((n, 64)).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    m = sb.Mlp(seed=3)
    p = m.get_params()
    xd, td = dev(X), dev(t)
    yd = torch.zeros(n, dtype=torch.float32, device="cuda")
    m.forward(xd.data_ptr(), n, yd.data_ptr())
    torch.cuda.synchronize()
    y = yd.cpu().numpy().astype(np.float64)
    loss_ref, g_ref, y_ref = R.gradients(p, X, t)
    assert np.max(np.abs(y - y_ref)) <= 1e-2 * np.max(np.abs(y_ref))
    g, loss = m.gradients(xd.data_ptr(), td.data_ptr(), n)
    assert abs(loss - loss_ref) <= 1e-3 * loss_ref
    for l, (ws, bs) in enumerate(R.layer_slices()):
        for sl, what in ((ws, "W"), (bs, "b")):
            err = np.linalg.norm(g[sl] - g_ref[sl]) / max(np.linalg.norm(g_ref[sl]), 1e-30)
            assert err <= 1e-2, f"layer {l} {what}: relative gradient error {err:.3e}"
    # the comparison detects a planted mistake in the restatement
    _, g_bad, _ = R.gradients(p, X, t, transpose_first_weight=True)
    ws0 = R.layer_slices()[0][0]
    assert np.linalg.norm(g[ws0] - g_bad[ws0]) / np.linalg.norm(g_bad[ws0]) > 0.1


def test_mlp_training_reduces_loss():
    rng = np.random.default_rng(9)
    n = 4096
    X = rng.uniform(-1, 1, (n, 64)).astype(np.float32)
    t = (np.sin(2 * X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * X[:, 3]).astype(np.float32)
    m = sb.Mlp(seed=4)
    xd, td = dev(X), dev(t)
    first = m.train_step(xd.data_ptr(), td.data_ptr(), n, 1e-3)
    for _ in range(300):
        last = m.train_step(xd.data_ptr(), td.data_ptr(), n, 1e-3)
    assert last < 0.1 * first, (first, last)


def _sweep(n_inst, seed):
    rng = np.random.default_rng(seed)
    made = [S.polar_sweep_instance(rng, max_panels=12, min_panels=6) for _ in range(n_inst)]
    return [m[0] for m in made], [m[1] for m in made]


def _conditions(cases, k, rng):
    """Per instance: freestream rotated by up to +-6 deg and scaled by 0.8-1.2."""
    out = []
    for c in cases:
        u = np.asarray(c["velocity"])
        a = rng.uniform(-6, 6) * math.pi / 180
        s = rng.uniform(0.8, 1.2)
        out.append(s * np.array([u[0] * math.cos(a) - u[1] * math.sin(a), u[0] * math.sin(a) + u[1] * math.cos(a), 0]))
    return out


def _first_step(batch, cases, vel, dts):
    batch.reset()
    fs = np.concatenate([S.uniform_freestream(s.nr_freestream_points(), v) for s, v in zip(batch.instances, vel)])
    return batch.do_step(times=None, time_steps=dts, freestream_velocity=fs)


def test_warm_start_lowers_iterations():
    setups, cases = _sweep(24, 31)
    batch = sb.Batch(setups)
    dts = np.array([c["time_step"] for c in cases])
    for s in batch.instances:
        s.set_solver_settings(max_iterations=3000, damping=0.05, residual_tolerance=0.0, strength_tolerance=1e-11)
    batch.set_record_features(True)
    rng = np.random.default_rng(2)
    Xs, Ts = [], []
    for _ in range(24):
        vel = _conditions(cases, 0, rng)
        res = _first_step(batch, cases, vel, dts)
        x, sc = batch.get_features()
        g = np.concatenate([r.force_input.circulation_strength for r in res])
        Xs.append(x)
        Ts.append((g / sc).astype(np.float32))
    X, T = np.concatenate(Xs), np.concatenate(Ts)
    m = sb.Mlp(seed=7)
    xd, td = dev(X), dev(T)
    for it in range(1500):
        m.train_step(xd.data_ptr(), td.data_ptr(), X.shape[0], 1e-3 if it < 1000 else 2e-4)
    # held-out conditions: tight reference, then cold and warm to 1e-6 of max |Gamma|
    vel = _conditions(cases, 0, np.random.default_rng(99))
    ref = _first_step(batch, cases, vel, dts)
    gref = [r.force_input.circulation_strength.copy() for r in ref]
    for s, g in zip(batch.instances, gref):
        s.set_solver_settings(max_iterations=3000, damping=0.05, residual_tolerance=0.0,
                              strength_tolerance=1e-6 * np.max(np.abs(g)))
    batch.set_warm_start(None, 0)
    cold = _first_step(batch, cases, vel, dts)
    batch.set_warm_start(m, 1)
    warm = _first_step(batch, cases, vel, dts)
    it_cold = np.array([r.iterations for r in cold])
    it_warm = np.array([r.iterations for r in warm])
    print(f"iterations cold {it_cold.mean():.1f} warm {it_warm.mean():.1f} fallbacks {batch.last_fallbacks()}")
    assert batch.last_fallbacks() == 0
    assert it_warm.mean() < 0.9 * it_cold.mean()
    for rc, rw, g in zip(cold, warm, gref):
        scale = np.max(np.abs(g))
        assert np.max(np.abs(rw.force_input.circulation_strength - g)) <= 1e-4 * scale
        assert np.max(np.abs(rc.force_input.circulation_strength - g)) <= 1e-4 * scale


def test_first_step_does_not_depend_on_wake_length():
    """The training set is made of first-step solves of instances that carry 8 wake rows:
    the first step sheds one row, so its circulation, forces and features equal those of the
    same geometry with its full wake length, bit for bit."""
    import json

    rng = np.random.default_rng(2)
    made = [S.polar_sweep_instance(rng) for _ in range(6)]
    long_setups = [m[0] for m in made]
    short_setups = []
    for js in long_setups:
        d = json.loads(js)
        d["simulation_settings"]["Dynamic"]["wake"]["nr_panels_per_line_element"] = 8
        short_setups.append(json.dumps(d))
    cases = [m[1] for m in made]
    dts = np.array([c["time_step"] for c in cases])
    out = []
    for setups in (long_setups, short_setups):
        b = sb.Batch(setups)
        b.set_record_features(True)
        fs = np.repeat(np.array([c["velocity"] for c in cases]), [s.nr_freestream_points() for s in b.instances],
                       axis=0)
        res = b.do_step(times=None, time_steps=dts, freestream_velocity=fs)
        out.append((res, b.get_features()))
    (rl, (xl, sl)), (rs, (xs, ss)) = out
    for i in range(6):
        compare_results(rs[i], rl[i], exact=True, where=f"instance {i}")
    assert np.array_equal(xl, xs) and np.array_equal(sl, ss)


def test_fused_inference_equals_layer_by_layer_forward():
    """Inference runs the whole network in one kernel; training runs it layer by layer,
    keeping what the backward pass needs.  Both use the same instruction in the same K
    order and the same rounding points: the training loss (k_head's squared errors summed
    by k_reduce_sq: 256 strided partial sums, then those in order) equals the same sum
    restated in float32 over the fused outputs, bit for bit."""
    rng = np.random.default_rng(11)
    n = 1000
    X = rng.standard_normal((n, 64)).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    m = sb.Mlp(seed=3)
    xd, td = dev(X), dev(t)
    yd = torch.zeros(n, dtype=torch.float32, device="cuda")
    m.forward(xd.data_ptr(), n, yd.data_ptr())
    torch.cuda.synchronize()
    y = yd.cpu().numpy()
    _, loss = m.gradients(xd.data_ptr(), td.data_ptr(), n)
    e = (y - t).astype(np.float32)
    sq = (e * e).astype(np.float32)
    total = np.float32(0.0)
    for lane in range(256):
        s = np.float32(0.0)
        for i in range(lane, n, 256):
            s = np.float32(s + sq[i])
        total = np.float32(total + s)
    assert np.float32(total / np.float32(n)) == np.float32(loss)
