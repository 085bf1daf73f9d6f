"""GPU parity of the one-workgroup Gaussian elimination (the Linearized solve's
linear system, stormath matrix/linalg.rs:41-104) against the oracle, on given
systems through sb_debug_gauss_device.

The solution must equal the oracle's bit for bit: the same pivot choice (largest
magnitude, the first of equal ones, a NaN at the pivot position kept), the same
a - f * p updates and the same ascending-column back substitution.  Cases: random
systems of every size class of the four register layouts (one to three 64-row
slices of the pivot search, two to four rows and 12-20 columns per thread) and of the
blocked form (16-column panels; partial and several panels, up to 1024 rows), systems with many equal
magnitudes (pivot ties), zero diagonals (row exchanges at every step), NaN and
infinite entries, and singular systems, whose failing step must be the oracle's.
"""
import ctypes
import re

import numpy as np
import pytest

import stormbird_amd as sb

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


def gpu_solve(A, b, variant):
    """variant 1-4: the one-workgroup register layouts; "blocked": the global-memory form."""
    L = sb.lib()
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.full(b.size, 7.0)
    if variant == "blocked":
        L.sb_debug_gauss_blocked_device.argtypes = [_dp, _dp, ctypes.c_size_t, _dp]
        L.sb_debug_gauss_blocked_device.restype = ctypes.c_int
        return L.sb_debug_gauss_blocked_device(_p(A), _p(b), b.size, _p(x)), x
    L.sb_debug_gauss_device.argtypes = [_dp, _dp, ctypes.c_size_t, _dp, ctypes.c_int]
    L.sb_debug_gauss_device.restype = ctypes.c_int
    rc = L.sb_debug_gauss_device(_p(A), _p(b), b.size, _p(x), variant)
    return rc, x


def oracle_solve(oracle, A, b):
    """(0, x), or (k + 1, None) when the oracle stops at the singular pivot of step k."""
    try:
        return 0, oracle.solve_gaussian_elimination(A, b)
    except ValueError as e:
        m = re.search(r"at row (\d+)", str(e))
        assert m, str(e)
        return int(m.group(1)) + 1, None


def check(oracle, A, b, variant, where=""):
    rc, x = gpu_solve(A, b, variant)
    assert rc >= 0, f"variant {variant} rejected n = {b.size} ({rc}) {where}"
    orc, ox = oracle_solve(oracle, A, b)
    assert rc == orc, f"singular step {rc} vs oracle {orc} {where}"
    if orc:
        assert np.all(x == 0.0)
        return
    same = (x.view(np.uint64) == ox.view(np.uint64)) | (np.isnan(x) & np.isnan(ox))
    bad = np.flatnonzero(~same)
    assert bad.size == 0, (f"variant {variant} n {b.size} {where}: {bad.size} entries differ, first {bad[0]}: "
                           f"{x[bad[0]]!r} vs {ox[bad[0]]!r}")


SIZES = {1: [1, 2, 3, 7, 8, 9, 17, 31, 32, 33, 40, 63, 64],
         2: [1, 5, 31, 32, 33, 63, 64, 65, 96, 100, 120, 127],
         3: [2, 33, 64, 97, 120, 128],
         4: [1, 9, 64, 65, 120, 128, 129, 137, 139],
         "blocked": [1, 2, 15, 16, 17, 33, 100, 150, 300, 1024]}


@pytest.mark.parametrize("variant", [1, 2, 3, 4, "blocked"])
def test_random_systems(oracle, variant):
    for n in SIZES[variant]:
        rng = np.random.default_rng((1000 * variant if variant != "blocked" else 5000) + n)
        A = rng.standard_normal((n, n)) + 0.1 * n * np.eye(n)
        b = rng.standard_normal(n)
        print(f"gauss variant {variant} n {n}", flush=True)
        check(oracle, A, b, variant, "random")
        # zero diagonal: an exchange at every step
        A0 = rng.standard_normal((n, n))
        np.fill_diagonal(A0, 0.0)
        if n > 1:
            check(oracle, A0, b, variant, "zero diagonal")


@pytest.mark.parametrize("variant", [2, 4, "blocked"])
def test_pivot_ties(oracle, variant):
    """Small integer entries: many equal magnitudes in the pivot column, so the
    first of them (in the exchanged order) must be taken, as linalg.rs:58-64 does."""
    for n in [4, 13, 64, 70, 120]:
        rng = np.random.default_rng(n)
        for trial in range(3):
            A = rng.integers(-2, 3, size=(n, n)).astype(np.float64)
            A += 3.0 * np.eye(n) * (trial == 2)
            b = rng.integers(-3, 4, size=n).astype(np.float64)
            check(oracle, A, b, variant, f"ties trial {trial}")
            check(oracle, -A, b, variant, f"ties negated {trial}")


@pytest.mark.parametrize("variant", [2, 4, "blocked"])
def test_nan_and_infinity(oracle, variant):
    for n in [3, 20, 70, 120]:
        rng = np.random.default_rng(7 + n)
        A = rng.standard_normal((n, n)) + 0.1 * n * np.eye(n)
        b = rng.standard_normal(n)
        An = A.copy()
        An[n // 2, 0] = np.nan  # a NaN candidate in the first pivot column: never chosen
        check(oracle, An, b, variant, "NaN candidate")
        Ak = A.copy()
        Ak[0, 0] = np.nan  # NaN at the pivot position: kept
        check(oracle, Ak, b, variant, "NaN pivot")
        Ai = A.copy()
        Ai[n - 1, 0] = np.inf  # an infinite pivot
        check(oracle, Ai, b, variant, "infinite pivot")
        bn = b.copy()
        bn[n // 3] = np.nan
        check(oracle, A, bn, variant, "NaN right-hand side")


@pytest.mark.parametrize("variant", [1, 2, 4, "blocked"])
def test_singular_systems_stop_where_the_oracle_does(oracle, variant):
    for n in ([2, 9, 40, 64] if variant == 1 else [2, 9, 65, 120]):
        rng = np.random.default_rng(n)
        A = rng.standard_normal((n, n))
        A[:, n // 2] = 2.0 * A[:, 0]  # dependent column
        check(oracle, A, rng.standard_normal(n), variant, "dependent column")
        Z = rng.standard_normal((n, n))
        Z[:, n - 1] = 0.0
        check(oracle, Z, rng.standard_normal(n), variant, "zero column")
        T = np.eye(n) * 1e-13  # pivot below 1e-12 at step 0
        check(oracle, T, np.ones(n), variant, "tiny pivot")


def test_sizes_beyond_a_variant_are_rejected():
    assert gpu_solve(np.eye(65), np.ones(65), 1)[0] == -1
    assert gpu_solve(np.eye(128), np.ones(128), 2)[0] == -1
    assert gpu_solve(np.eye(129), np.ones(129), 3)[0] == -1
    assert gpu_solve(np.eye(140), np.ones(140), 4)[0] == -1  # shared memory
    assert gpu_solve(np.eye(1025), np.ones(1025), "blocked")[0] == -1
