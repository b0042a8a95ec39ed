"""The CPU oracle's NNLS against closed forms, invariants and independent solvers.

The cases restate the reference's own known-answer tests (cpp/tests/sd/test_nnls.cpp, cited per
test) with its platform-stable LCG, then add what those leave open: scipy's NNLS on the stacked
ridge matrix (as tests/sd/test_l2nnls.py:42-54 does) and brute-force enumeration of active sets.
"""

import itertools

import numpy as np
import pytest
from scipy.optimize import nnls as scipy_nnls

import oracle

EPS = np.finfo(np.float64).eps


class PseudoRandom:
    """The reference tests' LCG: uniform numbers in [0, 1) (test_nnls.cpp:43-50)."""

    def __init__(self, state: int = 7):
        self.state = state

    def next(self) -> float:
        self.state = (self.state * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        return (self.state >> 11) * 2.0**-53


def tall() -> np.ndarray:
    """The full-rank positive (6, 3) design of test_nnls.cpp:19-28."""
    d = np.empty((6, 3))
    for r in range(6):
        for c in range(3):
            d[r, c] = 1.0 + ((r * 5 + c * 3) % 7) / 7.0 + (1.0 if r == c else 0.0)
    return d


def solve(design, data):
    return oracle.nnls(design, data)[0]


def test_identity_clips_negative_entries():  # test_nnls.cpp:54
    x = solve(np.eye(3), np.array([2.0, -1.0, 3.0]))
    np.testing.assert_allclose(x, [2.0, 0.0, 3.0], atol=1e-15)


def test_positive_solution_unchanged():  # test_nnls.cpp:60
    d = tall()
    truth = np.array([0.5, 1.5, 2.0])
    np.testing.assert_allclose(solve(d, d @ truth), truth, atol=1e-12)


def test_kkt_with_a_held_coefficient():  # test_nnls.cpp:66-84
    d = tall()
    data = d[:, 0] - 0.3 * d[:, 2]
    x = solve(d, data)
    assert x[2] == 0.0 and (x >= 0).all()
    g = d.T @ (data - d @ x)
    for c in range(3):
        if x[c] > 0:
            assert abs(g[c]) < 1e-12
        else:
            assert g[c] <= 1e-12


def test_tolerance_above_every_gradient_gives_zero():  # test_nnls.cpp:86
    d = tall()
    x, _ = oracle.nnls(d, d @ np.ones(3), tol=1e6)
    assert (x == 0).all()


def test_zero_column_never_enters():  # test_nnls.cpp:92
    d = tall()
    d[:, 1] = 0.0
    x = solve(d, d @ np.array([1.0, 5.0, 2.0]))
    assert x[1] == 0.0
    assert abs(x[0] - 1.0) <= 1e-12 and abs(x[2] - 2.0) <= 1e-12


def test_twenty_columns_recovered():  # test_nnls.cpp:101
    rows, cols = np.meshgrid(np.arange(60), np.arange(20), indexing="ij")
    d = np.exp(-0.05 * rows * (1.0 + cols / 4.0)) + (rows == cols)
    truth = np.linspace(0.5, 2.0, 20)
    np.testing.assert_allclose(solve(d, d @ truth), truth, atol=1e-10)


@pytest.mark.parametrize("noise", [0.0, 1e-12])
def test_dependent_columns_bounded(noise):  # test_nnls.cpp:116-147
    distinct = tall()
    d = np.empty((6, 4))
    d[:, :3] = distinct
    d[:, 3] = distinct[:, 0] + noise * np.where(np.arange(6) % 2 == 0, 1.0, -1.0)
    data = np.array([3.0, 1.0, 4.0, 1.0, 5.0, 9.0])
    x = solve(d, data)
    ref = solve(distinct, data)
    assert np.isfinite(x).all() and (x >= 0).all()
    assert np.abs(x).max() <= 10.0 * np.abs(ref).max()
    g = d.T @ (data - d @ x)
    for c in range(4):
        assert (abs(g[c]) < 1e-10) if x[c] > 0 else (g[c] <= 1e-10)
    assert np.linalg.norm(d @ x - data) <= np.linalg.norm(distinct @ ref - data) * (1 + 1e-10)


def test_duplicated_column_zero_tolerance():  # test_nnls.cpp:149-177
    rng = PseudoRandom()
    for _ in range(200):
        distinct = np.array([[rng.next() for _ in range(4)] for _ in range(30)])
        d = np.empty((30, 5))
        d[:, :4] = distinct
        d[:, 4] = distinct[:, 1]
        data = d @ np.array([0.5, 1.0, 0.0, 0.3, 1.0])
        data = data + np.array([0.05 * (rng.next() - 0.5) for _ in range(30)])
        x, _ = oracle.nnls(d, data, tol=0.0)
        ref, _ = oracle.nnls(distinct, data, tol=0.0)
        assert np.isfinite(x).all() and (x >= 0).all()
        assert np.linalg.norm(d @ x - data) <= np.linalg.norm(distinct @ ref - data) * (1 + 1e-8)


def test_default_tolerance_exact():  # test_nnls.cpp:179
    d = np.array([[1.0, -4.0], [2.0, 1.0], [-3.0, 0.5]])
    assert oracle.default_gradient_tolerance(d) == 10.0 * EPS * 6.0 * 3.0


def test_step_limit_reported():  # test_nnls.cpp:185-216
    rng = PseudoRandom()
    stopped = 0
    for _ in range(50):
        d = np.array([[rng.next() - 0.5 for _ in range(8)] for _ in range(12)])
        coef = np.array([rng.next() - 0.4 for _ in range(8)])
        data = d @ coef
        tol = oracle.default_gradient_tolerance(d)
        full, conv = oracle.nnls(d, data, tol=tol, max_inner=24)
        assert conv
        np.testing.assert_array_equal(full, oracle.nnls(d, data, tol=tol)[0])
        cut, conv0 = oracle.nnls(d, data, tol=tol, max_inner=0)
        if not conv0:
            stopped += 1
        else:
            np.testing.assert_array_equal(cut, full)
    assert stopped > 0


@pytest.mark.parametrize("ridge", [0.1, 1e-7])
@pytest.mark.parametrize("l1", [0.0, 0.01])
def test_ridge_matches_stacked_nnls(ridge, l1):  # test_nnls.cpp:218-248
    rng = PseudoRandom()
    kernel = np.array([[rng.next() for _ in range(40)] for _ in range(10)])
    assert oracle.ridge_normal_equations(kernel, ridge) == (ridge == 0.1)
    stacked = np.zeros((50, 40))
    stacked[:10] = kernel
    stacked[10:] = ridge * np.eye(40)
    for _ in range(5):
        signal = np.array([2.0 * rng.next() - 0.5 for _ in range(10)])
        data = np.concatenate([signal, np.full(40, -l1 / ridge)])
        got = oracle.ridge_nnls(kernel, signal, ridge, l1=l1, tol=1e-12)
        ref, _ = oracle.nnls(stacked, data, tol=1e-12)
        np.testing.assert_allclose(got, ref, atol=1e-9)


def test_ridge_zero_is_plain_nnls():  # test_nnls.cpp:250-269
    rng = PseudoRandom()
    kernel = np.array([[rng.next() for _ in range(8)] for _ in range(30)])
    assert oracle.ridge_normal_equations(kernel, 0.0)
    for _ in range(5):
        signal = np.array([2.0 * rng.next() - 0.5 for _ in range(30)])
        got = oracle.ridge_nnls(kernel, signal, 0.0, tol=1e-12)
        ref, _ = oracle.nnls(kernel, signal, tol=1e-12)
        np.testing.assert_allclose(got, ref, atol=1e-9)


def test_remove_cholesky_column_keeps_factor():  # test_nnls.cpp:271-292
    rng = PseudoRandom()
    cols = np.array([[rng.next() for _ in range(5)] for _ in range(8)])
    for position in range(5):
        factor = np.linalg.cholesky(cols.T @ cols)
        out = oracle.remove_cholesky_column(factor, position)
        kept = np.delete(cols, position, axis=1)
        lower = np.tril(out[:4, :4])
        assert np.abs(lower @ lower.T - kept.T @ kept).max() <= 1e-12
        assert (np.diag(lower) > 0).all()


def test_l1_without_ridge_raises():  # test_nnls.cpp:294
    rng = PseudoRandom()
    kernel = np.array([[rng.next() for _ in range(40)] for _ in range(10)])
    signal = np.array([2.0 * rng.next() - 0.5 for _ in range(10)])
    with pytest.raises(ValueError, match="l1 > 0 needs ridge > 0"):
        oracle.ridge_nnls(kernel, signal, 0.0, l1=0.5, tol=1e-12)


def test_ridge_objective_matches_scipy():
    """As tests/sd/test_l2nnls.py:42-54: the ridge solution minimises ||[K; rI] x - [s; 0]||
    as well as scipy's NNLS (an independent Lawson-Hanson implementation), to rel 1e-10."""
    rng = np.random.default_rng(11)
    for trial in range(20):
        m, n = 30 + trial, 12 + trial
        kernel = np.exp(-rng.uniform(0, 3, size=(m, n)))
        signal = kernel @ np.maximum(rng.normal(size=n), 0) + 0.01 * rng.normal(size=m)
        ridge = 0.1
        stacked = np.vstack([kernel, ridge * np.eye(n)])
        data = np.concatenate([signal, np.zeros(n)])
        x = oracle.ridge_nnls(kernel, signal, ridge, tol=1e-12)
        ref, _ = scipy_nnls(stacked, data)
        obj = np.linalg.norm(stacked @ x - data)
        obj_ref = np.linalg.norm(stacked @ ref - data)
        assert obj <= obj_ref * (1 + 1e-10)
        assert (x >= 0).all()
        assert np.array_equal(x > 0, ref > 1e-12)


def _brute_force_nnls(design, data):
    """The unique KKT point found by trying every active set."""
    n = design.shape[1]
    for k in range(n + 1):
        for subset in itertools.combinations(range(n), k):
            x = np.zeros(n)
            if subset:
                sub = design[:, subset]
                x[list(subset)] = np.linalg.lstsq(sub, data, rcond=None)[0]
            if (x < -1e-12).any():
                continue
            g = design.T @ (data - design @ x)
            inactive = [c for c in range(n) if c not in subset]
            if all(g[c] <= 1e-10 for c in inactive) and all(x[c] > 0 for c in subset):
                return x
    raise AssertionError("no KKT point")


def test_nnls_matches_brute_force_on_tiny_problems():
    rng = np.random.default_rng(5)
    for _ in range(40):
        n = int(rng.integers(1, 7))
        m = n + int(rng.integers(1, 6))
        design = rng.normal(size=(m, n))
        data = rng.normal(size=m)
        np.testing.assert_allclose(solve(design, data), _brute_force_nnls(design, data), atol=1e-10)


def test_least_squares_drops_dependent_columns():
    rng = np.random.default_rng(3)
    a = rng.normal(size=(20, 3))
    design = np.column_stack([a, a[:, 1]])
    data = rng.normal(size=20)
    x = oracle.least_squares(design, data)
    # One of the duplicate pair gets zero; the fit equals the full-rank one.
    assert (x[1] == 0.0) or (x[3] == 0.0)
    ref = np.linalg.lstsq(a, data, rcond=None)[0]
    np.testing.assert_allclose(design @ x, a @ ref, atol=1e-12)
    full = rng.normal(size=(15, 5))
    np.testing.assert_allclose(oracle.least_squares(full, data[:15]),
                               np.linalg.lstsq(full, data[:15], rcond=None)[0], atol=1e-12)


@pytest.mark.parametrize("m,n", [(610, 44), (1010, 300)])
def test_ridge_objective_matches_scipy_with_many_measurements(m, n):
    """The ridge solve at 610 and 1,010 measurements (where the GPU's 12- and 16-row-slot fits are
    compared with the oracle) minimises ||[K; rI] x - [s; 0]|| as scipy's NNLS does, to rel 1e-10."""
    from scipy.optimize import nnls as scipy_nnls

    rng = np.random.default_rng(m)
    kernel = np.exp(-rng.uniform(0, 3, size=(m, n)))
    signal = kernel @ np.maximum(rng.normal(size=n), 0) + 0.01 * rng.normal(size=m)
    ridge = 0.1
    got = oracle.ridge_nnls(kernel, signal, ridge, tol=1e-12)
    stacked = np.vstack([kernel, ridge * np.eye(n)])
    data = np.concatenate([signal, np.zeros(n)])
    ref, _ = scipy_nnls(stacked, data, maxiter=50 * n)
    obj = lambda x: np.sum((stacked @ x - data) ** 2)
    assert (got >= 0).all()
    assert obj(got) <= obj(ref) * (1 + 1e-10)
