"""lsqnonneg and ridge_l1_nnls (one problem on one wavefront) against the CPU FP64 oracle.

The device solves follow the oracle's order of operations (wave sums for the generic NNLS's
column sums and default tolerance, sequential sums for the ridge correlations and Gram matrix),
so the bar is identity.  The oracle is pinned by tests/test_oracle_nnls.py (the reference's NNLS
known answers, scipy on the stacked ridge matrix, brute force on tiny problems).  The reference's
own lsqnonneg tests (tests/sd/test_grl.py:211-226) and its stacked-ridge identity
(cpp/tests/sd/test_nnls.cpp:218-248) are restated on the device path.
"""

import numpy as np
import pytest

import oracle
from larmor_amd import _core

from test_oracle_nnls import PseudoRandom  # noqa: E402  (the reference tests' LCG)

pytestmark = pytest.mark.gpu

# (m, n): tall, square, wide, and past each column-slot and row-slot boundary of the kernels.
SHAPES = [(6, 3), (32, 3), (60, 8), (10, 40), (64, 64), (70, 65), (200, 70), (300, 5),
          (257, 30), (120, 330), (512, 90)]


def _problem(seed, m, n, nonneg=True):
    rng = np.random.default_rng(seed)
    c = rng.random((m, n)) if nonneg else rng.standard_normal((m, n))
    truth = np.where(rng.random(n) < 0.3, rng.random(n), 0.0)
    d = c @ truth + 0.05 * rng.standard_normal(m)
    return np.ascontiguousarray(c), d


@pytest.mark.parametrize("m,n", SHAPES)
@pytest.mark.parametrize("nonneg", [True, False])
def test_lsqnonneg_bit_identical(m, n, nonneg):
    print(f"lsqnonneg ({m}, {n}) nonneg={nonneg}", flush=True)
    c, d = _problem(m * 1000 + n, m, n, nonneg)
    for tol in (None, 1e-10, 1e-3):
        ref, converged = oracle.nnls(c, d, tol=tol)
        got, status = _core.lsqnonneg(c, d, tol, return_status=True)
        np.testing.assert_array_equal(got, ref)
        assert (status & 1 == 0) == converged


def test_lsqnonneg_kkt():
    """x >= 0; on the zero coefficients the gradient c^T (d - c x) is at most the tolerance,
    on the positive ones it is ~0 (the optimality conditions of the NNLS)."""
    for seed, (m, n) in enumerate([(40, 12), (200, 70), (30, 60)]):
        c, d = _problem(seed, m, n, nonneg=False)
        x = _core.lsqnonneg(c, d)
        assert (x >= 0).all()
        g = c.T @ (d - c @ x)
        tol = oracle.default_gradient_tolerance(c)
        assert (g[x == 0] <= tol + 1e-9).all()
        np.testing.assert_allclose(g[x > 0], 0.0, atol=1e-8)


def test_lsqnonneg_tol_x_zeroes_a_small_coefficient():  # tests/sd/test_grl.py:211-220
    b = np.repeat([0.0, 1000.0, 2000.0, 3000.0], 8)
    c = np.ascontiguousarray(np.exp(-np.outer(b, [0.7e-3, 3e-3])))
    d = c @ np.array([0.995, 0.005])
    assert (_core.lsqnonneg(c, d) > 0).all()
    x = _core.lsqnonneg(c, d, 1e-2)
    assert x[1] == 0
    np.testing.assert_allclose(x[0], 1.0, atol=1e-2)


def test_lsqnonneg_refuses_a_right_hand_side_of_another_length():  # test_grl.py:223-226
    with pytest.raises(ValueError, match="c must be"):
        _core.lsqnonneg(np.ones((32, 3)), np.ones(31))


def test_lsqnonneg_dtype_and_layout():
    with pytest.raises(TypeError):
        _core.lsqnonneg(np.ones((32, 3), dtype=np.float32), np.ones(32))
    with pytest.raises(ValueError, match="C-contiguous"):
        _core.lsqnonneg(np.asfortranarray(np.ones((32, 3))), np.ones(32))
    with pytest.raises(ValueError, match="exceeds 1024 measurements"):
        _core.lsqnonneg(np.ones((1025, 3)), np.ones(1025))


def test_lsqnonneg_empty_problems():
    assert _core.lsqnonneg(np.zeros((5, 0)), np.ones(5)).shape == (0,)
    np.testing.assert_array_equal(_core.lsqnonneg(np.zeros((0, 4)), np.zeros(0)), np.zeros(4))


def test_lsqnonneg_device_arrays_stay_on_the_device():
    import torch

    c, d = _problem(3, 50, 9)
    x = _core.lsqnonneg(torch.from_numpy(c).cuda(), torch.from_numpy(d).cuda())
    assert x.is_cuda
    np.testing.assert_array_equal(x.cpu().numpy(), oracle.nnls(c, d)[0])


@pytest.mark.parametrize("m,n", SHAPES)
@pytest.mark.parametrize("ridge,l1", [(0.0, 0.0), (0.1, 0.0), (0.1, 0.01), (1e-7, 0.0),
                                      (1e-7, 0.01), (2.0, 0.3)])
def test_ridge_l1_nnls_bit_identical(m, n, ridge, l1):
    print(f"ridge_l1_nnls ({m}, {n}) ridge={ridge} l1={l1}", flush=True)
    c, d = _problem(m * 7 + n, m, n, nonneg=False)
    for tol in (None, 1e-10):
        oracle_tol = oracle.default_gradient_tolerance(c) if tol is None else tol
        ref = oracle.ridge_nnls(c, d, ridge, l1=l1, tol=oracle_tol)
        got = _core.ridge_l1_nnls(c, d, ridge, l1, tol)
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("ridge", [0.1, 1e-7])
@pytest.mark.parametrize("l1", [0.0, 0.01])
def test_ridge_matches_stacked_lsqnonneg(ridge, l1):  # cpp/tests/sd/test_nnls.cpp:218-248
    rng = PseudoRandom()
    kernel = np.array([[rng.next() for _ in range(40)] for _ in range(10)])
    assert oracle.ridge_normal_equations(kernel, ridge) == (ridge == 0.1)
    stacked = np.zeros((50, 40))
    stacked[:10] = kernel
    stacked[10:] = ridge * np.eye(40)
    for _ in range(5):
        signal = np.array([2.0 * rng.next() - 0.5 for _ in range(10)])
        data = np.concatenate([signal, np.full(40, -l1 / ridge)])
        got = _core.ridge_l1_nnls(kernel, signal, ridge, l1, 1e-12)
        ref = _core.lsqnonneg(stacked, data, 1e-12)
        np.testing.assert_allclose(got, ref, atol=1e-9)


def test_ridge_zero_is_plain_nnls():  # cpp/tests/sd/test_nnls.cpp:250-269
    rng = PseudoRandom()
    kernel = np.array([[rng.next() for _ in range(8)] for _ in range(30)])
    for _ in range(5):
        signal = np.array([2.0 * rng.next() - 0.5 for _ in range(30)])
        got = _core.ridge_l1_nnls(kernel, signal, 0.0, 0.0, 1e-12)
        ref = _core.lsqnonneg(kernel, signal, 1e-12)
        np.testing.assert_allclose(got, ref, atol=1e-9)


def test_ridge_l1_nnls_refusals():  # cpp/src/python/sd.cpp:235-249
    c = np.ones((8, 3))
    d = np.ones(8)
    for ridge, l1 in [(-1.0, 0.0), (np.nan, 0.0), (0.1, -1.0), (0.1, np.inf)]:
        with pytest.raises(ValueError, match="ridge and l1 must be finite and >= 0."):
            _core.ridge_l1_nnls(c, d, ridge, l1)
    with pytest.raises(ValueError, match="l1 > 0 needs ridge > 0."):
        _core.ridge_l1_nnls(c, d, 0.0, 0.5)
    with pytest.raises(ValueError, match="c must be"):
        _core.ridge_l1_nnls(c, np.ones(7), 0.1, 0.0)
