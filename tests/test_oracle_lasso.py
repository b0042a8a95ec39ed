"""The elastic-net oracle (FodSolver.lasso, cpp/src/sd/elastic_net.cpp) pinned by the reference's
own checks (tests/sd/test_lasso.py): the optimality conditions on its six fixture problems, the
all-zero solution from lambda_max on, the least-squares limit of a tiny pure-L1 penalty and the
setting checks; plus scikit-learn's ElasticNet, an independent solver of the same objective
(without an intercept, on the standardised columns), and the multi-tissue invariants."""

from pathlib import Path

import numpy as np
import pytest

import oracle
from larmor_amd import _core, synthetic
from larmor_amd.sd.kernel import Kernel
from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
from larmor_amd.sd.solvers import Lasso
from larmor_amd.sd.sphere import DIRS_300, get_dirs

FIXTURE = np.load(Path(__file__).parent / "golden" / "lasso_fixture.npz")
PROBLEMS = [(str(FIXTURE[f"name_{i}"]), FIXTURE[f"X_{i}"].astype(np.float64), FIXTURE[f"y_{i}"],
             FIXTURE[f"lambda_{i}"]) for i in range(int(FIXTURE["n"]))]


def kkt_violation(x, y, b, reg, alpha=0.5):
    """Largest violation of the elastic net's optimality conditions (test_lasso.py:39-56)."""
    m = len(y)
    scales = np.sqrt(np.mean(x**2, axis=0))
    x0 = x / scales
    c = b * scales
    g = x0.T @ (y - x0 @ c) / m - reg * (1 - alpha) * c
    on = c != 0
    violation = np.maximum(np.abs(g) - reg * alpha, 0.0)
    violation[on] = np.abs(g[on] - reg * alpha * np.sign(c[on]))
    return float(violation.max())


@pytest.mark.parametrize(("name", "i"), [(p[0], i) for p in PROBLEMS for i in range(3)])
def test_fixture_problems_meet_the_optimality_conditions(name, i):
    _, x, y, lambdas = next(p for p in PROBLEMS if p[0] == name)
    reg = float(lambdas[i])
    b = oracle.fit_single_tissue_lasso(y[None], x, reg, 0.5, 1e-12, 10**7)[0]
    assert kkt_violation(x, y, b, reg) < 1e-6 * reg


def _lambda_max(x, y):
    x0 = x / np.sqrt(np.mean(x**2, axis=0))
    return float(np.abs(x0.T @ y).max() / (len(y) * 0.5))


@pytest.mark.parametrize(("factor", "all_zero"), [(1.0001, True), (0.99, False)])
def test_zero_from_lambda_max_on(factor, all_zero):
    _, x, y, _ = PROBLEMS[0]
    b = oracle.fit_single_tissue_lasso(y[None], x, factor * _lambda_max(x, y), 0.5, 1e-12, 10**7)
    assert np.all(b == 0) == all_zero


def test_tiny_pure_l1_is_least_squares():
    rng = np.random.default_rng(0)
    x = rng.uniform(0.0, 1.0, (60, 8))
    y = x @ rng.uniform(-1.0, 1.0, 8) + 0.01 * rng.standard_normal(60)
    expected, *_ = np.linalg.lstsq(x, y, rcond=None)
    b = oracle.fit_single_tissue_lasso(y[None], x, 1e-10, 1.0, 1e-14, 10**7)[0]
    np.testing.assert_allclose(b, expected, rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("alpha", [0.9, 0.5, 1.0])
def test_matches_scikit_learn(alpha):
    from sklearn.linear_model import ElasticNet

    rng = np.random.default_rng(3)
    x = rng.uniform(0.0, 1.0, (80, 30))
    y = x @ np.maximum(rng.normal(size=30), 0.0) + 0.05 * rng.standard_normal(80)
    cols, scales = oracle.elastic_net_design(x)
    np.testing.assert_allclose(scales, np.sqrt(np.mean(x**2, axis=0)), rtol=1e-15)
    np.testing.assert_allclose(cols, x / scales, rtol=1e-15)
    for reg in (0.1, 0.01, 0.001):
        b = oracle.fit_single_tissue_lasso(y[None], x, reg, alpha, 1e-13, 10**7)[0]
        ref = ElasticNet(alpha=reg, l1_ratio=alpha, fit_intercept=False, tol=1e-15,
                         max_iter=10**6).fit(cols, y).coef_
        np.testing.assert_allclose(b * scales, ref, atol=1e-8 * max(1.0, np.abs(ref).max()))


def test_zero_column_stays_out_and_limit_is_counted():
    rng = np.random.default_rng(4)
    x = rng.uniform(0.0, 1.0, (40, 12))
    x[:, 5] = 0.0
    y = rng.uniform(0.0, 1.0, (25, 12)) @ x.T
    b, sweeps = oracle.fit_single_tissue_lasso(y, x, 1e-4, 0.9, 1e-12, 10**6, sweeps=True)
    assert np.all(b[:, 5] == 0.0) and np.all(sweeps > 0)
    _, limited = oracle.fit_single_tissue_lasso(y, x, 1e-4, 0.9, 1e-12, 3, sweeps=True)
    assert np.all(limited == -3)


def test_threads_do_not_change_results():
    rng = np.random.default_rng(1)
    x = rng.uniform(0.0, 1.0, (40, 30))
    signals = rng.uniform(0.0, 1.0, (300, 30)) @ x.T
    a = oracle.fit_single_tissue_lasso(signals, x, 0.01, 0.5, 1e-8, 100, n_threads=1)
    b = oracle.fit_single_tissue_lasso(signals, x, 0.01, 0.5, 1e-8, 100, n_threads=3)
    np.testing.assert_array_equal(a, b)


def test_multi_tissue_fractions_are_a_partition():
    ph = synthetic.phantom((5, 5, 4), 10, [(1000.0, 40), (2500.0, 40)], seed=9, snr=25.0,
                           tissues=True)
    signals = synthetic.normalized_signals(ph.data, ph.bvals, ph.mask)[:60]
    rfs = [AnisotropicRF([1.7e-3, 0.3e-3, 0.3e-3], 0.0), IsotropicRF(0.8e-3), IsotropicRF(3e-3)]
    w = np.ones(len(ph.bvals))
    hr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=[get_dirs(45)[:-1]], weights=w)
    fods, fractions, rounds = oracle.fit_multi_tissue_lasso(signals, w, lr, hr, 0.01, 0.9, 1e-4,
                                                            1000, rounds=True)
    assert np.all(fractions >= 0)
    np.testing.assert_allclose(fractions.sum(axis=1), 1.0, atol=1e-12)
    assert np.all(rounds >= 1) and np.all(rounds <= 50)
    # one voxel alone equals the block (no cross-voxel state)
    f1, fr1 = oracle.fit_multi_tissue_lasso(signals[7:8], w, lr, hr, 0.01, 0.9, 1e-4, 1000)
    np.testing.assert_array_equal(f1[0], fods[7])
    np.testing.assert_array_equal(fr1[0], fractions[7])


INVALID = [({"reg": -0.1}, "reg"), ({"reg": float("nan")}, "reg"), ({"alpha": 0.0}, "alpha"),
           ({"alpha": 1.5}, "alpha"), ({"alpha": float("nan")}, "alpha"), ({"tol": 0.0}, "tol"),
           ({"tol": float("nan")}, "tol"), ({"max_iter": 0}, "max_iter")]


@pytest.mark.parametrize(("change", "match"), INVALID)
def test_lasso_refuses_invalid_settings(change, match):
    with pytest.raises(ValueError, match=match):
        Lasso(**change)


@pytest.mark.parametrize(("change", "match"), INVALID)
def test_binding_refuses_invalid_settings(change, match):
    settings = {"reg": 0.1, "alpha": 0.5, "tol": 1e-8, "max_iter": 100} | change
    with pytest.raises(ValueError, match=match):
        _core.FodSolver.lasso(np.ones((4, 3)), **settings)
