"""The CPU oracle's log-signal fit against closed forms, invariants, numpy and MRIToolkit.

The oracle restates the weighted linear least squares of larmor's fit_log_signal
(cpp/src/dwi/log_fit.cpp:28-62, 159-185) on a prepared design.  It is pinned here by:
 - its exponential against numpy's (within 1 ulp; exact at 0, 0 below the FP64 range);
 - the ordinary fit against numpy's least squares, and one weighted refit against numpy's least
   squares of the rows scaled by the weights' square roots (a different algorithm: QR, not the
   normal equations by LU);
 - exact recovery of a noise-free tensor-and-kurtosis signal, the intercept's shift under a
   scaled signal, and the ordinary fit kept when the weights underflow;
 - MRIToolkit's per-voxel weighted fit of the CFIN block (the reference's
   tests/sd/test_rf_kurtosis.py:96-124 bars: eigenvalues within 1e-3 of l1, kurtosis 1e-3).
"""

import numpy as np
import pytest

import oracle
from larmor_amd import _core
from larmor_amd.types import GradientTable

from test_wm_rf_host import fixture, mritoolkit_input, wm_design  # noqa: E402


def _design(seed, n=60, p=7):
    rng = np.random.default_rng(seed)
    x = np.column_stack([np.ones(n), rng.standard_normal((n, p - 1)) * 1e3])
    return x


def test_exp_matches_numpy():
    xs = np.concatenate([-np.logspace(-20, np.log10(745.0), 4000), [-708.4, -744.9, -745.2]])
    got = np.array([oracle.exp_nonpositive(x) for x in xs])
    ref = np.exp(xs)
    normal = ref > np.finfo(np.float64).tiny
    assert (np.abs(got[normal] - ref[normal]) <= np.spacing(ref[normal])).all()
    np.testing.assert_array_equal(got[~normal], ref[~normal])
    assert oracle.exp_nonpositive(0.0) == 1.0
    assert oracle.exp_nonpositive(-746.0) == 0.0 and oracle.exp_nonpositive(-np.inf) == 0.0
    assert np.isnan(oracle.exp_nonpositive(np.nan))


def test_ordinary_fit_is_least_squares():
    x = _design(0)
    rng = np.random.default_rng(1)
    y = rng.standard_normal((5, x.shape[0]))
    params, refits = oracle.fit_log_signal(*_core.log_signal_design(x), y, weighted=False)
    ref = np.linalg.lstsq(x, y.T, rcond=None)[0].T
    np.testing.assert_allclose(params, ref, rtol=1e-11, atol=1e-14)
    assert (refits == 0).all()


def test_one_weighted_refit_is_weighted_least_squares():
    x = _design(2, n=80, p=5) * np.array([1.0, 1e-3, 1e-3, 1e-3, 1e-3])
    rng = np.random.default_rng(3)
    y = x @ rng.standard_normal(5) + 0.1 * rng.standard_normal(80)
    params, refits = oracle.fit_log_signal(*_core.log_signal_design(x), y[None], weighted=True,
                                           max_reweightings=1)
    assert refits[0] == 1
    ordinary = np.linalg.lstsq(x, y, rcond=None)[0]
    pred = x @ ordinary
    root_w = np.exp(pred - pred.max())  # the square roots of exp(2 (pred - max))
    ref = np.linalg.lstsq(x * root_w[:, None], y * root_w, rcond=None)[0]
    np.testing.assert_allclose(params[0], ref, rtol=1e-9)


@pytest.mark.parametrize("with_kurtosis", [False, True])
def test_noise_free_signal_is_recovered(with_kurtosis):
    gtab = GradientTable(*mritoolkit_input(fixture())[1:])
    x = wm_design(gtab, with_kurtosis)
    truth = np.array([np.log(1000.0), 1.7e-3, 1e-4, -2e-4, 0.4e-3, 5e-5, 0.3e-3, 0.8][: x.shape[1]])
    params, refits = oracle.fit_log_signal(*_core.log_signal_design(x), (x @ truth)[None])
    np.testing.assert_allclose(params[0], truth, rtol=1e-8, atol=1e-12)
    assert 0 < refits[0] <= 20


def test_a_scaled_signal_shifts_only_the_intercept():
    gtab = GradientTable(*mritoolkit_input(fixture())[1:])
    x = wm_design(gtab, True)
    rng = np.random.default_rng(4)
    y = x @ np.array([5.0, 1.5e-3, 0, 0, 0.5e-3, 0, 0.4e-3, 0.7]) + 0.02 * rng.standard_normal(len(x))
    design = _core.log_signal_design(x)
    a, _ = oracle.fit_log_signal(*design, y[None])
    b, _ = oracle.fit_log_signal(*design, (y + np.log(7.0))[None])
    np.testing.assert_allclose(b[0, 1:], a[0, 1:], rtol=1e-8, atol=1e-13)
    np.testing.assert_allclose(b[0, 0], a[0, 0] + np.log(7.0), rtol=1e-12)


def test_underflowing_weights_keep_the_ordinary_fit():
    # A straight line falling by 400 per step: every weight but the first is exp(-800 k) = 0,
    # so the weighted normal matrix has rank 1 and the refit is refused.
    b = np.arange(8.0)
    x = np.column_stack([np.ones(8), b])
    y = -400.0 * b
    params, refits = oracle.fit_log_signal(*_core.log_signal_design(x), y[None])
    assert refits[0] == -1
    np.testing.assert_allclose(params[0], [0.0, -400.0], atol=1e-9)


def test_mritoolkit_voxel_fits():  # reference tests/sd/test_rf_kurtosis.py:96-124
    f = fixture()
    volume, bvals, bvecs = mritoolkit_input(f)
    gtab = GradientTable(bvals, bvecs)
    mask = f["mask"] > 0
    s = volume[mask].astype(np.float64)
    usable = (s > 0).all(axis=1)
    np.testing.assert_array_equal(usable, f["voxel_fitted"][mask] > 0)
    params, _ = oracle.fit_log_signal(*_core.log_signal_design(wm_design(gtab, True)),
                                      np.log(s[usable]))
    params = params.T
    tensors = params[[1, 2, 3, 2, 4, 5, 3, 5, 6]].T.reshape(-1, 3, 3)
    e_vals = np.linalg.eigvalsh(tensors)[:, ::-1]
    k = 1e-6 * params[7] / np.mean(e_vals**2, axis=1)
    ref_e = f["voxel_e_vals"][mask][usable]
    ref_k = f["voxel_k"][mask][usable]
    assert (np.abs(e_vals - ref_e).max(axis=1) <= 1e-3 * ref_e[:, 0]).all()
    np.testing.assert_allclose(k, ref_k, rtol=0, atol=1e-3)
