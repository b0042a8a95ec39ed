"""Host steps of the three-tissue response estimate (src/larmor/sd/tissue_responses.py) without
a GPU: the signal decay metric, the tensor FA and direction, the automatic threshold and its
golden-section search, the selections and the response rows."""

import numpy as np
import pytest

from larmor_amd.sd import tissue_responses as tr
from larmor_amd.sd.sphere import get_dirs
from larmor_amd.types import GradientTable


def _gtab():
    dirs = get_dirs(30)
    bvals = np.concatenate([np.zeros(4), np.repeat([1000.0, 2000.0], 30)])
    return GradientTable(bvals, np.vstack([np.zeros((4, 3)), dirs, dirs]))


def test_signal_decay_metric_is_the_mean_log_decay():
    gtab = _gtab()
    rows = np.where(gtab.b0_mask, 0, gtab.shell_index + 1)
    d = np.array([0.7e-3, 1.5e-3, 3e-3])
    signal = 1000.0 * np.exp(-np.outer(d, gtab.bvals))
    sdm, valid = tr._signal_decay_metric(signal, rows, 3)
    assert valid.all()
    # log(S0 / S_b) = b D, averaged over the shells by their 30 volumes each
    np.testing.assert_allclose(sdm, (1000.0 + 2000.0) / 2 * d, rtol=1e-6)
    bad = signal.copy()
    bad[0, rows == 2] = 0.0
    assert not tr._signal_decay_metric(bad, rows, 3)[1][0]
    assert tr._signal_decay_metric(signal * np.exp(50.0 * gtab.bvals / 1000.0)[None] ** 0, rows,
                                   3)[0].max() <= 10.0


def test_tensor_fa_and_direction_of_a_known_tensor():
    gtab = _gtab()
    rng = np.random.default_rng(0)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    evals = np.array([1.7e-3, 0.3e-3, 0.2e-3])
    tensor = q @ np.diag(evals) @ q.T
    g, b = gtab.bvecs, gtab.bvals
    signal = 1000.0 * np.exp(-b * np.einsum("mi,ij,mj->m", g, tensor, g))
    fa, direction = tr._tensor_fa_and_direction(signal[None], gtab)
    mean = evals.mean()
    want = np.sqrt(1.5 * ((evals - mean) ** 2).sum() / (evals ** 2).sum())
    assert fa[0] == pytest.approx(want, rel=1e-5)
    assert abs(direction[0] @ q[:, 0]) == pytest.approx(1.0, abs=1e-9)


def test_golden_section_and_threshold_split_two_modes():
    assert tr._golden_section_minimum(lambda x: (x - 0.3) ** 2, 0.0, 0.5, 1.0,
                                      tol=1e-8) == pytest.approx(0.3, abs=1e-6)
    rng = np.random.default_rng(1)
    values = np.concatenate([rng.normal(1.0, 0.3, 500), rng.normal(3.0, 0.3, 300)])
    threshold = tr._optimal_threshold(values)
    cost = tr._ThresholdCost(values.astype(np.float32))
    grid = np.linspace(values.min() + 0.01, values.max() - 0.01, 4001)
    best = min(cost(t) for t in grid)
    assert cost(threshold) <= best + 2e-3  # the search finds the best correlation
    assert 1.5 < threshold < 2.5  # between the two modes
    # the cost is minus the correlation of the values with the mask above the threshold
    mask = (values > 2.0).astype(float)
    assert cost(2.0) == pytest.approx(-np.corrcoef(values, mask)[0, 1], rel=1e-5)


def test_top_bottom_and_their_errors():
    v = np.array([0.0, 3.0, 1.0, 2.0, np.nan, 2.0])
    np.testing.assert_array_equal(tr._top(v, 2), [False, True, False, True, False, True])
    np.testing.assert_array_equal(tr._bottom(v, 1), [True, False, True, False, False, False])
    with pytest.raises(ValueError, match="too few voxels"):
        tr._top(v, 0)
    with pytest.raises(ValueError, match="Cannot select 9 of 4"):
        tr._bottom(v, 9)


def test_isotropic_response_and_shell_rows():
    gtab = _gtab()
    rows = np.where(gtab.b0_mask, 0, gtab.shell_index + 1)
    signal = np.tile(np.where(gtab.b0_mask, 100.0, 50.0), (3, 1))
    resp = tr._isotropic_response(signal, rows, 3)
    np.testing.assert_allclose(resp[:, 0], np.sqrt(4 * np.pi) * np.array([100.0, 50.0, 50.0]))
    assert tr.response_shells(gtab).tolist() == [0.0, 1000.0, 2000.0]
    assert tr.build_shell_response(np.ones((3, 1)), gtab).bvals.tolist() == [0.0, 1000.0, 2000.0]
    with pytest.raises(ValueError, match="give one row per shell"):
        tr.build_shell_response(np.ones((2, 1)), gtab)
