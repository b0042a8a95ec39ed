"""Host side of the CSD response estimate (src/larmor/sd/csd_rf.py) pinned without a GPU: the
search spheres, the integration weights, the first-lobe direction, the constrained response fit
against an independent QP solver and exact recovery, the selection and the setting checks."""

import numpy as np
import pytest
from scipy.optimize import minimize

from larmor_amd.sd import csd_rf
from larmor_amd.sd.sh import build_basis_matrix
from larmor_amd.sd.sphere import get_theta_phi
from larmor_amd.types import GradientTable

RESPONSE = np.array([1800.0, -700.0, 180.0, -35.0, 5.0])  # test_csd.py:35


def test_sphere_is_the_subdivided_icosahedron():
    v, nb = csd_rf._sphere()
    assert v.shape == (2562, 3) and nb.shape == (2562, 6)
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-15)
    t = (1 + np.sqrt(5)) / 2
    np.testing.assert_allclose(v[0], np.array([0.0, 1.0, t]) / np.sqrt(1 + t * t))
    np.testing.assert_allclose(v[4], np.array([t, 0.0, 1.0]) / np.sqrt(1 + t * t))
    degree = (nb != np.arange(2562)[:, None]).sum(axis=1)
    assert (degree == 5).sum() == 12 and (degree == 6).sum() == 2550
    for i in range(0, 2562, 97):
        for j in nb[i]:
            assert i in nb[j]
            assert j == i or v[i] @ v[j] > 0.99


def test_hemisphere_weights_integrate_even_harmonics():
    dirs, neighbours, w = csd_rf._hemisphere()
    assert dirs.shape == (1281, 3) and len(neighbours) == 1281
    assert (dirs @ dirs.T).min() > -1.0 + 1e-9  # no antipodal pair kept
    theta, phi = get_theta_phi(dirs)
    rng = np.random.default_rng(0)
    for lmax in (2, 8, 20):
        c = rng.normal(size=(lmax + 1) * (lmax + 2) // 2)
        integral = w @ (build_basis_matrix(theta, phi, lmax) @ c)
        assert integral == pytest.approx(2 * np.sqrt(np.pi) * c[0], rel=1e-9, abs=1e-9)


def _zonal_fod(axis, lmax=8):
    """SH (tournier) of a sharp zonal lobe pointing along axis."""
    dirs, _, w = csd_rf._hemisphere()
    theta, phi = get_theta_phi(np.vstack([dirs, -dirs]))
    b = build_basis_matrix(theta, phi, lmax)
    amp = np.exp(8.0 * ((np.vstack([dirs, -dirs]) @ axis) ** 2 - 1.0))
    return np.linalg.lstsq(b, amp, rcond=None)[0]


def test_first_lobe_direction_finds_the_larger_fibre():
    rng = np.random.default_rng(1)
    for _ in range(5):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        v = np.cross(u, rng.normal(size=3))
        v /= np.linalg.norm(v)
        fod = _zonal_fod(u) + 0.5 * _zonal_fod(v)
        d = csd_rf._first_lobe_direction(fod, 8)
        assert abs(d @ u) > np.cos(np.radians(1.0))


def test_constrained_least_squares_matches_a_qp_solver():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(40, 4))
    gram = a.T @ a
    projected = a.T @ rng.normal(size=40)
    constraints = rng.normal(size=(12, 4))
    x = csd_rf._constrained_least_squares(gram, projected, constraints)
    res = minimize(lambda y: y @ gram @ y - 2 * projected @ y, np.zeros(4),
                   jac=lambda y: 2 * gram @ y - 2 * projected, method="SLSQP",
                   constraints=[{"type": "ineq", "fun": lambda y: constraints @ y,
                                 "jac": lambda y: constraints}],
                   options={"ftol": 1e-15, "maxiter": 1000})
    assert np.all(constraints @ x >= -1e-9)
    np.testing.assert_allclose(x, res.x, atol=1e-6)
    # an inactive constraint set leaves the unconstrained minimiser
    free = csd_rf._constrained_least_squares(gram, projected, np.zeros((1, 4)))
    np.testing.assert_allclose(free, np.linalg.solve(gram, projected), rtol=1e-10)


def test_zonal_response_is_recovered_from_aligned_signals():
    rng = np.random.default_rng(3)
    bvecs = rng.normal(size=(60, 3))
    bvecs /= np.linalg.norm(bvecs, axis=1, keepdims=True)
    fibres = rng.normal(size=(25, 3))
    fibres /= np.linalg.norm(fibres, axis=1, keepdims=True)
    signals = np.array([csd_rf._zonal_basis(np.abs(bvecs @ f), 8) @ RESPONSE for f in fibres])
    got = csd_rf._fit_zonal_response(signals, bvecs, fibres, 10)
    np.testing.assert_allclose(got[:5], RESPONSE, rtol=1e-8, atol=1e-8 * RESPONSE[0])
    assert np.abs(got[5]) < 1e-8 * RESPONSE[0]


def test_top_keeps_ties_and_ignores_nan():
    cost = np.array([3.0, np.nan, 5.0, 3.0, 1.0])
    np.testing.assert_array_equal(csd_rf._top(cost, 2), [True, False, True, True, False])
    assert not csd_rf._top(np.full(4, np.nan), 2).any()
    assert csd_rf._top(cost, 10).sum() == 4


def test_settings_and_response_rows():
    with pytest.raises(ValueError, match="num_voxels must be at least 1"):
        csd_rf._check_settings(0, 10, 1, 10, 10)
    with pytest.raises(ValueError, match="num_candidates at least num_voxels"):
        csd_rf._check_settings(10, 5, 1, 10, 10)
    with pytest.raises(ValueError, match="max_iter >= 2"):
        csd_rf._check_settings(10, 100, 1, 1, 10)
    with pytest.raises(ValueError, match="lmax even"):
        csd_rf._check_settings(10, 100, 1, 10, 7)
    gtab = GradientTable(np.r_[0.0, [1000.0] * 30, [2000.0] * 30],
                         np.r_[np.zeros((1, 3)), np.tile(np.eye(3), (20, 1))])
    one = csd_rf.build_csd_response(RESPONSE, gtab)
    assert one.bvals.tolist() == [2000.0] and one.coefficients.shape == (1, 5)
    three = csd_rf.build_csd_response(np.ones((3, 5)), gtab)
    assert three.bvals.tolist() == [0.0, 1000.0, 2000.0]
    with pytest.raises(ValueError, match="give one row per shell"):
        csd_rf.build_csd_response(np.ones((2, 5)), gtab)
