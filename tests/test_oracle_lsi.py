"""The CPU oracle's least squares with inequality constraints (LSI) and the MSMT-CSD problem.

- the reference's known answers (cpp/tests/sd/test_lsi.cpp:31-125): a feasible least-squares
  solution is returned unchanged, identity constraints give the NNLS clip, a half-plane
  constraint projects to the origin, the QR fallback gives the Cholesky path's solution with or
  without a duplicated constraint row (the same LCG data), a rank-deficient design is refused
  unless a ridge is given;
- the MSMT-CSD checks of tests/sd/test_msmt.py: the forward model of a unit peak gives back the
  response, a WM + CSF mixture is recovered exactly, noisy fits keep every amplitude
  non-negative, and the KKT conditions hold (checked with scipy's NNLS, independent of the fit).
"""

import numpy as np
import pytest
from scipy.optimize import nnls

import oracle
from larmor_amd.sd.msmt import (
    NORM_REG,
    constraint_matrix,
    forward_matrix,
    msmt_ridge,
    prepare_msmt,
)
from larmor_amd.sd.rf import ZonalSHRF
from larmor_amd.sd.sh import build_basis_matrix, fit_sh
from larmor_amd.sd.sphere import DIRS_300, get_dirs, get_theta_phi
from larmor_amd.types import GradientTable

SHELLS = np.array([0.0, 1000.0, 2000.0, 3000.0])
FIBRE = np.array([0.6, 0.0, 0.8])


def tall():
    d = np.zeros((6, 3))
    for r in range(6):
        for c in range(3):
            d[r, c] = 1.0 + ((r * 5 + c * 3) % 7) / 7.0 + (1.0 if r == c else 0.0)
    return d


def test_feasible_least_squares_unchanged():  # test_lsi.cpp:31-39
    a = tall()
    truth = np.array([0.5, 1.5, 2.0])
    x, conv = oracle.fit_lsi((a @ truth)[None], a, np.eye(3), 0.0)
    assert conv[0]
    assert np.abs(x[0] - truth).max() < 1e-12


def test_identity_constraints_give_the_nnls_clip():  # test_lsi.cpp:41-47
    x, _ = oracle.fit_lsi(np.array([[2.0, -1.0, 3.0]]), np.eye(3), np.eye(3), 0.0)
    assert np.abs(x[0] - [2.0, 0.0, 3.0]).max() < 1e-12


def test_half_plane_projects_to_the_origin():  # test_lsi.cpp:49-58
    x, _ = oracle.fit_lsi(np.array([[0.5, 0.5]]), np.eye(2), np.array([[-1.0, -1.0]]), 0.0)
    assert np.abs(x[0]).max() < 1e-12


def lcg_problem():
    """test_lsi.cpp:60-82: the same generator, design (12, 4), data (12,), constraints (10, 4)."""
    state = 11
    mask = (1 << 64) - 1

    def nxt():
        nonlocal state
        state = (state * 6364136223846793005 + 1442695040888963407) & mask
        return (state >> 11) * 2.0**-53 - 0.5

    design = np.zeros((12, 4))
    data = np.zeros(12)
    for r in range(12):
        for c in range(4):
            design[r, c] = nxt() + (2.0 if r == c else 0.0)
        data[r] = nxt()
    constraints = np.zeros((10, 4))
    for r in range(10):
        for c in range(4):
            constraints[r, c] = nxt()
    return design, data, constraints


def test_qr_fallback_gives_the_cholesky_solution():  # test_lsi.cpp:60-105
    design, data, constraints = lcg_problem()
    duplicated = np.vstack([constraints, constraints[:1]])
    fast, conv = oracle.fit_lsi(data[None], design, constraints, 0.0)
    by_qr, conv_qr = oracle.solve_lsi_by_qr(data, design, constraints, 0.0)
    with_dup, conv_dup = oracle.solve_lsi_by_qr(data, design, duplicated, 0.0)
    assert conv[0] and conv_qr and conv_dup
    assert (constraints @ fast[0]).min() > -1e-12
    unconstrained = np.linalg.lstsq(design, data, rcond=None)[0]
    assert (constraints @ unconstrained).min() < -1e-3
    assert np.abs(fast[0] - by_qr).max() < 1e-12
    assert np.abs(fast[0] - with_dup).max() < 1e-12


def test_rank_deficient_design_refused_unless_ridge():  # test_lsi.cpp:107-113
    design = np.array([[1.0, 2.0], [2.0, 4.0], [3.0, 6.0]])
    with pytest.raises(ValueError, match="full column rank"):
        oracle.fit_lsi(np.zeros((1, 3)), design, np.eye(2), 0.0)
    oracle.fit_lsi(np.zeros((1, 3)), design, np.eye(2), 1e-3)


def test_design_matrices_are_consistent():
    design, data, constraints = lcg_problem()
    q_t, back, e, gram, sums = oracle.lsi_design(design, constraints, 0.3)
    stacked = np.vstack([design, 0.3 * np.eye(4)])
    # x = P R^-1 (z + Q^T [s; 0]) reproduces the ridge least squares at z = 0.
    x_ls = back @ (q_t @ data)
    np.testing.assert_allclose(x_ls, np.linalg.lstsq(stacked, np.r_[data, np.zeros(4)],
                                                     rcond=None)[0], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(e, axis=1), 1.0, rtol=1e-14)
    np.testing.assert_allclose(gram, e @ e.T, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(sums, np.abs(e).sum(axis=1), rtol=1e-14)
    # E is G P R^-1 up to each row's positive scale.
    raw = constraints @ back
    np.testing.assert_allclose(e, raw / np.linalg.norm(raw, axis=1, keepdims=True), rtol=1e-12)


def gradients():
    """test_msmt.py:22-38: two b0s and get_dirs(40) on each of three shells, rotated by 0.4 k."""
    dirs = get_dirs(40)
    bvecs = [np.zeros((2, 3))]
    for k in range(3):
        angle = 0.4 * k
        rot = np.array([[np.cos(angle), -np.sin(angle), 0.0], [np.sin(angle), np.cos(angle), 0.0],
                        [0.0, 0.0, 1.0]])
        bvecs.append(dirs @ rot.T)
    bvals = np.concatenate([[0.0, 0.0], np.repeat(SHELLS[1:], 40)])
    return GradientTable(bvals, np.vstack(bvecs))


def tensor_response(s0, parallel, perpendicular):
    nodes, weights = np.polynomial.legendre.leggauss(64)
    theta = np.arccos(nodes)
    rows = []
    for b in SHELLS:
        profile = s0 * np.exp(-b * (perpendicular + (parallel - perpendicular) * nodes**2))
        basis = build_basis_matrix(theta, np.zeros_like(theta), 8)
        rows.append(2 * np.pi * (weights * profile) @ basis[:, [0, 3, 10, 21, 36]])
    return ZonalSHRF(bvals=SHELLS, coefficients=np.array(rows))


def isotropic_response(s0, diffusivity):
    signal = s0 * np.exp(-SHELLS * diffusivity)
    return ZonalSHRF(bvals=SHELLS, coefficients=np.sqrt(4 * np.pi) * signal[:, None])


def test_forward_model_of_a_unit_peak_gives_back_the_response():  # test_msmt.py:55-66
    gtab = gradients()
    wm = tensor_response(1000.0, 1.7e-3, 0.3e-3)
    forward = forward_matrix([wm], [8], gtab.bvals, gtab.bvecs)
    delta = build_basis_matrix(*get_theta_phi(FIBRE[None]), 8)[0]
    np.testing.assert_allclose(forward @ delta,
                               wm.signal(gtab.bvals, gtab.bvecs, FIBRE[None])[:, 0], rtol=1e-10)


def msmt_problem():
    gtab = gradients()
    rfs = [tensor_response(1000.0, 1.7e-3, 0.2e-3), isotropic_response(2000.0, 3e-3)]
    forward = forward_matrix(rfs, [8, 0], gtab.bvals, gtab.bvecs)
    return gtab, rfs, forward, constraint_matrix([8, 0], DIRS_300), msmt_ridge(forward)


def test_wm_csf_mixture_recovered():  # test_msmt.py:69-88
    gtab, rfs, forward, g, ridge = msmt_problem()
    fod_true = fit_sh(0.7 * 9 / (4 * np.pi) * (DIRS_300 @ FIBRE) ** 8, DIRS_300, 8)
    signal = forward @ np.concatenate([fod_true, [0.3 / np.sqrt(4 * np.pi)]])
    x, conv = oracle.fit_lsi(signal[None], forward, g, ridge)
    assert conv[0]
    np.testing.assert_allclose(x[0, :45], fod_true, atol=1e-6)
    shares = np.maximum(x[0, [0, 45]] * np.sqrt(4 * np.pi), 0)
    np.testing.assert_allclose(shares / shares.sum(), (0.7, 0.3), atol=1e-6)


def test_noisy_fits_are_non_negative_and_meet_kkt():  # test_msmt.py:91-130
    gtab, rfs, forward, g, ridge = msmt_problem()
    rng = np.random.default_rng(1)
    noisy = 300.0 + 200.0 * rng.standard_normal((20, len(gtab.bvals)))
    x, conv, tr = oracle.fit_lsi(noisy, forward, g, ridge, trace=True)
    assert conv.all()
    ridge_squared = NORM_REG * np.max(np.sum(forward**2, axis=0))
    for signal, xv in zip(noisy, x):
        values = g @ xv
        assert values.min() > -1e-9
        active = values < 1e-9 * np.abs(values).max()
        assert active.sum() > 5
        gradient = forward.T @ (forward @ xv - signal) + ridge_squared * xv
        _, residual = nnls(g[active].T, gradient)
        assert residual < 1e-8 * np.linalg.norm(forward.T @ signal)


def test_refuses_more_isotropic_tissues_than_b_values_separate():  # test_msmt.py:133-146
    gtab = gradients()
    one_shell = gtab.bvals <= 1000
    rfs = [tensor_response(1000.0, 1.7e-3, 0.2e-3), isotropic_response(1300.0, 0.8e-3),
           isotropic_response(2000.0, 3e-3)]
    with pytest.raises(ValueError, match="determine only"):
        prepare_msmt(rfs, [8, 0, 0], gtab.bvals[one_shell], gtab.bvecs[one_shell], DIRS_300)


def _wide_msmt_problem(wm_lmax):
    """An MSMT problem with a WM response up to wm_lmax: 67 (lmax 10) or 121 (lmax 14)
    parameters, past the 63 of one lane slot."""
    gtab = gradients()
    nodes, weights = np.polynomial.legendre.leggauss(64)
    theta = np.arccos(nodes)
    basis = build_basis_matrix(theta, np.zeros_like(theta), wm_lmax)
    zonal = [l * (l + 1) // 2 for l in range(0, wm_lmax + 1, 2)]
    rows = [2 * np.pi * (weights * 1000.0 * np.exp(-b * (0.2e-3 + 1.5e-3 * nodes**2))) @ basis[:, zonal]
            for b in SHELLS]
    wm = ZonalSHRF(bvals=SHELLS, coefficients=np.array(rows))
    rfs = [wm, isotropic_response(2000.0, 3e-3)]
    forward = forward_matrix(rfs, [wm_lmax, 0], gtab.bvals, gtab.bvecs)
    return gtab, forward, constraint_matrix([wm_lmax, 0], DIRS_300), msmt_ridge(forward)


@pytest.mark.parametrize("wm_lmax", [10, 14])
def test_wide_problems_are_non_negative_and_meet_kkt(wm_lmax):
    """The KKT conditions of the LSI at 67 and 121 parameters (the GPU's two-slot kernel is compared
    with the oracle there): feasible, and the gradient of the objective a non-negative combination
    of the active constraints' rows."""
    gtab, forward, g, ridge = _wide_msmt_problem(wm_lmax)
    rng = np.random.default_rng(3)
    noisy = 300.0 + 200.0 * rng.standard_normal((6, len(gtab.bvals)))
    x, conv, tr = oracle.fit_lsi(noisy, forward, g, ridge, trace=True)
    assert conv.all() and tr[:, 3].max() > 40
    ridge_squared = NORM_REG * np.max(np.sum(forward**2, axis=0))
    for signal, xv in zip(noisy, x):
        values = g @ xv
        assert values.min() > -1e-7 * np.abs(values).max()
        active = values < 1e-7 * np.abs(values).max()
        gradient = forward.T @ (forward @ xv - signal) + ridge_squared * xv
        _, residual = nnls(g[active].T, gradient)
        assert residual < 1e-7 * np.linalg.norm(forward.T @ signal)
