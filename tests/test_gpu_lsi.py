"""The LSI kernels against the CPU FP64 oracle on the same inputs: bit-identical solutions and
converged flags.

The host design (pivoted QR of [A; ridge I], E, E E^T) and the device solves use the oracle's
order, so the bar is identity.  The oracle is pinned by tests/test_oracle_lsi.py (the
reference's LSI known answers, KKT via scipy, exact MSMT recovery).
"""

import numpy as np
import pytest

import oracle
from larmor_amd import _core
from larmor_amd.sd.msmt import constraint_matrix, forward_matrix, msmt_ridge
from larmor_amd.sd.msmt_csd import MSMTCSDModel
from larmor_amd.sd.sh import fit_sh
from larmor_amd.sd.sphere import DIRS_300
from larmor_amd.types import Volume

from test_oracle_lsi import (  # noqa: E402  (the reference tests' problem builders)
    FIBRE,
    gradients,
    isotropic_response,
    lcg_problem,
    tall,
    tensor_response,
)

pytestmark = pytest.mark.gpu


def _both(signals, a, g, ridge):
    ref_x, ref_c, tr = oracle.fit_lsi(signals, a, g, ridge, trace=True)
    design = _core.LsiDesign(a, g, ridge)
    x, c, status = _core.fit_lsi(np.ascontiguousarray(signals), design, 1, return_status=True)
    return (ref_x, ref_c, tr), (x, c, status)


def test_small_problems_bit_identical():
    a = tall()
    cases = [
        ((a @ np.array([0.5, 1.5, 2.0]))[None], a, np.eye(3), 0.0),
        (np.array([[2.0, -1.0, 3.0]]), np.eye(3), np.eye(3), 0.0),
        (np.array([[0.5, 0.5]]), np.eye(2), np.array([[-1.0, -1.0]]), 0.0),
    ]
    design, data, constraints = lcg_problem()
    cases.append((data[None], design, constraints, 0.0))
    cases.append((data[None], design, constraints, 0.3))
    for signals, am, gm, ridge in cases:
        (rx, rc, _), (x, c, _) = _both(signals, am, gm, ridge)
        np.testing.assert_array_equal(c, rc)
        np.testing.assert_array_equal(x, rx)


def _msmt(n_vox=400, seed=1, noise=200.0):
    gtab = gradients()
    rfs = [tensor_response(1000.0, 1.7e-3, 0.2e-3), isotropic_response(2000.0, 3e-3)]
    forward = forward_matrix(rfs, [8, 0], gtab.bvals, gtab.bvecs)
    g = constraint_matrix([8, 0], DIRS_300)
    rng = np.random.default_rng(seed)
    signals = 300.0 + noise * rng.standard_normal((n_vox, len(gtab.bvals)))
    return gtab, rfs, forward, g, msmt_ridge(forward), signals


def test_msmt_noisy_voxels_bit_identical():
    _, _, forward, g, ridge, signals = _msmt()
    print("lsi msmt", signals.shape, g.shape, flush=True)
    (rx, rc, tr), (x, c, status) = _both(signals, forward, g, ridge)
    np.testing.assert_array_equal(c, rc)
    np.testing.assert_array_equal(x, rx)
    np.testing.assert_array_equal(status != 0, tr[:, 0] != 0)
    assert (g @ x.T).min() > -1e-9


def test_qr_path_matches_oracle_and_the_cholesky_path():
    design, data, constraints = lcg_problem()
    ref_x, ref_c = oracle.solve_lsi_by_qr(data, design, constraints, 0.0)
    d = _core.LsiDesign(design, constraints, 0.0)
    x, c = _core.fit_lsi(data[None], d, 1, only_qr=True)
    assert c[0] == ref_c
    np.testing.assert_array_equal(x[0], ref_x)
    fast, _ = _core.fit_lsi(data[None], d, 1)
    assert np.abs(fast[0] - x[0]).max() < 1e-12
    dup = np.vstack([constraints, constraints[:1]])
    ref_dup, _ = oracle.solve_lsi_by_qr(data, design, dup, 0.0)
    xd, cd = _core.fit_lsi(data[None], _core.LsiDesign(design, dup, 0.0), 1, only_qr=True)
    np.testing.assert_array_equal(xd[0], ref_dup)
    assert np.abs(xd[0] - fast[0]).max() < 1e-12
    # The QR path over MSMT voxels too.
    _, _, forward, g, ridge, signals = _msmt(n_vox=40, seed=3)
    dm = _core.LsiDesign(forward, g, ridge)
    xq, cq = _core.fit_lsi(signals, dm, 1, only_qr=True)
    for v in range(len(signals)):
        rx, rcv = oracle.solve_lsi_by_qr(signals[v], forward, g, ridge)
        assert cq[v] == rcv
        np.testing.assert_array_equal(xq[v], rx)


def test_edge_cases():
    _, _, forward, g, ridge, signals = _msmt(n_vox=8)
    edge = np.vstack([np.zeros(signals.shape[1]), np.full(signals.shape[1], np.nan), signals])
    (rx, rc, _), (x, c, _) = _both(edge, forward, g, ridge)
    np.testing.assert_array_equal(c, rc)
    np.testing.assert_array_equal(x, rx)
    d = _core.LsiDesign(forward, g, ridge)
    empty, ce = _core.fit_lsi(np.zeros((0, signals.shape[1])), d, 1)
    assert empty.shape == (0, forward.shape[1]) and ce.shape == (0,)
    with pytest.raises(ValueError, match="measurements"):
        _core.fit_lsi(np.zeros((2, 5)), d, 1)
    with pytest.raises(ValueError, match="full column rank"):
        _core.LsiDesign(np.array([[1.0, 2.0], [2.0, 4.0], [3.0, 6.0]]), np.eye(2), 0.0)


def test_batch_invariance():
    _, _, forward, g, ridge, signals = _msmt(n_vox=200, seed=7)
    d = _core.LsiDesign(forward, g, ridge)
    alone, _ = _core.fit_lsi(np.ascontiguousarray(signals[5:7]), d, 1)
    mixed, _ = _core.fit_lsi(signals, d, 1)
    np.testing.assert_array_equal(alone, mixed[5:7])


def test_msmt_model_recovers_wm_csf_mixture():  # tests/sd/test_msmt.py:69-88
    gtab, rfs, forward, _, _, _ = _msmt(n_vox=1)
    fod_true = fit_sh(0.7 * 9 / (4 * np.pi) * (DIRS_300 @ FIBRE) ** 8, DIRS_300, 8)
    signal = forward @ np.concatenate([fod_true, [0.3 / np.sqrt(4 * np.pi)]])
    volume = Volume(np.tile(signal, (2, 1, 1, 1)), np.eye(4))
    fit = MSMTCSDModel([rfs[1], rfs[0]]).fit(volume, gtab, n_jobs=1)
    np.testing.assert_allclose(fit.tissue_fractions[0, 0, 0], (0.7, 0.3), atol=1e-6)
    np.testing.assert_allclose(fit.fods[0][0, 0, 0], fod_true, atol=1e-6)


SHELLS_WIDE = np.array([0.0, 1000.0, 2000.0, 3000.0])


def _wide_msmt(wm_lmax, n_vox, seed=1, n_dirs=90, full=False):
    """MSMT with a WM response up to wm_lmax and a CSF response: 67 (lmax 10) to 232 (lmax 20)
    parameters, on 2 b0 + 3 x n_dirs measurements; signals 300 +- 200 (noise), whose active sets
    reach n - 6 constraints, past 64 and 128 factor positions."""
    from larmor_amd import synthetic
    from larmor_amd.sd.rf import ZonalSHRF
    from larmor_amd.sd.sh import build_basis_matrix

    nodes, weights = np.polynomial.legendre.leggauss(64)
    theta = np.arccos(nodes)
    basis = build_basis_matrix(theta, np.zeros_like(theta), wm_lmax)
    zonal = [l * (l + 1) // 2 for l in range(0, wm_lmax + 1, 2)]
    rows = [2 * np.pi * (weights * 1000.0 * np.exp(-b * (0.2e-3 + 1.5e-3 * nodes**2))) @ basis[:, zonal]
            for b in SHELLS_WIDE]
    wm = ZonalSHRF(bvals=SHELLS_WIDE, coefficients=np.array(rows))
    csf = ZonalSHRF(bvals=SHELLS_WIDE,
                    coefficients=np.sqrt(4 * np.pi) * (2000.0 * np.exp(-SHELLS_WIDE * 3e-3))[:, None])
    dirs = synthetic.spiral_directions(n_dirs)
    bvecs = np.vstack([np.zeros((2, 3))] + [dirs] * 3)
    bvals = np.concatenate([[0.0, 0.0], np.repeat(SHELLS_WIDE[1:], n_dirs)])
    forward = forward_matrix([wm, csf], [wm_lmax, 0], bvals, bvecs)
    g = constraint_matrix([wm_lmax, 0], DIRS_300)
    signals = 300.0 + 200.0 * np.random.default_rng(seed).standard_normal((n_vox, len(bvals)))
    if full:
        return forward, g, msmt_ridge(forward), signals, (wm, csf), (bvals, bvecs)
    return forward, g, msmt_ridge(forward), signals


# 67, 92 and 121 parameters take two lane slots; 154 and 232 take four.
@pytest.mark.parametrize("wm_lmax,n_vox", [(10, 160), (12, 96), (14, 64), (16, 48), (20, 32)])
def test_lsi_wide_msmt_bit_identical(wm_lmax, n_vox):
    forward, g, ridge, signals = _wide_msmt(wm_lmax, n_vox)
    print(f"lsi wide wm lmax {wm_lmax}: {forward.shape[1]} parameters, {signals.shape}", flush=True)
    (rx, rc, tr), (x, c, status) = _both(signals, forward, g, ridge)
    if wm_lmax >= 12:  # the factor outgrows the first (and, from 154 parameters, the second) slot
        assert tr[:, 3].max() > 64 * (1 if forward.shape[1] < 128 else 2)
    np.testing.assert_array_equal(c, rc)
    np.testing.assert_array_equal(status != 0, tr[:, 0] != 0)
    np.testing.assert_array_equal(x, rx)
    # Feasible to the LDP's tolerance, which grows with the constraint count and the scale.
    assert (g @ x.T).min() > -1e-7


@pytest.mark.parametrize("wm_lmax", [10, 16])
def test_lsi_wide_qr_path_matches_oracle(wm_lmax):
    forward, g, ridge, signals = _wide_msmt(wm_lmax, 3, seed=4)
    print(f"lsi wide qr wm lmax {wm_lmax}: {forward.shape[1]} parameters", flush=True)
    d = _core.LsiDesign(forward, g, ridge)
    xq, cq = _core.fit_lsi(signals, d, 1, only_qr=True)
    fast, _ = _core.fit_lsi(signals, d, 1)
    for v in range(len(signals)):
        rx, rcv = oracle.solve_lsi_by_qr(signals[v], forward, g, ridge)
        assert cq[v] == rcv
        np.testing.assert_array_equal(xq[v], rx)
    # The two algorithms agree to the conditioning of a 67- or 154-parameter LDP.
    scale = np.abs(fast).max()
    assert np.abs(xq - fast).max() < 1e-6 * scale


def test_lsi_at_the_parameter_limit():
    # 255 parameters, the most the kernels take (four lane slots), on a random full-rank design
    # with 302 random constraints (x = 0 is always feasible); the factor grows past 128 positions.
    rng = np.random.default_rng(21)
    n, m, k = 255, 300, 302
    a = rng.normal(size=(m, n))
    g = rng.normal(size=(k, n))
    signals = rng.normal(size=(6, m))
    print(f"lsi at the limit: {n} parameters, {k} constraints", flush=True)
    (rx, rc, tr), (x, c, status) = _both(signals, a, g, 0.1)
    assert tr[:, 3].max() > 128
    np.testing.assert_array_equal(c, rc)
    np.testing.assert_array_equal(status != 0, tr[:, 0] != 0)
    np.testing.assert_array_equal(x, rx)


def test_lsi_wide_edge_cases_and_limit():
    forward, g, ridge, signals = _wide_msmt(12, 6, seed=9)
    edge = np.vstack([np.zeros(signals.shape[1]), np.full(signals.shape[1], np.nan), signals])
    (rx, rc, _), (x, c, _) = _both(edge, forward, g, ridge)
    np.testing.assert_array_equal(c, rc)
    np.testing.assert_array_equal(x, rx)
    d = _core.LsiDesign(forward, g, ridge)
    alone, _ = _core.fit_lsi(np.ascontiguousarray(signals[2:4]), d, 1)
    np.testing.assert_array_equal(alone, x[4:6])
    rng = np.random.default_rng(2)
    with pytest.raises(ValueError, match="at most 255 parameters, got 256"):
        _core.LsiDesign(rng.normal(size=(300, 256)), np.eye(256), 0.1)


def test_msmt_model_recovers_wm_csf_mixture_at_wm_lmax_10():
    # tests/sd/test_msmt.py:69-88 at WM lmax 10 (67 parameters).
    from larmor_amd.types import GradientTable

    forward, _, _, _, (wm, csf), (bvals, bvecs) = _wide_msmt(10, 1, full=True)
    fod_true = fit_sh(0.7 * 9 / (4 * np.pi) * (DIRS_300 @ FIBRE) ** 8, DIRS_300, 10)
    signal = forward @ np.concatenate([fod_true, [0.3 / np.sqrt(4 * np.pi)]])
    volume = Volume(np.tile(signal, (2, 1, 1, 1)), np.eye(4))
    fit = MSMTCSDModel([csf, wm], lmax=10).fit(volume, GradientTable(bvals, bvecs), n_jobs=1)
    np.testing.assert_allclose(fit.tissue_fractions[0, 0, 0], (0.7, 0.3), atol=1e-6)
    np.testing.assert_allclose(fit.fods[0][0, 0, 0], fod_true, atol=1e-6)
