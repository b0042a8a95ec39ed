"""Damped Richardson-Lucy (FodSolver.drl) and the GRL multi-tissue fit on the GPU against the CPU
oracle: bit-identical FODs, fractions and stopping pattern.

The oracle is pinned by tests/test_oracle_drl.py (the reference's dRL known answers, a numpy
statement of the update at rtol 1e-10, MATLAB's GRL block).  Every sum of the kernel runs in the
oracle's index order, so the bar is identity.
"""

import numpy as np
import pytest

import oracle
from larmor_amd import _core
from larmor_amd.sd.solvers import (
    DRL,
    MULTI_TISSUE_DRL_INIT,
    Stage,
    calibrate_drl_nu,
    fit_multi_tissue,
)

from test_oracle_drl import _kernels, _normalized, grl  # noqa: F401  (fixture and builders)

pytestmark = pytest.mark.gpu


def _products(k):
    return np.ascontiguousarray(k.T), np.ascontiguousarray(k.T @ k)


@pytest.mark.parametrize("nu, init", [(0.0, "uniform"), ("nu", "uniform"), ("nu", "floor")])
def test_single_tissue_bit_identical(grl, nu, init):  # noqa: F811
    _, hr = _kernels(grl)
    k_t, k_t_k = _products(hr.w_k_aniso)
    nu = float(grl["nu_h"]) if nu == "nu" else nu
    start = 1.0 / k_t.shape[0] if init == "uniform" else MULTI_TISSUE_DRL_INIT
    s = _normalized(grl)
    print("drl single", s.shape, nu, start, flush=True)
    ref, its = oracle.fit_single_tissue_drl(s, k_t, k_t_k, nu, 200, start, iterations=True)
    solver = _core.FodSolver.drl(k_t, k_t_k, nu=nu, n_iter=200, init=start)
    assert solver.is_drl and (solver.n_meas, solver.n_dirs) == hr.w_k_aniso.shape
    got, status = _core.fit_single_tissue(s, solver, 1, return_status=True)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(status != 0, its < 0)


def test_calibrated_nu_matches_the_oracle_and_matlab(grl):  # noqa: F811
    lr, hr = _kernels(grl)
    for kernel, matlab in ((lr, grl["nu_l"]), (hr, grl["nu_h"])):
        k = kernel.k
        signal = np.full((1, len(grl["bvals"])), np.exp(-grl["bvals"].max() * 0.7e-3))
        ref = 0.2 * oracle.fit_single_tissue_drl(signal, k.T.copy(), k.T @ k, 0.0, 200,
                                                 1.0 / k.shape[1]).max()
        nu = calibrate_drl_nu(kernel, grl["bvals"])
        assert nu == ref
        np.testing.assert_allclose(nu, matlab, rtol=1e-4)


def _stage(kernel, nu, n_iter):
    k_t, k_t_k = _products(kernel.w_k_aniso)
    return (k_t, k_t_k, nu, n_iter - 1, MULTI_TISSUE_DRL_INIT)


def test_grl_multi_tissue_bit_identical(grl):  # noqa: F811
    lr, hr = _kernels(grl)
    s = _normalized(grl)
    lr_rl = _stage(lr, float(grl["nu_l"]), 200)
    hr_rl = _stage(hr, float(grl["nu_h"]), 600)
    print("grl multi", s.shape, flush=True)
    ref_fods, ref_fr, ref_rounds = oracle.fit_multi_tissue_drl(
        s, lr.weights, lr, hr, lr_rl, hr_rl, fraction_tol=1e-3, final_tol=1e-2, rounds=True)
    lr_solver = _core.FodSolver.drl(*lr_rl[:2], nu=lr_rl[2], n_iter=lr_rl[3], init=lr_rl[4])
    hr_solver = _core.FodSolver.drl(*hr_rl[:2], nu=hr_rl[2], n_iter=hr_rl[3], init=hr_rl[4])
    fods, fr = fit_multi_tissue(s, lr, hr, lr_solver, hr_solver, 1, fraction_tol=1e-3)
    np.testing.assert_array_equal(fods, ref_fods)
    np.testing.assert_array_equal(fr, ref_fr)
    # MATLAB's run_grl.m on the same block, the reference's bars.
    np.testing.assert_allclose(fods, grl["fod"], rtol=1e-3)
    np.testing.assert_allclose(fr, grl["fractions"], atol=1e-3)


def test_drl_solver_stages_and_mixed_solvers(grl):  # noqa: F811
    """DRL()._prepare per stage (calibrated nu, n_iter - 1 from 1e-6 in the loop), and a GRL
    loop with an L2NNLS final solve, against the oracle."""
    lr, hr = _kernels(grl)
    s = _normalized(grl)[:64]
    lr_solver = DRL()._prepare(lr, grl["bvals"], Stage.LOW_RES)
    hr_solver = DRL()._prepare(hr, grl["bvals"], Stage.HIGH_RES)
    lr_nu, hr_nu = calibrate_drl_nu(lr, grl["bvals"]), calibrate_drl_nu(hr, grl["bvals"])
    ref_fods, ref_fr = oracle.fit_multi_tissue_drl(
        s, lr.weights, lr, hr, _stage(lr, lr_nu, 200), _stage(hr, hr_nu, 600), fraction_tol=1e-3)
    fods, fr = fit_multi_tissue(s, lr, hr, lr_solver, hr_solver, 1)
    np.testing.assert_array_equal(fods, ref_fods)
    np.testing.assert_array_equal(fr, ref_fr)


def test_batch_invariance_and_edges(grl):  # noqa: F811
    _, hr = _kernels(grl)
    k_t, k_t_k = _products(hr.w_k_aniso)
    s = _normalized(grl)
    solver = _core.FodSolver.drl(k_t, k_t_k, nu=float(grl["nu_h"]), n_iter=50, init=0.01)
    whole = _core.fit_single_tissue(s, solver, 1)
    part = _core.fit_single_tissue(np.ascontiguousarray(s[70:73]), solver, 1)
    np.testing.assert_array_equal(part, whole[70:73])
    empty = _core.fit_single_tissue(np.zeros((0, s.shape[1])), solver, 1)
    assert empty.shape == (0, k_t.shape[0])
    zero_iter = _core.FodSolver.drl(k_t, k_t_k, nu=0.1, n_iter=0, init=0.25)
    np.testing.assert_array_equal(_core.fit_single_tissue(s[:5], zero_iter, 1), 0.25)
    with pytest.raises(ValueError, match="shapes do not match"):
        _core.FodSolver.drl(k_t, np.ascontiguousarray(k_t_k[:, :-1]), nu=0.1, n_iter=10, init=0.1)
    with pytest.raises(ValueError, match="n_iter must be at least 0"):
        _core.FodSolver.drl(k_t, k_t_k, nu=0.1, n_iter=-1, init=0.1)
    with pytest.raises(ValueError, match="measurements"):
        _core.fit_single_tissue(np.ascontiguousarray(s[:, :-1]), solver, 1)


def _volume_case(n_rfs):
    from larmor_amd import synthetic
    from larmor_amd.dwi.inputs import estimate_s0, signal_rows
    from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
    from larmor_amd.types import GradientTable, Volume

    cfg = dict(synthetic.MFOD_CONFIG)
    cfg["shape"] = (8, 8, 5)
    ph = synthetic.phantom(**cfg)
    gtab = GradientTable(ph.bvals, ph.bvecs)
    volume = Volume(ph.data, np.eye(4))
    rfs = [AnisotropicRF([1.7e-3, 0.2e-3, 0.2e-3], 0.0), IsotropicRF(0.7e-3), IsotropicRF(3e-3)]
    rfs = rfs[:n_rfs]
    mask = ph.mask
    s0 = estimate_s0(volume, gtab, mask)[mask]
    signals = signal_rows(volume, np.nonzero(mask), slice(0, len(s0))) / s0[:, None]
    np.minimum(signals, 1.0, out=signals)
    return ph, gtab, volume, rfs, mask, np.ascontiguousarray(signals)


def test_sd_model_grl_default_matches_oracle_pipeline():
    """SDModel's default method (GRL: dRL, inner shells weighted 0.2) on a synthetic volume:
    the fractions equal the oracle's GRL fit bit for bit, the SH FODs its amplitudes projected
    (float32)."""
    from larmor_amd.sd.kernel import Kernel
    from larmor_amd.sd.model import SDModel, compute_shell_weights, low_resolution_dirs
    from larmor_amd.sd.sh import projection_matrix
    from larmor_amd.sd.solvers import FINAL_FRACTION_TOL_X
    from larmor_amd.sd.sphere import DIRS_300

    ph, gtab, volume, rfs, mask, signals = _volume_case(3)
    print("grl model", signals.shape, flush=True)
    fit = SDModel(rfs).fit(volume, gtab, mask=mask, n_jobs=1)
    w = compute_shell_weights(ph.bvals, 0.2)
    hr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=low_resolution_dirs(rfs, len(ph.bvals)), weights=w)
    nus = []
    for kernel in (lr, hr):
        k = kernel.k
        iso = np.full((1, len(ph.bvals)), np.exp(-ph.bvals.max() * 0.7e-3))
        nus.append(0.2 * oracle.fit_single_tissue_drl(iso, k.T.copy(), k.T @ k, 0.0, 200,
                                                      1.0 / k.shape[1]).max())
    fods, fractions = oracle.fit_multi_tissue_drl(
        signals, w, lr, hr, _stage(lr, nus[0], 200), _stage(hr, nus[1], 600),
        fraction_tol=1e-3, final_tol=FINAL_FRACTION_TOL_X)
    np.testing.assert_array_equal(fit.tissue_fractions[mask], fractions.astype(np.float32))
    sh = fods @ (projection_matrix(DIRS_300, 8).T * (300 / (4 * np.pi)))
    np.testing.assert_allclose(fit.fods[0][mask], sh.astype(np.float32), rtol=2e-6, atol=1e-6)


def test_sd_model_grl_single_tissue_matches_oracle():
    from larmor_amd.sd.kernel import Kernel
    from larmor_amd.sd.model import SDModel
    from larmor_amd.sd.sh import projection_matrix
    from larmor_amd.sd.sphere import DIRS_300

    ph, gtab, volume, rfs, mask, signals = _volume_case(1)
    fit = SDModel(rfs).fit(volume, gtab, mask=mask, n_jobs=1)
    kernel = Kernel(rfs, ph.bvals, ph.bvecs, dirs=DIRS_300)
    k = kernel.k
    iso = np.full((1, len(ph.bvals)), np.exp(-ph.bvals.max() * 0.7e-3))
    nu = 0.2 * oracle.fit_single_tissue_drl(iso, k.T.copy(), k.T @ k, 0.0, 200,
                                            1.0 / k.shape[1]).max()
    wk = kernel.w_k_aniso
    ref = oracle.fit_single_tissue_drl(signals, wk.T.copy(), wk.T @ wk, nu, 200, 1.0 / wk.shape[1])
    sh = ref @ (projection_matrix(DIRS_300, 8).T * (300 / (4 * np.pi)))
    assert fit.tissue_fractions is None
    np.testing.assert_allclose(fit.fods[0][mask], sh.astype(np.float32), rtol=2e-6, atol=1e-6)
