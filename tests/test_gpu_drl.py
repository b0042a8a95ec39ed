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
