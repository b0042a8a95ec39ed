"""The oracle's damped Richardson-Lucy (cpp/src/sd/drl.cpp:39-99) and GRL multi-tissue fit.

Pinned by:
- the known answers of the reference's cpp/tests/sd/test_drl.cpp on its toy kernel (a voxel
  converging before the short limit gives the same FOD with the long one, one does not; the
  converged FODs explain the signals within 2e-2; the floor holds; a voxel alone equals the
  same voxel in a block);
- an independent vectorised numpy statement of the published update (Dell'Acqua et al. 2010 as
  MRIToolkit runs it, the reference tests' own comparison, tests/sd/test_drl.py) at rtol 1e-10:
  the two differ only in summation order;
- MATLAB's run_grl.m on the 6 x 6 x 6 block of the reference's tests/data (grl_obj.mat,
  deconv_obj.mat, via tests/golden/grl_fixture.npz): the calibrated thresholds NN_L, NN_H at
  rtol 1e-4 and the GRL FODs (rtol 1e-3) and fractions (atol 1e-3) of every voxel, the
  reference's own bars (tests/sd/test_matlab_match.py:36-78).
"""

from pathlib import Path

import numpy as np
import pytest

import oracle
from larmor_amd.sd.kernel import Kernel
from larmor_amd.sd.model import normalize_by_b0
from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
from larmor_amd.sd.sphere import DIRS_300, get_dirs
from larmor_amd.types import GradientTable

GOLDEN = Path(__file__).parent / "golden"
FOD_FLOOR = 1e-6
MULTI_TISSUE_DRL_INIT = 1e-6


def toy_kernel():  # test_drl.cpp:22-31
    r = np.arange(12)[:, None]
    c = np.arange(8)[None, :]
    return 1.0 + ((r * 7 + c * 3) % 11) / 11.0


def toy_signals(kernel):  # test_drl.cpp:33-44
    fods = np.empty((5, 8))
    for v in range(5):
        for d in range(8):
            fods[v, d] = 1.0 + v if d == v % 8 else 0.05 * (1 + (v + d) % 3)
    return fods @ kernel.T / 12.0


def toy_fit(signals, n_iter):
    k = toy_kernel()
    return oracle.fit_single_tissue_drl(signals, k.T.copy(), k.T @ k, 0.05, n_iter, 1.0 / 8,
                                        iterations=True)


def test_toy_voxels_converge_at_different_iterations():
    s = toy_signals(toy_kernel())
    short, it_short = toy_fit(s, 1000)
    long, it_long = toy_fit(s, 5000)
    same = np.all(short == long, axis=1)
    assert 0 < same.sum() < 5
    assert np.array_equal(same, it_short > 0)
    assert (it_long > 0).all()


def test_toy_fods_explain_the_signals_and_keep_the_floor():
    k = toy_kernel()
    s = toy_signals(k)
    f, _ = toy_fit(s, 5000)
    predicted = f @ k.T
    error = np.abs(predicted - s).sum(axis=1) / s.sum(axis=1)
    assert (error < 2e-2).all()
    assert (f >= FOD_FLOOR).all()


def test_toy_voxel_alone_equals_the_block():
    s = toy_signals(toy_kernel())
    for n_iter in (1000, 5000):
        block, _ = toy_fit(s, n_iter)
        for v in range(5):
            alone, _ = toy_fit(s[v : v + 1], n_iter)
            np.testing.assert_array_equal(alone[0], block[v])


def numpy_drl(signal, k_t, k_t_k, nu, n_iter, init):
    """The damped Richardson-Lucy update written with numpy's vector operations."""
    f = np.full(k_t.shape[0], init)
    ks = k_t @ signal
    n = signal.shape[0]
    mu = max(0.0, 1.0 - 4.0 * np.std(signal) * np.sqrt(n / (n - 1)))
    previous = f.copy()
    eps = 1e-4 * np.max(signal)
    nu8 = nu**8
    for _ in range(n_iter):
        dem = k_t_k @ f
        f8 = f**8
        u = 1.0 - mu * (1.0 - f8 / (f8 + nu8))
        f = np.maximum(FOD_FLOOR, f * (1.0 + u * ((ks - dem) / (dem + 1e-16))))
        if np.abs(f - previous).sum() <= eps:
            break
        previous = f.copy()
    return f


@pytest.fixture(scope="module")
def grl():
    return dict(np.load(GOLDEN / "grl_fixture.npz"))


def _rfs():
    return [AnisotropicRF([1.7e-3, 0.2e-3, 0.2e-3], 0.0), IsotropicRF(0.7e-3), IsotropicRF(3e-3)]


def _kernels(grl):
    bvals, bvecs, w = grl["bvals"], grl["bvecs"], grl["shell_weight"]
    lr = Kernel(_rfs(), bvals, bvecs, dirs=[get_dirs(min(len(bvals), 45))], weights=w)
    hr = Kernel(_rfs(), bvals, bvecs, dirs=DIRS_300, weights=w)
    return lr, hr


def _normalized(grl):
    is_b0 = GradientTable(grl["bvals"], grl["bvecs"]).b0_mask
    return np.ascontiguousarray(normalize_by_b0(grl["signals"], is_b0))


@pytest.mark.parametrize("nu, init", [(0.0, "uniform"), ("nu", "uniform"), ("nu", "floor")])
def test_oracle_matches_the_numpy_update(grl, nu, init):
    _, hr = _kernels(grl)
    k = hr.w_k_aniso
    k_t, k_t_k = np.ascontiguousarray(k.T), np.ascontiguousarray(k.T @ k)
    nu = float(grl["nu_h"]) if nu == "nu" else nu
    start = 1.0 / k.shape[1] if init == "uniform" else MULTI_TISSUE_DRL_INIT
    s = _normalized(grl)[::7]
    got, its = oracle.fit_single_tissue_drl(s, k_t, k_t_k, nu, 200, start, iterations=True)
    for v in range(len(s)):
        np.testing.assert_allclose(got[v], numpy_drl(s[v], k_t, k_t_k, nu, 200, start),
                                   rtol=1e-10, atol=0)
    # Some voxels converge early and some run to the limit.
    assert (its > 0).any() or nu == 0.0
    assert (its < 0).any()


def _calibrated_nu(kernel, bvals):
    """calibrate_drl_nu (src/larmor/sd/solvers.py:37-76) with the oracle's dRL."""
    signal = np.full((1, len(bvals)), np.exp(-bvals.max() * 0.7e-3))
    k = kernel.k
    fod = oracle.fit_single_tissue_drl(signal, np.ascontiguousarray(k.T),
                                       np.ascontiguousarray(k.T @ k), 0.0, 200, 1.0 / k.shape[1])
    return 0.2 * fod.max()


def test_calibrated_nu_matches_matlab(grl):
    lr, hr = _kernels(grl)
    np.testing.assert_allclose(_calibrated_nu(lr, grl["bvals"]), grl["nu_l"], rtol=1e-4)
    np.testing.assert_allclose(_calibrated_nu(hr, grl["bvals"]), grl["nu_h"], rtol=1e-4)


def _matlab_stage(kernel, nu, n_iter):
    k = kernel.w_k_aniso
    return (np.ascontiguousarray(k.T), np.ascontiguousarray(k.T @ k), nu, n_iter - 1,
            MULTI_TISSUE_DRL_INIT)


def test_grl_fit_matches_matlab(grl):
    lr, hr = _kernels(grl)
    s = _normalized(grl)
    fods, fractions = oracle.fit_multi_tissue_drl(
        s, lr.weights, lr, hr, _matlab_stage(lr, float(grl["nu_l"]), 200),
        _matlab_stage(hr, float(grl["nu_h"]), 600), fraction_tol=1e-3, final_tol=1e-2)
    for v in range(len(s)):
        np.testing.assert_allclose(fods[v], grl["fod"][v], rtol=1e-3, err_msg=f"voxel {v}")
        np.testing.assert_allclose(fractions[v], grl["fractions"][v], atol=1e-3,
                                   err_msg=f"voxel {v}")
