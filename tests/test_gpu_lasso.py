"""The elastic-net kernel (FodSolver.lasso) on the GPU against the oracle on the same inputs.

The kernel and the oracle share one rounding order (wave sums for the correlations, in-order
residual updates, correctly rounded divisions), so the bar is bit-identity: FODs, fractions,
rounds, the status bit of a solve that max_iter stopped and the solver's unconverged count.
"""

from pathlib import Path

import numpy as np
import pytest

import oracle
from larmor_amd import _core, synthetic
from larmor_amd.dwi.inputs import estimate_s0, signal_rows
from larmor_amd.sd.kernel import Kernel
from larmor_amd.sd.model import SDModel, compute_shell_weights, low_resolution_dirs
from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
from larmor_amd.sd.solvers import Lasso
from larmor_amd.sd.sphere import DIRS_300, get_dirs
from larmor_amd.types import GradientTable, Volume

pytestmark = pytest.mark.gpu

RFS = [AnisotropicRF([1.7e-3, 0.3e-3, 0.3e-3], 0.0), IsotropicRF(0.8e-3), IsotropicRF(3e-3)]


def _setup(shells=None, seed=2):
    cfg = dict(synthetic.MFOD_CONFIG)
    cfg["shape"] = (8, 8, 6)
    cfg["seed"] = seed
    if shells is not None:
        cfg["shells"] = shells
    ph = synthetic.phantom(**cfg)
    signals = synthetic.normalized_signals(ph.data, ph.bvals, ph.mask)
    w = np.ones(len(ph.bvals))
    hr = Kernel(RFS, ph.bvals, ph.bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(RFS, ph.bvals, ph.bvecs, dirs=[get_dirs(45)[:-1]], weights=w)
    return signals, w, lr, hr


def _deconv(k):
    return _core.DeconvKernel(k.k_aniso, k.k_iso, k.w_k_iso,
                              [(b.start, b.stop) for b in k.aniso_blocks])


@pytest.mark.parametrize("resolution", ["lr", "hr"])
@pytest.mark.parametrize("shells", [None, [(700.0, 90), (1000.0, 90), (2000.0, 90), (3000.0, 90)]])
def test_single_tissue_bit_identical(resolution, shells):
    signals, _, lr, hr = _setup(shells)
    k = np.ascontiguousarray((lr if resolution == "lr" else hr).w_k_aniso)
    s = np.ascontiguousarray(signals[:96])
    print("lasso single", resolution, s.shape, k.shape, flush=True)
    solver = _core.FodSolver.lasso(k, 0.01, 0.9, 1e-4, 100_000)
    got, status = _core.fit_single_tissue(s, solver, 1, return_status=True)
    ref, sweeps = oracle.fit_single_tissue_lasso(s, k, 0.01, 0.9, 1e-4, 100_000, sweeps=True)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal((status & 1) != 0, sweeps < 0)
    assert solver.n_unconverged == int((sweeps < 0).sum())
    assert solver.is_lasso and not solver.is_drl


def test_max_iter_stops_and_counts():
    signals, _, _, hr = _setup()
    k = np.ascontiguousarray(hr.w_k_aniso)
    s = np.ascontiguousarray(signals[:64])
    solver = _core.FodSolver.lasso(k, 0.001, 0.5, 1e-10, 3)
    got, status = _core.fit_single_tissue(s, solver, 1, return_status=True)
    ref, sweeps = oracle.fit_single_tissue_lasso(s, k, 0.001, 0.5, 1e-10, 3, sweeps=True)
    np.testing.assert_array_equal(got, ref)
    assert np.all(sweeps == -3) and np.all((status & 1) != 0)
    assert solver.n_unconverged == 64
    _core.fit_single_tissue(s[:10], solver, 1)
    assert solver.n_unconverged == 74


FIXTURE = np.load(Path(__file__).parent / "golden" / "lasso_fixture.npz")


@pytest.mark.parametrize("i", range(int(FIXTURE["n"])))
def test_fixture_problems_bit_identical(i):
    """The reference's six elastic-net problems (60 to 600 columns, 133 and 198 rows) at its three
    penalties, alpha 0.5."""
    x = np.ascontiguousarray(FIXTURE[f"X_{i}"].astype(np.float64))
    y = FIXTURE[f"y_{i}"]
    for reg in FIXTURE[f"lambda_{i}"]:
        print("lasso fixture", str(FIXTURE[f"name_{i}"]), x.shape, reg, flush=True)
        solver = _core.FodSolver.lasso(x, float(reg), 0.5, 1e-8, 2_000)
        got = _core.fit_single_tissue(np.ascontiguousarray(y[None]), solver, 1)
        ref = oracle.fit_single_tissue_lasso(y[None], x, float(reg), 0.5, 1e-8, 2_000)
        np.testing.assert_array_equal(got, ref)


def test_multi_tissue_bit_identical():
    signals, w, lr, hr = _setup(seed=7)
    s = np.ascontiguousarray(signals[:120])
    print("lasso multi", s.shape, flush=True)
    lr_solver = _core.FodSolver.lasso(np.ascontiguousarray(lr.w_k_aniso), 0.01, 0.9, 1e-4, 100_000)
    hr_solver = _core.FodSolver.lasso(np.ascontiguousarray(hr.w_k_aniso), 0.01, 0.9, 1e-4, 100_000)
    fods, fr, rounds, _ = _core.fit_multi_tissue(
        s, w, _deconv(lr), lr_solver, _deconv(hr), hr_solver, 50, 1e-7, 1e-2, 1,
        return_diagnostics=True)
    ref_fods, ref_fr, ref_rounds = oracle.fit_multi_tissue_lasso(s, w, lr, hr, 0.01, 0.9, 1e-4,
                                                                 100_000, rounds=True)
    np.testing.assert_array_equal(rounds, ref_rounds)
    np.testing.assert_array_equal(fods, ref_fods)
    np.testing.assert_array_equal(fr, ref_fr)


def test_sd_model_with_lasso_matches_oracle_fractions():
    cfg = dict(synthetic.MFOD_CONFIG)
    cfg["shape"] = (8, 8, 5)
    ph = synthetic.phantom(**cfg)
    gtab = GradientTable(ph.bvals, ph.bvecs)
    volume = Volume(ph.data, np.eye(4))
    print("sd model lasso", ph.data.shape, flush=True)
    fit = SDModel(RFS, method="mFOD", solver=Lasso()).fit(volume, gtab, mask=ph.mask, n_jobs=1)
    mask = ph.mask
    s0 = estimate_s0(volume, gtab, mask)[mask]
    signals = signal_rows(volume, np.nonzero(mask), slice(0, len(s0))) / s0[:, None]
    np.minimum(signals, 1.0, out=signals)
    w = compute_shell_weights(ph.bvals, 1.0)
    hr = Kernel(RFS, ph.bvals, ph.bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(RFS, ph.bvals, ph.bvecs, dirs=low_resolution_dirs(RFS, len(ph.bvals)), weights=w)
    d = Lasso()
    _, fractions = oracle.fit_multi_tissue_lasso(np.ascontiguousarray(signals), w, lr, hr, d.reg,
                                                 d.alpha, d.tol, d.max_iter)
    np.testing.assert_array_equal(fit.tissue_fractions[mask], fractions.astype(np.float32))
    assert fit.fods[0].shape == mask.shape + (45,)
