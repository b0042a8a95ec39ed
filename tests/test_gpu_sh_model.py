"""The SH projection kernel and the whole SDModel fit on the GPU.

- `sh_project` (V_MFMA_F64_16X16X4_F64) against numpy's amps @ P: the MFMA accumulates in its own
  order, so the bar is 1e-13 of each row's scale, and c00 with exact_integral the row sum over
  sqrt(4 pi) to the same tolerance;
- `SDModel(method="mFOD").fit` against the same pipeline with the CPU oracle: b0 estimate and
  normalisation (shared host code), the oracle's multi-tissue fit, numpy's projection, the float32
  grids.  The fit is bit-identical, so fractions match exactly and FODs to float32 rounding of
  the 1e-13 projection difference.
"""

import numpy as np
import pytest

import oracle
from larmor_amd import _core, synthetic
from larmor_amd.dwi.inputs import estimate_s0, signal_rows
from larmor_amd.sd.kernel import Kernel
from larmor_amd.sd.model import SDModel, compute_shell_weights, low_resolution_dirs
from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
from larmor_amd.sd.sh import projection_matrix
from larmor_amd.sd.solvers import FINAL_FRACTION_TOL_X, L2NNLS, MAX_ROUNDS
from larmor_amd.sd.sphere import DIRS_300
from larmor_amd.types import GradientTable, Volume

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lmax", [2, 4, 6, 8, 10, 12])
@pytest.mark.parametrize("exact", [False, True])
def test_sh_project_matches_numpy(lmax, exact):
    rng = np.random.default_rng(lmax)
    amps = np.abs(rng.normal(size=(1013, 300)))
    amps[5] = 0.0
    p = projection_matrix(DIRS_300, lmax).T * (300 / (4 * np.pi))
    got = _core.sh_project(amps, np.ascontiguousarray(p), exact_integral=exact)
    ref = amps @ p
    if exact:
        ref[:, 0] = amps.sum(axis=1) / np.sqrt(4 * np.pi)
    scale = np.abs(amps).sum(axis=1, keepdims=True) * np.abs(p).max() + 1e-300
    assert np.all(np.abs(got - ref) <= 1e-13 * scale)
    assert (got[5] == 0).all()
    assert _core.sh_project(np.zeros((0, 300)), np.ascontiguousarray(p)).shape == (0, p.shape[1])


def test_sd_model_mfod_fit_matches_oracle_pipeline():
    cfg = dict(synthetic.MFOD_CONFIG)
    cfg["shape"] = (10, 10, 6)
    ph = synthetic.phantom(**cfg)
    gtab = GradientTable(ph.bvals, ph.bvecs)
    volume = Volume(ph.data, np.eye(4))
    rfs = [AnisotropicRF([1.7e-3, 0.3e-3, 0.3e-3], 0.0), IsotropicRF(0.8e-3), IsotropicRF(3.0e-3)]
    print("sd model", ph.data.shape, int(ph.mask.sum()), flush=True)
    fit = SDModel(rfs, method="mFOD", solver=L2NNLS()).fit(volume, gtab, mask=ph.mask, n_jobs=1)

    mask = ph.mask
    s0 = estimate_s0(volume, gtab, mask)[mask]
    voxels = np.nonzero(mask)
    signals = signal_rows(volume, voxels, slice(0, len(s0))) / s0[:, None]
    np.minimum(signals, 1.0, out=signals)
    w = compute_shell_weights(ph.bvals, 1.0)
    hr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=low_resolution_dirs(rfs, len(ph.bvals)), weights=w)
    fods, fractions = oracle.fit_multi_tissue(
        np.ascontiguousarray(signals), w, lr, hr, 0.0, 0.1, 1e-6, max_rounds=MAX_ROUNDS,
        fraction_tol=1e-7, final_tol=FINAL_FRACTION_TOL_X)
    p = projection_matrix(DIRS_300, 8).T * (300 / (4 * np.pi))
    sh = fods @ p
    sh[:, 0] = fods.sum(axis=1) / np.sqrt(4 * np.pi)

    got_fod = fit.fods[0]
    assert got_fod.dtype == np.float32 and got_fod.shape == mask.shape + (45,)
    np.testing.assert_allclose(got_fod[mask], sh.astype(np.float32), rtol=2e-6, atol=1e-6)
    assert (got_fod[~mask] == 0).all()
    np.testing.assert_array_equal(fit.tissue_fractions[mask], fractions.astype(np.float32))
