"""The CSD response estimate (estimate_csd_rf) on the GPU.

- The reference's own check (tests/sd/test_csd.py:105-131): single-fibre and crossing voxels made
  from a known response; the estimate converges, selects single fibres only and recovers the
  response within 1e-3 of its first coefficient.
- Parity: the same estimate with the per-iteration CSD fits and peak searches done by the oracle
  (its fit_csd and peaks_chunk on host amplitudes, as the reference computes them) selects the
  same voxels in the same number of iterations and gives the same response to rounding (the
  GPU's peak amplitudes differ from the oracle's in the last bits: transcendental functions and
  the device-side amplitude sums).
"""

import numpy as np
import pytest

import oracle
from larmor_amd import synthetic
from larmor_amd.dwi.inputs import signal_rows
from larmor_amd.sd import csd_rf
from larmor_amd.sd.csd_solver import CSDSolver, csd_problem
from larmor_amd.sd.rf import ZonalSHRF
from larmor_amd.sd.sh import build_basis_matrix
from larmor_amd.sd.sphere import get_theta_phi
from larmor_amd.types import GradientTable, Volume

pytestmark = pytest.mark.gpu

RESPONSE = np.array([1800.0, -700.0, 180.0, -35.0, 5.0])


def _gradients(n_dirs: int = 60) -> GradientTable:
    rng = np.random.default_rng(1)
    dirs = rng.normal(size=(n_dirs, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    return GradientTable(np.r_[0.0, [1000.0] * n_dirs], np.r_[np.zeros((1, 3)), dirs])


def _voxels(fibre_sets):
    gtab = _gradients()
    rf = ZonalSHRF(bvals=np.array([1000.0]), coefficients=RESPONSE[None])
    rows = [np.r_[2000.0, rf.signal(gtab.bvals[1:], gtab.bvecs[1:], f).mean(axis=1)]
            for f in fibre_sets]
    data = np.array(rows)
    return Volume(data.reshape(len(rows), 1, 1, -1), np.eye(4))


def test_estimated_response_is_that_of_the_single_fibres():
    rng = np.random.default_rng(3)
    singles = rng.normal(size=(60, 3))
    singles /= np.linalg.norm(singles, axis=1, keepdims=True)
    crossings = [np.array([d, np.cross(d, [0.0, 0.0, 1.0])]) for d in singles[:40]]
    crossings = [c / np.linalg.norm(c, axis=1, keepdims=True) for c in crossings]
    voxels = [d[None] for d in singles] + crossings
    order = rng.permutation(len(voxels))
    volume = _voxels([voxels[i] for i in order])
    is_single = order < len(singles)
    print("csd rf synthetic", volume.data.shape, flush=True)
    estimate = csd_rf.estimate_csd_rf(volume, _gradients(), np.ones(volume.data.shape[:3]),
                                      num_voxels=30, n_jobs=1)
    assert estimate.converged
    chosen = estimate.voxels[:, 0, 0]
    assert chosen.sum() == 30 and not (chosen & ~is_single).any()
    found = estimate.rf.coefficients[0]
    np.testing.assert_allclose(found[:5], RESPONSE, atol=1e-3 * RESPONSE[0])
    assert np.abs(found[5:]).max() < 1e-3 * RESPONSE[0]
    assert estimate.rf.bvals.tolist() == [1000.0]


def test_gpu_estimate_equals_the_oracle_pipeline(monkeypatch):
    cfg = dict(synthetic.CSD_CONFIG)
    cfg["shape"] = (12, 12, 8)
    ph = synthetic.phantom(**cfg)
    volume = Volume(ph.data, np.eye(4))
    gtab = GradientTable(ph.bvals, ph.bvecs)
    mask = np.ones(ph.data.shape[:3], dtype=bool)
    print("csd rf phantom", ph.data.shape, flush=True)
    gpu = csd_rf.estimate_csd_rf(volume, gtab, mask, num_voxels=40, n_jobs=1)

    def oracle_fods(volume, candidates, on_shell, bvecs, response, lmax, threads):
        problem = csd_problem(bvecs, response, lmax, CSDSolver())
        signals = np.maximum(signal_rows(volume, np.nonzero(candidates), slice(None), on_shell),
                             0.0)
        fods, _ = oracle.fit_csd(np.ascontiguousarray(signals), problem.initial, problem.forward,
                                 problem.constraint, problem.norm_weight, problem.threshold,
                                 problem.max_iter)
        return fods

    def oracle_peaks(fods, lmax):
        vertices, neighbours = csd_rf._sphere()
        theta, phi = get_theta_phi(vertices)
        on_sphere = np.ascontiguousarray(fods @ build_basis_matrix(theta, phi, lmax).T)
        _, amplitudes = oracle.peaks_chunk(np.ascontiguousarray(fods), lmax, on_sphere, vertices,
                                           neighbours, 2, 0.0, 0.0)
        peaks = np.nan_to_num(amplitudes, nan=0.0)
        peaks[fods[:, 0] <= 0] = 0.0
        return peaks

    monkeypatch.setattr(csd_rf, "_candidate_fods", oracle_fods)
    monkeypatch.setattr(csd_rf, "_two_largest_peaks", oracle_peaks)
    ref = csd_rf.estimate_csd_rf(volume, gtab, mask, num_voxels=40, n_jobs=1)
    assert gpu.num_iter == ref.num_iter and gpu.converged == ref.converged
    np.testing.assert_array_equal(gpu.voxels, ref.voxels)
    np.testing.assert_allclose(gpu.rf.coefficients, ref.rf.coefficients, rtol=1e-9,
                               atol=1e-12 * np.abs(ref.rf.coefficients).max())


def test_csd_model_without_a_response_estimates_one():
    from larmor_amd.sd.csd import CSDModel

    cfg = dict(synthetic.CSD_CONFIG)
    cfg["shape"] = (10, 10, 6)
    ph = synthetic.phantom(**cfg)
    volume = Volume(ph.data, np.eye(4))
    gtab = GradientTable(ph.bvals, ph.bvecs)
    mask = np.ones(ph.data.shape[:3], dtype=bool)
    print("csd model without response", ph.data.shape, flush=True)
    auto = CSDModel().fit(volume, gtab, mask, n_jobs=1)
    rf = csd_rf.estimate_csd_rf(volume, gtab, mask, n_jobs=1).rf
    given = CSDModel(rf).fit(volume, gtab, mask, n_jobs=1)
    np.testing.assert_array_equal(auto.fods[0], given.fods[0])
