"""The three-tissue response estimate (estimate_tissue_responses) on the GPU.

- The reference's own checks on its phantom (tests/sd/test_tissue_responses.py:25-136): each
  tissue's voxels come from that tissue, the responses follow the tissues' signal.
- Parity: the same estimate with every solve done by the oracle (its LSI fits for the two-tissue
  rankings and the response fit, its peak search on host amplitudes) selects the same voxels and
  gives the same responses to rounding.
- MSMTCSDModel without responses estimates them first.
"""

import types

import numpy as np
import pytest

import oracle
from larmor_amd import synthetic
from larmor_amd.sd import tissue_responses as tr
from larmor_amd.sd.msmt import constraint_matrix, forward_matrix, msmt_ridge
from larmor_amd.sd.peaks import Peaks, _build_icosphere, _get_amplitudes
from larmor_amd.types import GradientTable, Volume

pytestmark = pytest.mark.gpu

S0 = {"wm": 1000.0, "gm": 1300.0, "csf": 3000.0}


def _phantom():
    """The reference's phantom (tests/sd/test_tissue_responses.py:25-99), 26 x 26 x 16."""
    data, bvals, bvecs, mask, fractions, single = synthetic.three_tissue_phantom()
    volume = Volume(data, np.diag([2.5, 2.5, 2.5, 1.0]))
    return volume, GradientTable(bvals, bvecs), mask, fractions, single


@pytest.fixture(scope="module")
def phantom():
    return _phantom()


@pytest.fixture(scope="module")
def estimated(phantom):
    volume, gtab, mask, _, _ = phantom
    print("tissue responses", volume.data.shape, flush=True)
    return tr.estimate_tissue_responses(volume, gtab, mask, n_jobs=2)


def test_each_tissue_is_estimated_from_its_own_voxels(phantom, estimated):
    _, _, _, fractions, single = phantom
    assert (fractions["csf"][estimated.csf_voxels] > 0.9).all()
    assert (fractions["gm"][estimated.gm_voxels] > 0.75).all()
    assert (fractions["wm"][estimated.wm_voxels] > 0.9).all()
    assert single[estimated.wm_voxels].all()


def test_responses_follow_the_tissues_signal(estimated):
    root = np.sqrt(4 * np.pi)
    for rf, s0, d in ((estimated.csf, S0["csf"], 3e-3), (estimated.gm, S0["gm"], 0.8e-3)):
        want = root * s0 * np.exp(-np.array([0.0, 1000.0]) * d)
        np.testing.assert_allclose(rf.coefficients[:2, 0], want, rtol=0.03)
    assert estimated.wm.coefficients[0, 0] == pytest.approx(root * S0["wm"], rel=0.03)
    assert (estimated.wm.coefficients[1:, 1] < 0).all()
    assert estimated.wm.coefficients.shape == (4, 6)


def test_gpu_estimate_equals_the_oracle_pipeline(phantom, estimated, monkeypatch):
    volume, gtab, mask, _, _ = phantom

    def prepare(rfs, lmaxes, bvals, bvecs, dirs):
        forward = forward_matrix(rfs, lmaxes, bvals, bvecs)
        return forward, constraint_matrix(lmaxes, dirs), msmt_ridge(forward)

    def fit(signals, design, threads):
        forward, constraints, ridge = design
        return oracle.fit_lsi(np.ascontiguousarray(signals, dtype=np.float64), forward,
                              constraints, ridge)

    def peaks(coeffs, num_peaks, n_jobs):
        lmax = int(round((np.sqrt(8 * coeffs.shape[1] + 1) - 3) / 2))
        vertices, neighbours = _build_icosphere()
        amps = np.ascontiguousarray(_get_amplitudes(coeffs, lmax, False))
        dirs, found = oracle.peaks_chunk(np.ascontiguousarray(coeffs), lmax, amps, vertices,
                                         neighbours, num_peaks, 0.1, 0.0)
        return Peaks(directions=dirs, amplitudes=found)

    core = types.SimpleNamespace(
        LsiDesign=lambda a, g, ridge: (a, g, ridge),
        fit_lsi=lambda signals, design, n_threads: oracle.fit_lsi(signals, *design))
    monkeypatch.setattr(tr, "prepare_msmt", prepare)
    monkeypatch.setattr(tr, "fit_msmt", fit)
    monkeypatch.setattr(tr, "find_peaks", peaks)
    monkeypatch.setattr(tr, "_core", core)
    ref = tr.estimate_tissue_responses(volume, gtab, mask, n_jobs=2)
    for name in ("wm_voxels", "gm_voxels", "csf_voxels"):
        np.testing.assert_array_equal(getattr(estimated, name), getattr(ref, name))
    for name in ("wm", "gm", "csf"):
        got, want = getattr(estimated, name).coefficients, getattr(ref, name).coefficients
        np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


def test_msmt_model_without_responses_estimates_them(phantom, estimated):
    from larmor_amd.sd.msmt_csd import MSMTCSDModel

    volume, gtab, mask, _, _ = phantom
    auto = MSMTCSDModel().fit(volume, gtab, mask, n_jobs=2)
    given = MSMTCSDModel([estimated.wm, estimated.gm, estimated.csf]).fit(volume, gtab, mask,
                                                                           n_jobs=2)
    np.testing.assert_array_equal(auto.fods[0], given.fods[0])
    np.testing.assert_array_equal(auto.tissue_fractions, given.tissue_fractions)
