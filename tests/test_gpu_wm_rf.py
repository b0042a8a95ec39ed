"""The log-signal fit kernel and estimate_wm_rf on the GPU.

The kernel follows the oracle's order of operations (wave sums over measurements, the same
exponential and LU), so its fits are compared bit for bit.  estimate_wm_rf is checked against
MRIToolkit's estimate of the CFIN block and against the reference's synthetic known answers
(tests/sd/test_rf_kurtosis.py:126-400), restated here.
"""

import logging
import re

import numpy as np
import pytest

import oracle
from larmor_amd import _core
from larmor_amd.sd import AnisotropicRF, IsotropicRF, NoddiRF, build_response_functions, estimate_wm_rf
from larmor_amd.types import GradientTable, Volume

from test_wm_rf_host import fixture, mritoolkit_input, wm_design  # noqa: E402

pytestmark = pytest.mark.gpu

L1, L2, L3 = 1.7e-3, 0.3e-3, 0.2e-3
K_TRUE = 0.6
OBLIQUE = np.array([1.0, 1.0, 1.0]) / np.sqrt(3)
X_AXIS = np.array([1.0, 0.0, 0.0])
WLS_TOL = 1e-3
E_ATOL = 1e-9
K_RTOL = 1e-5
EXPECTED = [L1, (L2 + L3) / 2, (L2 + L3) / 2]


def _fit_both(x, log_s, weighted=True, max_rw=20, tol=1e-6):
    design = _core.log_signal_design(x)
    ref, ref_refits = oracle.fit_log_signal(*design, log_s, weighted, max_rw, tol)
    got, _, refits = _core.fit_log_signal(np.ascontiguousarray(x), np.zeros((0, x.shape[1])),
                                          np.ascontiguousarray(log_s), weighted, max_rw, tol, 1,
                                          return_refits=True)
    return (ref, ref_refits), (got, refits)


@pytest.mark.parametrize("n_meas,with_kurtosis", [(65, False), (133, True), (257, True),
                                                   (400, False)])
@pytest.mark.parametrize("weighted", [True, False])
def test_log_fit_bit_identical(n_meas, with_kurtosis, weighted):
    print(f"log fit n={n_meas} kurtosis={with_kurtosis} weighted={weighted}", flush=True)
    rng = np.random.default_rng(n_meas)
    bvals = np.concatenate([[0.0, 0.0], rng.choice([600.0, 1000.0, 2000.0], n_meas - 2)])
    dirs = rng.standard_normal((n_meas, 3))
    gtab = GradientTable(bvals, dirs / np.linalg.norm(dirs, axis=1, keepdims=True))
    x = wm_design(gtab, with_kurtosis)
    truth = np.array([np.log(800.0), 1.2e-3, 1e-4, 0.0, 0.5e-3, -1e-4, 0.4e-3, 0.7])
    truth = truth[: x.shape[1]]
    log_s = x @ truth + 0.05 * rng.standard_normal((3000, n_meas))
    (ref, ref_refits), (got, refits) = _fit_both(x, log_s, weighted)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(refits, ref_refits)


def test_log_fit_refused_refit_and_limits():
    b = np.arange(8.0)
    x = np.column_stack([np.ones(8), b])
    (ref, ref_refits), (got, refits) = _fit_both(x, (-400.0 * b)[None])
    assert refits[0] == -1 == ref_refits[0]
    np.testing.assert_array_equal(got, ref)
    for max_rw in (0, 1, 3):
        (ref, rr), (got, r) = _fit_both(x, np.array([[0.0, -1.1, -2.3, -2.9, -4.2, -5.0, -6.1, -6.8]]),
                                        max_rw=max_rw)
        np.testing.assert_array_equal(got, ref)
        assert r[0] == rr[0] <= max_rw


def test_log_fit_argument_checks():
    x = np.ones((8, 2))
    with pytest.raises(ValueError, match="x must be"):
        _core.fit_log_signal(np.ones((2, 3)), np.zeros((0, 3)), np.ones((1, 2)), True, 20, 1e-6, 1)
    with pytest.raises(ValueError, match="max_iter"):
        _core.fit_log_signal(x, np.zeros((0, 2)), np.ones((1, 8)), True, -1, 1e-6, 1)
    with pytest.raises(ValueError, match="n_threads"):
        _core.fit_log_signal(x, np.zeros((0, 2)), np.ones((1, 8)), True, 20, 1e-6, 0)
    with pytest.raises(ValueError, match="unconstrained"):
        _core.fit_log_signal(x, np.ones((1, 2)), np.ones((1, 8)), True, 20, 1e-6, 1)
    with pytest.raises(TypeError):
        _core.fit_log_signal(x.astype(np.float32), np.zeros((0, 2)), np.ones((1, 8)), True, 20,
                             1e-6, 1)


# The CFIN block and MRIToolkit's estimate (reference tests/sd/test_rf_kurtosis.py:96-185).


def _block():
    f = fixture()
    dwi, bvals, bvecs = mritoolkit_input(f)
    return f, Volume(dwi), GradientTable(bvals, bvecs), f["mask"] > 0


def _estimate_counting(caplog, volume, gtab, mask):
    with caplog.at_level(logging.INFO, logger="larmor_amd"):
        rf = estimate_wm_rf(volume, gtab, mask)
    found = re.findall(r"from (\d+) voxels", caplog.text)
    assert len(found) == 1, caplog.text
    return rf, int(found[0])


def test_block_estimate_matches_mritoolkit(caplog):
    f, volume, gtab, mask = _block()
    rf, n_selected = _estimate_counting(caplog, volume, gtab, mask)
    assert n_selected == f["correct_n"] == 9
    lambdas = f["correct_lambdas"]
    np.testing.assert_allclose(rf.eigen_vals, lambdas, rtol=0, atol=WLS_TOL * lambdas[0])
    # MRIToolkit's voxel K is on its mean(lambda^2) scale; the response's on MD^2.
    e = f["voxel_e_vals"][mask][f["voxel_fitted"][mask] > 0]
    k_md = f["voxel_k"][mask][f["voxel_fitted"][mask] > 0] * np.mean(e**2, axis=1) / np.mean(e, axis=1) ** 2
    fa = np.sqrt(1.5 * np.sum((e - e.mean(axis=1, keepdims=True)) ** 2, axis=1) / np.sum(e**2, axis=1))
    kept = (fa > 0.7) & (k_md > 0) & (k_md < 4)
    np.testing.assert_allclose(rf.mean_kurtosis, k_md[kept].mean(), rtol=0, atol=1e-3)


def test_single_shell_block_matches_mritoolkit(caplog):
    f, volume, gtab, mask = _block()
    keep = np.isin(np.round(gtab.bvals, -2), (0.0, 1000.0))
    volume = Volume(volume.data[..., keep])
    gtab = GradientTable(gtab.bvals[keep], gtab.bvecs[keep])
    rf, n_selected = _estimate_counting(caplog, volume, gtab, mask)
    assert n_selected == f["single_n"] == 20
    np.testing.assert_allclose(rf.eigen_vals, f["single_lambdas"], rtol=0,
                               atol=WLS_TOL * f["single_lambdas"][0])
    assert rf.mean_kurtosis == 0.0


def test_block_voxel_fits_equal_the_oracle():
    f, volume, gtab, mask = _block()
    s = volume.data[mask].astype(np.float64)
    log_s = np.log(s[(s > 0).all(axis=1)])
    (ref, _), (got, _) = _fit_both(wm_design(gtab, True), log_s)
    np.testing.assert_array_equal(got, ref)


# Synthetic known answers (reference tests/sd/test_rf_kurtosis.py:188-400).


def _gtab(shells, n_dirs=30):
    rng = np.random.default_rng(0)
    dirs = rng.normal(size=(n_dirs * len(shells), 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    return GradientTable(np.concatenate([[0.0, 0.0], np.repeat(shells, n_dirs)]),
                         np.vstack([np.zeros((2, 3)), dirs]))


def _signal(gtab, e_vals, k=K_TRUE, fibre=OBLIQUE):
    """ln S = ln 1000 - b g^T D g + b^2 k MD^2 / 6, D with e_vals, the first along fibre."""
    a = fibre / np.linalg.norm(fibre)
    helper = np.array([0.0, 0.0, 1.0]) if abs(a[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
    u = np.cross(a, helper)
    u /= np.linalg.norm(u)
    frame = np.column_stack([a, u, np.cross(a, u)])
    d = frame @ np.diag(e_vals) @ frame.T
    gdg = np.einsum("ni,ij,nj->n", gtab.bvecs, d, gtab.bvecs)
    b = gtab.bvals
    return 1000.0 * np.exp(-b * gdg + b**2 * k * np.mean(e_vals) ** 2 / 6)


def _volume(signals):
    return Volume(np.array(signals)[:, None, None])


def _all(volume):
    return np.ones(volume.data.shape[:3], dtype=bool)


def _with_fa(fa, l23=0.4e-3):
    l1 = l23 * (1 + np.sqrt(1 - (1 - fa**2) * (1 - 2 * fa**2))) / (1 - fa**2)
    return (float(l1), l23, l23)


@pytest.mark.parametrize("fibre", [X_AXIS, OBLIQUE], ids=["axis", "oblique"])
def test_known_tensor_and_kurtosis_recovered(fibre):
    gtab = _gtab((1000.0, 2000.0))
    volume = _volume([_signal(gtab, (L1, L2, L3), fibre=fibre)] * 3)
    rf = estimate_wm_rf(volume, gtab, _all(volume))
    np.testing.assert_allclose(rf.eigen_vals, EXPECTED, rtol=0, atol=E_ATOL)
    np.testing.assert_allclose(rf.mean_kurtosis, K_TRUE, rtol=K_RTOL)


def test_estimated_response_reproduces_its_signal():
    e_vals = (1.7e-3, 0.25e-3, 0.25e-3)
    gtab = _gtab((1000.0, 2000.0))
    voxel = _signal(gtab, e_vals)
    rf = estimate_wm_rf(_volume([voxel] * 3), gtab, _all(_volume([voxel] * 3)))
    predicted = rf.signal(gtab.bvals, gtab.bvecs, OBLIQUE[None])[:, 0]
    np.testing.assert_allclose(predicted, voxel / 1000.0, rtol=1e-6)


@pytest.mark.parametrize("shells", [(1000.0, 2000.0), (600.0, 1000.0)])
def test_two_shells_above_500_fit_the_kurtosis(shells):
    gtab = _gtab(shells)
    volume = _volume([_signal(gtab, (L1, L2, L3))])
    np.testing.assert_allclose(estimate_wm_rf(volume, gtab, _all(volume)).mean_kurtosis, K_TRUE,
                               rtol=K_RTOL)


@pytest.mark.parametrize("shells", [(1000.0,), (300.0, 1000.0), (990.0, 1000.0, 1010.0)])
def test_fewer_than_two_shells_above_500_give_zero_kurtosis(shells):
    gtab = _gtab(shells)
    volume = _volume([_signal(gtab, (L1, L2, L3))])
    assert estimate_wm_rf(volume, gtab, _all(volume)).mean_kurtosis == 0.0


def test_single_shell_fits_the_plain_tensor():
    gtab = _gtab((1000.0,))
    volume = _volume([_signal(gtab, (L1, L2, L3), k=0.0)])
    np.testing.assert_allclose(estimate_wm_rf(volume, gtab, _all(volume)).eigen_vals, EXPECTED,
                               rtol=0, atol=E_ATOL)


@pytest.mark.parametrize("other_fa,kept", [(0.69, False), (0.71, True)])
def test_only_voxels_with_fa_above_0_7_count(other_fa, kept):
    gtab = _gtab((1000.0, 2000.0))
    other = _with_fa(other_fa)
    volume = _volume([_signal(gtab, (L1, L2, L3))] * 2 + [_signal(gtab, other)] * 5)
    rf = estimate_wm_rf(volume, gtab, _all(volume))
    np.testing.assert_allclose(rf.eigen_vals, list(other) if kept else EXPECTED, rtol=0,
                               atol=E_ATOL)


def test_no_voxel_with_fa_above_0_7_is_refused():
    gtab = _gtab((1000.0, 2000.0))
    volume = _volume([_signal(gtab, _with_fa(0.69))] * 3)
    with pytest.raises(ValueError, match="FA > 0.7"):
        estimate_wm_rf(volume, gtab, _all(volume))


@pytest.mark.parametrize("bad_k", [-0.2, 4.5])
def test_voxel_kurtosis_outside_0_to_4_is_left_out(bad_k):
    gtab = _gtab((1000.0, 2000.0))
    volume = _volume([_signal(gtab, (L1, L2, L3))] * 2 + [_signal(gtab, (L1, L2, L3), k=bad_k)] * 3)
    np.testing.assert_allclose(estimate_wm_rf(volume, gtab, _all(volume)).mean_kurtosis, K_TRUE,
                               rtol=K_RTOL)


def test_kurtosis_averaged_over_the_selected_voxels_only():
    gtab = _gtab((1000.0, 2000.0))
    low_fa = _signal(gtab, _with_fa(0.5), k=1.5)
    volume = _volume([_signal(gtab, (L1, L2, L3))] * 2 + [low_fa] * 3)
    np.testing.assert_allclose(estimate_wm_rf(volume, gtab, _all(volume)).mean_kurtosis, K_TRUE,
                               rtol=K_RTOL)


def test_no_voxel_kurtosis_in_range_gives_zero_with_a_warning(caplog):
    gtab = _gtab((1000.0, 2000.0))
    volume = _volume([_signal(gtab, (L1, L2, L3), k=-0.2)] * 3)
    with caplog.at_level(logging.WARNING, logger="larmor_amd"):
        rf = estimate_wm_rf(volume, gtab, _all(volume))
    assert rf.mean_kurtosis == 0.0
    assert "kurtosis 0" in caplog.text
    np.testing.assert_allclose(rf.eigen_vals, EXPECTED, rtol=0, atol=E_ATOL)


@pytest.mark.parametrize("bad", [0.0, -5.0, np.nan, np.inf])
def test_voxels_with_a_bad_signal_are_skipped(bad):
    gtab = _gtab((1000.0, 2000.0))
    broken = _signal(gtab, _with_fa(0.9))
    broken[5] = bad
    volume = _volume([_signal(gtab, (L1, L2, L3))] * 2 + [broken] * 5)
    rf = estimate_wm_rf(volume, gtab, _all(volume))
    np.testing.assert_allclose(rf.eigen_vals, EXPECTED, rtol=0, atol=E_ATOL)
    np.testing.assert_allclose(rf.mean_kurtosis, K_TRUE, rtol=K_RTOL)


def test_build_response_functions():  # reference tests/sd/test_rf_kurtosis.py:400-
    gtab = _gtab((1000.0, 2000.0))
    volume = _volume([_signal(gtab, (L1, L2, L3))] * 3)
    mask = _all(volume)
    rfs = build_response_functions(volume, gtab, mask)
    assert len(rfs) == 1 and isinstance(rfs[0], AnisotropicRF)
    np.testing.assert_allclose(rfs[0].mean_kurtosis, K_TRUE, rtol=K_RTOL)
    rfs = build_response_functions(volume, gtab, mask, mean_kurtosis=0.0, gm_noddi=True,
                                   csf_diffusivity=3e-3)
    assert rfs[0].mean_kurtosis == 0.0
    assert isinstance(rfs[1], NoddiRF) and isinstance(rfs[2], IsotropicRF)
    rfs = build_response_functions(volume, gtab, mask, wm_eigen_vals=[1.7e-3, 2e-4, 2e-4],
                                   gm_diffusivity=0.8e-3, csf_diffusivity=3e-3)
    assert rfs[0].mean_kurtosis == 0.0
    assert [r.diffusivity for r in rfs[1:]] == [0.8e-3, 3e-3]
    with pytest.raises(ValueError, match="not both"):
        build_response_functions(volume, gtab, mask, gm_diffusivity=0.8e-3, gm_noddi=True)
