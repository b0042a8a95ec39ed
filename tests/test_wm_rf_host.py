"""Host parts of the white matter response and the FOD merge (no GPU).

The fixture (tests/golden/wm_rf_fixture.npz, made by make_wm_rf_fixture.py) holds the
reference's 8x8x8 CFIN block with MRIToolkit's per-voxel tensor fit and white matter response
(tests/data/wm_rf/wm_rf_ref.mat) and MRIToolkit's mFOD merge of two FODs
(tests/data/mfod/mfod_merge_ref.mat).  The merge tests restate the reference's
tests/sd/test_merge_fods.py.
"""

from functools import cache
from pathlib import Path

import numpy as np
import pytest

from larmor_amd.sd import merge_fods
from larmor_amd.sd.rf import _eigenvalues, _fractional_anisotropy, _wm_design
from larmor_amd.types import GradientTable

FIXTURE = Path(__file__).resolve().parent / "golden" / "wm_rf_fixture.npz"


@cache
def fixture() -> dict:
    with np.load(FIXTURE) as f:
        return {k: f[k] for k in f.files}


def mritoolkit_input(f: dict):
    """(dwi, bvals, bvecs) as MRIToolkit's pipeline saw the block: b-values rounded to the
    nearest 100 and, in each mask voxel whose smallest b0 is not above its largest weighted
    signal M, the b0s at or below M + 1 raised to M + 1 (the reference's
    tests/sd/test_rf_kurtosis.py:36-56 describes the same preparation)."""
    bvals = np.round(f["bvals"] / 100.0) * 100.0
    is_b0 = GradientTable(bvals, f["bvecs"]).b0_mask
    dwi = f["dwi"].copy()
    mask = f["mask"] > 0
    s = dwi[mask]
    top = s[:, ~is_b0].max(axis=1, keepdims=True)
    b0 = s[:, is_b0]
    lift = (b0.min(axis=1, keepdims=True) <= top) & (b0 <= top + 1)
    s[:, is_b0] = np.where(lift, top + 1, b0)
    dwi[mask] = s
    return dwi, bvals, f["bvecs"]


def wm_design(gtab: GradientTable, with_kurtosis: bool) -> np.ndarray:
    """ln S = ln S0 - sum_ij b g_i g_j D_ij (+ (b / 1000)^2 p_K / 6), the six distinct tensor
    elements as columns."""
    g, b = gtab.bvecs, gtab.bvals
    cols = [np.ones_like(b)]
    for i, j in ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)):
        cols.append(-(1 if i == j else 2) * b * g[:, i] * g[:, j])
    if with_kurtosis:
        cols.append((b * 1e-3) ** 2 / 6)
    return np.column_stack(cols)


def test_wm_design_columns():
    gtab = GradientTable(*mritoolkit_input(fixture())[1:])
    for with_kurtosis in (False, True):
        np.testing.assert_allclose(_wm_design(gtab, with_kurtosis), wm_design(gtab, with_kurtosis),
                                   rtol=1e-15, atol=1e-12)


def test_eigenvalues_and_fa():
    rng = np.random.default_rng(0)
    frames = np.linalg.qr(rng.standard_normal((20, 3, 3)))[0]
    ev = np.sort(rng.uniform(1e-4, 2e-3, (20, 3)), axis=1)[:, ::-1]
    tensors = frames @ (ev[:, :, None] * np.swapaxes(frames, 1, 2))
    elements = tensors[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]].T
    np.testing.assert_allclose(_eigenvalues(elements), ev, rtol=1e-10)
    # FA of [l1, l, l] is (l1 - l) / sqrt(l1^2 + 2 l^2).
    fa = _fractional_anisotropy(np.array([[1.7e-3, 0.3e-3, 0.3e-3], [0, 0, 0]]))
    np.testing.assert_allclose(fa, [1.4e-3 / np.sqrt(1.7e-3**2 + 2 * 0.3e-3**2), 0.0])


SHAPE = (2, 3, 4)  # last axis 4, the FODs' 6: a fraction on the wrong axis cannot pass
C = 6


def _fods():
    rng = np.random.default_rng(0)
    return [rng.normal(size=(*SHAPE, C)) for _ in range(2)]


def _fractions(wm, gm):
    return np.stack([wm, gm, 1.0 - wm - gm], axis=-1)


def test_merge_weighted_sum():  # test_merge_fods.py:29-39
    fods = _fods()
    rng = np.random.default_rng(1)
    wm, gm = rng.uniform(0.0, 0.5, SHAPE), rng.uniform(0.0, 0.5, SHAPE)
    merged = merge_fods(fods, _fractions(wm, gm), "ws")
    np.testing.assert_allclose(merged, wm[..., None] * fods[0] + gm[..., None] * fods[1])


def test_merge_majority_ignores_the_isotropic_fraction():  # test_merge_fods.py:42-54
    fods = _fods()
    wm, gm = np.full(SHAPE, 0.1), np.full(SHAPE, 0.2)
    wm[0] = 0.3
    merged = merge_fods(fods, _fractions(wm, gm), "majority")
    assert merged.shape == (*SHAPE, C)
    np.testing.assert_array_equal(merged[0], fods[0][0])
    np.testing.assert_array_equal(merged[1:], fods[1][1:])


@pytest.mark.parametrize("method", ["ws", "majority"])
def test_merge_single_fod_and_rows(method):  # test_merge_fods.py:57-76
    fod = _fods()[0]
    np.testing.assert_allclose(
        merge_fods([fod], np.stack([np.ones(SHAPE), np.zeros(SHAPE)], axis=-1), method), fod)
    fods = _fods()
    fractions = _fractions(np.full(SHAPE, 0.3), np.full(SHAPE, 0.2))
    fractions[0, 0, 0] = (0.1, 0.6, 0.3)
    grid = merge_fods(fods, fractions, method)
    rows = merge_fods([f.reshape(-1, C) for f in fods], fractions.reshape(-1, 3), method)
    np.testing.assert_allclose(rows, grid.reshape(-1, C))


def test_merge_refusals():  # test_merge_fods.py:79-94
    with pytest.raises(ValueError, match="majority"):
        merge_fods(_fods(), _fractions(np.zeros(SHAPE), np.zeros(SHAPE)), "mean")
    with pytest.raises(ValueError):
        merge_fods(_fods(), np.ones((*SHAPE, 1)), "ws")
    fods = _fods()
    fods[1] = fods[1][..., :3]
    with pytest.raises(ValueError):
        merge_fods(fods, _fractions(np.zeros(SHAPE), np.zeros(SHAPE)), "ws")
    with pytest.raises(ValueError, match="empty"):
        merge_fods([], np.ones((*SHAPE, 2)))


@pytest.mark.parametrize("method,key", [("ws", "merge_linear"), ("majority", "merge_majority")])
def test_merge_matches_mritoolkit(method, key):  # test_merge_fods.py:97-110
    f = fixture()
    merged = merge_fods([f["merge_fod1"], f["merge_fod2"]], f["merge_fractions"], method)
    np.testing.assert_allclose(merged, f[key], rtol=0, atol=1e-7)
