"""The GPU peak search, climbs and derivatives against the CPU oracle.

The kernels call the device's acos, atan2, sin and cos, which round differently from the host
library, so agreement is to rounding, not to the bit.  Stated tolerances:
- derivatives: 1e-12 of the series' scale;
- a climb's direction: 1e-6 rad; its amplitude: 1e-9 relative;
- a peak search: identical found/not-found pattern per slot, directions within 1e-5 rad,
  amplitudes within 1e-7 relative (a climb whose last step sits at the 1e-4 stopping threshold
  may take one more, quadratically small, step on one side).
The oracle is pinned by tests/test_oracle_peaks.py (MRtrix sh2peaks fixture, finite
differences, closed forms).
"""

import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from larmor_amd import _core
from larmor_amd.sd.peaks import _build_icosphere, _device_sphere, _get_amplitudes, find_peaks
from larmor_amd.sd.sh import build_basis_matrix, get_sh_indices
from larmor_amd.sd.sphere import get_theta_phi

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
DIR_TOL = 1e-5
AMP_RTOL = 1e-7


def _angle(a, b):
    c = np.abs(np.sum(a * b, axis=-1)) / (np.linalg.norm(a, axis=-1) * np.linalg.norm(b, axis=-1))
    return np.arccos(np.minimum(c, 1.0))


def _assert_peaks_close(got_d, got_a, ref_d, ref_a):
    np.testing.assert_array_equal(np.isfinite(got_a), np.isfinite(ref_a))
    found = np.isfinite(ref_a)
    assert _angle(got_d[found], ref_d[found]).max() < DIR_TOL
    np.testing.assert_allclose(got_a[found], ref_a[found], rtol=AMP_RTOL)
    assert np.isnan(got_d[~found]).all()


def _fibre_fods(n, lmax=8, seed=0, full=False):
    """Sums of 1-3 apodised fibre lobes (random directions and weights) plus noise."""
    rng = np.random.default_rng(seed)
    ls, _ = get_sh_indices(lmax, full)
    apod = np.exp(-0.02 * ls * (ls + 1))
    out = np.zeros((n, len(ls)))
    for v in range(n):
        k = rng.integers(1, 4)
        d = rng.normal(size=(k, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        w = rng.dirichlet(np.ones(k))
        b = build_basis_matrix(*get_theta_phi(d), lmax, full=full)
        out[v] = (w[:, None] * b * apod).sum(axis=0)
    return out + rng.normal(scale=0.01, size=out.shape)


def test_fixture_peaks_match_oracle_with_the_same_sphere_amplitudes():
    coeffs = np.load(GOLDEN / "peaks_fixture.npz")["fod"]
    vertices, neighbours = _build_icosphere()
    on_sphere = _get_amplitudes(coeffs, 8, False)
    print("peaks fixture", coeffs.shape, flush=True)
    ref_d, ref_a = oracle.peaks_chunk(coeffs, 8, on_sphere, vertices, neighbours, 3, 0.1, 0.0)
    got_d, got_a = _core.peaks_chunk(coeffs, 8, on_sphere, vertices, neighbours, 3, 0.1, 0.0, 1)
    _assert_peaks_close(got_d, got_a, ref_d, ref_a)


# lmax 14-20 run the lmax-20 instantiation with the lmax read at run time.
@pytest.mark.parametrize("lmax", [2, 4, 6, 8, 10, 12, 14, 16, 20])
def test_find_peaks_matches_oracle(lmax):
    coeffs = _fibre_fods(400, lmax, seed=lmax)
    vertices, neighbours = _build_icosphere()
    basis = build_basis_matrix(*get_theta_phi(vertices), lmax)
    on_sphere = oracle.sphere_amplitudes(coeffs, basis)
    print(f"find_peaks lmax {lmax}", flush=True)
    ref_d, ref_a = oracle.peaks_chunk(coeffs, lmax, on_sphere, vertices, neighbours, 3, 0.1, 0.0)
    got = find_peaks(coeffs)
    _assert_peaks_close(got.directions, got.amplitudes, ref_d, ref_a)


@pytest.mark.parametrize("lmax", [8, 11])
def test_find_peaks_full_series_matches_oracle(lmax):
    coeffs = _fibre_fods(300, lmax, seed=3, full=True)
    # One-sided lobes: add odd degrees so that +d and -d differ.
    ls, _ = get_sh_indices(lmax, True)
    coeffs[:, ls % 2 == 1] *= 3.0
    vertices, neighbours = _build_icosphere()
    basis = build_basis_matrix(*get_theta_phi(vertices), lmax, full=True)
    on_sphere = oracle.sphere_amplitudes(coeffs, basis)
    print(f"find_peaks full lmax {lmax}", flush=True)
    ref_d, ref_a = oracle.peaks_chunk(coeffs, lmax, on_sphere, vertices, neighbours, 6, 0.1, 0.0,
                                      full=True)
    got = find_peaks(coeffs)
    assert got.directions.shape == (300, 6, 3)
    _assert_peaks_close(got.directions, got.amplitudes, ref_d, ref_a)


def test_find_peaks_thresholds_mask_and_edge_cases():
    coeffs = _fibre_fods(60, 8, seed=9)
    coeffs[3] = np.nan
    coeffs[4] = 0.0
    mask = np.ones(60, dtype=bool)
    mask[7] = False
    vertices, neighbours = _build_icosphere()
    basis = build_basis_matrix(*get_theta_phi(vertices), 8)
    sel = coeffs[mask]
    on_sphere = oracle.sphere_amplitudes(sel, basis)
    ref_d, ref_a = oracle.peaks_chunk(sel, 8, on_sphere, vertices, neighbours, 5, 0.3, 0.2)
    got = find_peaks(coeffs, mask, num_peaks=5, relative_threshold=0.3, threshold=0.2)
    _assert_peaks_close(got.directions[mask], got.amplitudes[mask], ref_d, ref_a)
    assert np.isnan(got.amplitudes[7]).all() and np.isnan(got.amplitudes[3]).all()
    assert np.isnan(got.amplitudes[4]).all()
    empty = find_peaks(np.zeros((0, 45)))
    assert empty.directions.shape == (0, 3, 3)
    with pytest.raises(ValueError, match="at most 20"):
        find_peaks(np.zeros((2, 276)))


def test_climbs_and_derivatives_match_oracle():
    rng = np.random.default_rng(4)
    for lmax, full in ((8, False), (12, False), (5, True), (16, False), (20, False), (9, True),
                       (20, True)):
        n = len(get_sh_indices(lmax, full)[0])
        for _ in range(20):
            sh = rng.normal(size=n)
            theta, phi = rng.uniform(0, math.pi), rng.uniform(-math.pi, math.pi)
            ref = np.array(oracle.sh_derivatives(sh, lmax, theta, phi, full=full))
            got = np.array(_core.sh_derivatives(sh, lmax, theta, phi, full=full))
            np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
            start = rng.normal(size=3)
            start /= np.linalg.norm(start)
            rd, ra = oracle.climb(sh, lmax, start, full=full)
            gd, ga = _core.climb(sh, lmax, start, full=full)
            if np.isnan(ra):
                assert np.isnan(ga)
                continue
            assert _angle(gd, rd) < 1e-6
            assert ga == pytest.approx(ra, rel=1e-9, abs=1e-12)


def test_device_sphere_amplitudes_are_the_oracles():
    # find_peaks computes the sphere amplitudes on the device in the oracle's order; with the
    # same amplitudes the local maxima, and so the climbs' starts, are the same.
    coeffs = _fibre_fods(50, 8, seed=11)
    sphere = _device_sphere(8, False)
    d1, a1, _ = _core.find_peaks_device(coeffs, 8, sphere, 3, 0.1, 0.0)
    vertices, _ = _build_icosphere()
    basis = build_basis_matrix(*get_theta_phi(vertices), 8)
    d2, a2, _ = _core.find_peaks_device(coeffs, 8, sphere, 3, 0.1, 0.0,
                                        on_sphere=oracle.sphere_amplitudes(coeffs, basis))
    np.testing.assert_array_equal(a1, a2)
    np.testing.assert_array_equal(d1, d2)
