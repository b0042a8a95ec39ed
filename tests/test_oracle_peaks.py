"""The CPU oracle's peak search against what pins it.

- the series value against the tournier basis (scipy lpmv), and every first and second
  derivative against central differences, away from and at the poles
  (tests/sd/test_peaks.py:80-270);
- climbs: a single fibre is reached from starts up to 30 degrees away, a climb started on a peak
  stays, a crossing's peak is the maximum along the fibre plane (scipy's bounded minimiser), a
  climb that runs out of iterations is NaN (tests/sd/test_peaks.py:330-560);
- the icosphere: 642 unit vertices, every antipode, the 12 originals with 5 neighbours padded with
  themselves (tests/sd/test_peaks.py:600-640);
- the MRtrix sh2peaks fixture (tests/golden/peaks_fixture.npz, from the reference's
  tests/data/peaks): the largest peak within 0.01 degrees and 1e-5 relative, and every peak above
  the threshold with median angle < 0.01 degrees, 95th percentile < 0.1 degrees and none missed
  (tests/sd/test_peaks_reference.py:19-84).
"""

import math
from math import cos, pi, sin
from pathlib import Path

import numpy as np
import pytest
from scipy.optimize import minimize_scalar

import oracle
from larmor_amd.sd.peaks import _build_icosphere, _get_amplitudes
from larmor_amd.sd.sh import build_basis_matrix, get_sh_indices

GOLDEN = Path(__file__).resolve().parent / "golden"
H = 1e-5
RTOL_FD = 1e-6
LMAXES = [2, 8, 16]
LMAX_FOD = 8


def _random_sh(lmax, seed=0):
    return np.random.default_rng(seed).normal(size=len(get_sh_indices(lmax)[0]))


def _random_dirs(n=50, seed=1, margin=0.1):
    rng = np.random.default_rng(seed)
    return list(zip(rng.uniform(margin, pi - margin, n), rng.uniform(0, 2 * pi, n)))


def _all(sh, lmax, t, p):
    return np.array(oracle.sh_derivatives(sh, lmax, t, p))


def _d_dtheta(sh, lmax, t, p):
    return (_all(sh, lmax, t + H, p) - _all(sh, lmax, t - H, p)) / (2 * H)


def _d_dphi(sh, lmax, t, p):
    return (_all(sh, lmax, t, p + H) - _all(sh, lmax, t, p - H)) / (2 * H)


def _close_to_scale(actual, expected):
    np.testing.assert_allclose(actual, expected, rtol=0,
                               atol=RTOL_FD * float(np.max(np.abs(expected))))


@pytest.mark.parametrize("lmax", [0, 2, 8, 16])
def test_value_matches_tournier_basis(lmax):
    sh = _random_sh(lmax)
    rng = np.random.default_rng(1)
    theta, phi = rng.uniform(0, pi, 300), rng.uniform(0, 2 * pi, 300)
    f = np.array([_all(sh, lmax, t, p)[0] for t, p in zip(theta, phi)])
    expected = build_basis_matrix(theta, phi, lmax) @ sh
    np.testing.assert_allclose(f, expected, rtol=0, atol=1e-12 * np.abs(sh).sum())


@pytest.mark.parametrize("lmax", LMAXES)
@pytest.mark.parametrize("which", ["t", "p", "tt", "tp", "pp"])
def test_derivatives_match_finite_differences(lmax, which):
    sh = _random_sh(lmax)
    dirs = _random_dirs()
    index = {"t": 1, "p": 2, "tt": 3, "tp": 4, "pp": 5}[which]
    got = np.array([_all(sh, lmax, t, p)[index] for t, p in dirs])
    if which == "t":
        fd = np.array([_d_dtheta(sh, lmax, t, p)[0] for t, p in dirs])
    elif which == "p":
        fd = np.array([_d_dphi(sh, lmax, t, p)[0] / sin(t) for t, p in dirs])
    elif which == "tt":
        fd = np.array([_d_dtheta(sh, lmax, t, p)[1] for t, p in dirs])
    elif which == "tp":
        fd = np.array([_d_dphi(sh, lmax, t, p)[1] / sin(t) for t, p in dirs])
    else:
        fd = np.array([_d_dphi(sh, lmax, t, p)[2] / sin(t) for t, p in dirs])
    _close_to_scale(got, fd)


@pytest.mark.parametrize("lmax", LMAXES)
@pytest.mark.parametrize("theta", [0.0, 1e-6, pi - 1e-6, pi])
def test_poles(lmax, theta):
    sh = _random_sh(lmax)
    out = _all(sh, lmax, theta, 0.7)
    assert np.all(np.isfinite(out))
    expected = build_basis_matrix(np.array([theta]), np.array([0.7]), lmax) @ sh
    assert out[0] == pytest.approx(float(expected[0]), abs=1e-9 * np.abs(sh).sum())
    fd = _d_dtheta(sh, lmax, theta, 0.7)[0]
    assert out[1] == pytest.approx(fd, abs=RTOL_FD * max(abs(fd), 1.0))


def _unit(v):
    u = np.asarray(v, dtype=np.float64)
    return u / np.linalg.norm(u)


def _polar(v):
    return math.acos(max(-1.0, min(1.0, float(v[2])))), math.atan2(float(v[1]), float(v[0]))


def _amplitude(sh, v):
    t, p = _polar(v)
    return float((build_basis_matrix(np.array([t]), np.array([p]), LMAX_FOD) @ sh)[0])


def _fibre_sh(direction):
    ls, _ = get_sh_indices(LMAX_FOD)
    t, p = _polar(_unit(direction))
    return build_basis_matrix(np.array([t]), np.array([p]), LMAX_FOD)[0] * np.exp(-0.02 * ls * (ls + 1))


def _start(fibre, off_deg, azimuth):
    helper = np.array([1.0, 0.0, 0.0]) if abs(fibre[0]) < 0.9 else np.eye(3)[1]
    e1 = _unit(np.cross(fibre, helper))
    e2 = np.cross(fibre, e1)
    tangent = math.cos(azimuth) * e1 + math.sin(azimuth) * e2
    off = math.radians(off_deg)
    return _unit(math.cos(off) * fibre + math.sin(off) * tangent)


def _toward(a, b, angle):
    e = b - (a @ b) * a
    return _unit(math.cos(angle) * a + math.sin(angle) * _unit(e))


def _axis_deg(a, b):
    c = abs(float(a @ b)) / float(np.linalg.norm(a) * np.linalg.norm(b))
    return math.degrees(math.acos(min(1.0, c)))


FIBRES = [(0.0, 0.0, 1.0), (0.03, 0.02, 1.0), (1.0, 0.0, 0.0), (1.0, 1.0, 0.0), (0.3, 0.5, -1.0)]


@pytest.mark.parametrize("fibre", FIBRES)
@pytest.mark.parametrize("off_deg", [5.0, 15.0, 30.0])
@pytest.mark.parametrize("azimuth", [0.0, 2.5, 5.5])
def test_climb_reaches_single_fibre(fibre, off_deg, azimuth):
    d = _unit(fibre)
    sh = _fibre_sh(d)
    direction, f = oracle.climb(sh, LMAX_FOD, _start(d, off_deg, azimuth))
    assert _axis_deg(direction, d) < 1e-3
    assert float(np.linalg.norm(direction)) == pytest.approx(1.0, abs=1e-12)
    assert f == pytest.approx(_amplitude(sh, d), abs=1e-6)


@pytest.mark.parametrize("fibre", FIBRES)
def test_climb_started_on_the_peak_stays_there(fibre):
    d = _unit(fibre)
    sh = _fibre_sh(d)
    direction, f = oracle.climb(sh, LMAX_FOD, d)
    assert _axis_deg(direction, d) < 1e-3
    assert f == pytest.approx(_amplitude(sh, d), abs=1e-12)


@pytest.mark.parametrize("separation_deg", [60.0, 90.0])
@pytest.mark.parametrize("near_b", [False, True])
def test_climb_in_a_crossing_ends_on_the_nearby_peak(separation_deg, near_b):
    a = np.array([1.0, 0.0, 0.0])
    s = math.radians(separation_deg)
    b = np.array([math.cos(s), math.sin(s), 0.0])
    sh = _fibre_sh(a) + _fibre_sh(b)
    target, other = (b, a) if near_b else (a, b)
    res = minimize_scalar(lambda t: -_amplitude(sh, _toward(target, other, t)), bounds=(-0.2, 0.2),
                          method="bounded", options={"xatol": 1e-12})
    peak = _toward(target, other, float(res.x))
    direction, _ = oracle.climb(sh, LMAX_FOD, _start(target, 15.0, 1.0))
    assert _axis_deg(direction, peak) < 1e-3


def test_climb_that_runs_out_of_iterations_is_nan():
    d = _unit(FIBRES[4])
    direction, f = oracle.climb(_fibre_sh(d), LMAX_FOD, _start(d, 15.0, 1.0), max_iters=1)
    assert np.all(np.isnan(direction)) and math.isnan(f)
    direction, f = oracle.climb(_fibre_sh(d), LMAX_FOD, _start(d, 15.0, 1.0), tol=0.0)
    assert np.all(np.isnan(direction)) and math.isnan(f)


def test_icosphere():
    vertices, neighbours = _build_icosphere()
    assert vertices.shape == (642, 3) and neighbours.shape == (642, 6)
    np.testing.assert_allclose(np.linalg.norm(vertices, axis=1), 1.0, atol=1e-12)
    np.testing.assert_allclose(np.max(-vertices @ vertices.T, axis=1), 1.0, rtol=0, atol=1e-12)
    own = neighbours == np.arange(642)[:, None]
    distinct = np.array([len(set(row.tolist())) for row in neighbours])
    assert own[:12, 5].all() and (distinct[:12] == 6).all() and not own[12:].any()
    assert (distinct[12:] == 6).all()


def _fixture_peaks():
    z = np.load(GOLDEN / "peaks_fixture.npz")
    coeffs = z["fod"]
    vertices, neighbours = _build_icosphere()
    on_sphere = _get_amplitudes(coeffs, 8, False)
    dirs, amps = oracle.peaks_chunk(coeffs, 8, on_sphere, vertices, neighbours, 3, 0.1, 0.0)
    return coeffs, z["peaks"], dirs, amps


def test_largest_peak_matches_mrtrix_sh2peaks():
    _, ref, dirs, amps = _fixture_peaks()
    ref_first = ref[:, 0]
    has = np.isfinite(amps[:, 0])
    ref_has = np.isfinite(ref_first).all(axis=1)
    np.testing.assert_array_equal(has, ref_has)
    ref_amp = np.linalg.norm(ref_first[has], axis=1)
    cosine = np.abs(np.sum(dirs[has][:, 0] * (ref_first[has] / ref_amp[:, None]), axis=1))
    assert np.all(cosine >= math.cos(math.radians(0.01)))
    np.testing.assert_allclose(amps[has][:, 0], ref_amp, rtol=1e-5)


def test_every_peak_matches_mrtrix_sh2peaks_above_the_threshold():
    coeffs, ref, dirs, amps = _fixture_peaks()
    angles, missed = [], 0
    for i in range(len(coeffs)):
        stored = ref[i][np.isfinite(ref[i]).all(axis=1)]
        sd = stored / np.linalg.norm(stored, axis=1, keepdims=True)
        theta, phi = np.arccos(sd[:, 2]), np.arctan2(sd[:, 1], sd[:, 0])
        signed = build_basis_matrix(theta, phi, 8) @ coeffs[i]
        sd, signed = sd[signed > 0], signed[signed > 0]
        kept = signed >= 0.1 * signed.max() if len(signed) else signed > 0
        sd, signed = sd[kept], signed[kept]
        mine = dirs[i][np.isfinite(amps[i])]
        for d in mine:
            best = np.abs(sd @ d).max() if len(sd) else 0.0
            angles.append(math.degrees(math.acos(min(best, 1.0))))
        for d, a in zip(sd, signed):
            near = np.abs(mine @ d).max() if len(mine) else 0.0
            if a >= 0.101 * signed.max() and near < math.cos(math.radians(1.0)):
                missed += 1
    assert len(angles) > 500
    assert np.median(angles) < 0.01
    assert np.percentile(angles, 95) < 0.1
    assert missed == 0


def test_symmetric_peaks_point_up_and_are_sorted():
    _, _, dirs, amps = _fixture_peaks()
    found = np.isfinite(amps)
    assert (dirs[found][:, 2] >= 0).all()
    both = found[:, 1:] & found[:, :-1]
    assert (amps[:, 1:][both] <= amps[:, :-1][both]).all()
    assert not (found[:, 1:] & ~found[:, :-1]).any()
    assert (amps[found] > 0).all()
