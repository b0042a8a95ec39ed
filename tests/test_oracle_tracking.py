"""Pins the tracking oracle (oracle/track_oracle.cpp) without a GPU: the reference's known answers
for its random numbers and calibration (cpp/tests/tracking/test_rng.cpp, test_calibrate.cpp),
closed forms of the step loop, the SH amplitude against the basis pinned by MRtrix's sh2amp, and
MRtrix tckgen's own tracks of the reference's synthetic FOD (tests/golden/tracking_fixture.npz,
the bars of tests/tractography/test_det_tckgen.py and test_prob_tckgen.py)."""

import math

import numpy as np
import pytest
from scipy import stats

import oracle
from larmor_amd.sd.sh import build_basis_matrix
from tracking_cases import (CUTOFF, STEP, det_reference, fixture, halves, max_steps_for,
                            split_ours)

# ---------------------------------------------------------------------------------------------
# Random numbers (test_rng.cpp:22-33: values of the reference's numpy twin).


def test_streamline_draws_match_the_reference():
    key, u = oracle.rng_uniforms(42, 7, 2)
    assert key == 15985126342852841771
    assert u[0] == 0.8940814753778417
    assert u[1] == 0.5331080354362024


def test_search_draws_from_its_own_range():
    _, u = oracle.rng_uniforms(42, 7, 1, search=True)
    assert u[0] == 0.45263234128120544


def test_splitmix_draws_follow_the_published_generator():
    # SplitMix64 (Vigna, splitmix64.c): the k-th output of state s is the finaliser of
    # s + (k + 1) * 0x9E3779B97F4A7C15, and a uniform is its top 53 bits times 2^-53.
    def finalise(z):
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9 % 2**64
        z = (z ^ (z >> 27)) * 0x94D049BB133111EB % 2**64
        return z ^ (z >> 31)

    key = finalise((finalise(1234) + 99) % 2**64)
    got_key, u = oracle.rng_uniforms(1234, 99, 5)
    assert got_key == key
    expect = [(finalise((key + (k + 1) * 0x9E3779B97F4A7C15) % 2**64) >> 11) * 2.0**-53
              for k in range(5)]
    assert list(u) == expect


# ---------------------------------------------------------------------------------------------
# Calibration (test_calibrate.cpp).


def test_direction_grid_is_a_centre_and_hexagonal_rings():
    assert len(oracle.direction_grid(0.1, 1.0)) == 1
    assert len(oracle.direction_grid(1.0 + 1e-9, 1.0)) == 7
    assert len(oracle.direction_grid(2.0 + 1e-9, 1.0)) == 19
    grid = oracle.direction_grid(0.3, 0.1)
    assert np.any(np.all(grid == [0.0, 0.0, 1.0], axis=1))
    np.testing.assert_allclose(np.linalg.norm(grid, axis=1), 1.0, atol=1e-12)
    assert np.all(np.arccos(np.clip(grid[:, 2], -1, 1)) <= 0.3 + 1e-12)


# tckgen 3.0.8 -algorithm iFOD1 -info on an lmax-8 FOD (test_calibrate.cpp:51-68).
TCKGEN_CALIBRATION = [(15.0, 7, 1.21581316), (30.0, 7, 2.44004321), (45.0, 13, 4.48055172),
                      (60.0, 19, 3.26650047), (90.0, 37, 4.24406433)]


@pytest.mark.parametrize("angle,n_dirs,ratio", TCKGEN_CALIBRATION)
def test_calibration_matches_tckgen(angle, n_dirs, ratio):
    grid, r = oracle.calibrate(8, False, math.radians(angle))
    assert len(grid) == n_dirs
    assert abs(r - ratio) < 1e-4 * ratio
    assert np.any(np.all(grid == [0.0, 0.0, 1.0], axis=1))


@pytest.mark.parametrize("angle", [15.0, 30.0, 45.0, 60.0, 90.0])
def test_full_calibration_bounds_its_one_sided_lobe(angle):
    lmax = 8
    fibre = np.zeros((lmax + 1) ** 2)
    for l in range(lmax + 1):
        fibre[l * l + l] = math.sqrt((2 * l + 1) / (4 * math.pi))
    max_angle = math.radians(angle)
    grid, ratio = oracle.calibrate(lmax, True, max_angle)
    assert np.isfinite(ratio) and ratio >= 1.0
    cone = oracle.direction_grid(max_angle, 0.01)
    for tilt in [0.0, 5.0, 10.0, 20.0, 40.0]:
        t = math.radians(tilt)
        previous = np.array([math.sin(t), 0.0, math.cos(t)])
        bound = ratio * oracle.sh_amplitude(fibre, lmax, oracle.rotate_to(previous, grid),
                                            full=True).max()
        inside = oracle.sh_amplitude(fibre, lmax, oracle.rotate_to(previous, cone), full=True)
        assert inside.max() <= bound * (1 + 1e-9)


def test_library_calibration_is_the_oracles():
    """The C-ABI's host calibration (no GPU involved) equals the oracle's bit for bit."""
    from larmor_amd import _core

    for lmax, full, angle in [(8, False, 15.0), (8, False, 90.0), (6, True, 30.0),
                              (12, False, 45.0), (20, False, 20.0)]:
        g1, r1 = oracle.calibrate(lmax, full, math.radians(angle))
        g2, r2 = _core.calibrate_tracking(lmax, full, math.radians(angle))
        assert np.array_equal(g1, g2) and r1 == r2


# ---------------------------------------------------------------------------------------------
# Directions and amplitudes (maths.cpp rotate_to, harmonics.cpp amplitude).


def test_rotate_to_takes_z_to_the_reference_and_keeps_angles():
    rng = np.random.default_rng(3)
    for _ in range(20):
        r = rng.normal(size=3)
        r /= np.linalg.norm(r)
        d = rng.normal(size=(5, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        np.testing.assert_allclose(oracle.rotate_to(r, [0.0, 0.0, 1.0])[0], r, atol=1e-15)
        out = oracle.rotate_to(r, d)
        np.testing.assert_allclose(out @ out.T, d @ d.T, atol=1e-14)
        np.testing.assert_allclose(out @ r, d[:, 2], atol=1e-14)
    d = rng.normal(size=(4, 3))
    assert np.array_equal(oracle.rotate_to([0.0, 0.0, 1.0], d), d)
    assert np.array_equal(oracle.rotate_to([0.0, 0.0, -1.0], d), -d)


@pytest.mark.parametrize("lmax,full", [(2, False), (8, False), (12, False), (20, False),
                                       (1, True), (5, True), (8, True)])
def test_amplitude_is_the_basis_times_the_coefficients(lmax, full):
    rng = np.random.default_rng(lmax)
    n = (lmax + 1) ** 2 if full else (lmax + 1) * (lmax + 2) // 2
    c = rng.normal(size=n)
    d = rng.normal(size=(50, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[0] = [0.0, 0.0, 1.0]
    d[1] = [0.0, 0.0, -1.0]
    basis = build_basis_matrix(np.arccos(np.clip(d[:, 2], -1, 1)), np.arctan2(d[:, 1], d[:, 0]),
                               lmax, "tournier", full=full)
    np.testing.assert_allclose(oracle.sh_amplitude(c, lmax, d, full=full), basis @ c,
                               rtol=0, atol=1e-12 * np.abs(c).sum())


# ---------------------------------------------------------------------------------------------
# The step loop with the fixed-rule algorithm: closed forms on a 10^3 grid of unit voxels.

EYE = np.eye(4)
FIELD = np.zeros((10, 10, 10, 1), dtype=np.float32)


def rule_params(rule, stop=0, past_x=float("inf")):
    return oracle.TrackParams(algorithm=2, test_rule=rule, test_stop=stop,
                              test_refuses_past_x=past_x, max_angle_deg=180.0)


def run_test(rule, seeds, dirs, *, stop=0, max_steps=100, bidirectional=True, mask=None,
             past_x=float("inf"), step=1.0):
    return oracle.track_batch(FIELD, EYE, mask, np.asarray(seeds, float),
                              None if dirs is None else np.asarray(dirs, float),
                              np.arange(len(seeds)), 5, step, max_steps, bidirectional,
                              rule_params(rule, stop, past_x))


def test_a_bidirectional_line_ends_at_the_last_points_on_the_image():
    pos, counts, status, followed, _ = run_test(0, [[5.0, 5.0, 5.0]], [[1.0, 0.0, 0.0]])
    assert counts[0] == 10 and list(status[0]) == [2, 2]
    assert np.array_equal(pos[0, :10, 0], np.arange(9, -1, -1, dtype=np.float32))
    assert np.array_equal(followed[0], [1.0, 0.0, 0.0])


def test_one_way_starts_at_the_seed_and_leaves_its_start_untracked():
    pos, counts, status, _, _ = run_test(0, [[5.0, 5.0, 5.0]], [[1.0, 0.0, 0.0]],
                                         bidirectional=False)
    assert counts[0] == 5 and list(status[0]) == [0, 2]
    assert np.array_equal(pos[0, :5, 0], np.arange(5, 10, dtype=np.float32))


def test_max_steps_counts_the_steps_of_the_whole_streamline():
    pos, counts, status, _, _ = run_test(0, [[5.0, 5.0, 5.0]], [[1.0, 0.0, 0.0]], max_steps=3)
    assert counts[0] == 4 and list(status[0]) == [1, 1]
    assert np.array_equal(pos[0, :4, 0], [8, 7, 6, 5])


def test_the_mask_ends_a_streamline_at_its_last_point_inside():
    mask = np.zeros((10, 10, 10), dtype=np.float32)
    mask[2:8] = 1.0
    pos, counts, status, _, _ = run_test(0, [[5.0, 5.0, 5.0]], [[1.0, 0.0, 0.0]],
                                         mask=(mask, EYE))
    assert counts[0] == 6 and list(status[0]) == [3, 3]
    assert np.array_equal(pos[0, :6, 0], [7, 6, 5, 4, 3, 2])


def test_a_refused_step_adds_no_point_and_records_its_stop():
    pos, counts, status, _, _ = run_test(1, [[5.0, 5.0, 5.0]], [[1.0, 0.0, 0.0]], stop=4)
    assert counts[0] == 3 and list(status[0]) == [4, 4]
    assert np.array_equal(pos[0, :3, 0], [6, 5, 4])


def test_quarter_turns_walk_a_square():
    _, counts, status, _, _ = run_test(2, [[5.0, 5.0, 5.0]], [[1.0, 0.0, 0.0]],
                                       bidirectional=False, max_steps=8)
    pos, *_ = run_test(2, [[5.0, 5.0, 5.0]], [[1.0, 0.0, 0.0]], bidirectional=False, max_steps=8)
    assert counts[0] == 9 and list(status[0]) == [0, 1]
    expect = [[5, 5], [6, 5], [6, 6], [5, 6], [5, 5], [6, 5], [6, 6], [5, 6], [5, 5]]
    assert np.array_equal(pos[0, :9, :2], np.asarray(expect, dtype=np.float32))


def test_seeds_off_the_image_or_the_mask_or_without_a_direction_do_not_start():
    mask = np.ones((10, 10, 10), dtype=np.float32)
    mask[0] = 0.0
    seeds = [[-1.0, 5.0, 5.0], [0.2, 5.0, 5.0], [5.0, 5.0, 5.0], [8.0, 5.0, 5.0]]
    _, counts, status, followed, _ = run_test(0, seeds, None, mask=(mask, EYE), past_x=6.0)
    assert list(counts[[0, 1, 3]]) == [0, 0, 0]
    assert np.all(status[[0, 1, 3]] == 0) and np.all(followed[[0, 1, 3]] == 0)
    # The seed that starts follows the search's unit vector, two draws from counter 2^63.
    _, u = oracle.rng_uniforms(5, 2, 2, search=True)
    z, phi = 2 * u[0] - 1, 2 * math.pi * u[1]
    r = math.sqrt(1 - z * z)
    np.testing.assert_allclose(followed[2], [r * math.cos(phi), r * math.sin(phi), z], atol=1e-15)
    assert counts[2] > 1


def test_steps_draw_the_same_numbers_whether_the_direction_was_given_or_found():
    seeds = [[5.0, 5.0, 5.0]]
    p1, c1, s1, f1, _ = run_test(3, seeds, None, step=0.3)
    p2, c2, s2, f2, _ = run_test(3, seeds, f1, step=0.3)
    assert c1[0] == c2[0] > 2 and np.array_equal(p1, p2) and np.array_equal(s1, s2)


def test_a_batch_does_not_depend_on_the_number_of_threads():
    rng = np.random.default_rng(0)
    seeds = rng.uniform(0, 9, size=(64, 3))
    a = oracle.track_batch(FIELD, EYE, None, seeds, None, np.arange(64), 9, 0.4, 200, True,
                           rule_params(3), n_threads=1)
    b = oracle.track_batch(FIELD, EYE, None, seeds, None, np.arange(64), 9, 0.4, 200, True,
                           rule_params(3), n_threads=7)
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)


# ---------------------------------------------------------------------------------------------
# MRtrix tckgen's tracks of the reference's synthetic FOD.

# tckgen accumulates positions in float32; the reference's own bar is 1e-4 mm.
POS_TOL = 1e-4


def _det_run(z, name, max_length):
    refs, k, seeds = det_reference(z, name)
    params = oracle.TrackParams(algorithm=0, cutoff=CUTOFF, max_angle_deg=30.0)
    pos, counts, status, _, _ = oracle.track_batch(
        z["fod"], np.linalg.inv(z["fod_affine"]), (z["mask"], np.linalg.inv(z["mask_affine"])),
        seeds, np.tile([1.0, 0.0, 0.0], (len(seeds), 1)), np.arange(len(seeds)), 1, STEP,
        max_steps_for(max_length, STEP), True, params)
    ours = split_ours(pos, counts, seeds)
    theirs = [halves(r, int(j)) for r, j in zip(refs, k)]
    return ours, theirs, status


def test_det_matches_tckgen_point_by_point():
    z = fixture()
    ours, theirs, status = _det_run(z, "det_ref", 80.0)
    for (a, b), (ra, rb) in zip(ours, theirs):
        for x, y in [(a, ra), (b, rb)]:
            assert len(x) == len(y)
            np.testing.assert_allclose(x, y, rtol=0, atol=POS_TOL)
    # Every kind of end occurs: the image's edge, the mask, a low FOD and a sharp turn.
    assert set(np.unique(status).tolist()) == {2, 3, 4, 5}


def test_at_the_maximum_length_tckgen_takes_more_steps():
    z = fixture()
    ours, theirs, status = _det_run(z, "det_ref_maxlen", 10.0)
    steps = math.ceil(10.0 / STEP)
    assert np.all(status == 1)
    assert {len(a) for a, _ in ours} == {steps + 1} and {len(b) for _, b in ours} == {1}
    assert {len(a) for a, _ in theirs} == {steps + 3} and {len(b) for _, b in theirs} == {2}
    for (a, _), (ra, _) in zip(ours, theirs):
        np.testing.assert_allclose(a, ra[: len(a)], rtol=0, atol=POS_TOL)


KS_C = math.sqrt(-math.log(0.001 / 2) / 2)


def _prob_run(z):
    rows = z["prob_ref_seeds"]
    seeds = rows[:, 2:5]
    params = oracle.TrackParams(algorithm=1, cutoff=CUTOFF, max_angle_deg=15.0, trials=1000,
                                attempts=1000)
    return oracle.track_batch(
        z["fod"], np.linalg.inv(z["fod_affine"]), (z["mask"], np.linalg.inv(z["mask_affine"])),
        seeds, np.tile([1.0, 0.0, 0.0], (len(seeds), 1)), np.arange(len(seeds)), 1, STEP,
        max_steps_for(80.0, STEP), True, params)


def test_prob_sampling_and_ends_are_tckgens():
    z = fixture()
    _, counts, status, _, st = _prob_run(z)
    assert np.all((counts - 1) * STEP >= 5.0)  # every seed gives a streamline above 5 mm
    # Ends: tckgen's "diffusion model sub-threshold" is NO_DIRECTION, the others leave the mask.
    assert set(np.unique(status).tolist()) <= {2, 3, 6}
    percents = z["prob_info_no_direction_percent"]
    ours = 100 * np.count_nonzero(status == 6) / status.size
    assert abs(ours - percents.mean()) < 4 * percents.std(ddof=1)
    # Calibration and samples per step within tckgen's run-to-run spread.
    assert st["calibration_directions"] == 7
    means = z["prob_info_mean_samples_per_step"]
    mean = st["calibration_directions"] + st["accepted_index_sum"] / st["sampled_steps"]
    assert abs(mean - means.mean()) < 4 * means.std(ddof=1)
    between = z["prob_info_steps_per_truncation"]
    largest = z["prob_info_max_truncation"]
    if st["truncations"]:
        assert st["sampled_steps"] / st["truncations"] > np.nanmin(between) / 2
        assert st["max_truncation"] < 1 + 2 * (np.nanmax(largest) - 1)


def test_prob_lengths_have_tckgens_distribution():
    z = fixture()
    pos, counts, _, _, _ = _prob_run(z)
    rows = z["prob_ref_seeds"]
    seeds = rows[:, 2:5]
    refs = [halves(p, int(k)) for p, k in zip(
        [z["prob_ref_points"][a:b].astype(np.float64)
         for a, b in zip(z["prob_ref_offsets"][:-1], z["prob_ref_offsets"][1:])],
        rows[:, 1])]
    ours = split_ours(pos, counts, seeds)

    def length(p):
        return float(np.linalg.norm(np.diff(p, axis=0), axis=1).sum())

    for half in (0, 1):
        a = np.array([length(h[half]) for h in ours])
        b = np.array([length(h[half]) for h in refs])
        d = stats.ks_2samp(a, b).statistic
        n, m = len(a), len(b)
        assert d < KS_C * math.sqrt((n + m) / (n * m))
