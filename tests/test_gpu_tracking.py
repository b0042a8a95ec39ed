"""GPU parity of streamline tracking (lsd_track_batch through _core.track_batch) against the CPU
oracle (oracle/track_oracle.cpp), and against MRtrix tckgen's tracks of the reference's synthetic
FOD (tests/golden/tracking_fixture.npz).

Bars: the fixed-rule algorithm and the step loop are bit-identical to the oracle (positions,
counts, ends, directions); det and prob call acos, atan2, sin and cos, which the device library
rounds differently from glibc, so they match the oracle to rounding: the same number of points and
ends on at least 99% of the streamlines (all of them on the tckgen fixture), and points within
1e-6 mm where the counts agree; against tckgen, the reference's own 1e-4 mm.
"""

import math

import numpy as np
import pytest

import oracle
from larmor_amd import _core
from tracking_cases import (CUTOFF, STEP, det_reference, fibre_field, fixture, halves,
                            max_steps_for, random_seeds, split_ours)

pytestmark = pytest.mark.gpu

POINT_TOL = 1e-6  # mm, GPU against the oracle where the counts agree
POS_TOL = 1e-4    # mm, against tckgen (the reference's bar)


def _algo(kind, **kw):
    if kind == "det":
        return (_core.DetParams(kw.get("cutoff", CUTOFF), kw.get("angle", 30.0)),
                oracle.TrackParams(algorithm=0, full=int(kw.get("full", False)),
                                   cutoff=kw.get("cutoff", CUTOFF),
                                   max_angle_deg=kw.get("angle", 30.0)))
    if kind == "prob":
        return (_core.ProbParams(kw.get("cutoff", CUTOFF), kw.get("angle", 15.0),
                                 kw.get("trials", 1000), kw.get("attempts", 1000)),
                oracle.TrackParams(algorithm=1, full=int(kw.get("full", False)),
                                   cutoff=kw.get("cutoff", CUTOFF),
                                   max_angle_deg=kw.get("angle", 15.0),
                                   trials=kw.get("trials", 1000),
                                   attempts=kw.get("attempts", 1000)))
    rule = {"straight": 0, "refuse": 1, "quarter_turn": 2, "wander": 3}[kind]
    stop = kw.get("stop")
    past = kw.get("past_x", float("inf"))
    return (_core.TestAlgorithm(kind, stop, past),
            oracle.TrackParams(algorithm=2, test_rule=rule, test_stop=stop or 0,
                               test_refuses_past_x=past, max_angle_deg=180.0))


def _both(fod, affine, mask, seeds, dirs, idx, run_seed, step, max_steps, bidir, kind,
          full=False, **kw):
    gpu_algo, ref_algo = _algo(kind, full=full, **kw)
    inv = np.linalg.inv(affine)
    positions = np.zeros((len(seeds), max_steps + 1, 3), dtype=np.float32)
    g = _core.track_batch(fod, inv, mask, seeds, dirs, idx, run_seed, step, max_steps, bidir,
                          gpu_algo, positions, 1, full=full)
    r = oracle.track_batch(fod, inv, mask, seeds, dirs, idx, run_seed, step, max_steps, bidir,
                           ref_algo)
    return (positions, *g), r


def _written(positions, counts):
    keep = np.arange(positions.shape[1])[None, :, None] < counts[:, None, None]
    return np.where(keep, positions, 0.0)


def _assert_identical(g, r):
    pos, counts, status, followed, stats = g
    rpos, rcounts, rstatus, rfollowed, rstats = r
    assert np.array_equal(counts, rcounts)
    assert np.array_equal(status, rstatus)
    assert np.array_equal(followed, rfollowed)
    assert np.array_equal(_written(pos, counts), _written(rpos, rcounts))
    assert stats == rstats


def _agreement(g, r):
    """(share of streamlines with the same counts and ends, worst point distance among them)."""
    pos, counts, status, _, _ = g
    rpos, rcounts, rstatus, _, _ = r
    same = (counts == rcounts) & np.all(status == rstatus, axis=1)
    worst = 0.0
    for i in np.flatnonzero(same):
        if counts[i]:
            worst = max(worst, float(np.abs(pos[i, :counts[i]].astype(np.float64)
                                            - rpos[i, :counts[i]]).max()))
    return float(same.mean()), worst


# ---------------------------------------------------------------------------------------------
# The step loop (fixed-rule algorithm): bit-identical.

GRID = np.zeros((14, 12, 10, 1), dtype=np.float32)


def _oblique(scale=1.5):
    a = np.eye(4)
    t = math.radians(25.0)
    a[:3, :3] = scale * np.array([[math.cos(t), -math.sin(t), 0], [math.sin(t), math.cos(t), 0],
                                  [0, 0, 1.0]])
    a[:3, 3] = [-3.0, 2.0, 1.5]
    return a


@pytest.mark.parametrize("kind,kw", [("straight", {}), ("quarter_turn", {}), ("wander", {}),
                                     ("refuse", {"stop": 5}), ("wander", {"past_x": 4.0})])
@pytest.mark.parametrize("bidir", [True, False])
@pytest.mark.parametrize("given", [True, False])
def test_step_loop_is_bit_identical(kind, kw, bidir, given):
    affine = _oblique()
    mask = np.ones((10, 9, 8), dtype=np.float32)
    mask[2:4, :, :] = 0.0
    mask_affine = affine.copy()
    mask_affine[:3, :3] *= 1.3
    rng = np.random.default_rng(11)
    seeds = random_seeds(affine, GRID.shape[:3], 700, seed=5, margin=-1.5)
    dirs = rng.normal(size=(700, 3)) if given else None
    if dirs is not None:
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    idx = rng.permutation(10_000)[:700].astype(np.int64)
    g, r = _both(GRID, affine, (mask, np.linalg.inv(mask_affine)), seeds, dirs, idx, 77, 0.37,
                 150, bidir, kind, **kw)
    assert (g[1] > 1).sum() > 100
    if given:
        _assert_identical(g, r)
        return
    # The first-direction search draws its unit vector through cos and sin of a drawn azimuth:
    # the device library's rounding, so the directions agree to rounding and the steps follow.
    assert np.array_equal(g[1], r[1]) and np.array_equal(g[2], r[2])
    np.testing.assert_allclose(g[3], r[3], rtol=0, atol=1e-15)
    share, worst = _agreement(g, r)
    assert share == 1.0 and worst <= 1e-9


def test_step_loop_ends_at_the_maximum_length_and_spreads_few_seeds():
    affine = np.eye(4)
    seeds = np.array([[5.0, 5.0, 5.0], [6.0, 6.0, 6.0], [-2.0, 0.0, 0.0]])
    dirs = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [1.0, 0.0, 0.0]])
    g, r = _both(GRID, affine, None, seeds, dirs, np.arange(3), 1, 1.0, 3, True, "straight")
    _assert_identical(g, r)
    assert list(g[1]) == [4, 4, 0] and g[2][0].tolist() == [1, 1]


# ---------------------------------------------------------------------------------------------
# det and prob on the tckgen fixture.


def test_det_matches_the_oracle_and_tckgen_on_the_fixture():
    z = fixture()
    for name, max_length in [("det_ref", 80.0), ("det_ref_maxlen", 10.0)]:
        refs, k, seeds = det_reference(z, name)
        mask = (z["mask"], np.linalg.inv(z["mask_affine"]))
        g, r = _both(z["fod"], z["fod_affine"], mask, seeds, np.tile([1.0, 0, 0], (32, 1)),
                     np.arange(32), 1, STEP, max_steps_for(max_length, STEP), True, "det")
        share, worst = _agreement(g, r)
        print(f"{name}: same counts and ends {share:.4f}, worst point distance {worst:.2e} mm")
        assert share == 1.0 and worst <= POINT_TOL
        ours = split_ours(g[0], g[1], seeds)
        theirs = [halves(t, int(j)) for t, j in zip(refs, k)]
        for (a, b), (ra, rb) in zip(ours, theirs):
            for x, y in [(a, ra), (b, rb)]:
                n = min(len(x), len(y))
                np.testing.assert_allclose(x[:n], y[:n], rtol=0, atol=POS_TOL)


def test_prob_matches_the_oracle_on_the_fixture():
    z = fixture()
    rows = z["prob_ref_seeds"]
    seeds = rows[:, 2:5]
    mask = (z["mask"], np.linalg.inv(z["mask_affine"]))
    g, r = _both(z["fod"], z["fod_affine"], mask, seeds, np.tile([1.0, 0, 0], (len(seeds), 1)),
                 np.arange(len(seeds)), 1, STEP, max_steps_for(80.0, STEP), True, "prob")
    share, worst = _agreement(g, r)
    print(f"prob: same counts and ends {share:.4f}, worst point distance {worst:.2e} mm; "
          f"stats {g[4]} vs {r[4]}")
    assert share >= 0.99 and worst <= POINT_TOL
    assert g[4]["calibration_directions"] == r[4]["calibration_directions"] == 7
    # The sampling counts are sums over the streamlines: equal where every streamline is.
    if share == 1.0:
        for key in ("sampled_steps", "accepted_index_sum", "truncations"):
            assert g[4][key] == r[4][key]


# ---------------------------------------------------------------------------------------------
# det and prob on a synthetic field, with the first-direction search, symmetric and full.


@pytest.mark.parametrize("kind", ["det", "prob"])
# lmax 14 and 20 symmetric and 10 full run the lmax-20 instantiations (series in private memory).
@pytest.mark.parametrize("lmax,full", [(8, False), (4, False), (10, False), (6, True),
                                       (14, False), (20, False), (10, True)])
def test_synthetic_field_matches_the_oracle(kind, lmax, full):
    fod, affine = fibre_field(lmax=lmax, full=full)
    seeds = random_seeds(affine, fod.shape[:3], 600, seed=lmax + 10 * full)
    mask = np.ones(fod.shape[:3], dtype=np.float32)
    mask[:, :, -1] = 0.0
    kw = {"cutoff": 0.1}
    g, r = _both(fod, affine, (mask, np.linalg.inv(affine)), seeds, None, np.arange(600) * 3 + 1,
                 2024, 0.5, 300, True, kind, full=full, **kw)
    share, worst = _agreement(g, r)
    started = int((g[1] > 0).sum())
    print(f"{kind} lmax {lmax} full {full}: {started} started, same counts and ends "
          f"{share:.4f}, worst {worst:.2e} mm")
    assert started > 300
    assert share >= 0.99 and worst <= POINT_TOL


def test_results_do_not_depend_on_the_batch():
    fod, affine = fibre_field(lmax=8)
    seeds = random_seeds(affine, fod.shape[:3], 3000, seed=4)
    idx = np.arange(3000, dtype=np.int64) + 500
    inv = np.linalg.inv(affine)
    algo = _core.ProbParams(0.1, 20.0)
    full_pos = np.zeros((3000, 201, 3), dtype=np.float32)
    c1, s1, f1, _ = _core.track_batch(fod, inv, None, seeds, None, idx, 9, 0.5, 200, True, algo,
                                      full_pos, 1)
    part = np.arange(0, 3000, 7)
    part_pos = np.zeros((len(part) + 5, 201, 3), dtype=np.float32)
    c2, s2, f2, _ = _core.track_batch(fod, inv, None, seeds[part], None, idx[part], 9, 0.5, 200,
                                      True, algo, part_pos, 1)
    assert np.array_equal(c1[part], c2) and np.array_equal(s1[part], s2)
    assert np.array_equal(f1[part], f2)
    assert np.array_equal(_written(full_pos[part], c1[part]), _written(part_pos[: len(part)], c2))
    assert not part_pos[len(part):].any()  # rows past the batch untouched


def test_large_batch_agrees_with_the_oracle_on_a_sample():
    fod, affine = fibre_field(shape=(40, 40, 30), lmax=8, seed=3)
    n = 200_000
    seeds = random_seeds(affine, fod.shape[:3], n, seed=8, margin=0.0)
    idx = np.arange(n, dtype=np.int64)
    inv = np.linalg.inv(affine)
    positions = np.zeros((n, 201, 3), dtype=np.float32)
    counts, status, _, _ = _core.track_batch(fod, inv, None, seeds, None, idx, 3, 0.5, 200, True,
                                             _core.DetParams(0.1, 45.0), positions, 1)
    sample = np.random.default_rng(0).choice(n, 1500, replace=False)
    r = oracle.track_batch(fod, inv, None, seeds[sample], None, idx[sample], 3, 0.5, 200, True,
                           oracle.TrackParams(algorithm=0, cutoff=0.1, max_angle_deg=45.0))
    same = (counts[sample] == r[1]) & np.all(status[sample] == r[2], axis=1)
    print(f"200k det seeds: {int((counts > 0).sum())} started, mean points "
          f"{counts[counts > 0].mean():.1f}; sample agreement {same.mean():.4f}")
    assert same.mean() >= 0.99
    i = sample[same][counts[sample][same] > 0]
    j = np.flatnonzero(same)[counts[sample][same] > 0]
    worst = max(float(np.abs(positions[a, :counts[a]].astype(np.float64)
                             - r[0][b, :counts[a]]).max()) for a, b in zip(i, j))
    assert worst <= POINT_TOL


def test_gpu_limits_raise():
    fod = np.zeros((4, 4, 4, 105), dtype=np.float32)  # lmax 13 is odd for a symmetric series
    seeds = np.zeros((1, 3))
    pos = np.zeros((1, 11, 3), dtype=np.float32)
    with pytest.raises(ValueError, match="lmax must be at most 20, and even unless full"):
        _core.track_batch(fod, np.eye(4), None, seeds, None, np.zeros(1, np.int64), 1, 1.0, 10,
                          True, _core.DetParams(0.1, 30.0), pos, 1)
    fod = np.zeros((4, 4, 4, 276), dtype=np.float32)  # lmax 22
    with pytest.raises(ValueError, match="lmax must be at most 20"):
        _core.track_batch(fod, np.eye(4), None, seeds, None, np.zeros(1, np.int64), 1, 1.0, 10,
                          True, _core.DetParams(0.1, 30.0), pos, 1)


# ---------------------------------------------------------------------------------------------
# The Tracker end to end (the reference's tests/tractography/test_det_tckgen.py and
# test_prob_tckgen.py drive it the same way).


def _fixture_volumes():
    from larmor_amd.types import Volume

    z = fixture()
    return z, Volume(z["fod"], z["fod_affine"]), Volume(z["mask"], z["mask_affine"])


def test_tracker_det_matches_tckgen():
    from larmor_amd.tractography import EndStatus, Seeds, Tracker

    z, fod, mask = _fixture_volumes()
    refs, k, seeds = det_reference(z, "det_ref")
    run = Tracker("det", step_size=STEP, max_angle=30.0, cutoff=CUTOFF, min_length=5.0,
                  max_length=80.0).track(fod, None, Seeds(seeds, np.tile([1.0, 0, 0], (32, 1))),
                                         mask=mask, seed=1, n_jobs=1)
    t = run.tractogram
    assert run.num_streamlines == 32 and run.num_seeds == 32
    assert run.num_rejected == {"no_start": 0, "too_short": 0}
    for i, (r, j) in enumerate(zip(refs, k)):
        p = t.points[t.offsets[i]:t.offsets[i + 1]].astype(np.float64)
        at = int(np.flatnonzero(np.all(p == seeds[i].astype(np.float32), axis=1))[0])
        for x, y in zip(halves(p, at), halves(r, int(j))):
            assert len(x) == len(y)
            np.testing.assert_allclose(x, y, rtol=0, atol=POS_TOL)
    assert set(np.unique(t.streamline_data["end_status"]).tolist()) == {
        EndStatus.LEFT_VOLUME, EndStatus.LEFT_MASK, EndStatus.LOW_FOD, EndStatus.SHARP_TURN}
    assert run.parameters["max_steps"] == 400


def test_tracker_prob_sampling_is_tckgens():
    from larmor_amd.tractography import Seeds, Tracker

    z, fod, mask = _fixture_volumes()
    seeds = z["prob_ref_seeds"][:, 2:5]
    run = Tracker("prob", cutoff=CUTOFF, min_length=5.0, max_length=80.0).track(
        fod, None, Seeds(seeds, np.tile([1.0, 0, 0], (len(seeds), 1))), mask=mask, seed=1)
    assert run.num_streamlines == len(seeds)
    s = run.sampling
    means = z["prob_info_mean_samples_per_step"]
    assert s["calibration_directions"] == 7
    assert abs(s["mean_samples_per_step"] - means.mean()) < 4 * means.std(ddof=1)


def test_tracker_repeats_a_run_from_its_seed_file(tmp_path):
    from larmor_amd.io.tck import read_tck
    from larmor_amd.tractography import MaskSeeds, Tracker, read_seeds
    from larmor_amd.types import Volume

    fod_data, affine = fibre_field(lmax=8)
    fod = Volume(fod_data, affine)
    mask = Volume(np.ones(fod_data.shape[:3], np.float32), affine)
    tracker = Tracker("prob", step_size=0.5, max_angle=20.0, cutoff="auto", min_length=3.0,
                      max_length=40.0)
    run = tracker.track(fod, tmp_path / "t.tck", MaskSeeds(mask), num_streamlines=700,
                        mask=mask, seed=12345, output_seeds=tmp_path / "s.txt")
    assert run.num_streamlines == 700
    assert run.num_seeds == 700 + sum(run.num_rejected.values())
    first = read_tck(tmp_path / "t.tck")
    assert len(first) == 700
    seeds = read_seeds(tmp_path / "s.txt")
    assert len(seeds) == 700 and seeds.indices is not None and seeds.directions is not None
    again = tracker.track(fod, None, seeds, mask=mask, seed=12345)
    t = again.tractogram
    assert np.array_equal(t.offsets, first.offsets)
    assert np.array_equal(t.points, first.points)


def test_prob_near_the_z_axis_keeps_the_oracles_distribution():
    """On a field whose bundles run near +-z, prob's sampling frame (rotate_to, the rotation that
    takes +z to the previous direction) turns fast with the direction, so paths are hypersensitive
    there: a 1e-15 relative change of a start direction alters about 5-9% of the oracle's own
    streamlines.  The device's last-bit differences in sin and cos do the same, so the GPU is held
    to the oracle's distribution instead: lengths by a two-sample Kolmogorov-Smirnov test at
    alpha 0.001 and mean lengths within 4 standard errors.  det, whose climb contracts onto the
    peak, keeps its paths."""
    from scipy import stats

    from larmor_amd import synthetic

    fod, affine, mask = synthetic.tracking_field(shape=(48, 48, 30))
    mask_f = mask.astype(np.float32)
    inv = np.linalg.inv(affine)
    seeds = random_seeds(affine, fod.shape[:3], 1500, seed=21, margin=8.0)
    for kind, kw in [("prob", {"cutoff": 0.1}), ("det", {"cutoff": 0.1, "angle": 8.8})]:
        g, r = _both(fod, affine, (mask_f, inv), seeds, None, np.arange(1500), 5, 0.2, 500,
                     True, kind, **kw)
        share, worst = _agreement(g, r)
        a, b = g[1][g[1] > 0], r[1][r[1] > 0]
        d = stats.ks_2samp(a, b).statistic
        crit = math.sqrt(-math.log(0.001 / 2) / 2) * math.sqrt((len(a) + len(b)) / (len(a) * len(b)))
        se = math.sqrt(a.var() / len(a) + b.var() / len(b))
        print(f"{kind} near z: same paths {share:.4f}, KS {d:.3f} < {crit:.3f}, mean points "
              f"{a.mean():.1f} vs {b.mean():.1f}")
        assert d < crit and abs(a.mean() - b.mean()) < 4 * se
        if kind == "det":
            assert share >= 0.99 and worst <= POINT_TOL
