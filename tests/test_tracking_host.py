"""Host side of tracking, no GPU: the reference's argument checks and error texts (Tracker,
DetParams, ProbParams, TestAlgorithm, Seeds), seed draws and seed files, batching arithmetic, and
the TCK and NIfTI round trips the tracker and the fixture conversion use."""

import math

import numpy as np
import pytest

from larmor_amd import _core
from larmor_amd.io.nifti import read_nifti, write_nifti
from larmor_amd.io.tck import TckWriter, read_tck, write_tck
from larmor_amd.tractography import (EndStatus, MaskSeeds, Seeds, Tracker, read_seeds,
                                     seeds_per_voxel, write_seeds)
from larmor_amd.tractography import tracker as tracker_mod
from larmor_amd.types import Tractogram, Volume


def test_end_status_codes():
    assert [int(s) for s in EndStatus] == list(range(7))
    assert EndStatus.LOW_FOD == 4 and EndStatus.NO_DIRECTION == 6


@pytest.mark.parametrize("kwargs,message", [
    ({"step_size": 0.0}, "The step size must be positive and finite, got 0.0."),
    ({"min_length": -1.0}, "The min length must be positive and finite"),
    ({"max_length": float("inf")}, "The max length must be positive and finite"),
    ({"min_length": 5.0, "max_length": 5.0}, "Cannot track with min length >= max length."),
    ({"max_angle": 0.0}, r"Expected max angle to be in \(0, 180\], got 0.0"),
    ({"max_angle": 181.0}, r"Expected max angle to be in \(0, 180\]"),
    ({"cutoff": "half"}, "Expected a number or 'auto' as cutoff, got 'half'."),
    ({"cutoff": -0.1}, "The cutoff must be non-negative and finite, got -0.1."),
])
def test_tracker_refuses_bad_settings(kwargs, message):
    with pytest.raises(ValueError, match=message):
        Tracker("det", **kwargs)


def test_tracker_refuses_an_unknown_algorithm():
    with pytest.raises(ValueError, match="Unknown algorithm 'ifod2'; the known ones are: det, prob."):
        Tracker("ifod2")


def test_default_angles():
    det = Tracker("det")
    # The tightest curve at the default step (0.1 voxel) has radius 0.65 voxel: 2 asin(0.1/1.3).
    assert det._max_angle(0.2, 2.0) == pytest.approx(math.degrees(2 * math.asin(0.2 / 2.6)))
    assert det._max_angle(3.0, 1.0) == 180.0
    assert Tracker("prob")._max_angle(0.2, 2.0) == 15
    assert Tracker("prob", max_angle=20.0)._max_angle(0.2, 2.0) == 20.0


def test_algorithm_parameter_errors_match_the_binding():
    with pytest.raises(ValueError, match="the cutoff must be non-negative and finite, got -1"):
        _core.DetParams(-1.0, 30.0)
    with pytest.raises(ValueError, match="the cutoff must be non-negative and finite, got NaN"):
        _core.ProbParams(float("nan"), 30.0)
    with pytest.raises(ValueError, match=r"the maximum angle must be in \(0, 180\], got 0"):
        _core.DetParams(0.1, 0.0)
    with pytest.raises(ValueError, match="above 0.001 rad"):
        _core.ProbParams(0.1, 0.05)
    with pytest.raises(ValueError, match="trials and attempts must be at least 1, got 0 and 5"):
        _core.ProbParams(0.1, 15.0, 0, 5)
    with pytest.raises(ValueError, match="expected straight, quarter_turn, wander, or refuse"):
        _core.TestAlgorithm("refuse")
    with pytest.raises(ValueError, match="a stop is an EndStatus from 1 to 6, got 7"):
        _core.TestAlgorithm("refuse", 7)
    with pytest.raises(OverflowError):
        _core.TestAlgorithm("refuse", 300)
    with pytest.raises(TypeError, match="expected DetParams, ProbParams or TestAlgorithm"):
        _core.track_batch(np.zeros((2, 2, 2, 1), np.float32), np.eye(4), None, np.zeros((1, 3)),
                          None, np.zeros(1, np.int64), 1, 1.0, 5, True, "det",
                          np.zeros((1, 6, 3), np.float32), 1)


def test_seeds_check_and_normalise():
    s = Seeds(np.zeros((2, 3)), [[2.0, 0.0, 0.0], [0.0, 0.0, 1.0]], [3, 4])
    assert np.array_equal(s.directions, [[1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    assert s.indices.dtype == np.int64 and len(s) == 2
    with pytest.raises(ValueError, match=r"Expected positions of shape \(N, 3\)"):
        Seeds(np.zeros((2, 2)))
    with pytest.raises(ValueError, match="Found infinite/nan position."):
        Seeds([[np.nan, 0, 0]])
    with pytest.raises(ValueError, match="Found a zero direction"):
        Seeds(np.zeros((1, 3)), np.zeros((1, 3)))
    with pytest.raises(ValueError, match="Found negative indices"):
        Seeds(np.zeros((1, 3)), None, [-1])


def test_mask_seeds_draw_inside_their_voxels():
    data = np.zeros((5, 6, 7), np.float32)
    data[1, 2, 3] = data[4, 0, 6] = 1.0
    affine = np.diag([2.0, 2.0, 2.0, 1.0])
    affine[:3, 3] = [1.0, -2.0, 3.0]
    seeds = MaskSeeds(Volume(data, affine))
    assert np.array_equal(seeds.voxels, [[1, 2, 3], [4, 0, 6]])
    drawn = seeds.draw(500, np.random.default_rng(0))
    vox = (drawn.positions - affine[:3, 3]) / 2.0
    assert np.all(np.abs(vox - np.rint(vox)) <= 0.5)
    assert {tuple(v) for v in np.rint(vox).astype(int)} == {(1, 2, 3), (4, 0, 6)}
    # The same generator state draws the same seeds.
    again = seeds.draw(500, np.random.default_rng(0))
    assert np.array_equal(drawn.positions, again.positions)
    per = seeds_per_voxel(Volume(data, affine), 3, 7)
    assert len(per) == 6


def test_seed_files_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    s = Seeds(rng.normal(size=(5, 3)), rng.normal(size=(5, 3)), [2**60, 1, 2, 3, 4])
    write_seeds(s, tmp_path / "s.txt")
    back = read_seeds(tmp_path / "s.txt")
    assert np.array_equal(back.positions, s.positions)
    assert np.array_equal(back.directions, s.directions)
    assert np.array_equal(back.indices, s.indices)
    write_seeds(Seeds(s.positions), tmp_path / "p.txt")
    assert read_seeds(tmp_path / "p.txt").indices is None


def test_draw_size_and_keep_follow_the_reference_batching():
    prep = tracker_mod._Prepared(coeffs=None, full=False, fod_inv=None, mask=None, step_size=0.5,
                                 min_length=2.0, max_length=50.0, max_angle=30.0,
                                 min_radius=1.0, cutoff=0.1, params=None, max_steps=100,
                                 batch_size=1000)
    totals = tracker_mod._Counts()
    assert tracker_mod._draw_size(prep, totals, None) == 1000
    assert tracker_mod._draw_size(prep, totals, 10) == 256
    assert tracker_mod._draw_size(prep, totals, 200) == 800
    counts = np.array([0, 3, 10, 7, 1, 12])
    keep = tracker_mod._keep(counts, prep, 2, totals)
    # Rows 2 and 3 are kept ((count - 1) * 0.5 >= 2); the seeds after row 3 count as not drawn.
    assert list(keep) == [2, 3]
    assert (totals.seeds, totals.accepted, totals.no_start, totals.too_short) == (4, 2, 1, 1)
    assert tracker_mod.MAX_BATCH == 20_000 and tracker_mod.BUFFER_BYTES == 128 * 2**20


def test_tck_round_trip_and_batches(tmp_path):
    rng = np.random.default_rng(2)
    t = Tractogram(rng.normal(size=(9, 3)), [0, 4, 4, 9])
    write_tck(t, tmp_path / "a.tck", {"step_size": 0.2})
    back = read_tck(tmp_path / "a.tck")
    assert np.array_equal(back.points, t.points) and np.array_equal(back.offsets, t.offsets)
    with TckWriter(tmp_path / "b.tck") as w:
        w.write(Tractogram(t.points[:4], [0, 4]))
        w.write(Tractogram(t.points[4:], [0, 0, 5]))
    back = read_tck(tmp_path / "b.tck")
    assert np.array_equal(back.points, t.points) and np.array_equal(back.offsets, t.offsets)


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.float32, np.float64])
@pytest.mark.parametrize("suffix", [".nii", ".nii.gz"])
def test_nifti_round_trip(tmp_path, dtype, suffix):
    rng = np.random.default_rng(3)
    data = (rng.random((4, 5, 6, 2)) * 100).astype(dtype)
    t = math.radians(20)
    affine = np.eye(4)
    affine[:3, :3] = np.array([[math.cos(t), -math.sin(t), 0], [math.sin(t), math.cos(t), 0],
                               [0, 0, 1]]) @ np.diag([1.5, 2.0, -2.5])
    affine[:3, 3] = [10.0, -20.0, 5.0]
    write_nifti(Volume(data, affine, slope=2.0, intercept=1.0), tmp_path / f"v{suffix}")
    back = read_nifti(tmp_path / f"v{suffix}")
    assert back.data.dtype == dtype and np.array_equal(back.data, data)
    np.testing.assert_allclose(back.affine, affine, atol=1e-5)
    assert (back.slope, back.intercept) == (2.0, 1.0)
