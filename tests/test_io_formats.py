"""The on-disk formats without a GPU: NIfTI and MIF against MRtrix-written files of the reference's
fixture (tests/golden/mif/, made by mrconvert from source.nii; the generator's closed form of the
source image pins the NIfTI reader), MIF round trips, gradient files and response files."""

import gzip
import math
from pathlib import Path

import numpy as np
import pytest

from larmor_amd.io import (load_dwi, load_volume, read_bvals_bvecs, read_mif, read_mif_gradients,
                           read_mif_header, read_mrtrix_response, save_volume, write_bvals_bvecs,
                           write_mif)
from larmor_amd.types import GradientTable, Volume

MIF = Path(__file__).resolve().parent / "golden" / "mif"


def _source_affine():
    # tests/data/make_mif_ref.py of the reference: axes swapped and flipped, rotated 0.3 rad.
    c, s = math.cos(0.3), math.sin(0.3)
    a = np.eye(4)
    a[:3, :3] = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]]) @ np.array(
        [[0, -3.0, 0], [2.0, 0, 0], [0, 0, -2.5]])
    a[:3, 3] = [10.5, -20.25, 5.0]
    return a


def test_nifti_source_is_the_generators_closed_form():
    v = load_volume(MIF / "source.nii")
    want = (np.arange(2 * 3 * 4 * 5) * 250 - 15000).astype(np.int16).reshape(2, 3, 4, 5)
    assert v.data.dtype == np.int16 and np.array_equal(v.data, want)
    np.testing.assert_allclose(v.affine, _source_affine(), atol=1e-5)
    m = load_volume(MIF / "mask.nii")
    assert np.array_equal(m.data, (np.arange(24) * 7 % 5 == 0).astype(np.uint8).reshape(2, 3, 4))


def _assert_same_world(a: Volume, b: Volume):
    """Every voxel of a holds b's value at the same world position."""
    assert a.data.shape[3:] == b.data.shape[3:]
    idx = np.indices(a.data.shape[:3]).reshape(3, -1).T.astype(np.float64)
    world = idx @ a.affine[:3, :3].T + a.affine[:3, 3]
    inv = np.linalg.inv(b.affine)
    vox = world @ inv[:3, :3].T + inv[:3, 3]
    assert np.abs(vox - np.rint(vox)).max() < 1e-4
    vox = np.rint(vox).astype(int)
    av = a.as_float(dtype=np.float64).reshape(-1, *a.data.shape[3:])
    bv = b.as_float(dtype=np.float64)[vox[:, 0], vox[:, 1], vox[:, 2]]
    np.testing.assert_array_equal(av, bv)


@pytest.mark.parametrize("name,source,dtype", [
    ("layout_permuted.mif", "source.nii", np.int16),
    ("float32be.mif", "source.nii", np.float32),
    ("int16be.mif.gz", "source.nii", np.int16),
    ("mask_bit.mif", "mask.nii", np.bool_),
    ("split.mih", "source.nii", np.int16),
    ("dwi.mif", "source.nii", np.int16),
])
def test_mrtrix_file_reads_as_its_source(name, source, dtype):
    v = read_mif(MIF / name)
    assert v.data.dtype == dtype
    want = load_volume(MIF / source)
    if dtype == np.bool_:
        want = want.with_data(want.data.astype(bool))
    _assert_same_world(v, want)
    h = read_mif_header(MIF / name)
    assert h.shape == v.data.shape and np.array_equal(h.affine, v.affine)


def test_dw_scheme_equals_the_conversion_of_the_fsl_files():
    mrtrix = read_mif_gradients(MIF / "dwi.mif")
    ours = read_bvals_bvecs(MIF / "dwi.bval", MIF / "dwi.bvec",
                            affine=load_volume(MIF / "source.nii").affine)
    np.testing.assert_allclose(mrtrix.bvals, ours.bvals, rtol=1e-9)
    np.testing.assert_allclose(mrtrix.bvecs, ours.bvecs, rtol=0, atol=1e-6)
    volume, gtab = load_dwi(MIF / "dwi.mif")
    assert volume.data.shape == (3, 2, 4, 5) and np.array_equal(gtab.bvals, mrtrix.bvals)


DTYPES = ["i1", "u1", "i2", "u2", "i4", "u4", "i8", "u8", "f4", "f8", "?"]


def _values(dtype, shape):
    rng = np.random.default_rng(0)
    if dtype == "?":
        return rng.random(shape) > 0.5
    info = np.iinfo(dtype) if np.dtype(dtype).kind in "iu" else None
    if info is not None:
        return rng.integers(info.min, info.max, size=shape, dtype=dtype, endpoint=True)
    return rng.normal(size=shape).astype(dtype)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(4, 3, 5), (4, 3, 5, 6)])
@pytest.mark.parametrize("name", ["a.mif", "a.mif.gz", "a.mih"])
def test_mif_round_trip(tmp_path, dtype, shape, name):
    affine = _source_affine()
    vol = Volume(_values(dtype, shape), affine)
    write_mif(vol, tmp_path / name)
    back = read_mif(tmp_path / name)
    assert back.data.dtype == vol.data.dtype and np.array_equal(back.data, vol.data)
    np.testing.assert_allclose(back.affine, affine, rtol=0, atol=1e-12)
    assert (back.slope, back.intercept) == (1.0, 0.0)


def test_mif_layout_offset_bits_scaling_and_gradients(tmp_path):
    data = _values("i2", (2, 3, 4, 5))
    gt = GradientTable([0, 1000, 2000, 1000, 3000],
                       np.vstack([[0, 0, 0], np.eye(3), [0.6, 0.8, 0.0]]))
    write_mif(Volume(data, np.eye(4), slope=0.5, intercept=2.5), tmp_path / "a.mif.gz",
              gradients=gt)
    raw = gzip.decompress((tmp_path / "a.mif.gz").read_bytes())
    end = raw.index(b"\nEND\n") + 5
    offset = int(raw[:end].decode().split("file: . ")[1].split("\n")[0])
    assert offset % 4 == 0 and raw[end:offset] == bytes(offset - end)
    assert raw[offset:] == data.astype("<i2").tobytes(order="F")
    back = read_mif(tmp_path / "a.mif.gz")
    assert (back.slope, back.intercept) == (0.5, 2.5) and np.array_equal(back.data, data)
    table = read_mif_gradients(tmp_path / "a.mif.gz")
    np.testing.assert_array_equal(table.bvals, gt.bvals)
    np.testing.assert_array_equal(table.bvecs, gt.bvecs)
    bits = np.zeros((10, 1, 1), bool)
    bits[[0, 2, 3, 9], 0, 0] = True
    write_mif(Volume(bits, np.eye(4)), tmp_path / "b.mif")
    raw = (tmp_path / "b.mif").read_bytes()
    assert raw[-2:] == bytes([0b10110000, 0b01000000])
    with pytest.raises(ValueError, match="a gradient table needs a 4-D image"):
        write_mif(Volume(data[..., 0], np.eye(4)), tmp_path / "c.mif", gradients=gt)


def test_fsl_gradients_round_trip_through_the_world_frame(tmp_path):
    affine = _source_affine()
    rng = np.random.default_rng(4)
    d = rng.normal(size=(6, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    gt = GradientTable(np.array([0, 1000, 1000, 2000, 2000, 3000.0]), d)
    write_bvals_bvecs(gt, tmp_path / "b.bval", tmp_path / "b.bvec", affine=affine)
    back = read_bvals_bvecs(tmp_path / "b.bval", tmp_path / "b.bvec", affine=affine)
    np.testing.assert_allclose(back.bvecs, gt.bvecs, atol=1e-14)
    world = read_bvals_bvecs(tmp_path / "b.bval", tmp_path / "b.bvec", frame="world")
    assert not np.allclose(world.bvecs, gt.bvecs)


def test_response_file_and_volume_dispatch(tmp_path):
    (tmp_path / "wm.txt").write_text("# response\n1.0 -0.5 0.1\n0.8, -0.3, 0.05\n")
    r = read_mrtrix_response(tmp_path / "wm.txt")
    assert r.shape == (2, 3) and r[1, 1] == -0.3
    (tmp_path / "bad.txt").write_text("1 2 3\n4 5\n")
    with pytest.raises(ValueError, match="line 2: 2 numbers, but the first row has 3"):
        read_mrtrix_response(tmp_path / "bad.txt")
    v = Volume(np.arange(24, dtype=np.float32).reshape(2, 3, 4), _source_affine())
    for name in ["v.nii", "v.nii.gz", "v.mif", "v.mif.gz"]:
        save_volume(v, tmp_path / name)
        assert np.array_equal(load_volume(tmp_path / name).data, v.data)
    with pytest.raises(ValueError, match="Unsupported volume file"):
        save_volume(v, tmp_path / "v.img")
