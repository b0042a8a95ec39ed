"""Converts the reference's white-matter-response and FOD-merge fixtures to tests/golden/.

Sources (data only; reference tests/data/README.md:61, 69, 72):
  tests/data/dki/dwi.nii.gz, bvals, bvecs  the 8x8x8 CFIN block, 133 volumes, float32 NIfTI-1
  tests/data/wm_rf/wm_rf_ref.mat           MRIToolkit's per-voxel tensor fit of the block and the
                                           white matter response (MATLAB R2025a)
  tests/data/mfod/mfod_merge_ref.mat       MRIToolkit's two mFOD FODs, fractions and merged FODs
read with gzip + numpy (the NIfTI-1 header fields the block needs) and scipy.io.loadmat; no code
from the reference is imported or run.

    python tests/golden/make_wm_rf_fixture.py /root/reference
"""

import gzip
import struct
import sys
from pathlib import Path

import numpy as np
from scipy.io import loadmat

NIFTI_DTYPES = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64}


def read_nifti(path: Path) -> np.ndarray:
    """The data array of a little-endian NIfTI-1 file, scaled by scl_slope/scl_inter."""
    raw = gzip.decompress(path.read_bytes())
    if struct.unpack("<i", raw[:4])[0] != 348:
        raise ValueError(f"{path} is not a little-endian NIfTI-1 file")
    dims = struct.unpack("<8h", raw[40:56])
    datatype = struct.unpack("<h", raw[70:72])[0]
    offset = int(struct.unpack("<f", raw[108:112])[0])
    slope, inter = struct.unpack("<2f", raw[112:120])
    shape = dims[1:1 + dims[0]]
    dtype = np.dtype(NIFTI_DTYPES[datatype]).newbyteorder("<")
    count = int(np.prod(shape))
    data = np.frombuffer(raw, dtype=dtype, count=count, offset=offset).reshape(shape, order="F")
    if slope not in (0.0, 1.0) or inter != 0.0:
        data = data * np.float32(slope) + np.float32(inter)
    return np.ascontiguousarray(data)


def main(reference: str) -> None:
    data = Path(reference) / "tests" / "data"
    dwi = read_nifti(data / "dki" / "dwi.nii.gz").astype(np.float32)
    bvals = np.loadtxt(data / "dki" / "bvals").ravel()
    bvecs = np.loadtxt(data / "dki" / "bvecs")
    wm = loadmat(data / "wm_rf" / "wm_rf_ref.mat", simplify_cells=True)["ref"]
    merge = loadmat(data / "mfod" / "mfod_merge_ref.mat", simplify_cells=True)["ref"]
    out = Path(__file__).resolve().parent / "wm_rf_fixture.npz"
    np.savez_compressed(
        out,
        dwi=dwi,
        bvals=bvals,
        bvecs=bvecs if bvecs.shape[1] == 3 else bvecs.T,
        mask=np.asarray(wm["mask"], dtype=np.uint8),
        voxel_e_vals=np.asarray(wm["voxel"]["e_vals"], dtype=np.float64),
        voxel_k=np.asarray(wm["voxel"]["k"], dtype=np.float64),
        voxel_fitted=np.asarray(wm["voxel"]["fitted"], dtype=np.uint8),
        correct_lambdas=np.asarray(wm["correct"]["lambdas"], dtype=np.float64),
        correct_n=np.int64(wm["correct"]["n"]),
        correct_k=np.float64(wm["correct"]["K"]),
        single_lambdas=np.asarray(wm["single"]["correct"]["lambdas"], dtype=np.float64),
        single_n=np.int64(wm["single"]["correct"]["n"]),
        merge_fod1=np.asarray(merge["fod1"], dtype=np.float64),
        merge_fod2=np.asarray(merge["fod2"], dtype=np.float64),
        merge_fractions=np.asarray(merge["fractions"], dtype=np.float64),
        merge_linear=np.asarray(merge["linear"], dtype=np.float64),
        merge_majority=np.asarray(merge["majority"], dtype=np.float64),
    )
    print(out, out.stat().st_size)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
