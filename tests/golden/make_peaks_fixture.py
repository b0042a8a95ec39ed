"""Writes tests/golden/peaks_fixture.npz from the reference's MRtrix sh2peaks fixture.

Source (data, not code): the reference repository's tests/data/peaks/fod.mif (8 x 8 x 8 voxels of
lmax-8 FOD coefficients, float32, tournier basis) and peaks_ref.mif (MRtrix 3.0.8
`sh2peaks fod.mif peaks_ref.mif -num 3`: three peak vectors per voxel, each scaled by its
amplitude).  The two images store their voxels in different orders and orientations; both are put
on the FOD's voxel grid here, matched through their image transforms.

Output: fod (512, 45) float64 and peaks (512, 3, 3) float64, row v = voxel v of the FOD grid in C
order.  Run: python tests/golden/make_peaks_fixture.py [reference root]
"""

import sys
from pathlib import Path

import numpy as np


def read_mif(path: Path):
    """(data with axes in image order, voxel sizes, 3 x 4 transform) of an uncompressed MIF."""
    raw = path.read_bytes()
    end = raw.index(b"\nEND\n")
    header = raw[:end].decode().splitlines()
    assert header[0] == "mrtrix image"
    fields: dict[str, list[str]] = {}
    for line in header[1:]:
        key, value = line.split(":", 1)
        fields.setdefault(key.strip(), []).append(value.strip())
    dims = [int(x) for x in fields["dim"][0].split(",")]
    vox = [float(x) for x in fields["vox"][0].split(",")]
    layout = [x.strip() for x in fields["layout"][0].split(",")]
    assert fields["datatype"][0] == "Float32LE"
    transform = np.array([[float(x) for x in row.split(",")] for row in fields["transform"]])
    name, offset = fields["file"][0].split()
    assert name == "."
    data = np.frombuffer(raw, dtype="<f4", offset=int(offset), count=int(np.prod(dims)))
    ranks = [abs(int(x)) for x in layout]
    reversed_axes = [x.startswith("-") for x in layout]
    slow_to_fast = sorted(range(len(dims)), key=lambda a: -ranks[a])
    arr = data.reshape([dims[a] for a in slow_to_fast])
    arr = arr.transpose([slow_to_fast.index(a) for a in range(len(dims))])
    for a, rev in enumerate(reversed_axes):
        if rev:
            arr = np.flip(arr, axis=a)
    return np.ascontiguousarray(arr), np.array(vox[:3]), transform


def main() -> None:
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    fod, fod_vox, fod_t = read_mif(root / "tests/data/peaks/fod.mif")
    peaks, pk_vox, pk_t = read_mif(root / "tests/data/peaks/peaks_ref.mif")
    shape = fod.shape[:3]
    ijk = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)
    world = ijk * fod_vox @ fod_t[:, :3].T + fod_t[:, 3]
    pk_ijk = np.rint(np.linalg.solve(pk_t[:, :3], (world - pk_t[:, 3]).T).T / pk_vox).astype(int)
    assert (pk_ijk >= 0).all() and (pk_ijk < np.array(peaks.shape[:3])).all()
    coeffs = fod.reshape(-1, fod.shape[3]).astype(np.float64)
    vectors = peaks[pk_ijk[:, 0], pk_ijk[:, 1], pk_ijk[:, 2]].reshape(-1, 3, 3).astype(np.float64)
    out = Path(__file__).resolve().parent / "peaks_fixture.npz"
    np.savez_compressed(out, fod=coeffs, peaks=vectors)
    print(out, coeffs.shape, vectors.shape)


if __name__ == "__main__":
    main()
