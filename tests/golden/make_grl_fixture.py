"""Writes tests/golden/grl_fixture.npz from the reference's MATLAB GRL fixture.

Source (data, not code): the reference repository's tests/data/grl_obj.mat (MATLAB run_grl.m's
inputs on a 6 x 6 x 6 block: b-values, gradient directions, shell weights, the image, the mask,
its low/high-resolution kernels and the calibrated dRL thresholds NN_L, NN_H) and
tests/data/deconv_obj.mat (its GRL results: FOD amplitudes on the 300 directions and the three
tissue fractions per voxel).  Read with scipy.io.loadmat, which executes nothing from the file.

Output: bvals (33,), bvecs (33, 3), shell_weight (33,), signals (v, 33) of the masked voxels in
C order of the block, nu_l, nu_h, lr_kernel (33, 35), hr_kernel (33, 302), fod (v, 300),
fractions (v, 3).  Run: python tests/golden/make_grl_fixture.py [reference root]
"""

import sys
from pathlib import Path

import numpy as np
from scipy.io import loadmat


def main() -> None:
    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    grl = loadmat(root / "tests/data/grl_obj.mat", simplify_cells=True)["grl_obj"]
    deconv = loadmat(root / "tests/data/deconv_obj.mat", simplify_cells=True)["deconv_obj"]
    voxels = np.argwhere(grl["mask"] == 1)
    xs, ys, zs = voxels[:, 0], voxels[:, 1], voxels[:, 2]
    out = Path(__file__).with_name("grl_fixture.npz")
    np.savez_compressed(
        out,
        bvals=np.asarray(grl["bvals"], dtype=np.float64),
        bvecs=np.asarray(grl["bvecs"], dtype=np.float64),
        shell_weight=np.asarray(grl["shell_weight"], dtype=np.float64),
        signals=grl["img"][xs, ys, zs].astype(np.float64),
        nu_l=np.float64(grl["NN_L"]),
        nu_h=np.float64(grl["NN_H"]),
        lr_kernel=np.asarray(grl["LRKernel"], dtype=np.float64),
        hr_kernel=np.asarray(grl["HRKernel"], dtype=np.float64),
        fod=deconv["fod"][xs, ys, zs].astype(np.float64),
        fractions=deconv["fractions"][xs, ys, zs].astype(np.float64),
    )
    print(out, len(voxels), "voxels")


if __name__ == "__main__":
    main()
