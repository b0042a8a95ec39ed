"""Converts the reference's MRIToolkit mFOD fixture to tests/golden/mfod_fixture.npz.

Source (data only, CC0 CFIN multi-shell crop; reference tests/data/README.md:70-71):
  tests/data/mfod/mfod_kernel.mat  gradients, shell weights, LR/HR kernels (single precision)
  tests/data/mfod/mfod_ref.mat     512 voxels: signals, FODs (600 columns), fractions
read with scipy.io.loadmat (no code from the reference is imported or run).

    python tests/golden/make_mfod_fixture.py /root/reference
"""

import sys
from pathlib import Path

import numpy as np
from scipy.io import loadmat


def main(reference: str) -> None:
    base = Path(reference) / "tests" / "data" / "mfod"
    kernel = loadmat(base / "mfod_kernel.mat", simplify_cells=True)["kernel"]
    ref = loadmat(base / "mfod_ref.mat", simplify_cells=True)["ref"]
    out = Path(__file__).resolve().parent / "mfod_fixture.npz"
    np.savez_compressed(
        out,
        bvals=np.asarray(kernel["bvals"], dtype=np.float64),
        bvecs=np.asarray(kernel["bvecs"], dtype=np.float64),
        shell_weight=np.asarray(kernel["shell_weight"], dtype=np.float64),
        lr_kernel=np.asarray(kernel["LRKernel"], dtype=np.float32),
        hr_kernel=np.asarray(kernel["HRKernel"], dtype=np.float32),
        recon_dirs=np.asarray(kernel["recon_dirs"], dtype=np.float64),
        signal=np.asarray(ref["signal"], dtype=np.float32),
        fod=np.asarray(ref["fod"], dtype=np.float32),
        fractions=np.asarray(ref["fractions"], dtype=np.float32),
        ijk=np.asarray(ref["ijk"], dtype=np.int16),
    )
    print(out, out.stat().st_size)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
