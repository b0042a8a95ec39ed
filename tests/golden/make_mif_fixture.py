"""Copies the reference's MRtrix-written image fixture (data, not code) into tests/golden/mif/.

Source: the reference repository's tests/data/mif/: source.nii and mask.nii (NIfTI images with
swapped, flipped, slightly oblique axes), the MIF/MIH files MRtrix 3.0.8's mrconvert wrote from
them (permuted layout, float32 big endian, gzip int16 big endian, a bit mask, a split header and
data file, and a DWI with its FSL gradient files as dw_scheme), and the FSL gradient files.

Run: python tests/golden/make_mif_fixture.py [reference root]
"""

import shutil
import sys
from pathlib import Path

FILES = ["source.nii", "mask.nii", "layout_permuted.mif", "float32be.mif", "int16be.mif.gz",
         "mask_bit.mif", "split.mih", "split.dat", "dwi.mif", "dwi.bval", "dwi.bvec"]


def main(reference: Path) -> None:
    src = reference / "tests" / "data" / "mif"
    dest = Path(__file__).resolve().parent / "mif"
    dest.mkdir(exist_ok=True)
    for name in FILES:
        shutil.copyfile(src / name, dest / name)
    print(f"copied {len(FILES)} files to {dest}")


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference"))
