"""Writes tests/golden/tracking_fixture.npz from the reference's tckgen tracking fixture.

Source (data, not code), the reference repository's tests/data/tracking/:
  fod.nii.gz, mask.nii.gz   a synthetic lmax-8 FOD (16 x 16 x 5 voxels of 2 mm, x flipped,
                            tournier basis, float32) and its tracking mask;
  det_ref.tck, det_ref_maxlen.tck and their *_seeds.txt
                            MRtrix 3.0.8 `tckgen -algorithm SD_Stream` from 32 grid seeds along +x
                            (step 0.2 mm, angle 30, cutoff 0.3, min length 5, max length 80 / 10);
  prob_ref.tck, prob_ref_seeds.txt, prob_ref_info.json
                            `tckgen -algorithm iFOD1` from 256 grid seeds (defaults: step 0.2 mm,
                            angle 15, 1000 trials; cutoff 0.3, lengths 5-80) and the -info figures
                            of ten runs with different random seeds.
The images and tracks are read with this package's own NIfTI and TCK readers.

Run: python tests/golden/make_tracking_fixture.py [reference root]
"""

import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from larmor_amd.io.nifti import read_nifti  # noqa: E402
from larmor_amd.io.tck import read_tck  # noqa: E402


def main(reference: Path) -> None:
    data = reference / "tests" / "data" / "tracking"
    fod = read_nifti(data / "fod.nii.gz")
    mask = read_nifti(data / "mask.nii.gz")
    out = {
        "fod": np.ascontiguousarray(fod.as_float(), dtype=np.float32),
        "fod_affine": fod.affine,
        "mask": np.ascontiguousarray(mask.as_float(), dtype=np.float32),
        "mask_affine": mask.affine,
    }
    for name in ["det_ref", "det_ref_maxlen", "prob_ref"]:
        t = read_tck(data / f"{name}.tck")
        rows = np.loadtxt(data / f"{name}_seeds.txt", delimiter=",", usecols=range(5))
        out[f"{name}_points"] = t.points
        out[f"{name}_offsets"] = t.offsets
        out[f"{name}_seeds"] = rows  # track index, seed point index, seed x, y, z
    info = json.loads((data / "prob_ref_info.json").read_text())
    runs = info["runs"]

    def column(key):
        return np.array([np.nan if r[key] is None else r[key] for r in runs], dtype=np.float64)

    out["prob_info_calibration_directions"] = column("calibration_directions")
    out["prob_info_calibration_ratio"] = column("calibration_ratio")
    out["prob_info_mean_samples_per_step"] = column("mean_samples_per_step")
    out["prob_info_steps_per_truncation"] = column("steps_per_truncation")
    out["prob_info_max_truncation"] = column("max_truncation")
    out["prob_info_no_direction_percent"] = np.array(
        [r["termination_percent"]["Diffusion model sub-threshold"] for r in runs])
    dest = Path(__file__).resolve().parent / "tracking_fixture.npz"
    np.savez_compressed(dest, **out)
    print(f"wrote {dest} ({dest.stat().st_size} bytes)")


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference"))
