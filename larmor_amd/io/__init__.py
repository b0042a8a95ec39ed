"""On-disk formats of the path's inputs and outputs (mirrors larmor.io for them): NIfTI-1/2 and
MRtrix MIF images, FSL and MRtrix gradient tables, MRtrix response files and TCK tractograms, with
the `load_*` / `save_*` entry points that pick the format by file name."""

from pathlib import Path

from larmor_amd.io.gradients import (read_bvals_bvecs, read_mrtrix_gradients,
                                     read_mrtrix_response, write_bvals_bvecs)
from larmor_amd.io.mif import (has_mif_gradients, read_mif, read_mif_gradients,
                               read_mif_header, write_mif)
from larmor_amd.io.nifti import read_nifti, write_nifti
from larmor_amd.io.tck import TckWriter, read_tck, read_tck_header, write_tck

VOLUME_EXTENSIONS = {"nifti": (".nii", ".nii.gz"), "mif": (".mif", ".mif.gz", ".mih")}
TRACTOGRAM_EXTENSIONS = {"tck": (".tck",)}


def _volume_format(path: Path) -> str:
    name = path.name.lower()
    for kind, exts in VOLUME_EXTENSIONS.items():
        if name.endswith(exts):
            return kind
    known = ", ".join(e for exts in VOLUME_EXTENSIONS.values() for e in exts)
    raise ValueError(f"Unsupported volume file {path}; the known extensions are: {known}.")


def load_volume(path):
    """Reads a NIfTI or MIF image as a Volume (scaling kept as slope and intercept)."""
    path = Path(path)
    if not path.exists():
        raise FileNotFoundError(f"{path} does not exist.")
    return read_nifti(path) if _volume_format(path) == "nifti" else read_mif(path)


def save_volume(volume, path, gradients=None) -> None:
    """Writes a Volume as NIfTI or MIF by file name (a gradient table only into MIF)."""
    path = Path(path)
    if _volume_format(path) == "nifti":
        write_nifti(volume, path)
    else:
        write_mif(volume, path, gradients=gradients)


def load_dwi(path, bvals=None, bvecs=None, *, frame: str = "fsl"):
    """(volume, gradient table): the table from the bvals/bvecs files when given (frame "fsl"
    converted with the volume's affine, or "world"), else from a MIF header's dw_scheme."""
    p = Path(path)
    if (bvals is None) != (bvecs is None):
        raise ValueError("Give both the bvals and the bvecs file, or neither.")
    volume = load_volume(p)
    if volume.data.ndim != 4:
        raise ValueError(f"{p} has no fourth dimension; a gradient table needs a 4D image.")
    if bvals is not None:
        return volume, read_bvals_bvecs(bvals, bvecs, affine=volume.affine, frame=frame,
                                        n_volumes=volume.data.shape[3])
    if _volume_format(p) == "mif":
        table = read_mif_gradients(p)
        if table is not None:
            return volume, table
    raise ValueError(f"{p} needs its bvals and bvecs files: no gradient table in its header.")


def load_tractogram(path):
    path = Path(path)
    if path.suffix.lower() != ".tck":
        raise ValueError(f"Unsupported tractogram file {path}; the known extension is .tck.")
    return read_tck(path)


def save_tractogram(tractogram, path) -> None:
    path = Path(path)
    if path.suffix.lower() != ".tck":
        raise ValueError(f"Unsupported tractogram file {path}; the known extension is .tck.")
    write_tck(tractogram, path)


__all__ = ["TRACTOGRAM_EXTENSIONS", "VOLUME_EXTENSIONS", "TckWriter", "has_mif_gradients",
           "load_dwi", "load_tractogram", "load_volume", "read_bvals_bvecs", "read_mif",
           "read_mif_gradients", "read_mif_header", "read_mrtrix_gradients",
           "read_mrtrix_response", "read_nifti", "read_tck", "read_tck_header",
           "save_tractogram", "save_volume", "write_bvals_bvecs", "write_mif", "write_nifti",
           "write_tck"]
