"""Gradient tables and response files on disk (the contract of src/larmor/io/bvals_bvecs.py):
FSL bvals/bvecs (vectors in the image's voxel frame, x negated when the affine's rotation part has
a positive determinant -- FSL's convention), MRtrix `.b` tables (x y z b per row, world
coordinates) and MRtrix response files (one row of zonal coefficients per shell)."""

import re
import warnings
from pathlib import Path

import numpy as np

from larmor_amd.types.gradient_table import GradientTable

_SEPARATORS = re.compile(r"[\s,]+")


def _table(path) -> np.ndarray:
    path = Path(path)
    if not path.is_file():
        raise ValueError(f"{path} does not exist.")
    try:
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", UserWarning)
            data = np.loadtxt(path, dtype=np.float64)
    except (OSError, ValueError) as exc:
        raise ValueError(f"{path} cannot be read as a table of numbers: {exc}") from exc
    if data.size == 0:
        raise ValueError(f"{path} is empty.")
    return data


def fsl_to_world(affine: np.ndarray) -> np.ndarray:
    """The matrix taking FSL b-vectors of an image to world directions: the unit voxel axes,
    with x negated first when their determinant is positive."""
    linear = np.asarray(affine, dtype=np.float64)[:3, :3]
    axes = linear / np.linalg.norm(linear, axis=0)
    flip = np.eye(3)
    if np.linalg.det(axes) > 0:
        flip[0, 0] = -1.0
    return axes @ flip


def map_directions(vectors: np.ndarray, matrix: np.ndarray) -> np.ndarray:
    """matrix applied to each vector, rescaled to the vector's own length (shear-safe)."""
    mapped = vectors @ matrix.T
    old = np.linalg.norm(vectors, axis=-1, keepdims=True)
    new = np.linalg.norm(mapped, axis=-1, keepdims=True)
    return mapped * np.divide(old, new, out=np.zeros_like(new), where=new > 0)


def read_bvals_bvecs(bvals_path, bvecs_path, *, affine=None, frame: str = "fsl",
                     n_volumes: int | None = None) -> GradientTable:
    """Reads bvals and bvecs ((3, N) or (N, 3)) into a table in world coordinates; frame "fsl"
    needs the image's affine, "world" takes the vectors as they are."""
    if frame not in ("fsl", "world"):
        raise ValueError(f'Unsupported frame "{frame}"; expected "fsl" or "world".')
    if frame == "fsl" and affine is None:
        raise ValueError("Cannot convert b-vectors to RAS+ without a reference.")
    bvals = _table(bvals_path)
    if bvals.ndim == 2 and min(bvals.shape) > 1:
        raise ValueError(f"{bvals_path} has shape {bvals.shape}; a bvals file is one row or one "
                         "column.")
    bvals = bvals.ravel()
    bvecs = _table(bvecs_path)
    n = len(bvals)
    bvecs = np.atleast_2d(bvecs) if bvecs.shape == (3,) else bvecs
    if bvecs.shape == (3, n):
        bvecs = bvecs.T
    elif bvecs.shape != (n, 3):
        raise ValueError(f"{bvecs_path} has shape {bvecs.shape}, but {bvals_path} has {n} "
                         f"b-values: expected (3, {n}) or ({n}, 3).")
    if n_volumes is not None and n_volumes != n:
        raise ValueError(f"The gradient files have {n} entries, but the image has {n_volumes} "
                         "volumes.")
    if frame == "fsl":
        bvecs = map_directions(bvecs, fsl_to_world(affine))
    return GradientTable(bvals, bvecs)


def write_bvals_bvecs(gtab: GradientTable, bvals_path, bvecs_path, *, affine=None,
                      frame: str = "fsl") -> None:
    """Writes the table as FSL bvals (one row) and bvecs (3 rows), converting the world vectors to
    the image's FSL frame unless frame is "world"."""
    bvecs = np.asarray(gtab.bvecs, dtype=np.float64)
    if frame == "fsl":
        if affine is None:
            raise ValueError("Cannot convert b-vectors to the FSL frame without a reference.")
        bvecs = map_directions(bvecs, np.linalg.inv(fsl_to_world(affine)))
    np.savetxt(bvals_path, np.asarray(gtab.bvals)[None, :], fmt="%.17g")
    np.savetxt(bvecs_path, bvecs.T, fmt="%.17g")


def _rows(path, what: str) -> np.ndarray:
    path = Path(path)
    if not path.is_file():
        raise ValueError(f"{path} does not exist.")
    rows = []
    with path.open(encoding="utf-8", errors="replace") as f:
        for number, line in enumerate(f, start=1):
            text = line.partition("#")[0].strip()
            if not text:
                continue
            try:
                rows.append([float(x) for x in _SEPARATORS.split(text) if x])
            except ValueError as exc:
                raise ValueError(f"{path}, line {number}: {text!r} is not a row of numbers.") \
                    from exc
            if len(rows[-1]) != len(rows[0]):
                raise ValueError(f"{path}, line {number}: {len(rows[-1])} numbers, but the first "
                                 f"row has {len(rows[0])}.")
    if not rows:
        raise ValueError(f"{path} holds no {what} row.")
    return np.array(rows, dtype=np.float64)


def read_mrtrix_gradients(path, *, n_volumes: int | None = None) -> GradientTable:
    """An MRtrix `.b` table: rows x y z b, world directions."""
    table = _rows(path, "gradient")
    if table.shape[1] != 4:
        raise ValueError(f"{path}: {table.shape[1]} numbers per row; each row is x y z b.")
    if n_volumes is not None and n_volumes != len(table):
        raise ValueError(f"The gradient files have {len(table)} entries, but the image has "
                         f"{n_volumes} volumes.")
    return GradientTable(table[:, 3], table[:, :3])


def read_mrtrix_response(path) -> np.ndarray:
    """An MRtrix response file: (n_shells, n_degrees) zonal coefficients, one row per shell."""
    return _rows(path, "response")
