"""Seeds of a tracking (the contract of src/larmor/tractography/seed.py): fixed `Seeds`, random
`MaskSeeds`, `seeds_per_voxel`, and the seed text files.

Where a seed is drawn decides its streamline, so the draws are the reference's: a mask's voxels in
C order (numpy's argwhere), a voxel index per seed from `rng.integers`, then a uniform jitter in
[-0.5, 0.5) per axis, mapped to world coordinates by the mask's affine.
"""

import warnings
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from larmor_amd.types.volume import Volume


def _unit_directions(values, n: int) -> np.ndarray:
    directions = np.array(values, dtype=np.float64)
    if directions.shape != (n, 3):
        raise ValueError(f"Expected directions of shape ({n}, 3) to match the {n} "
                         f"positions, got {directions.shape}.")
    if not np.isfinite(directions).all():
        raise ValueError("Found infinite/nan direction.")
    norm = np.linalg.norm(directions, axis=1, keepdims=True)
    if np.any(norm == 0):
        raise ValueError("Found a zero direction, which cannot be followed.")
    off = np.abs(norm[:, 0] - 1) > 4 * np.finfo(np.float64).eps
    directions[off] /= norm[off]
    return directions


@dataclass(frozen=True)
class Seeds:
    """Seed positions (N, 3) world mm, optional start directions (N, 3; normalised) and optional
    global indices (N,) that key each streamline's random numbers."""

    positions: np.ndarray
    directions: np.ndarray | None = None
    indices: np.ndarray | None = None

    def __post_init__(self) -> None:
        positions = np.array(self.positions, dtype=np.float64)
        if positions.ndim != 2 or positions.shape[1] != 3:
            raise ValueError(f"Expected positions of shape (N, 3), got {positions.shape}.")
        if not np.isfinite(positions).all():
            raise ValueError("Found infinite/nan position.")
        object.__setattr__(self, "positions", positions)
        n = positions.shape[0]
        if self.directions is not None:
            object.__setattr__(self, "directions", _unit_directions(self.directions, n))
        if self.indices is not None:
            indices = np.asarray(self.indices)
            if indices.shape != (n,):
                raise ValueError(f"Expected indices of shape ({n},) to match the {n} positions, "
                                 f"got {indices.shape}.")
            if not np.issubdtype(indices.dtype, np.integer):
                raise ValueError(f"Expected integer indices, got {indices.dtype}.")
            if np.any(indices < 0):
                raise ValueError("Found negative indices in indices array.")
            object.__setattr__(self, "indices", indices.astype(np.int64))

    def __len__(self) -> int:
        return len(self.positions)


def _mask_voxels(mask: Volume) -> np.ndarray:
    if mask.data.ndim != 3:
        raise ValueError(f"Expected 3D mask, got mask of shape {mask.data.shape}")
    voxels = np.argwhere(mask.as_float(dtype=np.float64) > 0)
    if len(voxels) == 0:
        raise ValueError("Expected mask to contain values > 0, found only value <= 0.")
    return voxels


def _in_voxels(voxels: np.ndarray, affine: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    ijk = voxels + rng.uniform(-0.5, 0.5, size=(voxels.shape[0], 3))
    affine = np.asarray(affine, dtype=np.float64)
    return ijk @ affine[:3, :3].T + affine[:3, 3]


class SeedStrategy:
    """An endless source of seeds: `draw(n, rng)` returns n new seeds."""

    def draw(self, n: int, rng: np.random.Generator) -> Seeds:  # pragma: no cover
        raise NotImplementedError


class MaskSeeds(SeedStrategy):
    """Seeds at random positions in the voxels of a mask whose scaled value is above 0."""

    def __init__(self, mask: Volume) -> None:
        self.voxels = _mask_voxels(mask)
        self.affine = mask.affine

    def draw(self, n: int, rng: np.random.Generator) -> Seeds:
        pos = rng.integers(0, len(self.voxels), n)
        return Seeds(_in_voxels(self.voxels[pos], self.affine, rng))


# Second key of the stream seeds_per_voxel draws from with a run seed, apart from the tracker's.
PER_VOXEL_STREAM = 2**32


def seeds_per_voxel(mask: Volume, n: int, rng) -> Seeds:
    """n seeds at random positions in each voxel of the mask above 0 (voxels in C order)."""
    if n <= 0:
        raise ValueError(f"Cannot draw {n} seeds.")
    if isinstance(rng, int):
        rng = np.random.default_rng([rng, PER_VOXEL_STREAM])
    voxels = _mask_voxels(mask)
    return Seeds(_in_voxels(np.repeat(voxels, n, axis=0), mask.affine, rng))


def write_seeds(seeds: Seeds, path, *, header: bool = True) -> None:
    """Writes seeds as text rows [index] x y z [dx dy dz], values exact (%.17g)."""
    names = ["x", "y", "z"]
    if seeds.indices is not None:
        names = ["index", *names]
    if seeds.directions is not None:
        names += ["dx", "dy", "dz"]
    rows = np.empty(len(seeds), dtype=[(c, np.int64 if c == "index" else np.float64)
                                       for c in names])
    if seeds.indices is not None:
        rows["index"] = seeds.indices
    for i, c in enumerate("xyz"):
        rows[c] = seeds.positions[:, i]
        if seeds.directions is not None:
            rows["d" + c] = seeds.directions[:, i]
    fmt = ["%d" if c == "index" else "%.17g" for c in names]
    np.savetxt(path, rows, fmt=fmt, header=" ".join(names) if header else "")


_HEADERS = [["x", "y", "z"], ["index", "x", "y", "z"], ["x", "y", "z", "dx", "dy", "dz"],
            ["index", "x", "y", "z", "dx", "dy", "dz"]]


def read_seeds(path) -> Seeds:
    """Reads seeds written by `write_seeds`."""
    with open(path, encoding="utf-8") as f:
        header = f.readline().lstrip("#").split()
    if header not in _HEADERS:
        raise ValueError(f"{path} does not start with a seed header ('# [index] x y z "
                         f"[dx dy dz]'), got {header}.")
    has_index = header[0] == "index"
    has_dirs = "dx" in header
    with warnings.catch_warnings():
        warnings.filterwarnings("ignore", message=".*input contained no data")
        data = np.loadtxt(path, ndmin=2)
    if data.size == 0:
        return Seeds(np.empty((0, 3)), np.empty((0, 3)) if has_dirs else None,
                     np.empty(0, np.int64) if has_index else None)
    if data.shape[1] != len(header):
        raise ValueError(f"{path}: expected {len(header)} columns from the header, got "
                         f"{data.shape[1]}.")
    col = 1 if has_index else 0
    indices = None
    if has_index:
        indices = np.loadtxt(path, ndmin=2, usecols=0, dtype=np.int64).reshape(-1)
    return Seeds(data[:, col:col + 3], data[:, col + 3:col + 6] if has_dirs else None, indices)
