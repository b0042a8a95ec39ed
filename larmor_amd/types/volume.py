"""A 3-D or 4-D image with its affine and intensity scaling.

The part of larmor's `larmor.types.Volume` (src/larmor/types/volume.py) the deconvolution and
tracking paths read: `data`, `affine`, `slope`, `intercept`, `as_float` (the scaled image),
`voxel_sizes` and `grid`.
"""

from dataclasses import dataclass, field

import numpy as np


@dataclass(frozen=True)
class Grid:
    """A voxel grid: the voxel-to-world affine (4, 4) and the spatial shape."""

    affine: np.ndarray
    shape: tuple


@dataclass
class Volume:
    data: np.ndarray
    affine: np.ndarray = field(default_factory=lambda: np.eye(4))
    slope: float = 1.0
    intercept: float = 0.0

    def as_float(self, *, dtype=np.float32) -> np.ndarray:
        """The scaled image, data * slope + intercept, computed in `dtype` (float32 default)."""
        dtype = np.dtype(dtype)
        if not np.issubdtype(dtype, np.floating):
            raise ValueError(f"as_float needs a floating-point type, got {dtype}.")
        scaled = np.asarray(self.data).astype(dtype)
        if self.slope != 1.0 or self.intercept != 0.0:
            scaled = scaled * dtype.type(self.slope) + dtype.type(self.intercept)
        return scaled

    def with_data(self, data: np.ndarray) -> "Volume":
        return Volume(data=data, affine=self.affine.copy(), slope=1.0, intercept=0.0)

    @property
    def voxel_sizes(self) -> np.ndarray:
        """The voxel size along each voxel axis (3,), in mm."""
        return np.linalg.norm(np.asarray(self.affine, dtype=np.float64)[:3, :3], axis=0)

    @property
    def grid(self) -> Grid:
        return Grid(np.asarray(self.affine, dtype=np.float64), tuple(self.data.shape[:3]))
