"""A 3-D or 4-D image with its affine and intensity scaling.

The part of larmor's `larmor.types.Volume` (src/larmor/types/volume.py) the deconvolution path
reads: `data`, `affine`, `slope`, `intercept` and `as_float` (the scaled image, computed in
float32 as larmor does).
"""

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Volume:
    data: np.ndarray
    affine: np.ndarray = field(default_factory=lambda: np.eye(4))
    slope: float = 1.0
    intercept: float = 0.0

    def as_float(self) -> np.ndarray:
        """The scaled image in float32 (data * slope + intercept)."""
        scaled = self.data.astype(np.float32)
        if self.slope != 1.0 or self.intercept != 0.0:
            scaled = scaled * np.float32(self.slope) + np.float32(self.intercept)
        return scaled

    def with_data(self, data: np.ndarray) -> "Volume":
        return Volume(data=data, affine=self.affine.copy(), slope=1.0, intercept=0.0)
