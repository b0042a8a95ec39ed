"""A set of streamlines as one flat array of points and their offsets (the layout of larmor's
`larmor.types.Tractogram`, src/larmor/types/tractogram.py): streamline s is
`points[offsets[s]:offsets[s + 1]]`, points (P, 3) float32 world mm, offsets (S + 1,) int64."""

from dataclasses import dataclass, field

import numpy as np


@dataclass(eq=False)
class Tractogram:
    points: np.ndarray
    offsets: np.ndarray
    point_data: dict = field(default_factory=dict)
    streamline_data: dict = field(default_factory=dict)
    reference: object | None = None

    def __post_init__(self) -> None:
        self.points = np.asarray(self.points, dtype=np.float32).reshape(-1, 3)
        offsets = np.asarray(self.offsets)
        if offsets.size and not np.issubdtype(offsets.dtype, np.integer):
            raise ValueError(f"Offsets must be integers, got {offsets.dtype}.")
        self.offsets = offsets.astype(np.int64).reshape(-1)
        if len(self.offsets) == 0:
            raise ValueError("Offsets must hold at least the final 0.")
        if (self.offsets[0] != 0 or self.offsets[-1] != len(self.points)
                or np.any(np.diff(self.offsets) < 0)):
            raise ValueError("Offsets must start at 0, never decrease and end at the point count.")

    def __len__(self) -> int:
        return len(self.offsets) - 1

    def __getitem__(self, s: int) -> np.ndarray:
        return self.points[self.offsets[s]:self.offsets[s + 1]]

    @property
    def lengths(self) -> np.ndarray:
        """Points per streamline."""
        return np.diff(self.offsets)
