"""The diffusion weighting of each volume of a DWI.

Behaviour of larmor's `GradientTable` (src/larmor/types/gradient_table.py:19-231) as the fit API
uses it: on construction each b-value is scaled by its vector's squared length and the vector
normalised; NaN rows mean no direction; `b0_mask` marks b-values at or below b0_threshold; shells
group sorted weighted b-values, a gap above shell_epsilon starting a new one.
"""

from dataclasses import dataclass

import numpy as np

_UNIT_TOLERANCE = 1e-12
_MIN_LENGTH = 0.1


@dataclass
class GradientTable:
    bvals: np.ndarray
    bvecs: np.ndarray
    b0_threshold: float = 50.0
    shell_epsilon: float = 80.0

    def __post_init__(self) -> None:
        self.bvals = np.array(self.bvals, dtype=np.float64)
        self.bvecs = np.array(self.bvecs, dtype=np.float64)
        if self.bvals.ndim != 1:
            raise ValueError(f"Expected one-dimensional b-values, got shape {self.bvals.shape}.")
        n = len(self.bvals)
        if self.bvecs.shape != (n, 3):
            raise ValueError(
                f"Expected b-vectors of shape ({n}, 3) for {n} b-values, got {self.bvecs.shape}."
            )
        if n == 0:
            raise ValueError("A gradient table needs at least one volume, got none.")
        bad = ~np.isfinite(self.bvals)
        if bad.any():
            raise ValueError(f"Volumes {np.flatnonzero(bad).tolist()} have a NaN or infinite b-value.")
        neg = self.bvals < 0
        if neg.any():
            raise ValueError(f"Volumes {np.flatnonzero(neg).tolist()} have a negative b-value.")
        inf = np.isinf(self.bvecs).any(axis=1)
        if inf.any():
            raise ValueError(
                f"Volumes {np.flatnonzero(inf).tolist()} have an infinite b-vector component."
            )
        self.bvecs[np.isnan(self.bvecs).any(axis=1)] = 0.0
        lengths = np.linalg.norm(self.bvecs, axis=1)
        weighted = self.bvals > self.b0_threshold
        none = (lengths == 0) & weighted
        if none.any():
            raise ValueError(
                f"Volumes {np.flatnonzero(none).tolist()} have b above {self.b0_threshold} but no "
                "direction."
            )
        short = (lengths > 0) & (lengths < _MIN_LENGTH) & weighted
        if short.any():
            raise ValueError(
                f"Volumes {np.flatnonzero(short).tolist()} have b above {self.b0_threshold} but a "
                f"b-vector shorter than {_MIN_LENGTH}, which would scale their b-value to nearly 0."
            )
        rescale = (lengths > 0) & (np.abs(lengths - 1.0) > _UNIT_TOLERANCE)
        self.bvals[rescale] *= lengths[rescale] ** 2
        self.bvecs[rescale] /= lengths[rescale, None]
        self.bvals[lengths == 0] = 0.0

    @property
    def b0_mask(self) -> np.ndarray:
        return np.asarray(self.bvals <= self.b0_threshold, dtype=bool)

    @property
    def shell_index(self) -> np.ndarray:
        weighted = np.flatnonzero(~self.b0_mask)
        order = np.argsort(self.bvals[weighted], kind="stable")
        sorted_b = self.bvals[weighted][order]
        labels = np.concatenate([[0], np.cumsum(np.diff(sorted_b) > self.shell_epsilon)])
        index = np.full(len(self.bvals), -1)
        index[weighted[order]] = labels[: len(weighted)]
        return index

    @property
    def shells(self) -> np.ndarray:
        index = self.shell_index
        if index.max() < 0:
            return np.zeros(0)
        return np.array([self.bvals[index == k].mean() for k in range(index.max() + 1)])

    def count_shells(self, *, above: float = 0.0) -> int:
        return int(np.count_nonzero(self.shells > above))

    @property
    def b_matrix(self) -> np.ndarray:
        """(N, 6): the distinct entries of b g g^T, xx, xy, xz, yy, yz, zz, the off-diagonal ones
        doubled (they appear twice in g^T D g), so that g^T D g b = b_matrix . (Dxx, Dxy, ...)."""
        g = self.bvecs
        pairs = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))
        return np.stack([(1.0 if i == j else 2.0) * self.bvals * g[:, i] * g[:, j]
                         for i, j in pairs], axis=1)
