"""The deconvolution kernel: each response function's signal along each direction.

Mirrors `larmor.sd.kernel` (src/larmor/sd/kernel.py): `order_rfs` and `Kernel` with the
attributes `k`, `w_k`, `weights`, `rfs`, `num_iso`, `blocks`, `aniso_blocks`, `k_aniso`, `k_iso`,
`w_k_aniso`, `w_k_iso`.
"""

from collections.abc import Sequence

import numpy as np

from larmor_amd.sd.rf import IsotropicRF, ResponseFunction


def order_rfs(rfs: list[ResponseFunction]) -> list[ResponseFunction]:
    """The response functions with the anisotropic ones first, each group in the given order."""
    return [rf for rf in rfs if not isinstance(rf, IsotropicRF)] + [
        rf for rf in rfs if isinstance(rf, IsotropicRF)
    ]


class Kernel:
    """Kernel matrix, one row per measurement; anisotropic blocks first, isotropic columns last."""

    def __init__(
        self,
        rfs: list[ResponseFunction],
        bvals: np.ndarray,
        bvecs: np.ndarray,
        dirs: np.ndarray | Sequence[np.ndarray] | None = None,
        weights: np.ndarray | None = None,
    ) -> None:
        if dirs is None and any(not isinstance(rf, IsotropicRF) for rf in rfs):
            raise ValueError("dirs is required for an anisotropic response function.")
        self.rfs = order_rfs(rfs)
        self.num_iso = sum(isinstance(rf, IsotropicRF) for rf in self.rfs)
        n_aniso = len(self.rfs) - self.num_iso
        self.weights = np.ones(len(bvals)) if weights is None else weights
        empty = np.empty((0, 3))
        if dirs is None:
            per_rf = [empty] * len(self.rfs)
        elif isinstance(dirs, np.ndarray):
            per_rf = [dirs] * len(self.rfs)
        else:
            if len(dirs) != n_aniso:
                raise ValueError(
                    f"dirs has {len(dirs)} direction sets for {n_aniso} anisotropic "
                    "response functions."
                )
            per_rf = [*dirs] + [empty] * self.num_iso
        columns = []
        self.blocks: list[slice] = []
        start = 0
        for rf, rf_dirs in zip(self.rfs, per_rf, strict=True):
            block = rf.signal(bvals, bvecs, rf_dirs)
            columns.append(block)
            self.blocks.append(slice(start, start + block.shape[1]))
            start += block.shape[1]
        self.k = np.hstack(columns)
        self.w_k = self.k * self.weights[:, None]
        self.aniso_blocks = self.blocks[:n_aniso]
        self.k_aniso, self.k_iso = self._split(self.k)
        self.w_k_aniso, self.w_k_iso = self._split(self.w_k)

    def _split(self, k: np.ndarray) -> tuple[np.ndarray, np.ndarray | None]:
        if self.num_iso == 0:
            return k, None
        return k[:, : -self.num_iso], k[:, -self.num_iso :]
