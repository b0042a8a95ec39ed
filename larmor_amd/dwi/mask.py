"""Putting per-voxel values back on the grid (src/larmor/dwi/mask.py:unmask)."""

import numpy as np


def unmask(values: np.ndarray, mask: np.ndarray, fill_value: float = 0.0,
           dtype=np.float32) -> np.ndarray:
    """values (K,) or (K, D) of the mask's True voxels -> (X, Y, Z[, D]), fill outside."""
    dtype = values.dtype if dtype is None else dtype
    extra = values.shape[1:] if values.ndim > 1 else ()
    out = np.full(mask.shape + extra, fill_value, dtype=dtype)
    out[mask] = values.astype(dtype)
    return out
