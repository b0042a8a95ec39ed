"""Shell selection by b-value (src/larmor/dwi/shells.py)."""

import numpy as np

# Distance in s/mm^2 within which a b-value belongs to a shell (shells.py:SHELL_TOLERANCE).
SHELL_TOLERANCE = 100.0


def shell_mask(bvals: np.ndarray, shell: float, *, tol: float = SHELL_TOLERANCE) -> np.ndarray:
    """Marks the volumes whose b-value lies strictly within `tol` of `shell`."""
    return np.abs(np.asarray(bvals, dtype=np.float64) - shell) < tol
