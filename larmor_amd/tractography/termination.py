"""Why a streamline stops, and the automatic FOD amplitude cutoff (the contract of
src/larmor/tractography/termination.py)."""

from enum import IntEnum

import numpy as np

from larmor_amd.sd.peaks import find_peaks

# The automatic cutoff is a fifth of the mean largest peak over at most 5000 voxels with a
# positive DC term, drawn with a fixed seed (the reference's constants, so that the same FOD
# gives the same cutoff).
AUTO_CUTOFF_SEED = 1234
AUTO_CUTOFF_FRACTION = 0.2
AUTO_CUTOFF_VOXELS = 5000


class EndStatus(IntEnum):
    """Why one end of a streamline stopped (the codes of the tracking kernels' status)."""

    NONE = 0          # the end was not tracked (the seed end of a one-way streamline)
    MAX_LENGTH = 1    # the streamline reached the maximum length
    LEFT_VOLUME = 2   # the next point would leave the FOD image (not kept)
    LEFT_MASK = 3     # the next point would leave the tracking mask (not kept)
    LOW_FOD = 4       # the FOD amplitude fell below the cutoff
    SHARP_TURN = 5    # the next direction turned more than the maximum angle
    NO_DIRECTION = 6  # no next direction was found


def auto_cutoff(coeffs, *, basis: str = "tournier", mask=None) -> float:
    """0.2 times the mean largest peak amplitude of up to 5000 voxels (drawn with a fixed seed)
    whose DC term is positive, inside `mask` when given; the peaks come from the GPU search.

    Raises:
        ValueError: if no voxel (in the mask) has a positive DC term, or none drawn has a peak.
    """
    coeffs = np.asarray(coeffs)
    positive = coeffs[..., 0] > 0
    if mask is not None:
        positive &= np.asarray(mask, dtype=bool)
    candidates = np.flatnonzero(positive)
    if len(candidates) == 0:
        where = "inside the mask " if mask is not None else ""
        raise ValueError(f"No voxel of the FOD {where}has a positive DC term.")
    rng = np.random.default_rng(AUTO_CUTOFF_SEED)
    drawn = rng.choice(candidates, size=min(AUTO_CUTOFF_VOXELS, len(candidates)), replace=False)
    voxels = coeffs.reshape(-1, coeffs.shape[-1])[drawn]
    largest = find_peaks(voxels, basis=basis, num_peaks=1).amplitudes[:, 0]
    if np.all(np.isnan(largest)):
        raise ValueError("None of the FOD voxels drawn has a peak.")
    return float(AUTO_CUTOFF_FRACTION * np.nanmean(largest))
