"""Streamline tracking through a field of FODs on the GPU (mirrors larmor.tractography for the
tracker, its seeds and its end statuses)."""

from larmor_amd.tractography.seed import (MaskSeeds, Seeds, SeedStrategy, read_seeds,
                                          seeds_per_voxel, write_seeds)
from larmor_amd.tractography.termination import EndStatus, auto_cutoff
from larmor_amd.tractography.tracker import Tracker, TractographyRun

__all__ = ["EndStatus", "MaskSeeds", "SeedStrategy", "Seeds", "Tracker", "TractographyRun",
           "auto_cutoff", "read_seeds", "seeds_per_voxel", "write_seeds"]
