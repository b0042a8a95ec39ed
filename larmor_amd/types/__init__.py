"""The data types of the fit API (mirrors larmor.types for Volume, Grid, Tractogram and GradientTable)."""

from larmor_amd.types.gradient_table import GradientTable
from larmor_amd.types.tractogram import Tractogram
from larmor_amd.types.volume import Grid, Volume

__all__ = ["GradientTable", "Grid", "Tractogram", "Volume"]
