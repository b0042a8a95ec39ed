"""The data types of the fit API (mirrors larmor.types for Volume and GradientTable)."""

from larmor_amd.types.gradient_table import GradientTable
from larmor_amd.types.volume import Volume

__all__ = ["GradientTable", "Volume"]
