"""On-disk formats of the deconvolution's inputs and outputs (mirrors larmor.io for the path):
NIfTI-1/2 and MRtrix MIF images, TCK tractograms."""

from larmor_amd.io.nifti import read_nifti, write_nifti
from larmor_amd.io.tck import read_tck, read_tck_header, write_tck

__all__ = ["read_nifti", "read_tck", "read_tck_header", "write_nifti", "write_tck"]
