"""Spherical deconvolution: models, solvers, response functions, peaks and SH helpers.

The public names of larmor's `larmor.sd` (src/larmor/sd/__init__.py:1-78) that this package
builds.  Not built here: `filter_afod` (asymmetric FOD filtering) and the MRIToolkit FOD readers
(`MRIToolkitFOD`, `read_mritoolkit_fod`, `read_mritoolkit_fod_count`, `read_mritoolkit_sh`).
"""

from larmor_amd.sd.csd import CSDModel
from larmor_amd.sd.csd_rf import SingleFibreRF, build_csd_response, estimate_csd_rf
from larmor_amd.sd.csd_solver import CSDSolver
from larmor_amd.sd.method import SDMethod
from larmor_amd.sd.model import (PEAK_RELATIVE_THRESHOLD, SD_MODEL_PEAK_RELATIVE_THRESHOLD, SDFit,
                                 SDModel, merge_fods)
from larmor_amd.sd.msmt_csd import MSMTCSDModel
from larmor_amd.sd.peaks import Peaks, find_peaks
from larmor_amd.sd.rf import (AnisotropicRF, IsotropicRF, NoddiRF, ResponseFunction, ZonalSHRF,
                              build_response_functions, estimate_wm_rf)
from larmor_amd.sd.sh import convert_sh_basis, fit_sh, sh_layout
from larmor_amd.sd.solvers import DRL, L2NNLS, Lasso, SDSolver
from larmor_amd.sd.tissue_responses import (TissueResponses, build_shell_response,
                                            estimate_tissue_responses, response_shells)

__all__ = [
    "DRL",
    "L2NNLS",
    "Lasso",
    "AnisotropicRF",
    "CSDModel",
    "CSDSolver",
    "IsotropicRF",
    "MSMTCSDModel",
    "NoddiRF",
    "PEAK_RELATIVE_THRESHOLD",
    "Peaks",
    "ResponseFunction",
    "SDModel",
    "SDFit",
    "SDMethod",
    "SDSolver",
    "SD_MODEL_PEAK_RELATIVE_THRESHOLD",
    "SingleFibreRF",
    "TissueResponses",
    "ZonalSHRF",
    "build_csd_response",
    "build_response_functions",
    "build_shell_response",
    "convert_sh_basis",
    "estimate_csd_rf",
    "estimate_tissue_responses",
    "estimate_wm_rf",
    "find_peaks",
    "fit_sh",
    "merge_fods",
    "response_shells",
    "sh_layout",
]
