"""Least-squares fits of the log signal (src/larmor/dwi/fitting.py:16-76).

`fit_log_signal(design, s, weighted)` takes positive signals (n_voxels, N), fits log s on the
design (N, P) -- the ordinary least-squares fit, then for a weighted fit the iterated weighted
least squares of Veraart et al. (NeuroImage 81, 2013, 335-346), weights the predicted signal
squared, refitted until no predicted log signal moves by more than LOG_SIGNAL_TOL or after
MAX_REWEIGHTINGS refits -- and returns the parameters (P, n_voxels) and the constrained flags.
The fits run in liblarmor_sd.so's log-fit kernel, LOG_FIT_CHUNK voxels a launch.  Constrained
fits (the DTI/DKI models' G p >= 0) are not part of this build.
"""

import numpy as np

from larmor_amd import _core

LOG_SIGNAL_TOL = 1e-6
MAX_REWEIGHTINGS = 20
LOG_FIT_CHUNK = 1 << 20


def fit_log_signal(design, s, weighted: bool, constraints=None, n_threads: int = 1):
    """Parameters (P, n_voxels) and constrained flags (n_voxels,) of the log-signal fits."""
    design = np.ascontiguousarray(design, dtype=np.float64)
    p = design.shape[1]
    g = np.zeros((0, p)) if constraints is None else np.ascontiguousarray(constraints,
                                                                           dtype=np.float64)
    s = np.asarray(s)
    params = np.empty((p, len(s)))
    constrained = np.zeros(len(s), dtype=bool)
    for start in range(0, len(s), LOG_FIT_CHUNK):
        rows = slice(start, start + LOG_FIT_CHUNK)
        log_s = np.ascontiguousarray(np.log(s[rows]), dtype=np.float64)
        chunk, flags = _core.fit_log_signal(design, g, log_s, weighted, MAX_REWEIGHTINGS,
                                            LOG_SIGNAL_TOL, n_threads)
        params[:, rows] = chunk.T
        constrained[rows] = flags
    return params, constrained
