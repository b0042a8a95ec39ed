"""GPU kernels against the CPU FP64 oracle on the same seeded inputs.

The oracle (oracle/sd_oracle.cpp) computes in the same rounding order as the HIP kernels, so the
bar here is bit-identity: identical FOD amplitudes, fractions, rounds and active sets.
"""

import numpy as np
import pytest

import oracle
from larmor_amd import _core, synthetic
from larmor_amd.sd.kernel import Kernel
from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
from larmor_amd.sd.sphere import DIRS_300, get_dirs

pytestmark = pytest.mark.gpu


def _mfod_setup(shape=(10, 10, 8), seed=2):
    cfg = dict(synthetic.MFOD_CONFIG)
    cfg["shape"] = shape
    cfg["seed"] = seed
    ph = synthetic.phantom(**cfg)
    signals = synthetic.normalized_signals(ph.data, ph.bvals, ph.mask)
    rfs = [AnisotropicRF([1.7e-3, 0.3e-3, 0.3e-3], 0.0), IsotropicRF(0.8e-3), IsotropicRF(3e-3)]
    w = np.ones(len(ph.bvals))
    hr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=[get_dirs(min(len(ph.bvals), 45))[:-1]], weights=w)
    return signals, w, lr, hr


def _deconv(k):
    return _core.DeconvKernel(k.k_aniso, k.k_iso, k.w_k_iso,
                              [(b.start, b.stop) for b in k.aniso_blocks])


def test_multi_tissue_bit_identical_to_oracle():
    signals, w, lr, hr = _mfod_setup()
    n = min(len(signals), 300)
    s = np.ascontiguousarray(signals[:n])
    ref_fods, ref_fr, tr = oracle.fit_multi_tissue(s, w, lr, hr, 0.0, 0.1, 1e-6, trace=True)
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), 0.1, 1e-6)
    fods, fr, rounds, status = _core.fit_multi_tissue(
        s, w, _deconv(lr), lr_solver, _deconv(hr), hr_solver, 50, 1e-7, 1e-2, 1,
        return_diagnostics=True)
    np.testing.assert_array_equal(rounds, tr[:, 0])
    assert np.array_equal(fods > 0, ref_fods > 0), "active sets differ"
    np.testing.assert_array_equal(fods, ref_fods)
    np.testing.assert_array_equal(fr, ref_fr)


def test_single_tissue_bit_identical_to_oracle():
    signals, w, lr, hr = _mfod_setup(seed=5)
    s = np.ascontiguousarray(signals[:200])
    k = np.ascontiguousarray(hr.w_k_aniso)
    ref = oracle.fit_single_tissue(s, k, 0.1, 1e-6)
    solver = _core.FodSolver.l2nnls(k, 0.1, 1e-6)
    got = _core.fit_single_tissue(s, solver, 1)
    np.testing.assert_array_equal(got, ref)
