"""The C-ABI library loads and exports every symbol include/larmor_sd.h declares (no GPU)."""

import ctypes

import numpy as np
import pytest

from larmor_amd import _core, _lib


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    declared = _lib.declared_symbols()
    assert len(declared) >= 10
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.SIGNATURES)


def test_abi_version():
    assert _lib.load().lsd_abi_version() == _lib.ABI_VERSION == 6


def test_solver_argument_errors_match_reference():
    k = np.ones((4, 3))
    with pytest.raises(ValueError, match="reg must be at least 0"):
        _core.FodSolver.l2nnls(k, -1.0, 1e-6)
    with pytest.raises(ValueError, match="tol_x must be positive"):
        _core.FodSolver.l2nnls(k, 0.1, 0.0)
    with pytest.raises(TypeError):
        _core.FodSolver.l2nnls(k.astype(np.float32), 0.1, 1e-6)


def test_solver_properties_without_gpu():
    """Creating a solver copies the kernel and Gram to the device, so it needs one; without a
    GPU the call must fail loudly rather than fall back to the CPU."""
    import torch

    k = np.random.default_rng(0).random((10, 4))
    if torch.cuda.is_available():
        s = _core.FodSolver.l2nnls(k, 0.1, 1e-6)
        assert (s.n_dirs, s.n_meas, s.n_unconverged) == (4, 10, 0)
    else:
        with pytest.raises(RuntimeError):
            _core.FodSolver.l2nnls(k, 0.1, 1e-6)


def test_kernel_block_errors_match_reference():
    ka, ki = np.ones((5, 4)), np.ones((5, 2))
    with pytest.raises(ValueError, match=r"aniso_blocks must be non-empty ranges within 0..4"):
        _core.DeconvKernel(ka, ki, ki, [(0, 5)])
    with pytest.raises(ValueError, match="shapes do not match"):
        _core.DeconvKernel(ka, np.ones((4, 2)), ki, [(0, 4)])
