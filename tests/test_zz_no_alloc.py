"""The allocation-free multi-tissue fit (lsd_fit_multi_tissue_noalloc) against the allocating
one (lsd_fit_multi_tissue), both called through the C-ABI on the same device inputs: FODs,
fractions, rounds and status bit for bit, on two calls in a row (the second smaller, so it runs
on the workspace the first one left), then with a design whose ridge makes every solve a
pivoted-QR fit in the global-memory kernel, and a call too large for the workspace fails with a
clear error and leaves the library usable."""

import ctypes

import numpy as np
import pytest

from larmor_amd import _core, _lib, synthetic
from larmor_amd.sd.kernel import Kernel
from larmor_amd.sd.model import low_resolution_dirs
from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
from larmor_amd.sd.solvers import _deconv_kernel
from larmor_amd.sd.sphere import DIRS_300

pytestmark = pytest.mark.gpu

RFS = [AnisotropicRF([1.7e-3, 0.3e-3, 0.3e-3], 0.0), IsotropicRF(0.8e-3), IsotropicRF(3e-3)]


def _workload(n_voxels: int, seed: int):
    rng = np.random.default_rng(seed)
    bvals, bvecs = synthetic.acquisition(10, [(1000.0, 64), (2000.0, 64), (3000.0, 64)], rng)
    signals = synthetic.tissue_signals(n_voxels, bvals, bvecs, rng, snr=25.0)
    w = np.ones(len(bvals))
    hr = Kernel(RFS, bvals, bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(RFS, bvals, bvecs, dirs=low_resolution_dirs(RFS, len(bvals)), weights=w)
    return np.ascontiguousarray(signals, dtype=np.float64), w, lr, hr


def _fit(entry: str, dev, weights, lr, lr_solver, hr, hr_solver):
    import torch

    n, m = dev.shape
    fods = torch.full((n, hr.n_anisotropic), np.nan, dtype=torch.float64, device="cuda")
    fractions = torch.full((n, hr.n_blocks + hr.n_isotropic), np.nan, dtype=torch.float64,
                           device="cuda")
    rounds = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    status = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    rc = getattr(_lib.load(), entry)(
        dev.data_ptr(), n, m, weights.data_ptr(), lr._handle, lr_solver._handle, hr._handle,
        hr_solver._handle, 50, 1e-7, 1e-2, fods.data_ptr(), fractions.data_ptr(),
        rounds.data_ptr(), status.data_ptr(),
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    _lib.check(rc)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (fods, fractions, rounds, status)]


def _same(a, b, label):
    for name, x, y in zip(("fods", "fractions", "rounds", "status"), a, b):
        assert x.shape == y.shape
        assert np.array_equal(x, y), f"{label}: {name} differ, max {np.abs(x - y).max()}"


@pytest.mark.parametrize("hr_ridge", [0.1, 1e-7])
def test_noalloc_fit_equals_the_allocating_fit(hr_ridge):
    import torch

    lib = _lib.load()
    assert lib.lsd_noalloc_workspace_bytes() == 64 << 20
    signals, w, lr, hr = _workload(6000, seed=11)
    weights = torch.from_numpy(w).to("cuda")
    lr_k, hr_k = _deconv_kernel(lr), _deconv_kernel(hr)
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), hr_ridge, 1e-6)
    # ridge 1e-7 takes the QR-only path (the design fails the normal-equations test)
    assert hr_solver.normal_equations == (hr_ridge == 0.1)
    for n in (6000, 1500):  # two calls in a row, the second smaller
        print(f"no-alloc fit: hr ridge {hr_ridge}, {n} voxels", flush=True)
        dev = torch.from_numpy(signals[:n]).to("cuda")
        plain = _fit("lsd_fit_multi_tissue", dev, weights, lr_k, lr_solver, hr_k, hr_solver)
        free = _fit("lsd_fit_multi_tissue_noalloc", dev, weights, lr_k, lr_solver, hr_k,
                    hr_solver)
        assert np.isfinite(plain[0]).all() and (plain[2] >= 1).all()
        _same(plain, free, f"{n} voxels")


def test_noalloc_refuses_what_does_not_fit_and_keeps_working():
    import torch

    signals, w, lr, hr = _workload(200, seed=12)
    weights = torch.from_numpy(w).to("cuda")
    lr_k, hr_k = _deconv_kernel(lr), _deconv_kernel(hr)
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), 0.1, 1e-6)
    # 200,000 rows of zeros: far more per-voxel state than 64 MiB holds; refused before any launch
    big = torch.zeros((200_000, signals.shape[1]), dtype=torch.float64, device="cuda")
    print("no-alloc fit: oversized call", flush=True)
    with pytest.raises(ValueError, match="pass fewer voxels per call"):
        _fit("lsd_fit_multi_tissue_noalloc", big, weights, lr_k, lr_solver, hr_k, hr_solver)
    drl = _core.FodSolver.drl(np.ascontiguousarray(hr.w_k_aniso.T),
                              np.ascontiguousarray(hr.w_k_aniso.T @ hr.w_k_aniso), 0.1, 10, 1.0)
    dev = torch.from_numpy(signals).to("cuda")
    with pytest.raises(ValueError, match="ridge NNLS"):
        _fit("lsd_fit_multi_tissue_noalloc", dev, weights, lr_k, lr_solver, hr_k, drl)
    # 610 measurements: the allocating fit takes them (12 row slots), the allocation-free one stops
    # at 512
    rng = np.random.default_rng(13)
    bvals, bvecs = synthetic.acquisition(10, [(1000.0, 200), (2000.0, 200), (3000.0, 200)], rng)
    wide = Kernel(RFS, bvals, bvecs, dirs=DIRS_300, weights=np.ones(len(bvals)))
    wide_k = _deconv_kernel(wide)
    wide_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(wide.w_k_aniso), 0.1, 1e-6)
    wide_rows = torch.ones((4, len(bvals)), dtype=torch.float64, device="cuda")
    wide_w = torch.ones(len(bvals), dtype=torch.float64, device="cuda")
    print("no-alloc fit: 610 measurements", flush=True)
    with pytest.raises(ValueError, match="exceeds 512 measurements"):
        _fit("lsd_fit_multi_tissue_noalloc", wide_rows, wide_w, wide_k, wide_solver, wide_k,
             wide_solver)
    print("no-alloc fit: 200 voxels after the refusals", flush=True)
    plain = _fit("lsd_fit_multi_tissue", dev, weights, lr_k, lr_solver, hr_k, hr_solver)
    free = _fit("lsd_fit_multi_tissue_noalloc", dev, weights, lr_k, lr_solver, hr_k, hr_solver)
    _same(plain, free, "200 voxels")
