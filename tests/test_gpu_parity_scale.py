"""Parity at the benchmark workloads' scale, and the ridge solve's QR-only branch.

- The whole 96 x 96 x 60 three-tissue mFOD volume (10 b0 + 3 x 64, 221,184 masked voxels; the
  benchmark's workload): rounds, active sets, FODs and fractions of the GPU fit against the CPU
  FP64 oracle (on every host core, about 12 s on 16), with the agreement printed per quantity.
- The whole 32 x 32 x 20 single-shell CSD volume (20,480 voxels, lmax 8, 300 constraints).
- Ridge 1e-7, whose design fails the normal-equations test (the sum of squares of the kernel is
  above 1e8 ridge^2), so every fit is a pivoted-QR least squares in the global-memory kernel
  (cpp/src/sd/nnls.cpp:317-338): single-tissue and as the multi-tissue high-resolution solver.
"""

import numpy as np
import pytest

import oracle
from larmor_amd import _core, synthetic
from larmor_amd.sd.csd_solver import CSDSolver, csd_problem
from larmor_amd.sd.kernel import Kernel
from larmor_amd.sd.model import compute_shell_weights, low_resolution_dirs
from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
from larmor_amd.sd.solvers import _deconv_kernel
from larmor_amd.sd.sphere import DIRS_300

pytestmark = pytest.mark.gpu

RFS = [AnisotropicRF([1.7e-3, 0.3e-3, 0.3e-3], 0.0), IsotropicRF(0.8e-3), IsotropicRF(3e-3)]


def _agreement(label, got, ref):
    same = np.all(got == ref, axis=tuple(range(1, got.ndim))) if got.ndim > 1 else got == ref
    print(f"  {label}: {same.mean() * 100:.4f}% of {len(same)} voxels identical", flush=True)
    return same


@pytest.fixture(scope="module")
def mfod_volume():
    ph = synthetic.phantom(**synthetic.MFOD_CONFIG)
    signals = synthetic.normalized_signals(ph.data, ph.bvals, ph.mask)
    w = compute_shell_weights(ph.bvals, 1.0)
    hr = Kernel(RFS, ph.bvals, ph.bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(RFS, ph.bvals, ph.bvecs, dirs=low_resolution_dirs(RFS, len(ph.bvals)), weights=w)
    return signals, w, lr, hr


def test_mfod_volume_matches_oracle(mfod_volume):
    signals, w, lr, hr = mfod_volume
    assert len(signals) == 221_184
    s = np.ascontiguousarray(signals)
    print(f"mFOD volume: {s.shape}", flush=True)
    ref_fods, ref_fr, trace = oracle.fit_multi_tissue(s, w, lr, hr, 0.0, 0.1, 1e-6, trace=True)
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), 0.1, 1e-6)
    fods, fr, rounds, status = _core.fit_multi_tissue(
        s, w, _deconv_kernel(lr), lr_solver, _deconv_kernel(hr), hr_solver, 50, 1e-7, 1e-2, 1,
        return_diagnostics=True)
    checks = [
        _agreement("low-resolution rounds", rounds.astype(np.int64), trace[:, 0]),
        _agreement("high-resolution active sets", fods > 0, ref_fods > 0),
        _agreement("FOD amplitudes (bits)", fods, ref_fods),
        _agreement("fractions (bits)", fr, ref_fr),
    ]
    rel = np.linalg.norm(fods - ref_fods) / np.linalg.norm(ref_fods)
    print(f"  FOD relative L2 error against the oracle: {rel:.3e}; mean rounds "
          f"{rounds.mean():.2f}", flush=True)
    for same in checks:
        assert same.all()


def test_csd_volume_matches_oracle():
    ph = synthetic.phantom(**synthetic.CSD_CONFIG)
    on_shell = ph.bvals > 50
    signals = np.ascontiguousarray(
        np.maximum(ph.data.reshape(-1, len(ph.bvals))[:, on_shell].astype(np.float64), 0.0))
    assert signals.shape == (20_480, 60)
    p = csd_problem(ph.bvecs[on_shell], synthetic.tensor_zonal(1000.0), 8, CSDSolver())
    print(f"CSD volume: {signals.shape}", flush=True)
    ref = oracle.fit_csd(signals, p.initial, p.forward, p.constraint, p.norm_weight,
                         p.threshold, p.max_iter, return_solves=True)
    got = _core.fit_csd(signals, p.initial, p.forward, p.constraint, p.norm_weight,
                        p.threshold, p.max_iter, 1, return_solves=True)
    checks = [
        _agreement("converged flags", got[1], ref[1]),
        _agreement("Cholesky solves", got[2].astype(np.int64), ref[2].astype(np.int64)),
        _agreement("constrained sets (FOD bits)", got[0], ref[0]),
    ]
    print(f"  converged: {ref[1].mean() * 100:.2f}%", flush=True)
    for same in checks:
        assert same.all()


def _mfod_small(n, seed):
    rng = np.random.default_rng(seed)
    bvals, bvecs = synthetic.acquisition(10, [(1000.0, 64), (2000.0, 64), (3000.0, 64)], rng)
    s = synthetic.tissue_signals(n, bvals, bvecs, rng, snr=25.0).astype(np.float64)
    w = np.ones(len(bvals))
    hr = Kernel(RFS, bvals, bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(RFS, bvals, bvecs, dirs=low_resolution_dirs(RFS, len(bvals)), weights=w)
    return np.ascontiguousarray(s), w, lr, hr


def test_qr_only_ridge_single_tissue():
    s, _, _, hr = _mfod_small(300, seed=31)
    k = np.ascontiguousarray(hr.w_k_aniso)
    solver = _core.FodSolver.l2nnls(k, 1e-7, 1e-6)
    assert not solver.normal_equations
    assert not oracle.ridge_normal_equations(k, 1e-7)
    print(f"QR-only single tissue: {s.shape}", flush=True)
    got, status = _core.fit_single_tissue(s, solver, 1, return_status=True)
    ref = oracle.fit_single_tissue(s, k, 1e-7, 1e-6)
    _agreement("active sets", got > 0, ref > 0)
    np.testing.assert_array_equal(got, ref)
    # every voxel was a QR fit from the start, never a refused pivot of a Cholesky factor
    assert not (status & 4).any()


def test_qr_only_ridge_multi_tissue_high_resolution():
    s, w, lr, hr = _mfod_small(300, seed=32)
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), 1e-7, 1e-6)
    assert lr_solver.normal_equations and not hr_solver.normal_equations
    print(f"QR-only multi-tissue HR: {s.shape}", flush=True)
    fods, fr, rounds, _ = _core.fit_multi_tissue(
        s, w, _deconv_kernel(lr), lr_solver, _deconv_kernel(hr), hr_solver, 50, 1e-7, 1e-2, 1,
        return_diagnostics=True)
    ref_fods, ref_fr, trace = oracle.fit_multi_tissue(s, w, lr, hr, 0.0, 1e-7, 1e-6, trace=True)
    np.testing.assert_array_equal(rounds, trace[:, 0])
    np.testing.assert_array_equal(fods, ref_fods)
    np.testing.assert_array_equal(fr, ref_fr)
