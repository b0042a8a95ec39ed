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


def test_batch_invariance_alone_equals_in_the_mix():
    """A voxel's result does not depend on its batch mates (cpp/tests/sd/test_multi_tissue.cpp:56
    asks 1e-12; the GPU gives identity)."""
    signals, w, lr, hr = _mfod_setup(seed=7)
    s = np.ascontiguousarray(signals[:150])
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), 0.1, 1e-6)
    args = (w, _deconv(lr), lr_solver, _deconv(hr), hr_solver, 50, 1e-7, 1e-2, 1)
    fods, fr = _core.fit_multi_tissue(s, *args)
    for v in (0, 17, 149):
        f1, r1 = _core.fit_multi_tissue(np.ascontiguousarray(s[v : v + 1]), *args)
        np.testing.assert_array_equal(f1[0], fods[v])
        np.testing.assert_array_equal(r1[0], fr[v])
    rev = np.ascontiguousarray(s[::-1])
    f2, r2 = _core.fit_multi_tissue(rev, *args)
    np.testing.assert_array_equal(f2[::-1], fods)
    np.testing.assert_array_equal(r2[::-1], fr)


def test_edge_cases_empty_zero_and_saturated_signals():
    signals, w, lr, hr = _mfod_setup(seed=9)
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), 0.1, 1e-6)
    args = (w, _deconv(lr), lr_solver, _deconv(hr), hr_solver, 50, 1e-7, 1e-2, 1)
    empty = np.zeros((0, signals.shape[1]))
    f0, r0 = _core.fit_multi_tissue(empty, *args)
    assert f0.shape == (0, 300) and r0.shape == (0, 3)
    odd = np.ascontiguousarray(np.vstack([np.zeros(signals.shape[1]), np.ones(signals.shape[1]),
                                          signals[:3]]))
    fods, fr = _core.fit_multi_tissue(odd, *args)
    ref_fods, ref_fr = oracle.fit_multi_tissue(odd, w, lr, hr, 0.0, 0.1, 1e-6)
    np.testing.assert_array_equal(fods, ref_fods)
    np.testing.assert_array_equal(fr, ref_fr)
    assert (fods[0] == 0).all()


def test_mritoolkit_fixture_on_gpu():
    """The reference's 512-voxel MRIToolkit mFOD case (two anisotropic blocks, NODDI grey
    matter, 600 high-resolution columns with active sets up to ~220, which exercise the
    global-memory factor rerun): identical to the oracle, and within the reference's own
    tolerances of MRIToolkit (tests/sd/test_matlab_match.py:143-165)."""
    from pathlib import Path

    from larmor_amd.sd.rf import NoddiRF

    f = np.load(Path(__file__).resolve().parent / "golden" / "mfod_fixture.npz")
    bvals, bvecs = f["bvals"], f["bvecs"]
    rfs = [AnisotropicRF([1.7e-3, 0.2e-3, 0.2e-3], 0.0), NoddiRF(), IsotropicRF(3e-3)]
    w = np.ones(len(bvals))
    hr = Kernel(rfs, bvals, bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(rfs, bvals, bvecs, dirs=[get_dirs(min(len(bvals), rf.low_resolution_dirs))[:-1]
                                         for rf in rfs[:2]], weights=w)
    sig = f["signal"].astype(np.float64)
    s = np.ascontiguousarray(sig / sig[:, bvals <= bvals.min()].mean(axis=1, keepdims=True))
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), 0.1, 1e-6)
    fods, fr = _core.fit_multi_tissue(s, w, _deconv(lr), lr_solver, _deconv(hr), hr_solver, 50,
                                      1e-7, 1e-2, 1)
    ref_fods, ref_fr = oracle.fit_multi_tissue(s, w, lr, hr, 0.0, 0.1, 1e-6)
    np.testing.assert_array_equal(fods, ref_fods)
    np.testing.assert_array_equal(fr, ref_fr)
    mrt = f["fod"].astype(np.float64)
    assert (np.abs(fods - mrt).max(axis=1) / mrt.max(axis=1)).max() <= 1e-3
    np.testing.assert_allclose(fr, f["fractions"].astype(np.float64), rtol=0, atol=1e-4)


def test_profiling_counts_work_and_leaves_results_unchanged():
    from larmor_amd import _lib

    signals, w, lr, hr = _mfod_setup(seed=7)
    s = np.ascontiguousarray(signals[:256])
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), 0.1, 1e-6)
    args = (w, _deconv(lr), lr_solver, _deconv(hr), hr_solver, 50, 1e-7, 1e-2, 1)
    plain_fods, plain_fr = _core.fit_multi_tissue(s, *args)
    runs = []
    _lib.profile_enable(True)
    try:
        _lib.profile_collect()
        for _ in range(2):
            fods, fr = _core.fit_multi_tissue(s, *args)
            np.testing.assert_array_equal(fods, plain_fods)
            np.testing.assert_array_equal(fr, plain_fr)
            runs.append(_lib.profile_collect())
    finally:
        _lib.profile_enable(False)
    first, second = runs
    print("profiled", {k: v[1:] for k, v in first.items()}, flush=True)
    for name in ("fod_lr_corr", "fod_lr", "fraction_loop", "fod_hr", "fraction_final"):
        assert name in first and first[name][0] > 0.0 and first[name][1] > 0
    # The FOD kernels count their solves' work; the others count none; the count is a
    # property of the solves, so it repeats exactly.
    assert first["fod_lr"][2] > 0 and first["fod_hr"][2] > 0
    assert first["fraction_loop"][2] == 0 and first["fod_lr_corr"][2] == 0
    assert [v[1:] for v in first.values()] == [v[1:] for v in second.values()]


@pytest.mark.parametrize("n_lr", [20, 60, 70])
def test_multi_tissue_low_resolution_widths(n_lr):
    """Low-resolution designs of fewer than, more than one pass of (48) and more than 64 columns:
    the correlation pass's column passes and the multi-slot FOD kernel, bit-identical to the
    oracle."""
    cfg = dict(synthetic.MFOD_CONFIG)
    cfg["shape"] = (10, 10, 8)
    cfg["seed"] = 11
    ph = synthetic.phantom(**cfg)
    signals = synthetic.normalized_signals(ph.data, ph.bvals, ph.mask)
    rfs = [AnisotropicRF([1.7e-3, 0.3e-3, 0.3e-3], 0.0), IsotropicRF(0.8e-3), IsotropicRF(3e-3)]
    w = np.ones(len(ph.bvals))
    hr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=[get_dirs(n_lr + 1)[:-1]], weights=w)
    assert lr.w_k_aniso.shape[1] == n_lr
    s = np.ascontiguousarray(signals[:150])
    print("lr width", n_lr, s.shape, flush=True)
    ref_fods, ref_fr, tr = oracle.fit_multi_tissue(s, w, lr, hr, 0.0, 0.1, 1e-6, trace=True)
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), 0.1, 1e-6)
    fods, fr, rounds, _ = _core.fit_multi_tissue(
        s, w, _deconv(lr), lr_solver, _deconv(hr), hr_solver, 50, 1e-7, 1e-2, 1,
        return_diagnostics=True)
    np.testing.assert_array_equal(rounds, tr[:, 0])
    np.testing.assert_array_equal(fods, ref_fods)
    np.testing.assert_array_equal(fr, ref_fr)


@pytest.mark.parametrize("shells", [[(700.0, 90), (1000.0, 90), (2000.0, 90), (3000.0, 90)],
                                    [(700.0, 80), (1000.0, 80), (1500.0, 80), (2000.0, 80),
                                     (2500.0, 80), (3000.0, 80)]])
def test_multi_tissue_many_measurements(shells):
    """Acquisitions of 372 and 490 measurements (six and eight signal rows per lane), as the
    batched requests and the 4-shell brain-sized workload bring them, bit-identical to the
    oracle."""
    ph = synthetic.phantom((6, 6, 4), 10, shells, seed=13, snr=25.0, mask_share=None, tissues=True)
    signals = synthetic.normalized_signals(ph.data, ph.bvals, ph.mask)
    rfs = [AnisotropicRF([1.7e-3, 0.3e-3, 0.3e-3], 0.0), IsotropicRF(0.8e-3), IsotropicRF(3e-3)]
    w = np.ones(len(ph.bvals))
    hr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=[get_dirs(45)[:-1]], weights=w)
    s = np.ascontiguousarray(signals[:120])
    print("measurements", s.shape, flush=True)
    ref_fods, ref_fr, tr = oracle.fit_multi_tissue(s, w, lr, hr, 0.0, 0.1, 1e-6, trace=True)
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), 0.1, 1e-6)
    fods, fr, rounds, _ = _core.fit_multi_tissue(
        s, w, _deconv(lr), lr_solver, _deconv(hr), hr_solver, 50, 1e-7, 1e-2, 1,
        return_diagnostics=True)
    np.testing.assert_array_equal(rounds, tr[:, 0])
    np.testing.assert_array_equal(fods, ref_fods)
    np.testing.assert_array_equal(fr, ref_fr)


# Up to 3 anisotropic blocks and 6 fraction columns: two anisotropic tissues (300 high-resolution
# directions each) with four isotropic ones, and three (200 each) with three.
@pytest.mark.parametrize("n_aniso,n_hr,isos", [(2, 300, [0.3e-3, 0.8e-3, 1.5e-3, 3e-3]),
                                               (3, 200, [0.8e-3, 1.5e-3, 3e-3])])
def test_six_fraction_columns_bit_identical_to_oracle(n_aniso, n_hr, isos):
    cfg = dict(synthetic.MFOD_CONFIG)
    cfg["shape"] = (8, 8, 6)
    cfg["seed"] = 13
    ph = synthetic.phantom(**cfg)
    signals = np.ascontiguousarray(synthetic.normalized_signals(ph.data, ph.bvals, ph.mask)[:160])
    tensors = [[1.7e-3, 0.3e-3, 0.3e-3], [1.4e-3, 0.5e-3, 0.5e-3], [2.0e-3, 0.2e-3, 0.2e-3]]
    rfs = [AnisotropicRF(t, 0.0) for t in tensors[:n_aniso]] + [IsotropicRF(d) for d in isos]
    w = np.ones(len(ph.bvals))
    hr_dirs = get_dirs(n_hr + 1)[:-1] if n_hr != 300 else DIRS_300
    hr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=[hr_dirs] * n_aniso, weights=w)
    lr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=[get_dirs(45)[:-1]] * n_aniso, weights=w)
    print(f"{n_aniso} anisotropic blocks, {len(isos)} isotropic tissues: {signals.shape}",
          flush=True)
    ref_fods, ref_fr, tr = oracle.fit_multi_tissue(signals, w, lr, hr, 0.0, 0.1, 1e-6, trace=True)
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), 0.1, 1e-6)
    fods, fr, rounds, status = _core.fit_multi_tissue(
        signals, w, _deconv(lr), lr_solver, _deconv(hr), hr_solver, 50, 1e-7, 1e-2, 1,
        return_diagnostics=True)
    assert fr.shape == (len(signals), n_aniso + len(isos))
    np.testing.assert_array_equal(rounds, tr[:, 0])
    np.testing.assert_array_equal(fods, ref_fods)
    np.testing.assert_array_equal(fr, ref_fr)


def test_640_directions_the_column_limit():
    """640 FOD directions, the most the solve kernels take (ten column slots of every lane): the
    single-tissue fit (ridge 0.1 and the QR-only 1e-7) and the multi-tissue fit."""
    cfg = dict(synthetic.MFOD_CONFIG)
    cfg["shape"] = (6, 6, 4)
    cfg["seed"] = 7
    ph = synthetic.phantom(**cfg)
    signals = np.ascontiguousarray(synthetic.normalized_signals(ph.data, ph.bvals, ph.mask))
    rfs = [AnisotropicRF([1.7e-3, 0.3e-3, 0.3e-3], 0.0), IsotropicRF(0.8e-3), IsotropicRF(3e-3)]
    w = np.ones(len(ph.bvals))
    hr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=[get_dirs(640)], weights=w)
    lr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=[get_dirs(45)[:-1]], weights=w)
    k = np.ascontiguousarray(hr.w_k_aniso)
    assert k.shape[1] == 640
    print(f"640 directions: {signals.shape}", flush=True)
    for ridge in (0.1, 1e-7):
        ref = oracle.fit_single_tissue(signals, k, ridge, 1e-6)
        got = _core.fit_single_tissue(signals, _core.FodSolver.l2nnls(k, ridge, 1e-6), 1)
        np.testing.assert_array_equal(got, ref)
    ref_fods, ref_fr, tr = oracle.fit_multi_tissue(signals, w, lr, hr, 0.0, 0.1, 1e-6, trace=True)
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    fods, fr, rounds, status = _core.fit_multi_tissue(
        signals, w, _deconv(lr), lr_solver, _deconv(hr), _core.FodSolver.l2nnls(k, 0.1, 1e-6),
        50, 1e-7, 1e-2, 1, return_diagnostics=True)
    np.testing.assert_array_equal(rounds, tr[:, 0])
    np.testing.assert_array_equal(fods, ref_fods)
    np.testing.assert_array_equal(fr, ref_fr)
