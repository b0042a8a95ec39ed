"""The plain FP64 fits of 513 to 1,024 measurements (12 and 16 row slots per lane) against the
CPU FP64 oracle on the same inputs: bit-identical FODs, fractions and rounds.  (Acquisitions of
this size exist: the MGH Connectom protocol has 552 volumes.)  The other forms keep their
512-measurement limit and say so."""

import numpy as np
import pytest

import oracle
from larmor_amd import _core, synthetic
from larmor_amd.sd.kernel import Kernel
from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
from larmor_amd.sd.sphere import DIRS_300, get_dirs

pytestmark = pytest.mark.gpu


def _setup(shells, seed, shape=(6, 6, 4)):
    ph = synthetic.phantom(shape=shape, n_b0=10, shells=shells, seed=seed, snr=25.0,
                           mask_share=None, tissues=True)
    signals = synthetic.normalized_signals(ph.data, ph.bvals, ph.mask)
    rfs = [AnisotropicRF([1.7e-3, 0.3e-3, 0.3e-3], 0.0), IsotropicRF(0.8e-3), IsotropicRF(3e-3)]
    w = np.ones(len(ph.bvals))
    hr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(rfs, ph.bvals, ph.bvecs, dirs=[get_dirs(45)[:-1]], weights=w)
    return np.ascontiguousarray(signals), w, lr, hr


def _deconv(k):
    return _core.DeconvKernel(k.k_aniso, k.k_iso, k.w_k_iso,
                              [(b.start, b.stop) for b in k.aniso_blocks])


# 610 measurements: 12 row slots; 1,010: 16; 1,024: the limit, every row slot of every lane.
@pytest.mark.parametrize("shells", [[(1000.0, 200), (2000.0, 200), (3000.0, 200)],
                                    [(700.0, 250), (1000.0, 250), (2000.0, 250), (3000.0, 250)],
                                    [(700.0, 253), (1000.0, 253), (2000.0, 254), (3000.0, 254)]])
def test_multi_and_single_tissue_fits_bit_identical(shells):
    signals, w, lr, hr = _setup(shells, seed=len(shells))
    m = signals.shape[1]
    print(f"many measurements: {signals.shape}", flush=True)
    assert m > 512
    ref_fods, ref_fr, tr = oracle.fit_multi_tissue(signals, w, lr, hr, 0.0, 0.1, 1e-6, trace=True)
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), 0.1, 1e-6)
    fods, fr, rounds, status = _core.fit_multi_tissue(
        signals, w, _deconv(lr), lr_solver, _deconv(hr), hr_solver, 50, 1e-7, 1e-2, 1,
        return_diagnostics=True)
    np.testing.assert_array_equal(rounds, tr[:, 0])
    np.testing.assert_array_equal(fods, ref_fods)
    np.testing.assert_array_equal(fr, ref_fr)
    k = np.ascontiguousarray(hr.w_k_aniso)
    ref = oracle.fit_single_tissue(signals[:60], k, 0.1, 1e-6)
    got = _core.fit_single_tissue(np.ascontiguousarray(signals[:60]), _core.FodSolver.l2nnls(k, 0.1, 1e-6), 1)
    np.testing.assert_array_equal(got, ref)


def test_single_problem_nnls_with_many_rows():
    rng = np.random.default_rng(8)
    for m, n in ((700, 50), (1024, 120)):
        c = rng.normal(size=(m, n))
        d = c @ np.maximum(rng.normal(size=n), 0.0) + 0.1 * rng.normal(size=m)
        print(f"lsqnonneg {m} x {n}", flush=True)
        ref, _ = oracle.nnls(c, d)
        got = _core.lsqnonneg(c, d)
        np.testing.assert_array_equal(got, ref)


def test_limits_beyond_1024_and_for_the_other_forms():
    signals, w, lr, hr = _setup([(1000.0, 200), (2000.0, 200), (3000.0, 200)], seed=9,
                                shape=(3, 3, 2))
    lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
    hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), 0.1, 1e-6)
    with pytest.raises(ValueError, match="512 measurements"):
        _core.fit_multi_tissue(signals, w, _deconv(lr), lr_solver, _deconv(hr), hr_solver, 50,
                               1e-7, 1e-2, 1, form="mixed")
    c = np.zeros((1030, 10))
    with pytest.raises(ValueError, match="1024 measurements"):
        _core.lsqnonneg(c, np.zeros(1030), 1e-10)


def test_every_tuned_capacity_gives_the_same_bits_at_12_and_16_row_slots():
    """The low-resolution factor capacity of the 768- and 1,024-measurement row classes is tunable
    like the smaller ones, and every capacity gives the same bits."""
    from larmor_amd.sd import tuning

    for shells, seed in (([(1000.0, 200), (2000.0, 200), (3000.0, 200)], 11),
                         ([(700.0, 250), (1000.0, 250), (2000.0, 250), (3000.0, 250)], 12)):
        signals, w, lr, hr = _setup(shells, seed=seed, shape=(5, 5, 3))
        m = signals.shape[1]
        print(f"tuned capacities at {m} measurements", flush=True)
        lr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), 0.0, 1e-6)
        hr_solver = _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), 0.1, 1e-6)
        assert tuning.row_class(m) == (768 if m <= 768 else 1024)
        ref = None
        try:
            for cap in tuning.CAPACITIES:
                tuning.set_capacity(m, cap)
                assert tuning.capacity_in_force(m) == cap
                got = _core.fit_multi_tissue(signals, w, _deconv(lr), lr_solver, _deconv(hr),
                                             hr_solver, 50, 1e-7, 1e-2, 1,
                                             return_diagnostics=True)
                got = [np.asarray(x) for x in got]
                if ref is None:
                    ref = got
                for a, b in zip(got, ref):
                    np.testing.assert_array_equal(a, b)
        finally:
            tuning.set_capacity(m, 0)
        assert tuning.capacity_in_force(m) == tuning.DEFAULT_CAPACITY
    assert tuning.capacity_in_force(1030) == -1
    with pytest.raises(ValueError, match="1024 measurements"):
        tuning.row_class(1030)
