"""The CPU oracle's fractions and multi-tissue loop against the reference's known answers.

- median and compute_fractions closed forms (cpp/tests/sd/test_fractions.cpp:25-50) on the
  reference's toy kernel (cpp/tests/sd/toy_kernel.hpp:14-48);
- the multi-tissue invariants (cpp/tests/sd/test_multi_tissue.cpp:56-90), with the L2NNLS
  solvers of the mFOD method in place of the reference test's dRL;
- the MRIToolkit mFOD fixture (tests/sd/test_matlab_match.py:109-165): the kernels to 1e-6 and
  the fit on 512 voxels to FOD max-rel 1e-3 and fractions atol 1e-4.
"""

from pathlib import Path

import numpy as np
import pytest

import oracle
from larmor_amd.sd.kernel import Kernel
from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF, NoddiRF
from larmor_amd.sd.sphere import DIRS_300, get_dirs

GOLDEN = Path(__file__).resolve().parent / "golden"


class ToyKernel:
    """toy_kernel.hpp: 12 measurements on three shells, n_aniso columns and two decays."""

    M = 12

    def __init__(self, n_aniso: int):
        meas = np.arange(self.M)
        shell = (meas // 4).astype(np.float64)
        cols = np.arange(n_aniso)
        self.k_aniso = np.exp(
            -0.3 * shell[:, None] * (1.0 + ((meas[:, None] * 5 + cols[None, :] * 3) % 7) / 3.0)
        )
        self.k_iso = np.exp(-shell[:, None] * (0.4 + 1.2 * np.arange(2)[None, :]))
        self.weights = 1.0 + (meas % 3) * 0.5
        self.w_k_iso = self.weights[:, None] * self.k_iso
        self.w_k_aniso = self.weights[:, None] * self.k_aniso
        self.aniso_blocks = [slice(0, n_aniso)]


def test_median_odd_and_even():  # test_fractions.cpp:25
    assert oracle.median([3.0, 1.0, 2.0]) == 2.0
    assert oracle.median([4.0, 1.0, 3.0, 2.0]) == 2.5
    assert np.isnan(oracle.median(np.array([], dtype=np.float64)))
    # NaN sorts last.
    assert oracle.median([np.nan, 1.0, 2.0]) == 2.0


def test_pure_isotropic_signal():  # test_fractions.cpp:30
    k = ToyKernel(6)
    fr = oracle.compute_fractions(np.full(6, 1e-6), k.k_aniso, k.k_iso, k.aniso_blocks,
                                  k.k_iso[:, 1], lumped=True)
    np.testing.assert_allclose(fr, [0.0, 0.0, 1.0], atol=1e-10)


def test_median_threshold():  # test_fractions.cpp:39-50
    k = ToyKernel(6)
    fod = np.array([0.001, 0.002, 0.003, 1.0, 0.5, 0.004])
    kept = 1.0 * k.k_aniso[:, 3] + 0.5 * k.k_aniso[:, 4] + 0.004 * k.k_aniso[:, 5]
    signal = 2.0 * kept / kept.max()
    fr = oracle.compute_fractions(fod, k.k_aniso, k.k_iso, k.aniso_blocks, signal, lumped=True)
    np.testing.assert_allclose(fr, [2.0, 0.0, 0.0], atol=1e-10)


def _toy_signals(hr: ToyKernel) -> np.ndarray:  # test_multi_tissue.cpp:83-96
    s = np.empty((5, ToyKernel.M))
    for v in range(5):
        wm = (v + 1) / 5
        s[v] = (wm * hr.k_aniso[:, (2 * v) % 10] + (1 - wm) * 0.7 * hr.k_iso[:, 0]
                + (1 - wm) * 0.3 * hr.k_iso[:, 1])
    return s


def _toy_fit(signals):
    lr, hr = ToyKernel(6), ToyKernel(10)
    return oracle.fit_multi_tissue(signals, lr.weights, lr, hr, 0.0, 0.1, 1e-6, fraction_tol=1e-3,
                                   n_threads=2)


def test_block_equals_each_voxel_alone():  # test_multi_tissue.cpp:56-71
    s = _toy_signals(ToyKernel(10))
    fods, fr = _toy_fit(s)
    for v in range(5):
        f1, r1 = _toy_fit(s[v : v + 1])
        np.testing.assert_array_equal(fods[v], f1[0])
        np.testing.assert_array_equal(fr[v], r1[0])


def test_fractions_non_negative_sum_to_one_monotone():  # test_multi_tissue.cpp:73-90
    _, fr = _toy_fit(_toy_signals(ToyKernel(10)))
    assert (fr >= 0).all()
    np.testing.assert_allclose(fr.sum(axis=1), 1.0, atol=1e-12)
    assert (np.diff(fr[:, 0]) > 0).all()


@pytest.fixture(scope="module")
def mfod_fixture():
    f = np.load(GOLDEN / "mfod_fixture.npz")
    bvals, bvecs = f["bvals"], f["bvecs"]
    rfs = [AnisotropicRF([1.7e-3, 0.2e-3, 0.2e-3], 0.0), NoddiRF(), IsotropicRF(3e-3)]
    w = np.ones(len(bvals))
    hr = Kernel(rfs, bvals, bvecs, dirs=DIRS_300, weights=w)
    lr_dirs = [get_dirs(min(len(bvals), rf.low_resolution_dirs))[:-1] for rf in rfs[:2]]
    lr = Kernel(rfs, bvals, bvecs, dirs=lr_dirs, weights=w)
    sig = f["signal"].astype(np.float64)
    b0 = bvals <= bvals.min()
    signals = sig / sig[:, b0].mean(axis=1, keepdims=True)
    return f, lr, hr, w, signals


def test_mritoolkit_kernels(mfod_fixture):  # test_matlab_match.py:109-140
    f, lr, hr, _, _ = mfod_fixture
    np.testing.assert_allclose(hr.k, f["hr_kernel"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(lr.k, np.delete(f["lr_kernel"], [44, 134], axis=1), rtol=0,
                               atol=1e-6)


def test_mritoolkit_mfod_fit(mfod_fixture):  # test_matlab_match.py:143-165
    f, lr, hr, w, signals = mfod_fixture
    fods, fr = oracle.fit_multi_tissue(signals, w, lr, hr, 0.0, 0.1, 1e-6)
    ref = f["fod"].astype(np.float64)
    err = np.abs(fods - ref).max(axis=1) / ref.max(axis=1)
    assert err.max() <= 1e-3
    np.testing.assert_allclose(fr, f["fractions"].astype(np.float64), rtol=0, atol=1e-4)
