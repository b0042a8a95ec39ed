"""The CSD kernel against the CPU FP64 oracle on the same inputs: bit-identical FODs,
converged flags and solve counts.

The oracle (oracle/sd_oracle.cpp, solve_csd) computes in the kernel's rounding order; it is
pinned by tests/test_oracle_csd.py.  Inputs: the single-shell CSD workload (a 32 x 32 x 20
phantom, 5 b0 + 60 directions at b = 1000, 1-3 fibres, Rician SNR 20) with the tensor response
(1.7, 0.3, 0.3) um^2/ms as zonal SH, the reference's CSD defaults and the 300 built-in
constraint directions.
"""

import numpy as np
import pytest

import oracle
from larmor_amd import _core, synthetic
from larmor_amd.sd.csd import CSDModel
from larmor_amd.sd.csd_solver import CSDSolver, csd_problem, solve_csd
from larmor_amd.sd.rf import ZonalSHRF
from larmor_amd.types import GradientTable, Volume

pytestmark = pytest.mark.gpu


def _workload(shape=(16, 16, 8)):
    cfg = dict(synthetic.CSD_CONFIG)
    cfg["shape"] = shape
    ph = synthetic.phantom(**cfg)
    zonal = synthetic.tensor_zonal(1000.0)
    on_shell = ph.bvals > 50
    signals = ph.data.reshape(-1, len(ph.bvals))[:, on_shell].astype(np.float64)
    return ph, zonal, np.ascontiguousarray(np.maximum(signals, 0.0)), ph.bvecs[on_shell]


def _both(signals, p, max_iter=None):
    it = p.max_iter if max_iter is None else max_iter
    ref = oracle.fit_csd(signals, p.initial, p.forward, p.constraint, p.norm_weight,
                         p.threshold, it, return_solves=True)
    got = _core.fit_csd(signals, p.initial, p.forward, p.constraint, p.norm_weight,
                        p.threshold, it, 1, return_solves=True)
    return ref, got


def _assert_same(ref, got):
    np.testing.assert_array_equal(got[1], ref[1])
    np.testing.assert_array_equal(got[2], ref[2])
    np.testing.assert_array_equal(got[0], ref[0])


@pytest.mark.parametrize("lmax", [8, 6, 4, 2])
def test_csd_bit_identical_to_oracle(lmax):
    _, zonal, signals, bvecs = _workload()
    p = csd_problem(bvecs, zonal, lmax, CSDSolver())
    print(f"csd lmax {lmax}: {signals.shape}", flush=True)
    ref, got = _both(signals, p)
    assert ref[1].mean() > 0.99
    _assert_same(ref, got)


@pytest.mark.parametrize("max_iter", [0, 1, 2, 3])
def test_csd_iteration_limits(max_iter):
    _, zonal, signals, bvecs = _workload(shape=(8, 8, 4))
    p = csd_problem(bvecs, zonal, 8, CSDSolver(max_iter=max_iter))
    ref, got = _both(signals, p)
    _assert_same(ref, got)


def test_csd_threshold_and_custom_directions():
    _, zonal, signals, bvecs = _workload(shape=(8, 8, 4))
    dirs = synthetic.spiral_directions(128)
    p = csd_problem(bvecs, zonal, 8, CSDSolver(threshold=0.1, directions=dirs))
    ref, got = _both(signals, p)
    _assert_same(ref, got)


def test_csd_edge_cases():
    _, zonal, signals, bvecs = _workload(shape=(4, 4, 2))
    p = csd_problem(bvecs, zonal, 8, CSDSolver())
    edge = np.vstack([np.zeros(signals.shape[1]), np.full(signals.shape[1], 1e6),
                      np.full(signals.shape[1], np.nan), signals[:5]])
    ref, got = _both(np.ascontiguousarray(edge), p)
    _assert_same(ref, got)
    empty = np.zeros((0, signals.shape[1]))
    fods, conv = _core.fit_csd(empty, p.initial, p.forward, p.constraint, p.norm_weight,
                               p.threshold, p.max_iter, 1)
    assert fods.shape == (0, 45) and conv.shape == (0,)


def test_csd_batch_invariance():
    _, zonal, signals, bvecs = _workload()
    p = csd_problem(bvecs, zonal, 8, CSDSolver())
    alone = _core.fit_csd(np.ascontiguousarray(signals[7:9]), p.initial, p.forward,
                          p.constraint, p.norm_weight, p.threshold, p.max_iter, 1)[0]
    mixed = _core.fit_csd(signals, p.initial, p.forward, p.constraint, p.norm_weight,
                          p.threshold, p.max_iter, 1)[0]
    np.testing.assert_array_equal(alone, mixed[7:9])


def test_csd_model_fit_matches_oracle():
    ph, zonal, _, _ = _workload(shape=(12, 12, 6))
    gtab = GradientTable(ph.bvals, ph.bvecs)
    rf = ZonalSHRF(bvals=np.array([1000.0]), coefficients=zonal[None])
    mask = np.zeros(ph.data.shape[:3], dtype=bool)
    mask[2:10, 3:11, 1:5] = True
    fit = CSDModel(rf, lmax=8).fit(Volume(ph.data, np.eye(4)), gtab, mask=mask, n_jobs=1)
    on_shell = ph.bvals > 50
    rows = ph.data[mask][:, on_shell].astype(np.float32).astype(np.float64)
    p = csd_problem(ph.bvecs[on_shell], zonal, 8, CSDSolver())
    ref = oracle.fit_csd(np.maximum(rows, 0.0), p.initial, p.forward, p.constraint,
                         p.norm_weight, p.threshold, p.max_iter)[0]
    got = fit.fods[0]
    assert got.dtype == np.float32 and got.shape == ph.data.shape[:3] + (45,)
    np.testing.assert_array_equal(got[mask], ref.astype(np.float32))
    assert (got[~mask] == 0).all()


def test_solve_csd_keeps_device_tensors_on_device():
    import torch

    _, zonal, signals, bvecs = _workload(shape=(4, 4, 2))
    p = csd_problem(bvecs, zonal, 8, CSDSolver())
    dev = torch.from_numpy(signals).to("cuda")
    fods, conv = solve_csd(dev, p, 1)
    assert fods.device.type == "cuda"
    ref = oracle.fit_csd(signals, p.initial, p.forward, p.constraint, p.norm_weight,
                         p.threshold, p.max_iter)[0]
    np.testing.assert_array_equal(fods.cpu().numpy(), ref)


def _high_b_workload(shape, n_dirs=150, bval=3000.0, seed=5):
    ph = synthetic.phantom(shape=shape, n_b0=2, shells=[(bval, n_dirs)], seed=seed, snr=30.0,
                           mask_share=None, tissues=False)
    on_shell = ph.bvals > 50
    signals = ph.data.reshape(-1, len(ph.bvals))[:, on_shell].astype(np.float64)
    return np.ascontiguousarray(np.maximum(signals, 0.0)), ph.bvecs[on_shell]


# lmax 10-18: the wide form with the system's triangle in LDS (66-190 coefficients); lmax 20:
# its 231-coefficient triangle (208 KiB) exceeds a workgroup's LDS and lives in global memory.
@pytest.mark.parametrize("lmax,shape", [(10, (8, 8, 4)), (12, (8, 8, 2)), (14, (6, 6, 2)),
                                        (16, (6, 4, 2)), (18, (4, 4, 2)), (20, (4, 4, 2))])
def test_csd_wide_lmax_bit_identical_to_oracle(lmax, shape):
    signals, bvecs = _high_b_workload(shape)
    zonal = synthetic.tensor_zonal(3000.0, lmax=lmax)
    p = csd_problem(bvecs, zonal, lmax, CSDSolver())
    print(f"csd wide lmax {lmax}: {signals.shape}, {p.forward.shape[1]} coefficients", flush=True)
    ref, got = _both(signals, p)
    assert ref[1].mean() > 0.9 and ref[2].max() >= 2
    _assert_same(ref, got)


def test_csd_wide_any_coefficient_count():
    # A random well-posed problem whose coefficient count (50) has no register kernel, with
    # constraint directions in every slot class.
    rng = np.random.default_rng(11)
    m, c, c0 = 70, 50, 20
    forward = rng.normal(size=(m, c))
    initial = rng.normal(size=(c0, m)) * 0.05
    signals = np.ascontiguousarray(rng.normal(size=(300, m)))
    for k in (100, 300, 500):
        constraint = np.ascontiguousarray(rng.normal(size=(k, c)))
        print(f"csd wide c {c} k {k}", flush=True)
        ref = oracle.fit_csd(signals, initial, forward, constraint, 0.5, 0.0, 12,
                             return_solves=True)
        got = _core.fit_csd(signals, initial, forward, constraint, 0.5, 0.0, 12, 1,
                            return_solves=True)
        _assert_same(ref, got)


def test_csd_wide_at_the_coefficient_limit():
    # 496 coefficients (lmax 30), the largest system the kernels take: its 964 KiB triangle in
    # global memory, with constraint directions in the largest slot class.
    rng = np.random.default_rng(12)
    m, c, c0, k = 520, 496, 120, 500
    forward = rng.normal(size=(m, c))
    initial = rng.normal(size=(c0, m)) * 0.05
    constraint = np.ascontiguousarray(rng.normal(size=(k, c)))
    signals = np.ascontiguousarray(rng.normal(size=(6, m)))
    print(f"csd wide c {c} k {k}", flush=True)
    ref = oracle.fit_csd(signals, initial, forward, constraint, 0.5, 0.0, 6, return_solves=True)
    got = _core.fit_csd(signals, initial, forward, constraint, 0.5, 0.0, 6, 1, return_solves=True)
    assert ref[2].max() >= 2
    _assert_same(ref, got)


def test_csd_wide_batch_invariance_and_limit():
    signals, bvecs = _high_b_workload((4, 4, 2))
    p = csd_problem(bvecs, synthetic.tensor_zonal(3000.0, lmax=12), 12, CSDSolver())
    alone = _core.fit_csd(np.ascontiguousarray(signals[3:5]), p.initial, p.forward,
                          p.constraint, p.norm_weight, p.threshold, p.max_iter, 1)[0]
    mixed = _core.fit_csd(signals, p.initial, p.forward, p.constraint, p.norm_weight,
                          p.threshold, p.max_iter, 1)[0]
    np.testing.assert_array_equal(alone, mixed[3:5])
    with pytest.raises(ValueError, match="CSD on the GPU takes 1 to 496 coefficients"):
        _core.fit_csd(signals[:2], np.zeros((0, signals.shape[1])),
                      np.zeros((signals.shape[1], 497)), np.zeros((10, 497)), 1.0, 0.0, 5, 1)
