"""The batched multi-tissue fit (lsd_fit_multi_tissue_batched) on the GPU: volumes with their own
acquisitions fitted in one launch sequence give, volume by volume, the same bits (FODs,
fractions, rounds and status) as lsd_fit_multi_tissue on the volume alone -- which the oracle
pins (tests/test_gpu_parity.py) -- and the batch's job table and checks behave.
"""

import numpy as np
import pytest

from larmor_amd import _core, synthetic
from larmor_amd.sd.kernel import Kernel
from larmor_amd.sd.model import compute_shell_weights, low_resolution_dirs
from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
from larmor_amd.sd.solvers import FINAL_FRACTION_TOL_X, MAX_ROUNDS, _deconv_kernel
from larmor_amd.sd.sphere import DIRS_300

pytestmark = pytest.mark.gpu

RFS = [AnisotropicRF([1.7e-3, 0.3e-3, 0.3e-3], 0.0), IsotropicRF(0.8e-3), IsotropicRF(3.0e-3)]
FRACTION_TOL = 1e-3


def _design(bvals, bvecs, lr_ridge=0.0, hr_ridge=0.1):
    weights = compute_shell_weights(bvals, 0.2)
    hr = Kernel(RFS, bvals, bvecs, dirs=DIRS_300, weights=weights)
    lr = Kernel(RFS, bvals, bvecs, dirs=low_resolution_dirs(RFS, len(bvals)), weights=weights)
    return (weights, _deconv_kernel(lr), _core.FodSolver.l2nnls(np.ascontiguousarray(lr.w_k_aniso), lr_ridge, 1e-6),
            _deconv_kernel(hr), _core.FodSolver.l2nnls(np.ascontiguousarray(hr.w_k_aniso), hr_ridge, 1e-6))


def _jobs(seed, n_jobs, lo, hi, n_shells=(1, 3), n_dirs=(30, 64)):
    return synthetic.batched_jobs(n_jobs, seed=seed, min_voxels=lo, max_voxels=hi,
                                  n_shells=n_shells, n_dirs=n_dirs)


def _fit_alone(signals, design):
    weights, lr, lr_s, hr, hr_s = design
    return _core.fit_multi_tissue(signals, weights, lr, lr_s, hr, hr_s, MAX_ROUNDS, FRACTION_TOL,
                                  FINAL_FRACTION_TOL_X, 1, return_diagnostics=True)


def _batched(jobs, designs, max_outer=MAX_ROUNDS):
    import torch

    entries = []
    for job, (weights, lr, lr_s, hr, hr_s) in zip(jobs, designs):
        signals = torch.from_numpy(job.signals.astype(np.float64)).to("cuda")
        entries.append((signals, torch.as_tensor(weights, device="cuda"), lr, lr_s, hr, hr_s))
    out = _core.fit_multi_tissue_batched(entries, max_outer, FRACTION_TOL, FINAL_FRACTION_TOL_X,
                                         return_diagnostics=True)
    return [t.cpu().numpy() for t in out]


@pytest.mark.parametrize("seed,n_jobs", [(11, 5), (12, 9)])
def test_each_volume_equals_its_fit_alone(seed, n_jobs):
    # The volumes of a batch share their direction counts: more than 45 measurements give every
    # one the full 44 low-resolution directions.
    jobs = [j for j in _jobs(seed, n_jobs * 3, 300, 6000) if 45 < len(j.bvals) <= 256][:n_jobs]
    print(f"batched seed={seed}: {len(jobs)} volumes, m={[len(j.bvals) for j in jobs]}, "
          f"voxels={[len(j.signals) for j in jobs]}", flush=True)
    designs = [_design(j.bvals, j.bvecs) for j in jobs]
    fods, fractions, rounds, status = _batched(jobs, designs)
    start = 0
    for job, design in zip(jobs, designs):
        stop = start + len(job.signals)
        ref = _fit_alone(job.signals.astype(np.float64), design)
        np.testing.assert_array_equal(fods[start:stop], ref[0])
        np.testing.assert_array_equal(fractions[start:stop], ref[1])
        np.testing.assert_array_equal(rounds[start:stop], ref[2])
        np.testing.assert_array_equal(status[start:stop], ref[3])
        start = stop
    assert start == len(fods)


def test_larger_row_class_and_qr_only_designs():
    """Volumes of 257-384 measurements (six row slots), one of them with a high-resolution ridge
    small enough for the QR-only path (normal equations refused), and an empty volume."""
    jobs = [j for j in _jobs(21, 24, 400, 3000, n_shells=(3, 4), n_dirs=(70, 90))
            if 256 < len(j.bvals) <= 384][:3]
    assert len(jobs) == 3
    jobs.append(synthetic.Job(jobs[0].signals[:0], jobs[0].bvals, jobs[0].bvecs, jobs[0].lmax))
    designs = [_design(j.bvals, j.bvecs) for j in jobs[:1]]
    designs.append(_design(jobs[1].bvals, jobs[1].bvecs, hr_ridge=1e-7))
    designs += [_design(j.bvals, j.bvecs) for j in jobs[2:]]
    assert not designs[1][4].normal_equations
    fods, fractions, rounds, status = _batched(jobs, designs)
    start = 0
    for job, design in zip(jobs, designs):
        stop = start + len(job.signals)
        if stop > start:
            ref = _fit_alone(job.signals.astype(np.float64), design)
            for got, want in zip((fods, fractions, rounds, status), ref):
                np.testing.assert_array_equal(got[start:stop], want)
        start = stop


def test_round_limits():
    jobs = [j for j in _jobs(31, 9, 200, 1500) if 45 < len(j.bvals) <= 256][:3]
    designs = [_design(j.bvals, j.bvecs) for j in jobs]
    for max_outer in (0, 1, 3):
        fods, fractions, rounds, status = _batched(jobs, designs, max_outer)
        start = 0
        for job, (weights, lr, lr_s, hr, hr_s) in zip(jobs, designs):
            stop = start + len(job.signals)
            ref = _core.fit_multi_tissue(job.signals.astype(np.float64), weights, lr, lr_s, hr,
                                         hr_s, max_outer, FRACTION_TOL, FINAL_FRACTION_TOL_X, 1,
                                         return_diagnostics=True)
            for got, want in zip((fods, fractions, rounds, status), ref):
                np.testing.assert_array_equal(got[start:stop], want)
            start = stop


def test_batch_checks():
    import torch

    jobs = [j for j in _jobs(41, 30, 100, 400) if 45 < len(j.bvals) <= 256][:2]
    big = [j for j in _jobs(42, 30, 100, 400, n_shells=(4, 4), n_dirs=(80, 90))
           if len(j.bvals) > 256][:1]
    designs = [_design(j.bvals, j.bvecs) for j in jobs + big]
    entries = []
    for job, (weights, lr, lr_s, hr, hr_s) in zip(jobs + big, designs):
        signals = torch.from_numpy(job.signals.astype(np.float64)).to("cuda")
        entries.append((signals, torch.as_tensor(weights, device="cuda"), lr, lr_s, hr, hr_s))
    with pytest.raises(ValueError, match="row-slot class"):
        _core.fit_multi_tissue_batched(entries, 50, FRACTION_TOL, FINAL_FRACTION_TOL_X)
    with pytest.raises(ValueError, match="1 to 64 jobs"):
        _core.fit_multi_tissue_batched([], 50, FRACTION_TOL, FINAL_FRACTION_TOL_X)
    with pytest.raises(ValueError, match="1 to 64 jobs"):
        _core.fit_multi_tissue_batched(entries[:1] * 65, 50, FRACTION_TOL, FINAL_FRACTION_TOL_X)
    swapped = list(entries[0])
    swapped[2], swapped[4] = swapped[4], swapped[2]  # the LR and HR kernels exchanged
    with pytest.raises(ValueError, match="job 0"):
        _core.fit_multi_tissue_batched([tuple(swapped)], 50, FRACTION_TOL, FINAL_FRACTION_TOL_X)
    with pytest.raises(ValueError, match="max_outer"):
        _core.fit_multi_tissue_batched(entries[:1], -1, FRACTION_TOL, FINAL_FRACTION_TOL_X)


def test_captured_graph_replays_equal_the_eager_fit():
    """The batched fit captured as a HIP graph: replayed three times on new signals, with eager
    work and synchronisations between the replays, every replay equals the eager batched fit of
    the same signals bit for bit."""
    import torch

    jobs = [j for j in _jobs(31, 15, 300, 3000) if 45 < len(j.bvals) <= 256][:4]
    designs = [_design(j.bvals, j.bvecs) for j in jobs]
    entries = [(torch.from_numpy(j.signals.astype(np.float64)).to("cuda"),
                torch.as_tensor(d[0], device="cuda"), *d[1:]) for j, d in zip(jobs, designs)]
    graph = _core.BatchedFitGraph(entries, MAX_ROUNDS, FRACTION_TOL, FINAL_FRACTION_TOL_X)
    rng = np.random.default_rng(0)
    for replay in range(3):
        # New signals each time: the jobs' rows shuffled and scaled.
        signals = [e[0][torch.from_numpy(rng.permutation(e[0].shape[0])).to("cuda")]
                   * (1.0 + 0.01 * replay) for e in entries]
        eager = _core.fit_multi_tissue_batched(
            [(s, *e[1:]) for s, e in zip(signals, entries)], MAX_ROUNDS, FRACTION_TOL,
            FINAL_FRACTION_TOL_X, return_diagnostics=True)
        eager = [t.cpu().numpy() for t in eager]
        _ = torch.randn(512, 512, device="cuda") @ torch.randn(512, 512, device="cuda")
        torch.cuda.synchronize()
        out = [t.cpu().numpy() for t in graph.replay(signals)]
        for a, b in zip(out, eager):
            np.testing.assert_array_equal(a, b)


def test_every_tuned_capacity_gives_the_same_bits_and_the_tuner_keeps_the_fastest():
    """The low-resolution factor capacity (16, 20, 24 or 32 columns) changes which voxels finish
    in the global-memory solve, never a bit of the results; the tuner times each on the bucket's
    sample, keeps the fastest and puts it in force."""
    import torch

    from larmor_amd.sd import tuning

    jobs = [j for j in _jobs(41, 12, 3000, 20000) if 45 < len(j.bvals) <= 256][:2]
    designs = [_design(j.bvals, j.bvecs) for j in jobs]
    m = len(jobs[0].bvals)
    try:
        ref = None
        for cap in tuning.CAPACITIES:
            tuning.set_capacity(m, cap)
            assert tuning.capacity_in_force(m) == cap
            single = _fit_alone(jobs[0].signals.astype(np.float64), designs[0])
            batch = _batched(jobs, designs)
            got = [np.asarray(x) for x in single] + batch
            if ref is None:
                ref = got
            for a, b in zip(got, ref):
                np.testing.assert_array_equal(a, b)
        table = tuning.TuningTable()
        w, lr, lrs, hr, hrs = designs[0]
        chosen = tuning.autotune(torch.from_numpy(jobs[0].signals.astype(np.float64)).to("cuda"),
                                 w, lr, lrs, hr, hrs, fraction_tol=FRACTION_TOL,
                                 final_tol_x=FINAL_FRACTION_TOL_X, sample=4096, repeats=1,
                                 table=table)
        times = table.timings[table.bucket(m, lr.n_anisotropic)]
        print(f"tuned capacity {chosen} for m={m}: ms {times}")
        assert chosen == min(times, key=times.get) and tuning.capacity_in_force(m) == chosen
    finally:
        tuning.set_capacity(m, 0)
    with pytest.raises(ValueError, match="16, 20, 24 or 32 columns"):
        tuning.set_capacity(m, 40)
