"""The served form on the GPU: requests fitted through FodServer while others are in flight give
the same bits as the plain fit of the same rows alone (the reference's batch-invariance rule,
tests/sd/test_fit.py:87-119, cpp/tests/test_multi_tissue.cpp:56), under continuous admission and
under an HBM budget that admits one request at a time; served fractions equal the oracle's."""

import numpy as np
import pytest

from larmor_amd import synthetic
from larmor_amd.sd.model import SDModel, _to_sh, compute_max_lmax
from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
from larmor_amd.sd.serving import FodServer, request_bytes
from larmor_amd.sd.sh import projection_matrix
from larmor_amd.sd.solvers import L2NNLS, Lasso
from larmor_amd.sd.sphere import DIRS_300
from larmor_amd.types import GradientTable

pytestmark = pytest.mark.gpu

WM = [1.7e-3, 0.3e-3, 0.3e-3]


def _model(method: str, multi: bool = True) -> SDModel:
    rfs = [AnisotropicRF(WM, 0.0)]
    if multi:
        rfs += [IsotropicRF(0.8e-3), IsotropicRF(3.0e-3)]
    return SDModel(rfs, method=method, lmax=8)


def _alone(model: SDModel, job: synthetic.Job):
    """What SDModel.fit does with these rows on its own: _prepare's chunk fit, then the SH
    projection of each anisotropic block."""
    import torch

    solver = model.solver if model.solver is not None else model.method.default_solver()
    lmax = compute_max_lmax(job.bvals, job.lmax)
    hr, fit_chunk = model._prepare(solver, job.bvals, job.bvecs, 1)
    dev = torch.from_numpy(job.signals.astype(np.float64)).to("cuda")
    amplitudes, fractions = fit_chunk(dev)
    proj = torch.as_tensor(projection_matrix(DIRS_300, lmax, model.basis).T
                           * (len(DIRS_300) / (4.0 * np.pi)), dtype=torch.float64, device="cuda")
    exact = isinstance(solver, (L2NNLS, Lasso))
    sh = [_to_sh(amplitudes[:, b], proj, exact_integral=exact).cpu().numpy()
          for b in hr.aniso_blocks]
    torch.cuda.synchronize()
    return sh, None if fractions is None else fractions.cpu().numpy()


def _jobs(n, seed, lo, hi):
    jobs = synthetic.batched_jobs(n, seed=seed, min_voxels=lo, max_voxels=hi)
    # one more request on the first acquisition, to be served from the resident design
    extra = synthetic.batched_jobs(1, seed=seed + 1000, min_voxels=lo, max_voxels=hi)[0]
    rng = np.random.default_rng(seed)
    signals = synthetic.tissue_signals(len(extra.signals), jobs[0].bvals, jobs[0].bvecs, rng,
                                       snr=25.0)
    jobs.append(synthetic.Job(signals, jobs[0].bvals, jobs[0].bvecs, jobs[0].lmax))
    return jobs


def _check_same(served, alone):
    sh, fractions = alone
    assert len(served.sh) == len(sh)
    for a, b in zip(served.sh, sh):
        assert a.shape == b.shape
        assert np.array_equal(a, b), np.abs(a - b).max()
    if fractions is None:
        assert served.fractions is None
    else:
        assert np.array_equal(served.fractions, fractions), np.abs(served.fractions - fractions).max()


def _serve(model, jobs, **kwargs):
    print(f"serving {len(jobs)} requests, {sum(len(j.signals) for j in jobs)} voxels, {kwargs}",
          flush=True)
    with FodServer(model, **kwargs) as server:
        first = [server.submit(j.signals, GradientTable(j.bvals, j.bvecs), lmax=j.lmax)
                 for j in jobs[: len(jobs) // 2]]
        # the rest arrive while the first ones are on the device
        rest = [server.submit(j.signals, GradientTable(j.bvals, j.bvecs), lmax=j.lmax)
                for j in jobs[len(jobs) // 2:]]
        results = [f.result(timeout=120) for f in first + rest]
    return server, results


def test_served_mfod_equals_alone_and_overlaps():
    model = _model("mFOD")
    jobs = _jobs(7, seed=21, lo=500, hi=6000)
    server, results = _serve(model, jobs, slots=3)
    assert server.max_in_flight >= 2
    for job, fit in zip(jobs, results):
        assert fit.lmax == compute_max_lmax(job.bvals, job.lmax)
        assert fit.n_voxels == len(job.signals)
        _check_same(fit, _alone(model, job))
    summary = server.summary()
    assert summary["requests"] == len(jobs)
    assert summary["p99_ms"] >= summary["p50_ms"] > 0


def test_served_fractions_equal_oracle():
    import oracle
    from larmor_amd.sd.kernel import Kernel
    from larmor_amd.sd.model import compute_shell_weights, low_resolution_dirs

    model = _model("mFOD")
    job = synthetic.batched_jobs(1, seed=31, min_voxels=400, max_voxels=400)[0]
    _, (fit,) = _serve(model, [job], slots=2)
    w = compute_shell_weights(job.bvals, model.inner_shell_weight)
    hr = Kernel(model.rfs, job.bvals, job.bvecs, dirs=DIRS_300, weights=w)
    lr = Kernel(model.rfs, job.bvals, job.bvecs,
                dirs=low_resolution_dirs(model.rfs, len(job.bvals)), weights=w)
    _, ref_fractions = oracle.fit_multi_tissue(job.signals.astype(np.float64), w, lr, hr, 0.0, 0.1,
                                               1e-6)
    assert np.array_equal(fit.fractions, ref_fractions)


def test_served_grl_and_single_tissue_equal_alone():
    for model in (_model("GRL"), _model("mFOD", multi=False)):
        jobs = _jobs(3, seed=41, lo=200, hi=800)
        server, results = _serve(model, jobs, slots=2)
        for job, fit in zip(jobs, results):
            _check_same(fit, _alone(model, job))


def test_budget_admits_one_at_a_time():
    model = _model("mFOD")
    jobs = _jobs(4, seed=51, lo=500, hi=3000)
    sizes = [request_bytes(model, len(j.signals), len(j.bvals), compute_max_lmax(j.bvals, j.lmax),
                           True) for j in jobs]
    budget = 2 * min(sizes) - 1  # any two requests exceed it, each alone fits
    assert budget >= max(sizes)
    server, results = _serve(model, jobs, slots=4, hbm_budget=budget)
    assert server.max_in_flight == 1
    assert server.peak_bytes <= budget
    for job, fit in zip(jobs, results):
        _check_same(fit, _alone(model, job))
    big = synthetic.batched_jobs(1, seed=52, min_voxels=200_000, max_voxels=200_000)[0]
    need = request_bytes(model, len(big.signals), len(big.bvals),
                         compute_max_lmax(big.bvals, model.lmax), True)
    with FodServer(model, hbm_budget=need - 1) as small:
        with pytest.raises(ValueError, match="more than the server's budget"):
            small.submit(big.signals, GradientTable(big.bvals, big.bvecs))


def test_request_errors():
    model = _model("mFOD")
    job = synthetic.batched_jobs(1, seed=61, min_voxels=100, max_voxels=100)[0]
    gtab = GradientTable(job.bvals, job.bvecs)
    with FodServer(model) as server:
        with pytest.raises(ValueError, match="do not match the gradient table"):
            server.submit(job.signals[:, :-1], gtab)
        with pytest.raises(TypeError, match="float32 or float64"):
            server.submit(job.signals.astype(np.int32), gtab)
        with pytest.raises(ValueError, match="lmax must be even"):
            server.submit(job.signals, gtab, lmax=7)
        assert server.submit(job.signals, gtab).result(timeout=60).n_voxels == 100
        # A request the kernels refuse (more than 1024 measurements) fails alone: its future
        # holds the library's error and the next request is served.
        rng = np.random.default_rng(62)
        big_b = np.concatenate([np.zeros(10), np.full(1020, 1000.0)])
        big_v = np.vstack([np.zeros((10, 3)), synthetic.spiral_directions(1020)])
        big = synthetic.tissue_signals(20, big_b, big_v, rng, snr=25.0)
        refused = server.submit(big, GradientTable(big_b, big_v))
        with pytest.raises(ValueError, match="1024 measurements"):
            refused.result(timeout=60)
        assert server.submit(job.signals, gtab).result(timeout=60).n_voxels == 100
    with pytest.raises(RuntimeError, match="closed"):
        server.submit(job.signals, gtab)


def test_non_finite_rows_and_an_empty_request():
    """Rows with a non-finite value are left out of the fit and come back NaN (as SDModel.fit
    leaves such voxels out); the other rows equal the plain fit of the finite rows alone; an
    empty request is served with empty results."""
    model = _model("mFOD")
    job = synthetic.batched_jobs(1, seed=71, min_voxels=300, max_voxels=300)[0]
    gtab = GradientTable(job.bvals, job.bvecs)
    sig = job.signals.copy()
    sig[[3, 100]] = np.nan
    sig[200, 5] = np.inf
    finite = np.isfinite(sig).all(axis=1)
    with FodServer(model) as server:
        fit = server.submit(sig, gtab, lmax=job.lmax).result(timeout=60)
        empty = server.submit(sig[:0], gtab, lmax=job.lmax).result(timeout=60)
    assert fit.n_voxels == 300 and fit.sh[0].shape[0] == 300
    assert np.isnan(fit.sh[0][~finite]).all() and np.isnan(fit.fractions[~finite]).all()
    sh, fractions = _alone(model, synthetic.Job(job.signals[finite], job.bvals, job.bvecs,
                                                job.lmax))
    assert np.array_equal(fit.sh[0][finite], sh[0])
    assert np.array_equal(fit.fractions[finite], fractions)
    assert empty.n_voxels == 0 and empty.sh[0].shape == (0, sh[0].shape[1])
    assert empty.fractions.shape == (0, 3)


def test_queued_requests_on_one_acquisition_are_fitted_together():
    """With one stream, requests queue; those on the head's acquisition are merged into one fit,
    and each still gets the same bits as its plain fit alone."""
    model = _model("mFOD")
    base = synthetic.batched_jobs(2, seed=81, min_voxels=300, max_voxels=1500)
    rng = np.random.default_rng(81)
    jobs = []
    for k in range(8):
        acq = base[k % 2]
        n = int(rng.integers(100, 1200))
        jobs.append(synthetic.Job(synthetic.tissue_signals(n, acq.bvals, acq.bvecs, rng, snr=25.0),
                                  acq.bvals, acq.bvecs, acq.lmax))
    server, results = _serve(model, jobs, slots=1)
    assert server.max_merged >= 3
    for job, fit in zip(jobs, results):
        assert fit.n_voxels == len(job.signals)
        _check_same(fit, _alone(model, job))


def test_batched_and_sliced_serving_equals_alone():
    """With batched fits, queued requests on different acquisitions share one launch sequence
    (lsd_fit_multi_tissue_batched), and requests above slice_voxels are cut into parts fitted as
    separate jobs; each result still has the bits of the plain fit of its rows alone."""
    model = _model("mFOD")
    jobs = [j for j in synthetic.batched_jobs(14, seed=91, min_voxels=200, max_voxels=4000,
                                              n_shells=(1, 3), n_dirs=(30, 64))
            if len(j.bvals) > 45][:8]
    server, results = _serve(model, jobs, slots=1, batch_jobs=16, slice_voxels=900,
                             merge_voxels=1 << 20)
    assert server.max_merged >= 3
    for job, fit in zip(jobs, results):
        assert fit.n_voxels == len(job.signals)
        assert fit.lmax == compute_max_lmax(job.bvals, job.lmax)
        _check_same(fit, _alone(model, job))
