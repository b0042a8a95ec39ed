"""Host side of the served and batched forms (no GPU): the request memory estimate, the latency
summary, the seeded batched workload, and that the server refuses to start without a GPU."""

import numpy as np
import pytest

from larmor_amd import synthetic
from larmor_amd.sd.model import SDModel, compute_max_lmax
from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
from larmor_amd.sd.serving import ServedFit, request_bytes, summarize

WM = [1.7e-3, 0.3e-3, 0.3e-3]


def _model(multi=True):
    rfs = [AnisotropicRF(WM, 0.0)] + ([IsotropicRF(0.8e-3), IsotropicRF(3e-3)] if multi else [])
    return SDModel(rfs, method="mFOD")


def test_request_bytes_is_affine_in_voxels():
    model = _model()
    a = request_bytes(model, 1000, 202, 8, False)
    b = request_bytes(model, 2000, 202, 8, False)
    c = request_bytes(model, 3000, 202, 8, False)
    assert b - a == c - b > 0
    per_voxel = (b - a) // 1000
    # signals, HR amplitudes + correlations (300 each), LR FOD + correlations (44 each), SH (45),
    # fractions, isotropic state, lists: 8*202 + 16*300 + 16*44 + 8*(45 + 3 + 6) + 20 bytes
    assert per_voxel == 8 * 202 + 16 * 300 + 16 * 44 + 8 * (45 + 3 + 6) + 20
    assert request_bytes(model, 1000, 202, 8, True) - a == 4 * 202 * 1000
    single = request_bytes(_model(multi=False), 1000, 202, 8, False)
    assert single < a


def test_summarize_percentiles():
    fits = [ServedFit([], None, 8, 100 * (i + 1), arrival=0.1 * i, admitted=0.1 * i,
                      finished=0.1 * i + 0.001 * (i + 1)) for i in range(100)]
    s = summarize(fits)
    lat = np.arange(1, 101, dtype=float)
    assert s["requests"] == 100
    assert s["voxels"] == 100 * 101 // 2 * 100
    assert s["p50_ms"] == pytest.approx(np.percentile(lat, 50))
    assert s["p99_ms"] == pytest.approx(np.percentile(lat, 99))
    assert s["max_ms"] == pytest.approx(100.0)
    span = fits[-1].finished - fits[0].arrival
    assert s["voxel_solves_per_s"] == pytest.approx(s["voxels"] / span)
    assert summarize([])["requests"] == 0


def test_batched_jobs_follow_the_workload():
    jobs = synthetic.batched_jobs(12, seed=7, min_voxels=50, max_voxels=400)
    again = synthetic.batched_jobs(12, seed=7, min_voxels=50, max_voxels=400)
    for j, k in zip(jobs, again):
        assert np.array_equal(j.signals, k.signals) and np.array_equal(j.bvals, k.bvals)
    for job in jobs:
        n, m = job.signals.shape
        assert 50 <= n <= 400 and job.signals.dtype == np.float32
        assert job.lmax in (6, 8, 10)
        assert np.sum(job.bvals == 0) == 10
        shells = np.unique(job.bvals[job.bvals > 0])
        assert 1 <= len(shells) <= 4 and np.all(np.diff(shells) > 100)
        assert shells.min() >= 700 and shells.max() <= 3000
        for b in shells:
            assert 30 <= np.sum(job.bvals == b) <= 90
        assert np.allclose(np.linalg.norm(job.bvecs[job.bvals > 0], axis=1), 1.0)
        assert np.all(job.signals <= 1.0) and np.all(job.signals > 0)
        # b0 rows average to 1 unless capped
        assert np.all(job.signals[:, :10].mean(axis=1) <= 1.0 + 1e-6)
        assert compute_max_lmax(job.bvals, job.lmax) <= job.lmax


def test_server_needs_a_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from larmor_amd.sd.serving import FodServer

    with pytest.raises(RuntimeError, match="needs a HIP GPU"):
        FodServer(_model())


def test_running_records_compare_by_identity():
    """Requests finish out of order: removing the second of two running records must not
    compare the first one's tensors with it."""
    import torch

    from larmor_amd.sd.serving import _Running

    a = _Running(None, None, 0.0, keep=[torch.zeros(3, 261)])
    b = _Running(None, None, 0.0, keep=[torch.zeros(3, 52)])
    running = [a, b]
    running.remove(b)
    assert running == [a] and a != _Running(None, None, 0.0, keep=a.keep)


def test_expand_places_fitted_rows():
    from larmor_amd.sd.serving import _expand

    finite = np.array([True, False, True, True, False])
    out = _expand(np.arange(6.0).reshape(3, 2), finite)
    assert out.shape == (5, 2)
    np.testing.assert_array_equal(out[finite], np.arange(6.0).reshape(3, 2))
    assert np.isnan(out[~finite]).all()


def test_request_bytes_counts_the_drl_buffers():
    rfs = [AnisotropicRF(WM, 0.0), IsotropicRF(0.8e-3), IsotropicRF(3e-3)]
    grl = SDModel(rfs, method="GRL")
    mfod = SDModel(rfs, method="mFOD")
    def per_1000(model):
        return request_bytes(model, 2000, 202, 8, False) - request_bytes(model, 1000, 202, 8, False)

    assert per_1000(grl) - per_1000(mfod) == 1000 * 8 * (2 * 202 + 4 * (300 + 44))


def test_request_bytes_covers_the_library_rerun_scratch():
    """The admission estimate holds the rerun slabs the library allocates per fit (both
    resolutions' global-memory slabs live at once)."""
    from larmor_amd import _lib

    lib = _lib.load()
    scratch = lib.lsd_fit_scratch_bytes(202, 44, 300)
    assert scratch > 2e9  # 1024 slabs of the 300-direction solve alone are ~2.4 GB
    assert request_bytes(_model(), 0, 202, 8, False) >= scratch
    assert request_bytes(_model(multi=False), 0, 202, 8, False) >= lib.lsd_fit_scratch_bytes(
        202, 0, 300)
    assert lib.lsd_fit_scratch_bytes(600, 44, 300) > scratch  # 12 row slots: larger slabs
    assert lib.lsd_fit_scratch_bytes(1100, 44, 300) == -1


class _Fut:
    """Stand-ins for the scheduler's device objects."""


def _stub_server(monkeypatch):
    import collections
    import queue
    import threading

    from larmor_amd.sd import serving

    monkeypatch.setattr(serving, "_torch", lambda: _Fut())
    srv = serving.FodServer.__new__(serving.FodServer)
    srv._incoming = queue.SimpleQueue()
    srv._pending = collections.deque()
    srv._running = []
    srv._completing = None
    srv._streams = []
    srv._closed = threading.Event()
    srv._error = None
    srv._served = []
    srv._running_bytes = 0
    srv._design_bytes = 0
    srv.hbm_budget = 1 << 40
    srv.merge_voxels = 1 << 20
    srv.batch_jobs = 1
    srv.slice_voxels = 0
    srv.max_in_flight = 0
    srv.max_merged = 1
    srv.peak_bytes = 0
    srv.poll_interval = 1e-4
    return srv


def _request(key=b"k", rows=4):
    from concurrent.futures import Future

    import torch

    from larmor_amd.sd.serving import _Request

    return _Request(torch.zeros((rows, 3)), np.zeros(3), np.zeros((3, 3)), 8, key, 1, Future(),
                    0.0, n_rows=rows)


def test_an_error_while_completing_fails_every_merged_member(monkeypatch):
    """The scheduler's error handler settles every member of every batch, including the batch it
    was completing when the error happened and the requests merged into a batch's head."""
    from larmor_amd.sd.serving import _Member, _Part, _Running

    srv = _stub_server(monkeypatch)

    class Done:
        def query(self):
            return True

    class Boom:
        def numpy(self):
            raise RuntimeError("device copy failed")

    heads = [_request(), _request()]
    members = [_request(), _request()]
    batches = []
    for head, member in zip(heads, members):
        r = _Running(object(), Done(), 0.0)
        r.members = [_Member(_Part(head, 0, 4, 1), None, 0, 4, sh_host=[Boom()]),
                     _Member(_Part(member, 0, 4, 1), None, 4, 8, sh_host=[Boom()])]
        batches.append(r)
    srv._running = batches
    srv._closed.set()
    srv._incoming.put(None)
    queued = _request()
    srv._pending.append(_Part(queued, 0, 4, 1))
    srv._loop()
    for req in heads + members + [queued]:
        assert req.future.done()
        with pytest.raises(RuntimeError, match="device copy failed"):
            req.future.result()


def test_cancelled_requests_are_skipped_and_running_ones_cannot_be_cancelled(monkeypatch):
    from larmor_amd.sd.serving import _Part, _settle

    srv = _stub_server(monkeypatch)
    srv._streams = [object()]
    srv._design = lambda key, bvals, bvecs, lmax: _Fut()
    launched = []
    srv._launch = lambda torch, r: launched.append(r)
    cancelled_head, head, cancelled_member, member = (_request() for _ in range(4))
    assert cancelled_head.future.cancel() and cancelled_member.future.cancel()
    srv._pending.extend(_Part(r, 0, 4, 1) for r in (cancelled_head, head, cancelled_member, member))

    class Torch:
        class cuda:
            Event = object

    assert srv._admit(Torch)
    (batch,) = launched
    assert [m.part.request for m in batch.members] == [head, member]
    assert head.future.running() and member.future.running()
    assert not head.future.cancel()  # admitted: a client can no longer cancel it
    # settling a cancelled or already settled future does not raise
    _settle(cancelled_head.future, 1)
    _settle(head.future, 2)
    _settle(head.future, exc=RuntimeError("late"))
    assert head.future.result() == 2


def test_tuning_row_classes_and_buckets():
    from larmor_amd.sd import tuning

    assert [tuning.row_class(m) for m in (1, 256, 257, 384, 385, 512, 513, 768, 769, 1024)] == \
        [256, 256, 384, 384, 512, 512, 768, 768, 1024, 1024]
    with pytest.raises(ValueError, match="1024 measurements"):
        tuning.row_class(1025)
    assert tuning.TuningTable().bucket(610, 44) == (768, 44)
    assert tuning.DEFAULT_CAPACITY in tuning.CAPACITIES
