"""Serving spherical-deconvolution requests on one GPU: continuous batching under an HBM budget.

A request is one volume's masked voxels ready for the solve: the b0-normalised signals capped at
1, one row per voxel (the rows `SDModel.fit` sends to the device, src/larmor/sd/model.py:369-372),
with the volume's gradient table and the lmax asked for.  `FodServer(model)` fits each request
with the same kernels, solver settings and SH projection as `model.fit` (model.py:171-391), so a
served result is bit-identical to the plain fit of the same rows alone, whatever else is in
flight: every kernel of the path computes a voxel from that voxel's row and the shared design
only (the reference's own batch-invariance rule, tests/sd/test_fit.py:87-119,
cpp/tests/test_multi_tissue.cpp:56).

Scheduling.  Requests queue in arrival order.  One scheduler thread owns every GPU call.  It
admits the head of the queue when one of `slots` HIP streams is free and the request's device
memory (`request_bytes`) fits in what the budget has left; the request's upload, its whole fit
(the low-resolution rounds, the high-resolution solve, the fractions, the SH projection) and the
download of its results are enqueued on that stream, and an event marks its end.  Requests that
arrive while others run are admitted as soon as a stream and memory free up, and their kernels
overlap on the device with the ones already running.  Queued requests on the head request's
acquisition (same design) are fitted together with it, as one launch sequence over their
concatenated rows, and their results split back.  The scheduler never waits on the device: it
polls the events, so the host keeps admitting while the GPU works.

Designs.  The per-acquisition objects (the low- and high-resolution kernels and their Gram
matrices in the solvers, the shell weights and the SH projection matrix) are built once per
(b-values, b-vectors, lmax) and kept resident on the device, up to `max_designs` of them; their
bytes count against the budget.  Building one is host work (the Gram matrix is summed on the
host), done on the scheduler thread before the request that needs it is admitted, or ahead of
time by `prepare`.

Latency of a request = time from `submit` (its arrival) until its results are in host memory and
its future is resolved; `FodServer.summary()` reports p50/p99 over the requests served.
"""

from __future__ import annotations

import collections
import queue
import threading
import time
from concurrent.futures import Future, InvalidStateError
from dataclasses import dataclass, field, replace

import numpy as np

from larmor_amd import _core, _lib
from larmor_amd.sd.kernel import Kernel
from larmor_amd.sd.model import (SDModel, _to_sh, compute_max_lmax, compute_shell_weights,
                                  low_resolution_dirs)
from larmor_amd.sd.rf import IsotropicRF
from larmor_amd.sd.sh import projection_matrix
from larmor_amd.sd.solvers import (DRL, FINAL_FRACTION_TOL_X, L2NNLS, MAX_ROUNDS, Lasso, Stage,
                                   _deconv_kernel)
from larmor_amd.sd.sphere import DIRS_300
from larmor_amd.types import GradientTable

# Device memory a fit needs beyond its inputs and outputs that no per-voxel term covers, besides
# the global-memory rerun slabs the library reports (lsd_fit_scratch_bytes): counters and the
# allocator's rounding.
_FIXED_BYTES = 16 << 20


def _torch():
    return _core._torch()


def _n_sh(lmax: int) -> int:
    return (lmax + 1) * (lmax + 2) // 2


@dataclass
class ServedFit:
    """One request's results: SH coefficients per anisotropic tissue (n, C) and the tissue
    fractions (n, n_tissues) (None for a single-tissue model), float64 like SDModel's chunks,
    with the request's timestamps (time.perf_counter seconds)."""

    sh: list[np.ndarray]
    fractions: np.ndarray | None
    lmax: int
    n_voxels: int
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def request_bytes(model: SDModel, n_voxels: int, n_meas: int, lmax: int,
                  float32_input: bool) -> int:
    """Device bytes one request holds while it runs: its signals (plus the float32 upload when
    the rows come as float32), the fit's per-voxel state and outputs (the high-resolution
    amplitudes and correlations, the low-resolution FODs and correlations, fractions, rounds,
    status, voxel lists, and a damped Richardson-Lucy solver's position buffers) and the SH
    coefficients, plus the rerun slabs of the ridge NNLS solves (the library's figure for the
    measurement count and direction counts) and a fixed allowance."""
    n_aniso = sum(1 for rf in model.rfs if not isinstance(rf, IsotropicRF))
    n_iso = len(model.rfs) - n_aniso
    n_hr = len(DIRS_300) * n_aniso
    n_lr = sum(len(d) for d in low_resolution_dirs(model.rfs, n_meas)) if model.is_multi_tissue else 0
    per_voxel = 8 * n_meas + (4 * n_meas if float32_input else 0)
    per_voxel += 8 * (2 * n_hr + 2 * n_lr)  # amplitudes + correlations, HR and LR
    per_voxel += 8 * (n_aniso * _n_sh(lmax) + len(model.rfs) + 2 * (1 + n_iso))
    per_voxel += 4 * 5  # rounds, status, the overflow list and the two voxel lists
    solver = model.solver if model.solver is not None else model.method.default_solver()
    fixed = _FIXED_BYTES
    if isinstance(solver, DRL):
        # damped Richardson-Lucy keeps, per stage, the signal and four c-length vectors per voxel
        per_voxel += 8 * (2 * n_meas + 4 * (n_hr + n_lr))
    elif isinstance(solver, L2NNLS):
        # -1 for sizes no kernel takes: such a request fails in the fit with the library's
        # message, alone, like any request the kernels refuse.
        fixed += max(0, _lib.load().lsd_fit_scratch_bytes(n_meas, n_lr, n_hr))
    return int(n_voxels) * per_voxel + fixed


def _settle(future: Future, result=None, exc: BaseException | None = None) -> None:
    """Resolves a future with a result or an exception unless it is done already (its client
    cancelled it while it queued, or an earlier error settled it); never raises."""
    try:
        if exc is not None:
            future.set_exception(exc)
        else:
            future.set_result(result)
    except InvalidStateError:
        pass


class _Design:
    """The device-resident objects of one (gradient table, lmax): what SDModel._prepare builds
    for a fit (model.py:289-319), kept for every request on that acquisition."""

    def __init__(self, model: SDModel, solver, bvals: np.ndarray, bvecs: np.ndarray, lmax: int,
                 form: str = "fp64"):
        torch = _torch()
        self.form = form
        self.lmax = compute_max_lmax(bvals, lmax)
        self.n_meas = len(bvals)
        self.multi = model.is_multi_tissue
        self.exact_integral = isinstance(solver, (L2NNLS, Lasso))
        # Designs with the same batch key can share one batched fit (lsd_fit_multi_tissue_batched:
        # ridge NNLS, the FP64 form, equal direction counts and row-slot class); None: fitted alone.
        self.batch_key = None
        if self.multi:
            weights = compute_shell_weights(bvals, model.inner_shell_weight)
            hr = Kernel(model.rfs, bvals, bvecs, dirs=DIRS_300, weights=weights)
            lr = Kernel(model.rfs, bvals, bvecs, dirs=low_resolution_dirs(model.rfs, len(bvals)),
                        weights=weights)
            self.lr_solver = solver._prepare(lr, bvals, Stage.LOW_RES)
            self.hr_solver = solver._prepare(hr, bvals, Stage.HIGH_RES)
            self.lr_kernel = _deconv_kernel(lr)
            self.hr_kernel = _deconv_kernel(hr)
            self.n_lr = lr.k_aniso.shape[1]
            self.weights = torch.as_tensor(np.ascontiguousarray(lr.weights, dtype=np.float64),
                                           device="cuda")
            self.fraction_tol = model.method._fraction_tolerance
            self.blocks = hr.aniso_blocks
            matrices = [hr.k, lr.k, hr.w_k_aniso, lr.w_k_aniso]
            grams = hr.k_aniso.shape[1] ** 2 + lr.k_aniso.shape[1] ** 2
            if isinstance(solver, L2NNLS) and form == "fp64" and self.n_meas <= 512:
                self.batch_key = (_core.batch_class(self.n_meas), lr.k_aniso.shape[1],
                                  hr.k_aniso.shape[1], len(hr.aniso_blocks), lr.k_iso.shape[1])
        else:
            kernel = Kernel(model.rfs, bvals, bvecs, dirs=DIRS_300)
            self.solver = solver._prepare(kernel, bvals, Stage.SINGLE)
            self.blocks = kernel.aniso_blocks
            matrices = [kernel.w_k_aniso]
            grams = kernel.k_aniso.shape[1] ** 2
        proj = projection_matrix(DIRS_300, self.lmax, model.basis).T * (len(DIRS_300) / (4.0 * np.pi))
        self.projection = torch.as_tensor(proj, dtype=torch.float64, device="cuda")
        self.n_fractions = len(model.rfs) if self.multi else 0
        self.bytes = 8 * (sum(a.size for a in matrices) + grams + proj.size)

    def enqueue(self, signals):
        """Enqueues the fit of device signals (n, m) float64 on the current stream; returns the
        device SH coefficients per anisotropic block and the fractions (or None)."""
        if self.multi:
            fods, fractions = _core.fit_multi_tissue(
                signals, self.weights, self.lr_kernel, self.lr_solver, self.hr_kernel,
                self.hr_solver, MAX_ROUNDS, self.fraction_tol, FINAL_FRACTION_TOL_X, 1,
                form=self.form)
        else:
            fods, fractions = _core.fit_single_tissue(signals, self.solver, 1), None
        return self.project(fods), fractions

    def project(self, fods):
        """The SH coefficients of each anisotropic block of device amplitudes (n, 300 blocks)."""
        return [_to_sh(fods[:, block], self.projection, exact_integral=self.exact_integral)
                for block in self.blocks]


@dataclass(eq=False)  # compared by identity
class _Request:
    host: object  # pinned CPU tensor (n, m), float32 or float64
    bvals: np.ndarray
    bvecs: np.ndarray
    lmax: int
    key: tuple
    nbytes: int
    future: Future
    arrival: float
    n_rows: int = 0  # rows submitted, non-finite ones included
    finite: np.ndarray | None = None  # which rows were fitted; None when all were
    # Filled as its parts are fitted: the results of the fitted rows, the parts still out, when
    # its first part was admitted, the lmax of its design.
    sh: list | None = None
    fractions: np.ndarray | None = None
    parts_left: int = 1
    admitted: float | None = None
    fit_lmax: int = 0


@dataclass(eq=False)  # compared by identity
class _Part:
    """Rows [start, stop) of a request's pinned signals: the unit the scheduler fits."""

    request: _Request
    start: int
    stop: int
    nbytes: int


@dataclass(eq=False)  # compared by identity
class _Member:
    """A part in a running fit: its design and its rows [start, stop) in the fit's outputs, with
    the pinned host copies of its results."""

    part: _Part
    design: _Design
    start: int
    stop: int
    sh_host: list = field(default_factory=list)
    fractions_host: object = None


@dataclass(eq=False)  # compared by identity
class _Running:
    stream: object
    event: object
    admitted: float
    members: list = field(default_factory=list)
    keep: list = field(default_factory=list)  # device tensors alive until the event completes
    batched: bool = False  # one job per member (lsd_fit_multi_tissue_batched)


class FodServer:
    """Fits deconvolution requests as they arrive, several at once on one GPU.

    model: the SDModel whose fit is served (its responses, method, solver, shell weighting, lmax
    and basis); slots: fits in flight at once, one HIP stream each (the device runs four
    hardware queues per process, so more than four share them); hbm_budget: device bytes that
    admitted requests and resident designs may hold together; max_designs: acquisitions kept
    resident (least recently used ones go first, never one in use); merge_voxels: the rows one
    fit may hold -- queued parts on the head's acquisition, or with batching on any acquisition
    of the same batch key, join the head's fit up to it; batch_jobs: parts (each with its own
    acquisition) one batched fit may hold (1: no batched fits); slice_voxels: requests with more
    rows are cut into parts of at most this many, fitted as separate work items, so a large
    request does not hold the device while small ones wait (0: no cuts); form: the multi-tissue
    fit's form, "fp64" (the reference's computation, results bit-identical to the plain fit) or
    "mixed" (the mixed-precision form, lsd_fit_multi_tissue_mixed; never batched).
    """

    def __init__(self, model: SDModel, *, slots: int = 4, hbm_budget: float = 200e9,
                 max_designs: int = 256, merge_voxels: int = 1 << 18, batch_jobs: int = 1,
                 slice_voxels: int = 0, poll_interval: float = 1e-4, form: str = "fp64",
                 tuning=None) -> None:
        if form not in ("fp64", "mixed"):
            raise ValueError(f"form must be 'fp64' or 'mixed', got {form!r}")
        self.form = form
        if slots < 1:
            raise ValueError(f"slots must be at least 1, got {slots}")
        if not hbm_budget > 0:
            raise ValueError(f"hbm_budget must be positive, got {hbm_budget}")
        if max_designs < 1:
            raise ValueError(f"max_designs must be at least 1, got {max_designs}")
        if not 1 <= batch_jobs <= _core.MAX_BATCH_JOBS:
            raise ValueError(f"batch_jobs must be in 1..{_core.MAX_BATCH_JOBS}, got {batch_jobs}")
        if slice_voxels < 0:
            raise ValueError(f"slice_voxels must be at least 0, got {slice_voxels}")
        torch = _torch()
        self.model = model
        self.solver = model.solver if model.solver is not None else model.method.default_solver()
        self.slots = slots
        self.hbm_budget = int(hbm_budget)
        self.max_designs = max_designs
        self.poll_interval = poll_interval
        self._streams = [torch.cuda.Stream() for _ in range(slots)]
        self._incoming: queue.SimpleQueue = queue.SimpleQueue()
        self._pending: collections.deque[_Part] = collections.deque()
        self._running: list[_Running] = []
        self._completing: _Running | None = None  # the fit _complete is resolving
        self._designs: collections.OrderedDict[tuple, _Design] = collections.OrderedDict()
        self._design_lock = threading.Lock()
        self._running_bytes = 0  # held by running parts (scheduler thread only)
        self._design_bytes = 0  # held by resident designs (under the design lock)
        self._closed = threading.Event()
        self._served: list[ServedFit] = []
        self.max_in_flight = 0
        self.max_merged = 1
        self.merge_voxels = int(merge_voxels)
        self.batch_jobs = int(batch_jobs)
        self.slice_voxels = int(slice_voxels)
        # A larmor_amd.sd.tuning.TuningTable: each multi-tissue fit runs with its bucket's tuned
        # kernel shapes (None: the defaults).
        self.tuning = tuning
        self.peak_bytes = 0
        self._error: BaseException | None = None
        self._thread = threading.Thread(target=self._loop, name="fod-server", daemon=True)
        self._thread.start()

    # -- client side -------------------------------------------------------------------------

    def submit(self, signals, gtab: GradientTable, *, lmax: int | None = None) -> Future:
        """Queues the fit of signals (n_voxels, n_meas): b0-normalised rows, float32 or float64,
        numpy or a CPU tensor.  Returns a Future resolving to a ServedFit.  The rows are copied
        into pinned host memory on the caller's thread, so the caller may reuse its array; rows
        with a non-finite value are not fitted and their results are NaN."""
        torch = _torch()
        arrival = time.perf_counter()
        if self._closed.is_set():
            raise RuntimeError("the server is closed")
        if self._error is not None:
            raise RuntimeError("the server stopped after an error") from self._error
        lmax = self.model.lmax if lmax is None else int(lmax)
        if lmax < 2 or lmax % 2 != 0:
            raise ValueError(f"lmax must be even and at least 2, got {lmax}")
        bvals = np.asarray(gtab.bvals, dtype=np.float64)
        bvecs = np.asarray(gtab.bvecs, dtype=np.float64)
        host = torch.as_tensor(signals)
        if host.ndim != 2:
            raise ValueError(f"signals must be 2-D (voxels, measurements), got shape {tuple(host.shape)}")
        if host.dtype not in (torch.float32, torch.float64):
            raise TypeError(f"signals must be float32 or float64, got {host.dtype}")
        if host.device.type != "cpu":
            raise ValueError("signals must be in host memory")
        n, m = host.shape
        if m != len(bvals):
            raise ValueError(f"signals ({n}, {m}) do not match the gradient table's {len(bvals)} "
                             "volumes")
        fit_lmax = compute_max_lmax(bvals, lmax)
        f32 = host.dtype == torch.float32
        nbytes = request_bytes(self.model, n, m, fit_lmax, f32)
        if nbytes > self.hbm_budget:
            raise ValueError(f"the request needs about {nbytes} bytes of device memory, more than "
                             f"the server's budget of {self.hbm_budget}")
        # Voxels with a non-finite signal are left out of the fit and get NaN results, as
        # SDModel.fit leaves them out (model.py:171-224).
        finite = torch.isfinite(host).all(dim=1)
        keep = None
        if not bool(finite.all()):
            keep = finite.numpy().copy()
            host = host[finite]
        pinned = torch.empty(host.shape, dtype=host.dtype, pin_memory=True)
        pinned.copy_(host)
        key = (bvals.tobytes(), bvecs.tobytes(), lmax)
        future: Future = Future()
        req = _Request(pinned, bvals, bvecs, lmax, key, nbytes, future, arrival, n_rows=n,
                       finite=keep)
        rows = int(pinned.shape[0])
        n_parts = 1
        if self.slice_voxels and rows > self.slice_voxels:
            n_parts = -(-rows // self.slice_voxels)
        bounds = [rows * k // n_parts for k in range(n_parts + 1)]
        req.parts_left = n_parts
        parts = [_Part(req, a, b, request_bytes(self.model, b - a, m, fit_lmax, f32) if n_parts > 1
                       else nbytes) for a, b in zip(bounds[:-1], bounds[1:])]
        self._incoming.put(parts)
        return future

    def prepare(self, gtab: GradientTable, *, lmax: int | None = None):
        """Builds and keeps resident the design of an acquisition ahead of its requests (on the
        caller's thread), so that the first request on it does not wait for the build; returns
        it."""
        lmax = self.model.lmax if lmax is None else int(lmax)
        bvals = np.asarray(gtab.bvals, dtype=np.float64)
        bvecs = np.asarray(gtab.bvecs, dtype=np.float64)
        return self._design((bvals.tobytes(), bvecs.tobytes(), lmax), bvals, bvecs, lmax)

    def autotune(self, gtab: GradientTable, signals, *, lmax: int | None = None,
                 sample: int = 65536):
        """Tunes the kernel shapes of the bucket of an acquisition (its row class and
        low-resolution directions) on a sample of `signals`, once per bucket, into the server's
        tuning table (made on first use); call it before submitting, on the caller's thread."""
        from larmor_amd.sd import tuning

        if self.tuning is None:
            self.tuning = tuning.TuningTable()
        d = self.prepare(gtab, lmax=lmax)
        if not d.multi or self.tuning.bucket(d.n_meas, d.n_lr) in self.tuning.choice:
            return None
        return tuning.autotune(signals, d.weights, d.lr_kernel, d.lr_solver, d.hr_kernel,
                               d.hr_solver, fraction_tol=d.fraction_tol, sample=sample,
                               table=self.tuning)

    def close(self, wait: bool = True) -> None:
        """Stops taking requests; with wait, returns once every queued request is served."""
        self._closed.set()
        self._incoming.put(None)
        if wait:
            self._thread.join()
            self._designs.clear()

    def __enter__(self) -> "FodServer":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    @property
    def served(self) -> list[ServedFit]:
        return list(self._served)

    def summary(self) -> dict:
        """Latency percentiles and throughput over the requests served so far."""
        return summarize(self._served)

    # -- scheduler thread --------------------------------------------------------------------

    def _loop(self) -> None:
        torch = _torch()
        try:
            while True:
                closing = self._drain()
                progress = self._complete()
                progress = self._admit(torch) or progress
                if closing and not self._pending and not self._running:
                    break
                if not progress:
                    if not self._running and not self._pending:
                        self._drain(block=True)
                    else:
                        time.sleep(self.poll_interval)
        except BaseException as exc:  # noqa: BLE001 - every waiting future must hear of it
            self._error = exc
            fits = list(self._running)
            if self._completing is not None:
                fits.append(self._completing)
            for r in fits:
                for member in r.members:
                    _settle(member.part.request.future, exc=exc)
            for part in self._pending:
                _settle(part.request.future, exc=exc)
            while True:
                try:
                    parts = self._incoming.get_nowait()
                except queue.Empty:
                    break
                for part in parts or ():
                    _settle(part.request.future, exc=exc)

    def _drain(self, block: bool = False) -> bool:
        """Moves arrived requests' parts to the queue; returns True once close() was called."""
        while True:
            try:
                parts = self._incoming.get(timeout=0.05) if block else self._incoming.get_nowait()
            except queue.Empty:
                return self._closed.is_set()
            block = False
            if parts is None:
                continue
            self._pending.extend(parts)

    def _complete(self) -> bool:
        done = [r for r in self._running if r.event.query()]
        for r in done:
            self._completing = r
            self._running.remove(r)
            finished = time.perf_counter()
            results = [([t.numpy() for t in m.sh_host],
                        None if m.fractions_host is None else m.fractions_host.numpy())
                       for m in r.members]
            r.keep.clear()
            self._streams.append(r.stream)
            for member, (sh, fractions) in zip(r.members, results):
                part, req = member.part, member.part.request
                self._running_bytes -= part.nbytes
                if req.future.done():  # an earlier part failed
                    continue
                if req.sh is None:
                    rows = int(req.host.shape[0])
                    req.sh = [np.empty((rows,) + a.shape[1:], dtype=a.dtype) for a in sh]
                    if fractions is not None:
                        req.fractions = np.empty((rows,) + fractions.shape[1:], dtype=fractions.dtype)
                    req.fit_lmax = member.design.lmax
                for whole, a in zip(req.sh, sh):
                    whole[part.start:part.stop] = a
                if fractions is not None:
                    req.fractions[part.start:part.stop] = fractions
                req.parts_left -= 1
                if req.parts_left > 0:
                    continue
                sh_all, fr_all = req.sh, req.fractions
                if req.finite is not None:
                    sh_all = [_expand(a, req.finite) for a in sh_all]
                    fr_all = None if fr_all is None else _expand(fr_all, req.finite)
                fit = ServedFit(sh=sh_all, fractions=fr_all, lmax=req.fit_lmax,
                                n_voxels=req.n_rows, arrival=req.arrival, admitted=req.admitted,
                                finished=finished)
                req.sh, req.fractions = None, None
                self._served.append(replace(fit, sh=[], fractions=None))  # timings only
                _settle(req.future, fit)
            self._completing = None
        return bool(done)

    def _design(self, key: tuple, bvals: np.ndarray, bvecs: np.ndarray, lmax: int) -> _Design:
        with self._design_lock:
            design = self._designs.get(key)
            if design is not None:
                self._designs.move_to_end(key)
                return design
            in_use = {id(m.design) for r in list(self._running) for m in r.members}
            while len(self._designs) >= self.max_designs:
                victim = next((k for k, d in self._designs.items() if id(d) not in in_use), None)
                if victim is None:
                    break
                self._design_bytes -= self._designs.pop(victim).bytes
            design = _Design(self.model, self.solver, bvals, bvecs, lmax, self.form)
            self._designs[key] = design
            self._design_bytes += design.bytes
            return design

    def _start(self, part: _Part) -> bool:
        """Marks a part's request as running when its first part is admitted; False when the
        request is settled already (cancelled by its client, or failed in an earlier part)."""
        req = part.request
        if req.admitted is None:
            # From here the client can no longer cancel it (False: it was cancelled meanwhile).
            if not req.future.set_running_or_notify_cancel():
                return False
            req.admitted = time.perf_counter()
            return True
        return not req.future.done()

    def _part_design(self, part: _Part):
        """The part's design, or None (its request settled with the error) when it cannot be
        built."""
        req = part.request
        try:
            return self._design(req.key, req.bvals, req.bvecs, req.lmax)
        except Exception as exc:  # a request the design cannot be built for
            _settle(req.future, exc=exc)
            return None

    def _admit(self, torch) -> bool:
        admitted = False
        while self._pending and self._streams:
            part = self._pending[0]
            if part.request.future.cancelled() or (part.request.admitted is not None
                                                   and part.request.future.done()):
                self._pending.popleft()
                continue
            design = self._part_design(part)
            if design is None:
                self._pending.popleft()
                continue
            # Strict arrival order: the head waits for memory, later parts wait behind it.  A
            # server with nothing running always admits the head (submit() refused any request
            # larger than the whole budget).
            used = self._running_bytes + self._design_bytes
            if self._running and used + part.nbytes > self.hbm_budget:
                break
            self._pending.popleft()
            if not self._start(part):
                continue
            members = self._gather(part, design, used)
            stream = self._streams.pop()
            running = _Running(stream, torch.cuda.Event(), time.perf_counter())
            running.batched = (self.batch_jobs > 1 and len(members) > 1
                               and design.batch_key is not None)
            start = 0
            for p, d in members:
                stop = start + (p.stop - p.start)
                running.members.append(_Member(p, d, start, stop))
                start = stop
            try:
                self._launch(torch, running)
            except Exception as exc:  # invalid request caught by the C-ABI checks
                self._streams.append(stream)
                for p, _ in members:
                    _settle(p.request.future, exc=exc)
                continue
            self._running_bytes += sum(p.nbytes for p, _ in members)
            self._running.append(running)
            self.max_in_flight = max(self.max_in_flight, len(self._running))
            self.max_merged = max(self.max_merged, len(members))
            self.peak_bytes = max(self.peak_bytes, self._running_bytes + self._design_bytes)
            admitted = True
        return admitted

    def _gather(self, head: _Part, design: _Design, used: int) -> list:
        """The head part and the queued parts that join its fit: with batching, parts of any
        acquisition whose design has the head's batch key (one job each, up to batch_jobs);
        otherwise parts on the head's acquisition and lmax with the same input dtype (one fit over
        their concatenated rows).  Both up to merge_voxels rows and within the budget.  Either
        way each voxel gets the same bits as its request alone: no kernel of the path combines
        voxels."""
        members = [(head, design)]
        rows = head.stop - head.start
        nbytes = head.nbytes
        batching = self.batch_jobs > 1 and design.batch_key is not None
        limit = self.batch_jobs if batching else None
        if self.merge_voxels <= rows:
            return members
        kept = collections.deque()
        while self._pending:
            other = self._pending.popleft()
            req = other.request
            if req.future.cancelled() or (req.admitted is not None and req.future.done()):
                continue
            n = other.stop - other.start
            joins = False
            other_design = None
            if rows + n <= self.merge_voxels and used + nbytes + other.nbytes <= self.hbm_budget \
                    and (limit is None or len(members) < limit):
                if batching:
                    other_design = self._part_design(other)
                    if other_design is None:
                        continue
                    joins = other_design.batch_key == design.batch_key
                else:
                    joins = (req.key == head.request.key
                             and req.host.dtype == head.request.host.dtype)
                    other_design = design
            if joins and self._start(other):
                members.append((other, other_design))
                rows += n
                nbytes += other.nbytes
            elif not joins:
                kept.append(other)
        self._pending = kept
        return members

    def _launch(self, torch, r: _Running) -> None:
        d_head = r.members[0].design
        if self.tuning is not None and d_head.multi:
            self.tuning.use(d_head.n_meas, d_head.n_lr)
        with torch.cuda.stream(r.stream):
            if r.batched:
                jobs = []
                for m in r.members:
                    req, d = m.part.request, m.design
                    dev = req.host[m.part.start:m.part.stop].to("cuda", non_blocking=True)
                    r.keep.append(dev)
                    if dev.dtype != torch.float64:
                        dev = dev.to(torch.float64)
                        r.keep.append(dev)
                    jobs.append((dev, d.weights, d.lr_kernel, d.lr_solver, d.hr_kernel,
                                 d.hr_solver))
                d0 = r.members[0].design
                fods, fractions = _core.fit_multi_tissue_batched(
                    jobs, MAX_ROUNDS, d0.fraction_tol, FINAL_FRACTION_TOL_X)
                r.keep.extend([fods, fractions])
                outputs = [(m.design.project(fods[m.start:m.stop]), fractions[m.start:m.stop])
                           for m in r.members]
            else:
                d = r.members[0].design
                if len(r.members) == 1:
                    p = r.members[0].part
                    dev = p.request.host[p.start:p.stop].to("cuda", non_blocking=True)
                else:
                    first = r.members[0].part.request.host
                    dev = torch.empty((r.members[-1].stop, first.shape[1]), dtype=first.dtype,
                                      device="cuda")
                    for m in r.members:
                        dev[m.start:m.stop].copy_(m.part.request.host[m.part.start:m.part.stop],
                                                  non_blocking=True)
                r.keep.append(dev)
                if dev.dtype != torch.float64:
                    dev = dev.to(torch.float64)
                    r.keep.append(dev)
                sh, fractions = d.enqueue(dev)
                r.keep.extend(sh)
                if fractions is not None:
                    r.keep.append(fractions)
                outputs = [([t[m.start:m.stop] for t in sh],
                            None if fractions is None else fractions[m.start:m.stop])
                           for m in r.members]
            for m, (sh, fractions) in zip(r.members, outputs):
                for t in sh:
                    r.keep.append(t)
                    host = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
                    host.copy_(t, non_blocking=True)
                    m.sh_host.append(host)
                if fractions is not None:
                    m.fractions_host = torch.empty(fractions.shape, dtype=fractions.dtype,
                                                   pin_memory=True)
                    m.fractions_host.copy_(fractions, non_blocking=True)
            r.event.record(r.stream)


def _expand(values: np.ndarray, finite: np.ndarray) -> np.ndarray:
    """Rows of the fitted voxels placed at their positions among all submitted rows, NaN
    elsewhere."""
    out = np.full((len(finite),) + values.shape[1:], np.nan)
    out[finite] = values
    return out


def summarize(served: list[ServedFit]) -> dict:
    """p50/p99 latency (ms), requests, voxels and voxel solves/s from the first arrival to the
    last result."""
    if not served:
        return {"requests": 0, "voxels": 0, "p50_ms": None, "p99_ms": None,
                "voxel_solves_per_s": None}
    lat = np.array([f.latency for f in served]) * 1e3
    voxels = sum(f.n_voxels for f in served)
    span = max(f.finished for f in served) - min(f.arrival for f in served)
    return {
        "requests": len(served),
        "voxels": voxels,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "max_ms": float(lat.max()),
        "voxel_solves_per_s": voxels / span if span > 0 else None,
    }
