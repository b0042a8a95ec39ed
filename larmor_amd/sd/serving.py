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
        if self.multi:
            weights = compute_shell_weights(bvals, model.inner_shell_weight)
            hr = Kernel(model.rfs, bvals, bvecs, dirs=DIRS_300, weights=weights)
            lr = Kernel(model.rfs, bvals, bvecs, dirs=low_resolution_dirs(model.rfs, len(bvals)),
                        weights=weights)
            self.lr_solver = solver._prepare(lr, bvals, Stage.LOW_RES)
            self.hr_solver = solver._prepare(hr, bvals, Stage.HIGH_RES)
            self.lr_kernel = _deconv_kernel(lr)
            self.hr_kernel = _deconv_kernel(hr)
            self.weights = torch.as_tensor(np.ascontiguousarray(lr.weights, dtype=np.float64),
                                           device="cuda")
            self.fraction_tol = model.method._fraction_tolerance
            self.blocks = hr.aniso_blocks
            matrices = [hr.k, lr.k, hr.w_k_aniso, lr.w_k_aniso]
            grams = hr.k_aniso.shape[1] ** 2 + lr.k_aniso.shape[1] ** 2
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
        sh = [_to_sh(fods[:, block], self.projection, exact_integral=self.exact_integral)
              for block in self.blocks]
        return sh, fractions


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


@dataclass(eq=False)  # compared by identity
class _Running:
    request: _Request
    design: _Design
    stream: object
    event: object
    admitted: float
    keep: list = field(default_factory=list)  # device tensors alive until the event completes
    sh_host: list = field(default_factory=list)
    fractions_host: object = None
    # The requests fitted together: each with its rows' (start, stop) in the merged fit.
    members: list = field(default_factory=list)


class FodServer:
    """Fits deconvolution requests as they arrive, several at once on one GPU.

    model: the SDModel whose fit is served (its responses, method, solver, shell weighting, lmax
    and basis); slots: fits in flight at once, one HIP stream each (the device runs four
    hardware queues per process, so more than four share them); hbm_budget: device bytes that
    admitted requests and resident designs may hold together; max_designs: acquisitions kept
    resident (least recently used ones go first, never one in use); merge_voxels: queued requests
    on the head request's acquisition are fitted with it as one launch sequence up to this many
    rows in all; form: the multi-tissue fit's form, "fp64" (the reference's computation, results
    bit-identical to the plain fit) or "mixed" (the mixed-precision form, lsd_fit_multi_tissue_mixed).
    """

    def __init__(self, model: SDModel, *, slots: int = 4, hbm_budget: float = 200e9,
                 max_designs: int = 256, merge_voxels: int = 1 << 18,
                 poll_interval: float = 1e-4, form: str = "fp64") -> None:
        if form not in ("fp64", "mixed"):
            raise ValueError(f"form must be 'fp64' or 'mixed', got {form!r}")
        self.form = form
        if slots < 1:
            raise ValueError(f"slots must be at least 1, got {slots}")
        if not hbm_budget > 0:
            raise ValueError(f"hbm_budget must be positive, got {hbm_budget}")
        if max_designs < 1:
            raise ValueError(f"max_designs must be at least 1, got {max_designs}")
        torch = _torch()
        self.model = model
        self.solver = model.solver if model.solver is not None else model.method.default_solver()
        self.slots = slots
        self.hbm_budget = int(hbm_budget)
        self.max_designs = max_designs
        self.poll_interval = poll_interval
        self._streams = [torch.cuda.Stream() for _ in range(slots)]
        self._incoming: queue.SimpleQueue = queue.SimpleQueue()
        self._pending: collections.deque[_Request] = collections.deque()
        self._running: list[_Running] = []
        self._completing: _Running | None = None  # the batch _complete is resolving
        self._designs: collections.OrderedDict[tuple, _Design] = collections.OrderedDict()
        self._design_lock = threading.Lock()
        self._running_bytes = 0  # held by running requests (scheduler thread only)
        self._design_bytes = 0  # held by resident designs (under the design lock)
        self._closed = threading.Event()
        self._served: list[ServedFit] = []
        self.max_in_flight = 0
        self.max_merged = 1
        self.merge_voxels = int(merge_voxels)
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
        nbytes = request_bytes(self.model, n, m, fit_lmax, host.dtype == torch.float32)
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
        self._incoming.put(_Request(pinned, bvals, bvecs, lmax, key, nbytes, future, arrival,
                                    n_rows=n, finite=keep))
        return future

    def prepare(self, gtab: GradientTable, *, lmax: int | None = None) -> None:
        """Builds and keeps resident the design of an acquisition ahead of its requests (on the
        caller's thread), so that the first request on it does not wait for the build."""
        lmax = self.model.lmax if lmax is None else int(lmax)
        bvals = np.asarray(gtab.bvals, dtype=np.float64)
        bvecs = np.asarray(gtab.bvecs, dtype=np.float64)
        self._design((bvals.tobytes(), bvecs.tobytes(), lmax), bvals, bvecs, lmax)

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
            batches = list(self._running)
            if self._completing is not None:
                batches.append(self._completing)
            for r in batches:
                for req in [r.request, *(member for member, _, _ in r.members)]:
                    _settle(req.future, exc=exc)
            for req in self._pending:
                _settle(req.future, exc=exc)
            while True:
                try:
                    req = self._incoming.get_nowait()
                except queue.Empty:
                    break
                if req is not None:
                    _settle(req.future, exc=exc)

    def _drain(self, block: bool = False) -> bool:
        """Moves arrived requests to the queue; returns True once close() was called."""
        while True:
            try:
                req = self._incoming.get(timeout=0.05) if block else self._incoming.get_nowait()
            except queue.Empty:
                return self._closed.is_set()
            block = False
            if req is None:
                continue
            self._pending.append(req)

    def _complete(self) -> bool:
        done = [r for r in self._running if r.event.query()]
        for r in done:
            self._completing = r
            self._running.remove(r)
            finished = time.perf_counter()
            all_sh = [t.numpy() for t in r.sh_host]
            all_fractions = None if r.fractions_host is None else r.fractions_host.numpy()
            r.keep.clear()
            self._streams.append(r.stream)
            for req, start, stop in r.members:
                sh = [a[start:stop] for a in all_sh]
                fractions = None if all_fractions is None else all_fractions[start:stop]
                if req.finite is not None:
                    sh = [_expand(a, req.finite) for a in sh]
                    fractions = None if fractions is None else _expand(fractions, req.finite)
                fit = ServedFit(sh=sh, fractions=fractions, lmax=r.design.lmax,
                                n_voxels=req.n_rows, arrival=req.arrival, admitted=r.admitted,
                                finished=finished)
                self._running_bytes -= req.nbytes
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
            in_use = {id(r.design) for r in list(self._running)}
            while len(self._designs) >= self.max_designs:
                victim = next((k for k, d in self._designs.items() if id(d) not in in_use), None)
                if victim is None:
                    break
                self._design_bytes -= self._designs.pop(victim).bytes
            design = _Design(self.model, self.solver, bvals, bvecs, lmax, self.form)
            self._designs[key] = design
            self._design_bytes += design.bytes
            return design

    def _admit(self, torch) -> bool:
        admitted = False
        while self._pending and self._streams:
            req = self._pending[0]
            if req.future.cancelled():
                self._pending.popleft()
                continue
            try:
                design = self._design(req.key, req.bvals, req.bvecs, req.lmax)
            except Exception as exc:  # a request the design cannot be built for
                self._pending.popleft()
                _settle(req.future, exc=exc)
                continue
            # Strict arrival order: the head waits for memory, later requests wait behind it.
            # A server with nothing running always admits the head (submit() refused any
            # request larger than the whole budget).
            used = self._running_bytes + self._design_bytes
            if self._running and used + req.nbytes > self.hbm_budget:
                break
            self._pending.popleft()
            # From here the client can no longer cancel it (False: it was cancelled meanwhile).
            if not req.future.set_running_or_notify_cancel():
                continue
            batch = self._merge(req, used)
            stream = self._streams.pop()
            running = _Running(req, design, stream, torch.cuda.Event(), time.perf_counter())
            start = 0
            for member in batch:
                stop = start + int(member.host.shape[0])
                running.members.append((member, start, stop))
                start = stop
            try:
                self._launch(torch, running)
            except Exception as exc:  # invalid request caught by the C-ABI checks
                self._streams.append(stream)
                for member in batch:
                    _settle(member.future, exc=exc)
                continue
            self._running_bytes += sum(member.nbytes for member in batch)
            self._running.append(running)
            self.max_in_flight = max(self.max_in_flight, len(self._running))
            self.max_merged = max(self.max_merged, len(batch))
            self.peak_bytes = max(self.peak_bytes, self._running_bytes + self._design_bytes)
            admitted = True
        return admitted

    def _merge(self, head: _Request, used: int) -> list:
        """The head request and the queued requests on the same acquisition and lmax (same design,
        same input dtype) that fit with it under merge_voxels rows and the budget: they run as one
        fit, which gives each voxel the same bits as its request alone (no kernel of the path
        combines voxels)."""
        batch = [head]
        rows = int(head.host.shape[0])
        nbytes = head.nbytes
        if self.merge_voxels <= rows:
            return batch
        kept = collections.deque()
        while self._pending:
            other = self._pending.popleft()
            n = int(other.host.shape[0])
            if other.future.cancelled():
                continue
            if (other.key == head.key and other.host.dtype == head.host.dtype
                    and rows + n <= self.merge_voxels
                    and used + nbytes + other.nbytes <= self.hbm_budget):
                if other.future.set_running_or_notify_cancel():
                    batch.append(other)
                    rows += n
                    nbytes += other.nbytes
            else:
                kept.append(other)
        self._pending = kept
        return batch

    def _launch(self, torch, r: _Running) -> None:
        req, design = r.request, r.design
        with torch.cuda.stream(r.stream):
            if len(r.members) == 1:
                dev = req.host.to("cuda", non_blocking=True)
            else:
                rows = r.members[-1][2]
                dev = torch.empty((rows, req.host.shape[1]), dtype=req.host.dtype, device="cuda")
                for member, start, stop in r.members:
                    dev[start:stop].copy_(member.host, non_blocking=True)
            r.keep.append(dev)
            if dev.dtype != torch.float64:
                dev = dev.to(torch.float64)
                r.keep.append(dev)
            sh, fractions = design.enqueue(dev)
            r.keep.extend(sh)
            for t in sh:
                host = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
                host.copy_(t, non_blocking=True)
                r.sh_host.append(host)
            if fractions is not None:
                r.keep.append(fractions)
                r.fractions_host = torch.empty(fractions.shape, dtype=fractions.dtype,
                                               pin_memory=True)
                r.fractions_host.copy_(fractions, non_blocking=True)
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
