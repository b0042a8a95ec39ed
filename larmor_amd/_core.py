"""The compiled core of the spherical-deconvolution path, on the GPU.

Mirrors the spherical-deconvolution part of larmor's nanobind module `larmor._core`
(cpp/src/python/sd.cpp:398-543, stub src/larmor/_core.pyi): the same names, arguments, result
layout and ValueError/TypeError messages, so larmor's Python fit code can call it unchanged.  The
work runs in liblarmor_sd.so's HIP kernels; device memory and streams come from torch.

Arrays may be numpy (host; copied to the device and the results copied back, like the
reference's copies in and out of C++) or torch CUDA tensors (device; results stay on the
device).  `n_threads` is validated as the reference does and otherwise unused: the GPU grid is
sized to the device.
"""

import ctypes
import sys

import numpy as np

from larmor_amd import _lib


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("larmor_amd needs a HIP GPU (torch.cuda.is_available() is False).")
    return torch


def _check_f64(array, name: str) -> None:
    """The reference binds with .noconvert(): another dtype is a TypeError, not a copy."""
    dtype = getattr(array, "dtype", None)
    if dtype == np.float64:
        return
    try:
        import torch

        if dtype == torch.float64:
            return
    except ImportError:  # pragma: no cover
        pass
    raise TypeError(f"{name} must be a float64 array, got {dtype}")


def _is_c_contiguous(array) -> bool:
    if isinstance(array, np.ndarray):
        return array.flags.c_contiguous
    return array.is_contiguous()


def _require_threads(n_threads: int) -> None:
    if n_threads < 1:
        raise ValueError("n_threads must be at least 1")


def _host_matrix(a, name: str) -> np.ndarray:
    """A design matrix copied to a C-contiguous float64 host array (the reference copies them)."""
    _check_f64(a, name)
    if not isinstance(a, np.ndarray):
        a = a.detach().cpu().numpy()
    if a.ndim != 2:
        raise ValueError(f"{name} must be 2-D, got shape {a.shape}")
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _stream_ptr(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _release(obj, destroy: str, _sys=sys, _libmod=_lib) -> None:
    """Frees a device object; skipped during interpreter shutdown, when modules (and the HIP
    runtime's state) may already be gone and the process frees everything anyway.  The modules
    are bound as defaults because shutdown may clear this module's globals first."""
    if _sys.is_finalizing():
        return
    handle = getattr(obj, "_handle", None)
    lib = getattr(_libmod, "_lib", None)
    if handle and lib is not None:
        getattr(lib, destroy)(handle)
        obj._handle = None


class FodSolver:
    """An FOD solver with its matrices resident on the device: ridge NNLS (FodSolver.l2nnls,
    sd.cpp:89-99), damped Richardson-Lucy (FodSolver.drl, sd.cpp:127-145) or the elastic net
    (FodSolver.lasso, sd.cpp:101-123)."""

    def __init__(self, handle, m: int, n: int):
        self._handle = handle
        self._m = m
        self._n = n

    @staticmethod
    def l2nnls(k, reg: float, tol_x: float) -> "FodSolver":
        kernel = _host_matrix(k, "k")
        lib = _lib.load()
        handle = ctypes.c_void_p()
        _lib.check(
            lib.lsd_solver_create_l2nnls(
                _ptr(kernel), kernel.shape[0], kernel.shape[1], float(reg), float(tol_x),
                ctypes.byref(handle),
            )
        )
        return FodSolver(handle, kernel.shape[0], kernel.shape[1])

    @staticmethod
    def drl(k_t, k_t_k, nu: float, n_iter: int, init: float) -> "FodSolver":
        """Damped Richardson-Lucy from K^T (c, m) and K^T K (c, c), starting from init in every
        direction, at most n_iter iterations per voxel (sd.cpp:127-145)."""
        kt = _host_matrix(k_t, "k_t")
        g = _host_matrix(k_t_k, "k_t_k")
        lib = _lib.load()
        handle = ctypes.c_void_p()
        _lib.check(
            lib.lsd_solver_create_drl(
                _ptr(kt), kt.shape[0], kt.shape[1], _ptr(g), g.shape[0], g.shape[1], float(nu),
                int(n_iter), float(init), ctypes.byref(handle),
            )
        )
        return FodSolver(handle, kt.shape[1], kt.shape[0])

    @staticmethod
    def lasso(k, reg: float, alpha: float, tol: float, max_iter: int) -> "FodSolver":
        """Elastic net on the columns of k (m, c) scaled to unit root mean square: b minimises
        (1 / 2m) ||s - k0 b||^2 + reg (alpha ||b||_1 + (1 - alpha) / 2 ||b||^2) by coordinate
        descent until no coefficient moves by tol relative to 1 + its size over a full sweep, or
        max_iter sweeps; the FOD is b over the scales and may be negative (sd.cpp:101-123)."""
        kernel = _host_matrix(k, "k")
        lib = _lib.load()
        handle = ctypes.c_void_p()
        _lib.check(
            lib.lsd_solver_create_lasso(
                _ptr(kernel), kernel.shape[0], kernel.shape[1], float(reg), float(alpha),
                float(tol), int(max_iter), ctypes.byref(handle),
            )
        )
        return FodSolver(handle, kernel.shape[0], kernel.shape[1])

    @property
    def is_drl(self) -> bool:
        return bool(_lib.load().lsd_solver_is_drl(self._handle))

    @property
    def is_lasso(self) -> bool:
        return bool(_lib.load().lsd_solver_is_lasso(self._handle))

    @property
    def n_dirs(self) -> int:
        return self._n

    @property
    def n_meas(self) -> int:
        return self._m

    @property
    def n_unconverged(self) -> int:
        """Solves so far that max_iter stopped; always 0 except for the elastic net.  Waits for
        the device."""
        out = ctypes.c_int64()
        _lib.check(_lib.load().lsd_solver_n_unconverged(self._handle, ctypes.byref(out)))
        return int(out.value)

    @property
    def normal_equations(self) -> bool:
        return bool(_lib.load().lsd_solver_normal_equations(self._handle))

    def __del__(self, _release=_release):
        _release(self, "lsd_solver_destroy")


class DeconvKernel:
    """One resolution of a multi-tissue kernel, resident on the device (sd.cpp:240-277)."""

    def __init__(self, k_aniso, k_iso, w_k_iso, aniso_blocks):
        ka = _host_matrix(k_aniso, "k_aniso")
        ki = _host_matrix(k_iso, "k_iso")
        wki = _host_matrix(w_k_iso, "w_k_iso")
        m = ka.shape[0]
        if ki.shape[0] != m or wki.shape[0] != m or wki.shape[1] != ki.shape[1]:
            raise ValueError("the kernel matrices' shapes do not match")
        blocks = np.ascontiguousarray(np.asarray(list(aniso_blocks), dtype=np.int64).reshape(-1, 2))
        lib = _lib.load()
        handle = ctypes.c_void_p()
        _lib.check(
            lib.lsd_tissue_kernel_create(
                _ptr(ka), _ptr(ki), _ptr(wki), m, ka.shape[1], ki.shape[1],
                blocks.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(blocks),
                ctypes.byref(handle),
            )
        )
        self._handle = handle
        self.m = m
        self.n_anisotropic = ka.shape[1]
        self.n_isotropic = ki.shape[1]
        self.n_blocks = len(blocks)

    def __del__(self, _release=_release):
        _release(self, "lsd_tissue_kernel_destroy")


def _device_signals(signals, torch):
    """Returns (device tensor, was_host) for a float64 C-contiguous 2-D signal array."""
    _check_f64(signals, "signals")
    if signals.ndim != 2:
        raise ValueError(f"signals must be 2-D, got shape {tuple(signals.shape)}")
    if not _is_c_contiguous(signals):
        raise ValueError("the array must be C-contiguous")
    if isinstance(signals, np.ndarray):
        return torch.from_numpy(signals).to("cuda"), True
    if signals.device.type != "cuda":
        return signals.to("cuda"), True
    return signals, False


def _out(tensor, host: bool):
    return tensor.cpu().numpy() if host else tensor


def fit_single_tissue(signals, solver: FodSolver, n_threads: int, *, return_status: bool = False):
    """Fits the FOD (n, c) of every row of signals (n, m) with one solver (sd.cpp:145-166)."""
    torch = _torch()
    dev, host = _device_signals(signals, torch)
    n, m = dev.shape
    if m != solver.n_meas:
        raise ValueError(
            f"signals ({n}, {m}) do not have the solver's {solver.n_meas} measurements"
        )
    _require_threads(n_threads)
    fods = torch.empty((n, solver.n_dirs), dtype=torch.float64, device="cuda")
    status = torch.zeros(n, dtype=torch.int32, device="cuda")
    lib = _lib.load()
    _lib.check(
        lib.lsd_fit_single_tissue(
            dev.data_ptr(), n, m, solver._handle, fods.data_ptr(), status.data_ptr(),
            _stream_ptr(torch),
        )
    )
    if return_status:
        return _out(fods, host), _out(status, host)
    return _out(fods, host)


def _single_problem(c, d, torch):
    """(c, d) on the device for one NNLS problem: c (m, n) and d (m,), float64, c C-contiguous."""
    _check_f64(c, "c")
    _check_f64(d, "d")
    if c.ndim != 2 or d.ndim != 1:
        raise TypeError(f"c must be 2-D and d 1-D, got shapes {tuple(c.shape)} and "
                        f"{tuple(d.shape)}")
    if not _is_c_contiguous(c):
        raise ValueError("the array must be C-contiguous")
    dev_c, host = _as_device(c, torch)
    dev_d, _ = _as_device(d, torch)
    return dev_c, dev_d, host


def _solve_single(entry: str, c, d, penalties: tuple, tol_x, return_status: bool):
    torch = _torch()
    dev_c, dev_d, host = _single_problem(c, d, torch)
    m, n = dev_c.shape
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    tol = 0.0 if tol_x is None else float(tol_x)
    _lib.check(
        getattr(_lib.load(), entry)(
            dev_c.data_ptr(), m, n, dev_d.data_ptr(), dev_d.shape[0], *penalties, tol,
            0 if tol_x is None else 1, x.data_ptr(), status.data_ptr(), _stream_ptr(torch),
        )
    )
    if return_status:
        return _out(x, host), int(status.item())
    return _out(x, host)


def lsqnonneg(c, d, tol_x: float | None = None, *, return_status: bool = False):
    """Solves min ||c x - d|| with x >= 0 for c (m, n) and d (m,) (sd.cpp:209-229).  A
    coefficient enters only when its gradient exceeds tol_x; None takes
    10 eps norm(c, 1) max(m, n).  return_status also gives the status bits (1: the inner-step
    limit stopped the solve)."""
    return _solve_single("lsd_lsqnonneg", c, d, (), tol_x, return_status)


def ridge_l1_nnls(c, d, ridge: float, l1: float, tol_x: float | None = None, *,
                  return_status: bool = False):
    """Solves min ||c x - d||^2 + ridge^2 ||x||^2 + 2 l1 sum(x) with x >= 0 (sd.cpp:231-258);
    ridge = l1 = 0 is plain NNLS by the normal equations.  Raises ValueError for a negative or
    non-finite ridge or l1, or l1 > 0 with ridge 0."""
    return _solve_single("lsd_ridge_l1_nnls", c, d, (float(ridge), float(l1)), tol_x,
                         return_status)


def fit_multi_tissue(
    signals,
    weights,
    lr: DeconvKernel,
    lr_solver: FodSolver,
    hr: DeconvKernel,
    hr_solver: FodSolver,
    max_outer: int,
    fraction_tol: float,
    final_tol_x: float,
    n_threads: int,
    *,
    return_diagnostics: bool = False,
    form: str = "fp64",
):
    """Fits the FODs (n, ca_hr) and fractions (n, n_blocks + ci) of every row of signals.

    Replaces `_core.fit_multi_tissue` (sd.cpp:310-358).  With return_diagnostics, also returns
    the low-resolution rounds run per voxel and the status bits (include/larmor_sd.h).  form
    "fp64" is the reference's computation (bit-identical to the oracle); "mixed" is the
    mixed-precision form (lsd_fit_multi_tissue_mixed: BF16 matrix-core correlations, then FP64
    refinement rounds, warm-started FP64 solves), which matches it to the loop tolerance.
    """
    if form not in ("fp64", "mixed"):
        raise ValueError(f"form must be 'fp64' or 'mixed', got {form!r}")
    torch = _torch()
    dev, host = _device_signals(signals, torch)
    _check_f64(weights, "weights")
    n, m = dev.shape
    w = torch.as_tensor(np.ascontiguousarray(weights) if isinstance(weights, np.ndarray) else weights)
    w = w.to(device="cuda", dtype=torch.float64).contiguous()
    if w.numel() != m or hr.m != m or lr.m != m:
        raise ValueError(
            f"signals ({n}, {m}), weights {w.numel()} and the kernels' {hr.m} measurements "
            "do not match"
        )
    _require_threads(n_threads)
    nf = hr.n_blocks + hr.n_isotropic
    fods = torch.empty((n, hr.n_anisotropic), dtype=torch.float64, device="cuda")
    fractions = torch.empty((n, nf), dtype=torch.float64, device="cuda")
    rounds = torch.zeros(n, dtype=torch.int32, device="cuda")
    status = torch.zeros(n, dtype=torch.int32, device="cuda")
    lib = _lib.load()
    entry = lib.lsd_fit_multi_tissue if form == "fp64" else lib.lsd_fit_multi_tissue_mixed
    _lib.check(
        entry(
            dev.data_ptr(), n, m, w.data_ptr(), lr._handle, lr_solver._handle, hr._handle,
            hr_solver._handle, int(max_outer), float(fraction_tol), float(final_tol_x),
            fods.data_ptr(), fractions.data_ptr(), rounds.data_ptr(), status.data_ptr(),
            _stream_ptr(torch),
        )
    )
    out = (_out(fods, host), _out(fractions, host))
    if return_diagnostics:
        out += (_out(rounds, host), _out(status, host))
    return out


class _BatchJob(ctypes.Structure):
    """lsd_batch_job (include/larmor_sd.h)."""

    _fields_ = [("signals", ctypes.c_void_p), ("n_voxels", ctypes.c_int64),
                ("m", ctypes.c_int64), ("weights", ctypes.c_void_p), ("lr", ctypes.c_void_p),
                ("lr_solver", ctypes.c_void_p), ("hr", ctypes.c_void_p),
                ("hr_solver", ctypes.c_void_p)]


MAX_BATCH_JOBS = 64


def batch_class(m: int) -> int:
    """The row-slot class of m measurements: the jobs of one batched fit share it."""
    for limit in (256, 384, 512):
        if m <= limit:
            return limit
    raise ValueError(f"more than 512 measurements ({m})")


def fit_multi_tissue_batched(jobs, max_outer: int, fraction_tol: float, final_tol_x: float, *,
                             return_diagnostics: bool = False):
    """The multi-tissue fit of several volumes in one launch sequence (lsd_fit_multi_tissue_batched).

    jobs: (signals, weights, lr, lr_solver, hr, hr_solver) per volume, signals (n_j, m_j) float64
    and weights (m_j,) as CUDA tensors, the kernels and solvers of the volume's own acquisition.
    Returns device tensors holding the volumes' voxels one volume after the other: FODs
    (sum n_j, ca_hr) and fractions (sum n_j, n_blocks + ci), and with return_diagnostics the
    rounds and status bits.  Each volume's rows equal fit_multi_tissue of that volume alone."""
    torch = _torch()
    if not 1 <= len(jobs) <= MAX_BATCH_JOBS:
        raise ValueError(f"a batch holds 1 to {MAX_BATCH_JOBS} jobs, got {len(jobs)}")
    table = (_BatchJob * len(jobs))()
    keep = []
    total = 0
    for slot, (signals, weights, lr, lr_solver, hr, hr_solver) in zip(table, jobs):
        _check_f64(signals, "signals")
        if not (signals.is_cuda and signals.ndim == 2 and signals.is_contiguous()):
            raise ValueError("each job's signals must be a C-contiguous 2-D CUDA tensor")
        w = weights.to(device="cuda", dtype=torch.float64).contiguous()
        keep.extend([signals, w])
        n, m = signals.shape
        slot.signals, slot.n_voxels, slot.m = signals.data_ptr(), n, m
        slot.weights = w.data_ptr()
        slot.lr, slot.lr_solver = lr._handle, lr_solver._handle
        slot.hr, slot.hr_solver = hr._handle, hr_solver._handle
        total += n
    hr0 = jobs[0][4]
    nf = hr0.n_blocks + hr0.n_isotropic
    fods = torch.empty((total, hr0.n_anisotropic), dtype=torch.float64, device="cuda")
    fractions = torch.empty((total, nf), dtype=torch.float64, device="cuda")
    rounds = torch.zeros(total, dtype=torch.int32, device="cuda")
    status = torch.zeros(total, dtype=torch.int32, device="cuda")
    _lib.check(
        _lib.load().lsd_fit_multi_tissue_batched(
            ctypes.cast(table, ctypes.c_void_p), len(jobs), int(max_outer), float(fraction_tol),
            float(final_tol_x), fods.data_ptr(), fractions.data_ptr(), rounds.data_ptr(),
            status.data_ptr(), _stream_ptr(torch),
        )
    )
    out = (fods, fractions)
    if return_diagnostics:
        out += (rounds, status)
    return out



class BatchedFitGraph:
    """A batched multi-tissue fit (lsd_fit_multi_tissue_batched) captured once as a HIP graph and
    replayed on new signals of the same shapes.

    The fit's launch sequence has no host synchronisation (every round's voxel list lives on the
    device), so it captures whole, its stream-ordered scratch becoming the graph's own memory.
    The graph holds device copies of each job's signals and weights and the outputs; `replay`
    copies the new signals in (or takes them already there), replays, and returns the output
    tensors (overwritten by the next replay).  Each replay equals the eager batched fit of the
    same signals bit for bit."""

    def __init__(self, jobs, max_outer: int, fraction_tol: float, final_tol_x: float):
        torch = _torch()
        if not 1 <= len(jobs) <= MAX_BATCH_JOBS:
            raise ValueError(f"a batch holds 1 to {MAX_BATCH_JOBS} jobs, got {len(jobs)}")
        self.signals = []
        self._keep = []
        self._table = (_BatchJob * len(jobs))()
        total = 0
        for slot, (signals, weights, lr, lr_solver, hr, hr_solver) in zip(self._table, jobs):
            _check_f64(signals, "signals")
            dev = torch.as_tensor(signals, dtype=torch.float64, device="cuda").contiguous().clone()
            w = torch.as_tensor(weights, dtype=torch.float64, device="cuda").contiguous().clone()
            if dev.ndim != 2:
                raise ValueError("each job's signals must be 2-D (voxels, measurements)")
            self.signals.append(dev)
            self._keep.extend([w, lr, lr_solver, hr, hr_solver])
            slot.signals, slot.n_voxels, slot.m = dev.data_ptr(), dev.shape[0], dev.shape[1]
            slot.weights = w.data_ptr()
            slot.lr, slot.lr_solver = lr._handle, lr_solver._handle
            slot.hr, slot.hr_solver = hr._handle, hr_solver._handle
            total += dev.shape[0]
        hr0 = jobs[0][4]
        nf = hr0.n_blocks + hr0.n_isotropic
        self.fods = torch.empty((total, hr0.n_anisotropic), dtype=torch.float64, device="cuda")
        self.fractions = torch.empty((total, nf), dtype=torch.float64, device="cuda")
        self.rounds = torch.empty(total, dtype=torch.int32, device="cuda")
        self.status = torch.empty(total, dtype=torch.int32, device="cuda")
        self._args = (int(max_outer), float(fraction_tol), float(final_tol_x))
        lib = _lib.load()

        def run():
            return lib.lsd_fit_multi_tissue_batched(
                ctypes.cast(self._table, ctypes.c_void_p), len(jobs), *self._args,
                self.fods.data_ptr(), self.fractions.data_ptr(), self.rounds.data_ptr(),
                self.status.data_ptr(), _stream_ptr(torch))

        # One eager run on a side stream first (argument checks, first-use set-up such as the
        # device constants), then the capture.
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            _lib.check(run())
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            rc = run()
        _lib.check(rc)

    def replay(self, signals=None):
        """Fits `signals` (one array per job, the captured shapes; None: the signals already in
        self.signals) and returns (fods, fractions, rounds, status) as device tensors."""
        if signals is not None:
            if len(signals) != len(self.signals):
                raise ValueError(f"expected {len(self.signals)} jobs' signals, got {len(signals)}")
            for dst, src in zip(self.signals, signals):
                if tuple(src.shape) != tuple(dst.shape):
                    raise ValueError(f"signals of shape {tuple(src.shape)} do not match the "
                                     f"captured {tuple(dst.shape)}")
                dst.copy_(_torch().as_tensor(src, dtype=dst.dtype).to(dst.device),
                          non_blocking=True)
        self.graph.replay()
        return self.fods, self.fractions, self.rounds, self.status

def log_signal_design(x):
    """(xt, pinv, scales) of a log-signal design x (n, p): its columns scaled to unit norm (a
    zero column kept, its norm taken as 1), transposed (p, n); their pseudo-inverse (p, n); the
    norms (make_log_signal_design, cpp/src/dwi/log_fit.cpp:134-149)."""
    x = np.asarray(x, dtype=np.float64)
    scales = np.sqrt(np.einsum("ij,ij->j", x, x))
    scales[scales == 0.0] = 1.0
    unit = x / scales
    return (np.ascontiguousarray(unit.T), np.ascontiguousarray(np.linalg.pinv(unit)),
            np.ascontiguousarray(scales))


def fit_log_signal(x, g, log_s, weighted: bool, max_iter: int, tol: float, n_threads: int, *,
                   return_refits: bool = False):
    """Fits each row of log_s (v, n) by least squares on the design x (n, p): the ordinary fit,
    then, with weighted, refits weighted by the predicted signal squared until no predicted log
    signal moves by more than tol, at most max_iter of them (cpp/src/python/dwi.cpp:41-82).
    Returns (params (v, p), constrained (v,)).  Only unconstrained designs are built: g must
    have no rows.  return_refits also gives the refits each voxel kept (-1: the ordinary fit)."""
    _check_f64(x, "x")
    _check_f64(g, "g")
    _check_f64(log_s, "log_s")
    for name, a in (("x", x), ("g", g), ("log_s", log_s)):
        if a.ndim != 2:
            raise TypeError(f"{name} must be 2-D, got shape {tuple(a.shape)}")
        if not _is_c_contiguous(a):
            raise ValueError("the array must be C-contiguous")
    n, p = x.shape
    if n < p or g.shape[1] != p or log_s.shape[1] != n:
        raise ValueError(
            f"x must be (n, p) with n >= p, g (m, p) and log_s (v, n): got x ({n}, {p}), g "
            f"({g.shape[0]}, {g.shape[1]}), log_s ({log_s.shape[0]}, {log_s.shape[1]})"
        )
    if max_iter < 0:
        raise ValueError("max_iter must be at least 0")
    _require_threads(n_threads)
    if g.shape[0] != 0:
        raise ValueError("only unconstrained log-signal fits are built: g must have 0 rows")
    torch = _torch()
    host_x = x if isinstance(x, np.ndarray) else x.detach().cpu().numpy()
    xt, pinv, scales = (torch.from_numpy(a).to("cuda") for a in log_signal_design(host_x))
    dev, host = _device_signals(log_s, torch)
    v = dev.shape[0]
    params = torch.empty((v, p), dtype=torch.float64, device="cuda")
    refits = torch.zeros(v, dtype=torch.int32, device="cuda")
    _lib.check(
        _lib.load().lsd_fit_log_signal(
            xt.data_ptr(), pinv.data_ptr(), scales.data_ptr(), n, p, dev.data_ptr(), v,
            int(bool(weighted)), int(max_iter), float(tol), params.data_ptr(), refits.data_ptr(),
            _stream_ptr(torch),
        )
    )
    constrained = np.zeros(v, dtype=bool)
    if return_refits:
        return _out(params, host), constrained, _out(refits, host)
    return _out(params, host), constrained


def correlations(signals, weights, kernel: DeconvKernel, solver: FodSolver, iso, *,
                 precision: str = "fp64"):
    """K^T a of every row of signals (n, m), a = max(w s - W_iso p, 0) with p the rows of iso
    (n, ci): "fp64" the fits' FP64 pass, "bf16" the mixed-precision form's matrix-core pass.
    Returns (n, n_dirs) float64 (numpy for numpy inputs)."""
    torch = _torch()
    if precision not in ("fp64", "bf16"):
        raise ValueError(f"precision must be 'fp64' or 'bf16', got {precision!r}")
    dev, host = _device_signals(signals, torch)
    n, m = dev.shape
    w = torch.as_tensor(np.ascontiguousarray(weights, dtype=np.float64)).to("cuda")
    p = torch.as_tensor(np.ascontiguousarray(iso, dtype=np.float64)).to("cuda")
    if p.shape != (n, kernel.n_isotropic) or w.numel() != m:
        raise ValueError(f"iso must be ({n}, {kernel.n_isotropic}) and weights ({m},)")
    out = torch.empty((n, solver.n_dirs), dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().lsd_correlations(
        dev.data_ptr(), n, m, w.data_ptr(), kernel._handle, solver._handle, p.data_ptr(),
        0 if precision == "fp64" else 1, out.data_ptr(), _stream_ptr(torch)))
    return _out(out, host)


def fit_multi_tissue_fp32(signals, weights, lr: DeconvKernel, lr_solver: FodSolver,
                          hr: DeconvKernel, hr_solver: FodSolver, max_outer: int,
                          fraction_tol: float, final_tol_x: float, n_threads: int, *,
                          return_diagnostics: bool = False):
    """The plain FP32 form of fit_multi_tissue: float32 signals (n, m) in, float32 FODs and
    fractions out, every kernel computing in single precision (the solvers' and kernels' float32
    copies).  Same arguments and checks as fit_multi_tissue otherwise."""
    torch = _torch()
    dtype = getattr(signals, "dtype", None)
    if dtype not in (np.float32, torch.float32):
        raise TypeError(f"signals must be a float32 array, got {dtype}")
    if signals.ndim != 2:
        raise ValueError(f"signals must be 2-D, got shape {tuple(signals.shape)}")
    if not _is_c_contiguous(signals):
        raise ValueError("the array must be C-contiguous")
    host = isinstance(signals, np.ndarray) or signals.device.type != "cuda"
    dev = torch.as_tensor(signals).to("cuda")
    n, m = dev.shape
    w = torch.as_tensor(np.ascontiguousarray(weights) if isinstance(weights, np.ndarray) else weights)
    w = w.to(device="cuda", dtype=torch.float32).contiguous()
    if w.numel() != m or hr.m != m or lr.m != m:
        raise ValueError(
            f"signals ({n}, {m}), weights {w.numel()} and the kernels' {hr.m} measurements "
            "do not match"
        )
    _require_threads(n_threads)
    nf = hr.n_blocks + hr.n_isotropic
    fods = torch.empty((n, hr.n_anisotropic), dtype=torch.float32, device="cuda")
    fractions = torch.empty((n, nf), dtype=torch.float32, device="cuda")
    rounds = torch.zeros(n, dtype=torch.int32, device="cuda")
    status = torch.zeros(n, dtype=torch.int32, device="cuda")
    _lib.check(
        _lib.load().lsd_fit_multi_tissue_fp32(
            dev.data_ptr(), n, m, w.data_ptr(), lr._handle, lr_solver._handle, hr._handle,
            hr_solver._handle, int(max_outer), float(fraction_tol), float(final_tol_x),
            fods.data_ptr(), fractions.data_ptr(), rounds.data_ptr(), status.data_ptr(),
            _stream_ptr(torch),
        )
    )
    out = (_out(fods, host), _out(fractions, host))
    if return_diagnostics:
        out += (_out(rounds, host), _out(status, host))
    return out


class CsdProblem:
    """A CSD problem resident on the device: the linear start, forward model, weighted
    constraint basis and the Gram matrix F^T F + norm_weight I (csd.cpp:11-21)."""

    def __init__(self, initial, forward, constraint, norm_weight: float, threshold: float,
                 max_iter: int):
        ini = _host_matrix(initial, "initial")
        fw = _host_matrix(forward, "forward")
        con = _host_matrix(constraint, "constraint")
        lib = _lib.load()
        handle = ctypes.c_void_p()
        _lib.check(lib.lsd_csd_create(
            _ptr(ini), ini.shape[0], ini.shape[1], _ptr(fw), fw.shape[1], _ptr(con),
            con.shape[0], float(norm_weight), float(threshold), int(max_iter),
            ctypes.byref(handle)))
        self._handle = handle
        self.n_coefficients = fw.shape[1]
        self.n_meas = fw.shape[0]

    def __del__(self, _release=_release):
        _release(self, "lsd_csd_destroy")


def fit_csd(signals, initial, forward, constraint, norm_weight: float, threshold: float,
            max_iter: int, n_threads: int, *, return_solves: bool = False):
    """Fits every row of signals (n, m) by constrained spherical deconvolution (sd.cpp:171-206).

    Returns the FODs (n, c) and whether each voxel's constrained set repeated (n,) bool; with
    return_solves also the Cholesky solves run per voxel.
    """
    torch = _torch()
    for a, name in ((initial, "initial"), (forward, "forward"), (constraint, "constraint")):
        _check_f64(a, name)
    dev, host = _device_signals(signals, torch)
    n, m = dev.shape
    ini_shape, fw_shape, con_shape = tuple(initial.shape), tuple(forward.shape), tuple(constraint.shape)
    if (len(ini_shape) != 2 or len(fw_shape) != 2 or len(con_shape) != 2 or m != fw_shape[0]
            or ini_shape[1] != fw_shape[0] or ini_shape[0] > fw_shape[1]
            or con_shape[1] != fw_shape[1]):
        def tup(t):
            return "(" + ", ".join(str(x) for x in t) + ")"
        raise ValueError(
            f"shapes do not match: signals ({n}, {m}), initial {tup(ini_shape)}, forward "
            f"{tup(fw_shape)}, constraint {tup(con_shape)}"
        )
    if max_iter < 0:
        raise ValueError("max_iter must be at least 0")
    _require_threads(n_threads)
    problem = CsdProblem(initial, forward, constraint, norm_weight, threshold, max_iter)
    c = problem.n_coefficients
    fods = torch.empty((n, c), dtype=torch.float64, device="cuda")
    converged = torch.zeros(n, dtype=torch.uint8, device="cuda")
    solves = torch.zeros(n, dtype=torch.int32, device="cuda")
    lib = _lib.load()
    _lib.check(lib.lsd_fit_csd(dev.data_ptr(), n, m, problem._handle, fods.data_ptr(),
                               converged.data_ptr(), solves.data_ptr(), _stream_ptr(torch)))
    conv = converged.to(torch.bool)
    out = (_out(fods, host), _out(conv, host))
    if return_solves:
        out += (_out(solves, host),)
    return out


# The largest lmax of the reference's compiled SH code (cpp/src/sh/harmonics.hpp MAX_LMAX).  The
# GPU kernels take every lmax up to it: exact instantiations to 12 (symmetric) and 8 (full), one
# lmax-20 instantiation per layout for the rest, with the lmax read at run time.
MAX_LMAX = 20


def _n_sh(lmax: int, full: bool) -> int:
    return (lmax + 1) ** 2 if full else (lmax + 1) * (lmax + 2) // 2


def _check_lmax(lmax: int, full: bool, n_coeffs: int) -> None:
    """check_lmax (cpp/src/python/sh.cpp:41-54)."""
    if lmax < 0 or lmax > MAX_LMAX or (not full and lmax % 2 != 0):
        raise ValueError(f"lmax must be at most {MAX_LMAX}, and even unless full, got {lmax}")
    expected = _n_sh(lmax, full)
    if expected != n_coeffs:
        raise ValueError(f"lmax {lmax} has {expected} coefficients, got {n_coeffs}")


class Sphere:
    """A search sphere resident on the device: vertices, neighbours and (optionally) the tournier
    basis at the vertices for the device-side amplitudes."""

    def __init__(self, vertices, neighbours, basis=None):
        vert = np.ascontiguousarray(vertices, dtype=np.float64)
        nb = np.ascontiguousarray(neighbours, dtype=np.int64)
        b = None if basis is None else np.ascontiguousarray(basis, dtype=np.float64)
        lib = _lib.load()
        handle = ctypes.c_void_p()
        _lib.check(lib.lsd_sphere_create(
            _ptr(vert), vert.shape[0], nb.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            None if b is None else _ptr(b), 0 if b is None else b.shape[1], ctypes.byref(handle)))
        self._handle = handle
        self.n_vertices = vert.shape[0]
        self.n_coeffs = 0 if b is None else b.shape[1]

    def __del__(self, _release=_release):
        _release(self, "lsd_sphere_destroy")


def _as_device(a, torch):
    if isinstance(a, np.ndarray):
        return torch.from_numpy(np.ascontiguousarray(a)).to("cuda"), True
    if a.device.type != "cuda":
        return a.contiguous().to("cuda"), True
    return a.contiguous(), False


def find_peaks_device(coeffs, lmax: int, sphere: Sphere, n_peaks: int, relative_threshold: float,
                      threshold: float, *, full: bool = False, on_sphere=None):
    """The peaks of every row of coeffs (n, C) on the device; amplitudes on the sphere from
    on_sphere (n, V) or, when None, from the sphere's basis.  Returns (directions (n, P, 3),
    amplitudes (n, P), voxels whose distinct peaks overflowed)."""
    torch = _torch()
    dev, host = _as_device(coeffs, torch)
    n, c = dev.shape
    _check_lmax(lmax, full, c)
    if n_peaks < 1:
        raise ValueError("n_peaks must be at least 1")
    osph = None
    if on_sphere is not None:
        osph, _ = _as_device(on_sphere, torch)
    dirs = torch.empty((n, n_peaks, 3), dtype=torch.float64, device="cuda")
    amps = torch.empty((n, n_peaks), dtype=torch.float64, device="cuda")
    overflow = torch.zeros(1, dtype=torch.int32, device="cuda")
    lib = _lib.load()
    _lib.check(lib.lsd_find_peaks(
        dev.data_ptr(), n, c, int(lmax), int(bool(full)),
        None if osph is None else osph.data_ptr(), sphere._handle, int(n_peaks),
        float(relative_threshold), float(threshold), dirs.data_ptr(), amps.data_ptr(),
        overflow.data_ptr(), _stream_ptr(torch)))
    return _out(dirs, host), _out(amps, host), int(overflow.item())


def peaks_chunk(coeffs, lmax: int, on_sphere, vertices, neighbours, n_peaks: int,
                relative_threshold: float, threshold: float, n_threads: int, full: bool = False):
    """Finds the peaks of every voxel of a chunk (cpp/src/python/sh.cpp:112-198): coeffs (n, C)
    tournier, on_sphere (n, V), vertices (V, 3), neighbours (V, 6).  Returns directions
    (n, n_peaks, 3) and amplitudes (n, n_peaks), largest first, NaN-padded."""
    _check_f64(coeffs, "coeffs")
    _check_f64(on_sphere, "on_sphere")
    if not _is_c_contiguous(coeffs) or not _is_c_contiguous(on_sphere):
        raise ValueError("the array must be C-contiguous")
    _check_lmax(lmax, full, coeffs.shape[1])
    nv = vertices.shape[0]
    if (vertices.shape[1] != 3 or neighbours.shape != (nv, 6)
            or on_sphere.shape[0] != coeffs.shape[0] or on_sphere.shape[1] != nv):
        def tup(t):
            return "(" + ", ".join(str(x) for x in t) + ")"
        raise ValueError(
            f"shapes do not match: coeffs {tup(coeffs.shape)}, on_sphere {tup(on_sphere.shape)}, "
            f"vertices {tup(vertices.shape)}, neighbours {tup(neighbours.shape)}"
        )
    nb = np.asarray(neighbours)
    if nb.size and (nb.min() < 0 or nb.max() >= nv):
        raise ValueError("a neighbour index is outside the sphere")
    if n_peaks < 1:
        raise ValueError("n_peaks must be at least 1")
    _require_threads(n_threads)
    sphere = Sphere(np.asarray(vertices, dtype=np.float64), nb)
    dirs, amps, _ = find_peaks_device(coeffs, lmax, sphere, n_peaks, relative_threshold,
                                      threshold, full=full, on_sphere=on_sphere)
    return dirs, amps


def sh_derivatives(sh, lmax: int, theta: float, phi: float, full: bool = False):
    """The series' value and derivatives at (theta, phi) (cpp/src/python/sh.cpp:67-83):
    (f, df/dtheta, df/dphi / sin, d2f/dtheta2, d2f/dtheta dphi / sin, d2f/dphi2 / sin^2)."""
    torch = _torch()
    s = np.ascontiguousarray(np.asarray(sh, dtype=np.float64))
    _check_lmax(lmax, full, s.shape[0])
    dev = torch.from_numpy(s[None]).to("cuda")
    th = torch.tensor([float(theta)], dtype=torch.float64, device="cuda")
    ph = torch.tensor([float(phi)], dtype=torch.float64, device="cuda")
    out = torch.empty((1, 6), dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().lsd_sh_derivatives(dev.data_ptr(), 1, s.shape[0], int(lmax),
                                              int(bool(full)), th.data_ptr(), ph.data_ptr(),
                                              out.data_ptr(), _stream_ptr(torch)))
    return tuple(float(x) for x in out.cpu().numpy()[0])


def climb(sh, lmax: int, start, max_iters: int = 50, tol: float = 1e-4, max_step: float = 0.2,
          full: bool = False):
    """Climbs from start to the nearest peak (cpp/src/python/sh.cpp:85-108): the peak's
    direction (3,) and amplitude, both NaN if the climb did not stop."""
    torch = _torch()
    s = np.ascontiguousarray(np.asarray(sh, dtype=np.float64))
    _check_lmax(lmax, full, s.shape[0])
    if max_iters < 0:
        raise ValueError("max_iters must be at least 0")
    dev = torch.from_numpy(s[None]).to("cuda")
    st = torch.tensor(np.asarray(start, dtype=np.float64).reshape(1, 3), device="cuda")
    d = torch.empty((1, 3), dtype=torch.float64, device="cuda")
    a = torch.empty(1, dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().lsd_sh_climb(dev.data_ptr(), 1, s.shape[0], int(lmax), int(bool(full)),
                                        st.data_ptr(), int(max_iters), float(tol),
                                        float(max_step), d.data_ptr(), a.data_ptr(),
                                        _stream_ptr(torch)))
    return d.cpu().numpy()[0], float(a.item())


class LsiDesign:
    """min ||a x - s||^2 + ridge^2 ||x||^2 subject to g x >= 0, prepared once on the host and kept
    on the device (cpp/src/python/sd.cpp:360-366, 521-532)."""

    def __init__(self, a, g, ridge: float):
        am = _host_matrix(a, "a")
        gm = _host_matrix(g, "g")
        lib = _lib.load()
        handle = ctypes.c_void_p()
        _lib.check(lib.lsd_lsi_create(_ptr(am), am.shape[0], am.shape[1], _ptr(gm), gm.shape[0],
                                      gm.shape[1], float(ridge), ctypes.byref(handle)))
        self._handle = handle
        self.n_meas = am.shape[0]
        self.n_params = am.shape[1]
        self.n_constraints = gm.shape[0]

    def __del__(self, _release=_release):
        _release(self, "lsd_lsi_destroy")


def fit_lsi(signals, design: LsiDesign, n_threads: int, *, return_status: bool = False,
            only_qr: bool = False):
    """The constrained solution (v, n) of every row of signals (v, m) and whether each voxel's
    NNLS converged (v,) (sd.cpp:368-394, 534-542).  With return_status also the per-voxel
    status (LSD_STATUS_QR_FALLBACK where the QR-based NNLS finished the voxel); only_qr runs the
    QR-based NNLS for every voxel (the reference's solve_lsi_by_qr)."""
    torch = _torch()
    dev, host = _device_signals(signals, torch)
    v, m = dev.shape
    if m != design.n_meas:
        raise ValueError(
            f"signals ({v}, {m}) do not have the design's {design.n_meas} measurements"
        )
    _require_threads(n_threads)
    x = torch.empty((v, design.n_params), dtype=torch.float64, device="cuda")
    converged = torch.zeros(v, dtype=torch.uint8, device="cuda")
    status = torch.zeros(v, dtype=torch.int32, device="cuda")
    lib = _lib.load()
    fn = lib.lsd_fit_lsi_by_qr if only_qr else lib.lsd_fit_lsi
    _lib.check(fn(dev.data_ptr(), v, m, design._handle, x.data_ptr(), converged.data_ptr(),
                  status.data_ptr(), _stream_ptr(torch)))
    out = (_out(x, host), _out(converged.to(torch.bool), host))
    if return_status:
        out += (_out(status, host),)
    return out


def sh_project(amplitudes, projection, *, exact_integral: bool = False):
    """SH coefficients (n, c) = amplitudes (n, k) @ projection (k, c) on the device; with
    exact_integral, c00 = sum(amplitudes) / sqrt(4 pi) (model.py:412-440)."""
    torch = _torch()
    amps, host = _as_device(amplitudes, torch)
    proj, _ = _as_device(projection, torch)
    _check_f64(amps, "amplitudes")
    _check_f64(proj, "projection")
    n, k = amps.shape
    if proj.shape[0] != k:
        raise ValueError(f"amplitudes ({n}, {k}) and projection {tuple(proj.shape)} do not match")
    c = proj.shape[1]
    out = torch.empty((n, c), dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().lsd_sh_project(amps.data_ptr(), n, k, proj.data_ptr(), c,
                                          int(bool(exact_integral)), out.data_ptr(),
                                          _stream_ptr(torch)))
    return _out(out, host)


# ---------------------------------------------------------------------------------------------
# Streamline tracking (cpp/src/python/tracking.cpp).

def _number_text(value: float) -> str:
    """The binding's number text: fixed notation, NaN and inf spelled out."""
    if value != value:
        return "NaN"
    if value in (float("inf"), float("-inf")):
        return "inf" if value > 0 else "-inf"
    text = np.format_float_positional(float(value), trim="-")
    return text


def _require_cutoff(cutoff: float) -> None:
    if not (np.isfinite(cutoff) and cutoff >= 0.0):
        raise ValueError(f"the cutoff must be non-negative and finite, got {_number_text(cutoff)}")


class DetParams:
    """Parameters of deterministic tracking (cpp/src/python/tracking.cpp:180-193, 409-426):
    cutoff, the FOD amplitude below which a streamline stops, and max_angle, the largest turn
    between steps in degrees.  lmax comes from the FOD."""

    def __init__(self, cutoff: float, max_angle: float):
        cutoff, max_angle = float(cutoff), float(max_angle)
        _require_cutoff(cutoff)
        if not (0.0 < max_angle <= 180.0):
            raise ValueError(f"the maximum angle must be in (0, 180], got {_number_text(max_angle)}")
        self.cutoff = cutoff
        self.max_angle = max_angle

    def _params(self, full: bool) -> "_lib.TrackParams":
        return _lib.TrackParams(algorithm=0, full=int(full), cutoff=self.cutoff,
                                max_angle_deg=self.max_angle, trials=1, attempts=1)


class ProbParams:
    """Parameters of probabilistic tracking (cpp/src/python/tracking.cpp:195-224, 428-450):
    cutoff, the cone's half-angle max_angle in degrees, trials candidates per step and attempts
    random directions per seed.  The rejection-sampling bound is calibrated from the FOD's lmax."""

    def __init__(self, cutoff: float, max_angle: float, trials: int = 1000, attempts: int = 1000):
        cutoff, max_angle = float(cutoff), float(max_angle)
        _require_cutoff(cutoff)
        if not (max_angle * (np.pi / 180.0) > 0.001 and max_angle <= 180.0):
            raise ValueError("the maximum angle must be above 0.001 rad (0.057 degrees) and at "
                             f"most 180, got {_number_text(max_angle)}")
        if trials < 1 or attempts < 1:
            raise ValueError(f"trials and attempts must be at least 1, got {trials} and {attempts}")
        self.cutoff = cutoff
        self.max_angle = max_angle
        self.trials = int(trials)
        self.attempts = int(attempts)

    def _params(self, full: bool) -> "_lib.TrackParams":
        return _lib.TrackParams(algorithm=1, full=int(full), cutoff=self.cutoff,
                                max_angle_deg=self.max_angle, trials=self.trials,
                                attempts=self.attempts)


_TEST_RULES = {"straight": 0, "refuse": 1, "quarter_turn": 2, "wander": 3}


class TestAlgorithm:
    """The fixed-rule algorithm that tests the step loop (cpp/src/python/tracking.cpp:117-138,
    452-478): "straight", "quarter_turn" (90 degrees about z per step), "wander" (a random walk of
    three draws per step), or "refuse" with stop, an EndStatus from 1 to 6.  Its first-direction
    search draws a random unit vector and fails where the seed's x > search_refuses_past_x."""

    __test__ = False  # not a pytest class

    def __init__(self, step: str, stop: int | None = None,
                 search_refuses_past_x: float = float("inf")):
        if stop is not None and not (0 <= int(stop) <= 255):
            raise OverflowError("out of range integral type conversion attempted")
        if stop is None and step in ("straight", "quarter_turn", "wander"):
            pass
        elif stop is not None and step == "refuse":
            if not 1 <= int(stop) <= 6:
                raise ValueError(f"a stop is an EndStatus from 1 to 6, got {int(stop)}")
        else:
            stop_text = f"Some({int(stop)})" if stop is not None else "None"
            raise ValueError("expected straight, quarter_turn, wander, or refuse with a stop; got "
                             f'"{step}" with stop {stop_text}')
        self.step = step
        self.stop = None if stop is None else int(stop)
        self.search_refuses_past_x = float(search_refuses_past_x)

    def _params(self, full: bool) -> "_lib.TrackParams":
        return _lib.TrackParams(algorithm=2, full=int(full), cutoff=0.0, max_angle_deg=180.0,
                                trials=1, attempts=1, test_rule=_TEST_RULES[self.step],
                                test_stop=self.stop or 0,
                                test_refuses_past_x=self.search_refuses_past_x)


def calibrate_tracking(lmax: int, full: bool, max_angle_rad: float):
    """The rejection-sampling calibration of prob (cpp/src/tracking/calibrate.cpp:63-108): the
    grid of directions around +z (n, 3) and the bound's ratio."""
    lib = _lib.load()
    n = ctypes.c_int64()
    ratio = ctypes.c_double()
    _lib.check(lib.lsd_track_calibrate(int(lmax), int(bool(full)), float(max_angle_rad), None, 0,
                                       ctypes.byref(n), ctypes.byref(ratio)))
    grid = np.zeros((n.value, 3))
    _lib.check(lib.lsd_track_calibrate(int(lmax), int(bool(full)), float(max_angle_rad),
                                       _ptr(grid), n.value, ctypes.byref(n), ctypes.byref(ratio)))
    return grid, ratio.value


def _tuple_text(shape) -> str:
    return "(" + ", ".join(str(x) for x in shape) + ")"


def _inverse_affine(m, name: str) -> np.ndarray:
    _check_f64(m, name)
    a = np.asarray(m.cpu().numpy() if hasattr(m, "cpu") else m)
    if a.ndim != 2 or a.shape != (4, 4):
        raise ValueError(f"{name} must be (4, 4), got {_tuple_text(a.shape[:2])}")
    return np.ascontiguousarray(a, dtype=np.float64)


def _require_dtype(array, dtype, name: str) -> None:
    torch_types = {np.float32: "float32", np.float64: "float64", np.int64: "int64"}
    if isinstance(array, np.ndarray):
        if array.dtype != dtype:
            raise TypeError(f"{name} must be {np.dtype(dtype).name}, got {array.dtype}")
        return
    if str(array.dtype).replace("torch.", "") != torch_types[dtype]:
        raise TypeError(f"{name} must be {np.dtype(dtype).name}, got {array.dtype}")


def track_batch(fod, fod_inv_affine, mask, seeds, directions, indices, run_seed: int,
                step_size: float, max_steps: int, bidirectional: bool, algorithm, positions,
                n_threads: int, full: bool = False):
    """Tracks a batch of seeds into the rows of positions (cpp/src/python/tracking.cpp:264-404).

    fod (X, Y, Z, C) float32 tournier, all degrees if full; mask None or ((X, Y, Z) float32,
    (4, 4) world-to-voxel); seeds (n, 3) float64 world mm, directions (n, 3) or None for the
    algorithm's first-direction search, indices (n,) int64 keying each streamline's random
    numbers; positions (batch >= n, max_steps + 1, 3) float32, filled in place (rows past n and
    points past a row's count are left as they are).  Arrays may be numpy (copied to the device
    and back) or torch CUDA tensors (used in place).  Returns counts (n,) int64, status (n, 2)
    uint8, followed (n, 3) float64 and the sampling counts -- torch tensors when positions is one.
    """
    if not isinstance(algorithm, (DetParams, ProbParams, TestAlgorithm)):
        raise TypeError("expected DetParams, ProbParams or TestAlgorithm for algorithm, got "
                        f"{type(algorithm).__name__}")
    torch = _torch()
    _require_dtype(fod, np.float32, "fod")
    if not _is_c_contiguous(fod):
        raise ValueError("the array must be C-contiguous")
    if len(fod.shape) != 4:
        raise ValueError(f"fod must be (X, Y, Z, C), got {_tuple_text(fod.shape)}")
    fod_inv = _inverse_affine(fod_inv_affine, "fod_inv_affine")
    mask_values, mask_inv = None, None
    if mask is not None:
        mask_values, mask_inv_in = mask
        _require_dtype(mask_values, np.float32, "mask")
        if not _is_c_contiguous(mask_values):
            raise ValueError("the array must be C-contiguous")
        mask_inv = _inverse_affine(mask_inv_in, "mask inverse affine")
    _require_dtype(seeds, np.float64, "seeds")
    if directions is not None:
        _require_dtype(directions, np.float64, "directions")
    _require_dtype(indices, np.int64, "indices")
    n = seeds.shape[0]
    bad_dirs = directions is not None and (len(directions.shape) != 2 or directions.shape[0] != n
                                           or directions.shape[1] != 3)
    if len(seeds.shape) != 2 or seeds.shape[1] != 3 or bad_dirs or indices.shape[0] != n:
        dir_text = "None" if directions is None else f"Some({_tuple_text(directions.shape[:2])})"
        raise ValueError("expected seeds and directions (n, 3) and indices (n,): got seeds "
                         f"{_tuple_text(seeds.shape[:2])}, directions {dir_text}, indices "
                         f"{indices.shape[0]}")
    idx_min = (int(indices.min()) if isinstance(indices, np.ndarray) else int(indices.min().item())) \
        if n else 0
    if idx_min < 0:
        raise ValueError("seed indices must be non-negative")
    if not (np.isfinite(step_size) and step_size > 0 and max_steps > 0):
        raise ValueError(f"need a positive step size and max_steps, got {_number_text(step_size)} "
                         f"and {max_steps}")
    _require_dtype(positions, np.float32, "positions")
    if not _is_c_contiguous(positions):
        raise ValueError("positions must be C-contiguous")
    pshape = tuple(positions.shape)
    if len(pshape) != 3 or pshape[0] < n or pshape[1] != max_steps + 1 or pshape[2] != 3:
        raise ValueError(f"positions must be (at least {n}, {max_steps + 1}, 3), got "
                         f"{_tuple_text(pshape)}")
    _require_threads(n_threads)

    params = algorithm._params(full)
    d_fod, _ = _as_device(fod, torch)
    d_mask = None
    if mask_values is not None:
        d_mask, _ = _as_device(mask_values, torch)
    d_seeds, _ = _as_device(seeds, torch)
    d_dirs = None if directions is None else _as_device(directions, torch)[0]
    d_idx, _ = _as_device(indices, torch)
    on_host = isinstance(positions, np.ndarray)
    d_pos = torch.empty((n, max_steps + 1, 3), dtype=torch.float32, device="cuda") if on_host \
        else positions
    counts = torch.empty(n, dtype=torch.int64, device="cuda")
    status = torch.empty((n, 2), dtype=torch.uint8, device="cuda")
    followed = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    stats = _lib.TrackStats()
    fs = fod.shape
    ms = (0, 0, 0) if mask_values is None else tuple(mask_values.shape)
    _lib.check(_lib.load().lsd_track_batch(
        d_fod.data_ptr(), fs[0], fs[1], fs[2], fs[3], _ptr(fod_inv),
        None if d_mask is None else d_mask.data_ptr(), ms[0], ms[1], ms[2],
        None if mask_inv is None else _ptr(mask_inv),
        d_seeds.data_ptr(), None if d_dirs is None else d_dirs.data_ptr(), d_idx.data_ptr(), n,
        int(run_seed) & 0xFFFFFFFFFFFFFFFF, float(step_size), int(max_steps),
        int(bool(bidirectional)), ctypes.byref(params), d_pos.data_ptr(), pshape[0],
        counts.data_ptr(), status.data_ptr(), followed.data_ptr(), ctypes.byref(stats),
        _stream_ptr(torch)))
    stats_out = {"sampled_steps": stats.sampled_steps,
                 "accepted_index_sum": stats.accepted_index_sum,
                 "truncations": stats.truncations, "max_truncation": stats.max_truncation,
                 "calibration_directions": stats.calibration_directions}
    if not on_host:
        return counts, status, followed, stats_out
    counts_h = counts.cpu().numpy()
    if n:
        rows = d_pos.cpu().numpy()
        written = np.arange(max_steps + 1)[None, :, None] < counts_h[:, None, None]
        np.copyto(positions[:n], rows, where=written)
    return counts_h, status.cpu().numpy(), followed.cpu().numpy(), stats_out
