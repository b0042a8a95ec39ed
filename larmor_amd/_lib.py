"""Loads liblarmor_sd.so (the HIP kernels and their C-ABI) and declares its entry points.

The library is built in-tree by `larmor_amd.build.build_library()`.  There is no other compute
path: when the library or a GPU is missing, every fit raises.
"""

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "liblarmor_sd.so"
ABI_VERSION = 6  # LSD_ABI_VERSION of include/larmor_sd.h these bindings follow
HEADER_PATH = Path(__file__).resolve().parents[1] / "include" / "larmor_sd.h"

_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p

# name: (restype, argtypes)
SIGNATURES = {
    "lsd_last_error": (ctypes.c_char_p, []),
    "lsd_abi_version": (ctypes.c_int, []),
    "lsd_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "lsd_profile_collect": (
        ctypes.c_int,
        [ctypes.c_char_p, ctypes.c_int64, _d, _i64, _d, ctypes.c_int64, _i64],
    ),
    "lsd_solver_create_l2nnls": (
        ctypes.c_int,
        [_d, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
         ctypes.POINTER(_vp)],
    ),
    "lsd_solver_create_drl": (
        ctypes.c_int,
        [_d, ctypes.c_int64, ctypes.c_int64, _d, ctypes.c_int64, ctypes.c_int64,
         ctypes.c_double, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(_vp)],
    ),
    "lsd_solver_is_drl": (ctypes.c_int, [_vp]),
    "lsd_solver_create_lasso": (
        ctypes.c_int,
        [_d, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
         ctypes.c_int64, ctypes.POINTER(_vp)],
    ),
    "lsd_solver_is_lasso": (ctypes.c_int, [_vp]),
    "lsd_solver_n_unconverged": (ctypes.c_int, [_vp, _i64]),
    "lsd_solver_destroy": (None, [_vp]),
    "lsd_solver_n_dirs": (ctypes.c_int64, [_vp]),
    "lsd_solver_n_meas": (ctypes.c_int64, [_vp]),
    "lsd_solver_normal_equations": (ctypes.c_int, [_vp]),
    "lsd_tissue_kernel_create": (
        ctypes.c_int,
        [_d, _d, _d, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _i64, ctypes.c_int64,
         ctypes.POINTER(_vp)],
    ),
    "lsd_tissue_kernel_destroy": (None, [_vp]),
    "lsd_fit_single_tissue": (
        ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp, _vp, _vp]
    ),
    "lsd_fit_multi_tissue": (
        ctypes.c_int,
        [_vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, ctypes.c_int64,
         ctypes.c_double, ctypes.c_double, _vp, _vp, _vp, _vp, _vp],
    ),
    "lsd_correlations": (
        ctypes.c_int,
        [_vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp],
    ),
    "lsd_fit_multi_tissue_mixed": (
        ctypes.c_int,
        [_vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, ctypes.c_int64,
         ctypes.c_double, ctypes.c_double, _vp, _vp, _vp, _vp, _vp],
    ),
    "lsd_fit_multi_tissue_fp32": (
        ctypes.c_int,
        [_vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, ctypes.c_int64,
         ctypes.c_double, ctypes.c_double, _vp, _vp, _vp, _vp, _vp],
    ),
    "lsd_fit_multi_tissue_noalloc": (
        ctypes.c_int,
        [_vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, ctypes.c_int64,
         ctypes.c_double, ctypes.c_double, _vp, _vp, _vp, _vp, _vp],
    ),
    "lsd_noalloc_workspace_bytes": (ctypes.c_int64, []),
    "lsd_fit_scratch_bytes": (ctypes.c_int64, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]),
    "lsd_lsqnonneg": (
        ctypes.c_int,
        [_vp, ctypes.c_int64, ctypes.c_int64, _vp, ctypes.c_int64, ctypes.c_double, ctypes.c_int,
         _vp, _vp, _vp],
    ),
    "lsd_ridge_l1_nnls": (
        ctypes.c_int,
        [_vp, ctypes.c_int64, ctypes.c_int64, _vp, ctypes.c_int64, ctypes.c_double,
         ctypes.c_double, ctypes.c_double, ctypes.c_int, _vp, _vp, _vp],
    ),
    "lsd_fit_multi_tissue_batched": (
        ctypes.c_int,
        [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _vp, _vp, _vp,
         _vp, _vp],
    ),
    "lsd_fit_log_signal": (
        ctypes.c_int,
        [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_int64, _vp, ctypes.c_int64, ctypes.c_int,
         ctypes.c_int64, ctypes.c_double, _vp, _vp, _vp],
    ),
    "lsd_csd_create": (
        ctypes.c_int,
        [_d, ctypes.c_int64, ctypes.c_int64, _d, ctypes.c_int64, _d, ctypes.c_int64,
         ctypes.c_double, ctypes.c_double, ctypes.c_int64, ctypes.POINTER(_vp)],
    ),
    "lsd_csd_destroy": (None, [_vp]),
    "lsd_fit_csd": (
        ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp]
    ),
    "lsd_sphere_create": (
        ctypes.c_int,
        [_d, ctypes.c_int64, _i64, _d, ctypes.c_int64, ctypes.POINTER(_vp)],
    ),
    "lsd_sphere_destroy": (None, [_vp]),
    "lsd_find_peaks": (
        ctypes.c_int,
        [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, _vp, _vp,
         ctypes.c_int64, ctypes.c_double, ctypes.c_double, _vp, _vp, _vp, _vp],
    ),
    "lsd_sh_derivatives": (
        ctypes.c_int,
        [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, _vp, _vp, _vp, _vp],
    ),
    "lsd_sh_climb": (
        ctypes.c_int,
        [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, _vp, ctypes.c_int64,
         ctypes.c_double, ctypes.c_double, _vp, _vp, _vp],
    ),
    "lsd_lsi_create": (
        ctypes.c_int,
        [_d, ctypes.c_int64, ctypes.c_int64, _d, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
         ctypes.POINTER(_vp)],
    ),
    "lsd_lsi_destroy": (None, [_vp]),
    "lsd_lsi_n_meas": (ctypes.c_int64, [_vp]),
    "lsd_lsi_n_params": (ctypes.c_int64, [_vp]),
    "lsd_lsi_n_constraints": (ctypes.c_int64, [_vp]),
    "lsd_fit_lsi": (
        ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp]
    ),
    "lsd_fit_lsi_by_qr": (
        ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp]
    ),
    "lsd_sh_project": (
        ctypes.c_int,
        [_vp, ctypes.c_int64, ctypes.c_int64, _vp, ctypes.c_int64, ctypes.c_int, _vp, _vp],
    ),
}




class TrackParams(ctypes.Structure):
    """lsd_track_params of include/larmor_sd.h."""

    _fields_ = [
        ("algorithm", ctypes.c_int32),
        ("full", ctypes.c_int32),
        ("cutoff", ctypes.c_double),
        ("max_angle_deg", ctypes.c_double),
        ("trials", ctypes.c_int64),
        ("attempts", ctypes.c_int64),
        ("test_rule", ctypes.c_int32),
        ("test_stop", ctypes.c_int32),
        ("test_refuses_past_x", ctypes.c_double),
    ]


class TrackStats(ctypes.Structure):
    """lsd_track_stats of include/larmor_sd.h."""

    _fields_ = [
        ("sampled_steps", ctypes.c_uint64),
        ("accepted_index_sum", ctypes.c_uint64),
        ("truncations", ctypes.c_uint64),
        ("max_truncation", ctypes.c_double),
        ("calibration_directions", ctypes.c_uint64),
    ]


SIGNATURES["lsd_tune_lr_capacity"] = (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64])
SIGNATURES["lsd_tuned_lr_capacity"] = (ctypes.c_int64, [ctypes.c_int64])
SIGNATURES["lsd_track_calibrate"] = (
    ctypes.c_int,
    [ctypes.c_int64, ctypes.c_int, ctypes.c_double, _d, ctypes.c_int64, _i64, _d],
)
SIGNATURES["lsd_track_batch"] = (
    ctypes.c_int,
    [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _d,
     _vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _d,
     _vp, _vp, _vp, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double, ctypes.c_int64,
     ctypes.c_int, ctypes.POINTER(TrackParams), _vp, ctypes.c_int64, _vp, _vp, _vp,
     ctypes.POINTER(TrackStats), _vp],
)


class LibraryMissing(RuntimeError):
    """liblarmor_sd.so is not built (run __graft_entry__.build() or larmor_amd.build)."""


def load() -> ctypes.CDLL:
    """Returns the loaded library, declaring every entry point's signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise LibraryMissing(
            f"{LIB_PATH} is not built; build it with `python -m larmor_amd.build`."
        )
    # torch ships its own HIP runtime (torch/lib/libamdhip64.so, SONAME libamdhip64.so.7).
    # Importing torch first makes the loader bind this library to that same runtime, so torch's
    # allocations and streams and these kernels share one HIP context.
    import torch  # noqa: F401

    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    version = lib.lsd_abi_version()
    if version != ABI_VERSION:
        raise LibraryMissing(
            f"{LIB_PATH} has ABI version {version}, these bindings need {ABI_VERSION}; rebuild it "
            "with `python -m larmor_amd.build`."
        )
    _lib = lib
    return lib


def check(status: int) -> None:
    """Raises ValueError (invalid argument) or RuntimeError (HIP failure) for a non-zero status."""
    if status == 0:
        return
    message = load().lsd_last_error().decode()
    if status == 1:
        raise ValueError(message)
    raise RuntimeError(message)


def profile_enable(on: bool = True) -> None:
    """Turns per-launch event timing on or off."""
    check(load().lsd_profile_enable(1 if on else 0))


def profile_collect() -> dict[str, tuple[float, int, float]]:
    """{kernel name: (summed ms, launches, counted FP64 work)} since the last collect; waits for
    the events.  The work is 0 for kernels that do not count theirs."""
    lib = load()
    names = ctypes.create_string_buffer(4096)
    ms = (ctypes.c_double * 64)()
    counts = (ctypes.c_int64 * 64)()
    work = (ctypes.c_double * 64)()
    n = ctypes.c_int64()
    check(lib.lsd_profile_collect(names, 4096, ms, counts, work, 64, ctypes.byref(n)))
    labels = names.value.decode().split("\n")
    return {labels[i]: (ms[i], counts[i], work[i]) for i in range(n.value)}


def declared_symbols() -> list[str]:
    """The function names include/larmor_sd.h declares."""
    import re

    text = HEADER_PATH.read_text()
    return sorted(set(re.findall(r"\b(lsd_[a-z0-9_]+)\s*\(", text)))
