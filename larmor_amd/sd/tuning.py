"""Autotuning of the multi-tissue fit's kernel shapes per workload bucket (no counterpart in the
reference, which has one CPU code path).

The knob is the low-resolution solve's LDS factor capacity (16, 20, 24 or 32 active columns):
a larger factor fits more voxels' active sets in LDS but leaves fewer waves per SIMD (8, 8, 7, 6)
and more registers to spill at the lower end; a voxel whose active set outgrows it finishes in
the global-memory solve with the same arithmetic, so every capacity gives the same bits and the
choice is speed only.  Which one is fastest depends on the bucket -- the row class of the
acquisition (m <= 256, 384, 512, 768, 1024 measurements) and the number of low-resolution directions
(27, 44 or 65 for lmax 6, 8, 10), which set the active-set sizes -- so `autotune` times each on a
sample of the bucket's own voxels and keeps the fastest; `use` puts a bucket's choice in force
before its fits (the library reads it when it enqueues a fit's launches).
"""

import time
from dataclasses import dataclass, field

import numpy as np

from larmor_amd import _core, _lib

CAPACITIES = (16, 20, 24, 32)
DEFAULT_CAPACITY = 20


def row_class(m: int) -> int:
    """The measurement-count class of the solve kernels (rows per lane 4, 6, 8, 12, 16)."""
    for bound in (256, 384, 512, 768, 1024):
        if m <= bound:
            return bound
    raise ValueError(f"problem size exceeds 1024 measurements, got {m}")


@dataclass
class TuningTable:
    """The chosen capacity per bucket (row class, low-resolution directions) and the timings
    (ms per fit of the sample) it was chosen from."""

    choice: dict = field(default_factory=dict)
    timings: dict = field(default_factory=dict)

    def bucket(self, m: int, n_lr: int) -> tuple:
        return (row_class(m), int(n_lr))

    def use(self, m: int, n_lr: int) -> int:
        """Puts the bucket's capacity in force (the default for an untuned bucket); returns it."""
        cap = self.choice.get(self.bucket(m, n_lr), DEFAULT_CAPACITY)
        _lib.check(_lib.load().lsd_tune_lr_capacity(int(m), int(cap)))
        return cap

    def as_rows(self) -> list[dict]:
        return [{"row_class": b[0], "n_lr": b[1], "capacity": c,
                 "ms_by_capacity": self.timings.get(b, {})} for b, c in sorted(self.choice.items())]


TABLE = TuningTable()


def set_capacity(m: int, capacity: int) -> None:
    _lib.check(_lib.load().lsd_tune_lr_capacity(int(m), int(capacity)))


def capacity_in_force(m: int) -> int:
    return int(_lib.load().lsd_tuned_lr_capacity(int(m)))


def autotune(signals, weights, lr, lr_solver, hr, hr_solver, *, max_rounds: int = 50,
             fraction_tol: float = 1e-7, final_tol_x: float = 1e-2, sample: int = 65536,
             repeats: int = 2, capacities=CAPACITIES, table: TuningTable = TABLE,
             seed: int = 0) -> int:
    """Times the multi-tissue fit of a random sample of `signals` (device or host rows of one
    acquisition) at each capacity (a warm-up, then the best of `repeats` timed fits), records the
    fastest for the bucket in `table`, puts it in force and returns it."""
    import torch

    m = signals.shape[1]
    n_lr = lr.n_anisotropic
    rows = signals if isinstance(signals, torch.Tensor) else torch.from_numpy(
        np.ascontiguousarray(signals, dtype=np.float64))
    rows = rows.to("cuda", dtype=torch.float64)
    if rows.shape[0] > sample:
        pick = np.sort(np.random.default_rng(seed).choice(rows.shape[0], sample, replace=False))
        rows = rows[torch.from_numpy(pick).to("cuda")].contiguous()
    w = torch.as_tensor(weights, dtype=torch.float64, device="cuda")
    times = {}
    for cap in capacities:
        set_capacity(m, cap)
        best = float("inf")
        for r in range(repeats + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _core.fit_multi_tissue(rows, w, lr, lr_solver, hr, hr_solver, max_rounds,
                                   fraction_tol, final_tol_x, 1)
            torch.cuda.synchronize()
            if r > 0:
                best = min(best, (time.perf_counter() - t0) * 1e3)
        times[cap] = best
    chosen = min(times, key=times.get)
    bucket = table.bucket(m, n_lr)
    table.choice[bucket] = chosen
    table.timings[bucket] = times
    set_capacity(m, chosen)
    return chosen
