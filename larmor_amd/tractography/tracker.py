"""The tracker: streamlines through a field of FODs from seeds, batch by batch on the GPU.

Keeps the contract of larmor's `Tracker` (src/larmor/tractography/tracker.py): the same
parameters, defaults, validation and error texts, the same batching of seeds (so that a run seed
gives the same seeds and the same streamlines), the same run record.  What differs is where the
work runs: the FOD, the mask and the points buffer stay in HBM for the whole run, each batch is
one launch of the tracking kernel (`lsd_track_batch`), and only the points of the streamlines kept
are copied back to the host.
"""

import contextlib
import json
import secrets
from dataclasses import dataclass, fields
from math import asin, ceil, degrees, isfinite, radians, sin
from pathlib import Path

import numpy as np

from larmor_amd import _core
from larmor_amd.io.tck import TckWriter
from larmor_amd.sd.model import n_threads
from larmor_amd.sd.sh import convert_sh_basis, sh_layout
from larmor_amd.tractography.seed import MaskSeeds, Seeds, SeedStrategy, write_seeds
from larmor_amd.tractography.termination import auto_cutoff
from larmor_amd.types.tractogram import Tractogram
from larmor_amd.types.volume import Volume

VERSION = "larmor_amd 0.2"


@dataclass(frozen=True)
class _Algorithm:
    """A tracking algorithm: its name, default angle (or the smallest radius of curvature, in
    voxels, the default angle keeps) and the maker of its kernel parameters."""

    name: str
    default_max_angle: float | None
    make_params: object
    default_min_radius: float | None = None


# Smallest radius of curvature, in voxels, that the deterministic default angle allows.
DET_MIN_RADIUS = 0.65

ALGORITHMS = {
    "det": _Algorithm("det", None, _core.DetParams, DET_MIN_RADIUS),
    "prob": _Algorithm("prob", 15, _core.ProbParams),
}

# Batching of the seeds: the reference's figures, kept so that a run seed draws the same seeds
# batch by batch (a MaskSeeds batch is drawn from (run seed, batch number) and its size).
BUFFER_BYTES = 128 * 2**20
MAX_BATCH = 20_000
SEEDS_PER_STREAMLINE = 1000
DRAW_PER_MISSING = 4
MIN_DRAW = 256


@dataclass(frozen=True)
class TractographyRun:
    """The record of one tracking: what was written, the settings and the counts."""

    path: Path | None
    seeds_path: Path | None
    parameters: dict
    num_seeds: int
    num_streamlines: int
    num_rejected: dict
    sampling: dict | None = None
    inputs: dict | None = None
    tractogram: Tractogram | None = None

    def write_json(self, path) -> None:
        data = {f.name: getattr(self, f.name) for f in fields(self)}
        del data["tractogram"]
        data["path"] = None if self.path is None else str(self.path)
        data["seeds_path"] = None if self.seeds_path is None else str(self.seeds_path)
        Path(path).write_text(json.dumps(data, indent=4), encoding="utf-8")


@dataclass
class _Counts:
    seeds: int = 0
    accepted: int = 0
    no_start: int = 0
    too_short: int = 0
    sampled_steps: int = 0
    accepted_index_sum: int = 0
    truncations: int = 0
    max_truncation: float = 0.0
    calibration_directions: int = 0

    def add_sampling(self, stats: dict) -> None:
        self.sampled_steps += int(stats["sampled_steps"])
        self.accepted_index_sum += int(stats["accepted_index_sum"])
        self.truncations += int(stats["truncations"])
        self.max_truncation = max(self.max_truncation, float(stats["max_truncation"]))
        self.calibration_directions = int(stats["calibration_directions"])

    def sampling(self) -> dict | None:
        """Mean samples per step (calibration directions + mean accepted candidate index),
        steps per truncation and the largest truncation; None if no step was sampled."""
        if self.sampled_steps == 0:
            return None
        return {
            "mean_samples_per_step": self.calibration_directions
            + self.accepted_index_sum / self.sampled_steps,
            "steps_per_truncation": self.sampled_steps / self.truncations
            if self.truncations else None,
            "max_truncation": self.max_truncation if self.truncations else None,
            "truncations": self.truncations,
            "sampled_steps": self.sampled_steps,
            "calibration_directions": self.calibration_directions,
        }


@dataclass(frozen=True)
class _Prepared:
    coeffs: np.ndarray
    full: bool
    fod_inv: np.ndarray
    mask: tuple | None
    step_size: float
    min_length: float
    max_length: float
    max_angle: float
    min_radius: float
    cutoff: float
    params: object
    max_steps: int
    batch_size: int


def _check_positive(name: str, value) -> None:
    if value is not None and not (isfinite(value) and value > 0):
        raise ValueError(f"The {name} must be positive and finite, got {value}.")


class Tracker:
    """Tracks streamlines through a field of FODs with "det" (deterministic) or "prob"
    (probabilistic) tracking; lengths in mm.  Without a max angle, "det" keeps the tightest curve
    at 0.65 voxel radius and "prob" uses 15 degrees."""

    def __init__(self, algorithm, *, step_size: float | None = None,
                 max_angle: float | None = None, cutoff="auto", min_length: float | None = None,
                 max_length: float | None = None, bidirectional: bool = True,
                 progress: bool = False) -> None:
        if isinstance(algorithm, str):
            if algorithm not in ALGORITHMS:
                known = ", ".join(sorted(ALGORITHMS)) or "none yet"
                raise ValueError(f"Unknown algorithm {algorithm!r}; the known ones are: {known}.")
            algo = ALGORITHMS[algorithm]
        else:
            algo = algorithm
        _check_positive("step size", step_size)
        _check_positive("min length", min_length)
        _check_positive("max length", max_length)
        if min_length is not None and max_length is not None and min_length >= max_length:
            raise ValueError("Cannot track with min length >= max length.")
        if max_angle is not None and not (0 < max_angle <= 180):
            raise ValueError(f"Expected max angle to be in (0, 180], got {max_angle}")
        if isinstance(cutoff, str):
            if cutoff != "auto":
                raise ValueError(f"Expected a number or 'auto' as cutoff, got {cutoff!r}.")
        elif not (isfinite(cutoff) and cutoff >= 0):
            raise ValueError(f"The cutoff must be non-negative and finite, got {cutoff}.")
        self.algorithm = algo
        self.step_size = step_size
        self.min_length = min_length
        self.max_length = max_length
        self.max_angle = max_angle
        self.cutoff = cutoff
        self.bidirectional = bidirectional
        self.progress = progress

    def track(self, fod: Volume, output, seeds, *, num_streamlines: int | None = None,
              mask: Volume | None = None, basis: str = "tournier", seed: int | None = None,
              output_seeds=None, n_jobs: int = -1) -> TractographyRun:
        """Tracks from `seeds` (fixed `Seeds`, or an endless `SeedStrategy` with
        num_streamlines) and writes the accepted streamlines to `output` (.tck), or keeps them in
        the returned record when output is None."""
        if num_streamlines is None and not isinstance(seeds, Seeds):
            raise ValueError("Seeds from an endless source need num_streamlines, or tracking "
                             "would never end.")
        if num_streamlines is not None and num_streamlines <= 0:
            raise ValueError(f"Cannot seed {num_streamlines} streamlines.")
        if mask is not None and mask.data.ndim != 3:
            raise ValueError(f"Expected 3D mask, got shape {mask.data.shape}")
        if mask is not None:
            mask = mask.with_data(mask.as_float(dtype=np.float32))
        if seed is not None and not 0 <= seed < 2**64:
            raise ValueError(f"The run seed must be in [0, 2^64), got {seed}.")
        if output is not None and Path(output).suffix.lower() != ".tck":
            raise ValueError(f"Cannot write tractograms as {Path(output).suffix}; use .tck")
        threads = n_threads(n_jobs)
        _check_alignment(fod, seeds, mask)
        run_seed = int(seed) if seed is not None else secrets.randbits(64)
        prep = self._prepare(fod, basis, mask)
        counts, tractogram = self._run_batches(prep, fod, seeds, num_streamlines, run_seed,
                                               output, output_seeds)
        parameters = {
            "algorithm": self.algorithm.name,
            "step_size": prep.step_size,
            "max_angle": prep.max_angle,
            "min_radius": prep.min_radius,
            "cutoff": prep.cutoff,
            "cutoff_mode": "auto" if self.cutoff == "auto" else "given",
            "min_length": prep.min_length,
            "max_length": prep.max_length,
            "bidirectional": self.bidirectional,
            "max_steps": prep.max_steps,
            "batch_size": prep.batch_size,
            "seed": run_seed,
            "num_threads": threads,
            "num_streamlines_target": num_streamlines,
            "basis": basis,
            "mask": mask is not None,
            "larmor_version": VERSION,
        }
        return TractographyRun(
            path=None if output is None else Path(output),
            seeds_path=None if output_seeds is None else Path(output_seeds),
            parameters=parameters, num_seeds=counts.seeds, num_streamlines=counts.accepted,
            num_rejected={"no_start": counts.no_start, "too_short": counts.too_short},
            sampling=counts.sampling(), tractogram=tractogram)

    def _prepare(self, fod: Volume, basis: str, mask: Volume | None) -> _Prepared:
        lmax, full = sh_layout(fod.data.shape[-1])
        if lmax > _core.MAX_LMAX:
            raise ValueError(f"The FOD's lmax must be at most {_core.MAX_LMAX}, got {lmax}.")
        coeffs = np.ascontiguousarray(
            convert_sh_basis(fod.as_float(dtype=np.float32), basis, "tournier"), dtype=np.float32)
        voxel = float(np.mean(fod.voxel_sizes))
        step_size = self.step_size if self.step_size is not None else 0.1 * voxel
        min_length = self.min_length if self.min_length is not None else 5 * voxel
        max_length = self.max_length if self.max_length is not None else 100 * voxel
        if min_length >= max_length:
            raise ValueError(f"Invalid params: Got min_length {min_length} and max_length "
                             f"{max_length}.")
        max_angle = self._max_angle(step_size, voxel)
        min_radius = step_size / (2 * sin(radians(max_angle) / 2))
        cutoff = (auto_cutoff(coeffs, mask=None if mask is None else _mask_on_grid(mask, fod))
                  if self.cutoff == "auto" else float(self.cutoff))
        # Rounded first, so that rounding errors do not add a step.
        max_steps = ceil(round(max_length / step_size, 9))
        mask_arrays = None if mask is None else (
            np.ascontiguousarray(mask.data, dtype=np.float32),
            np.linalg.inv(mask.affine).astype(np.float64))
        row = max_steps + 1
        return _Prepared(
            coeffs=coeffs, full=full, fod_inv=np.linalg.inv(fod.affine).astype(np.float64),
            mask=mask_arrays, step_size=float(step_size), min_length=float(min_length),
            max_length=float(max_length), max_angle=float(max_angle),
            min_radius=float(min_radius), cutoff=float(cutoff),
            params=self.algorithm.make_params(cutoff, max_angle), max_steps=max_steps,
            batch_size=max(1, min(MAX_BATCH, BUFFER_BYTES // (row * 3 * 4))))

    def _max_angle(self, step_size: float, voxel: float) -> float:
        if self.max_angle is not None:
            return self.max_angle
        if self.algorithm.default_max_angle is not None:
            return self.algorithm.default_max_angle
        ratio = step_size / (2 * self.algorithm.default_min_radius * voxel)
        return 180.0 if ratio >= 1 else degrees(2 * asin(ratio))

    def _run_batches(self, prep: _Prepared, fod: Volume, seeds, num_streamlines, run_seed: int,
                     output, output_seeds):
        import torch

        row = prep.max_steps + 1
        # Resident for the whole run: the FOD, the mask and one batch of points.
        d_fod = torch.from_numpy(prep.coeffs).to("cuda")
        d_mask = None if prep.mask is None else (torch.from_numpy(prep.mask[0]).to("cuda"),
                                                 prep.mask[1])
        d_pos = torch.empty((prep.batch_size, row, 3), dtype=torch.float32, device="cuda")
        totals = _Counts()
        gathered = _Gathered()
        batch_index = 0
        batch = _next_batch(seeds, totals, _draw_size(prep, totals, num_streamlines), run_seed, 0)
        writer = (TckWriter(output, _tck_fields(prep)) if output is not None
                  else contextlib.nullcontext())
        seeds_file = (open(output_seeds, "w", encoding="utf-8") if output_seeds is not None
                      else contextlib.nullcontext())
        with writer as tck_out, seeds_file as seeds_out:
            while len(batch) > 0:
                d_seeds = torch.from_numpy(batch.positions).to("cuda")
                d_dirs = (None if batch.directions is None
                          else torch.from_numpy(batch.directions).to("cuda"))
                d_idx = torch.from_numpy(batch.indices).to("cuda")
                counts_d, status_d, followed_d, stats = _core.track_batch(
                    d_fod, prep.fod_inv, d_mask, d_seeds, d_dirs, d_idx, run_seed,
                    prep.step_size, prep.max_steps, self.bidirectional, prep.params, d_pos, 1,
                    full=prep.full)
                counts = counts_d.cpu().numpy()
                status = status_d.cpu().numpy()
                totals.add_sampling(stats)
                remaining = None if num_streamlines is None else num_streamlines - totals.accepted
                keep = _keep(counts, prep, remaining, totals)
                points = _kept_points(d_pos, counts_d, keep, torch)
                piece = Tractogram(points, np.concatenate([[0], np.cumsum(counts[keep])]),
                                   streamline_data={"end_status": status[keep].copy()},
                                   reference=fod.grid)
                if tck_out is not None:
                    tck_out.write(piece)
                else:
                    gathered.add(piece)
                if seeds_out is not None:
                    followed = followed_d.cpu().numpy()
                    write_seeds(Seeds(batch.positions[keep], followed[keep], batch.indices[keep]),
                                seeds_out, header=batch_index == 0)
                if _should_stop(totals, num_streamlines):
                    break
                batch_index += 1
                batch = _next_batch(seeds, totals, _draw_size(prep, totals, num_streamlines),
                                    run_seed, batch_index)
        return totals, (None if output is not None else gathered.tractogram(fod))


def _tck_fields(prep: _Prepared) -> dict:
    return {"step_size": prep.step_size, "max_angle": prep.max_angle, "threshold": prep.cutoff,
            "min_dist": prep.min_length, "max_dist": prep.max_length}


def _kept_points(d_pos, counts_d, keep: np.ndarray, torch) -> np.ndarray:
    """The points of the kept rows, packed in row order on the device, then copied back."""
    if len(keep) == 0:
        return np.empty((0, 3), np.float32)
    rows = torch.from_numpy(keep.astype(np.int64)).to("cuda")
    lengths = counts_d[rows]
    written = torch.arange(d_pos.shape[1], device="cuda")[None, :] < lengths[:, None]
    return d_pos[rows][written].cpu().numpy()


def _draw_size(prep: _Prepared, totals: _Counts, num_streamlines) -> int:
    if num_streamlines is None:
        return prep.batch_size
    left = SEEDS_PER_STREAMLINE * num_streamlines - totals.seeds
    missing = num_streamlines - totals.accepted
    wanted = max(MIN_DRAW, DRAW_PER_MISSING * missing)
    return max(0, min(prep.batch_size, left, wanted))


def _next_batch(seeds, totals: _Counts, size: int, run_seed: int, batch_index: int) -> Seeds:
    """The next `size` seeds, each with its global index: fixed seeds sliced in order, an
    endless source drawn with a generator of (run seed, batch number)."""
    done = totals.seeds
    if isinstance(seeds, Seeds):
        part = slice(done, done + size)
        batch = Seeds(seeds.positions[part],
                      None if seeds.directions is None else seeds.directions[part],
                      None if seeds.indices is None else seeds.indices[part])
    else:
        batch = seeds.draw(size, np.random.default_rng([run_seed, batch_index]))
    if batch.indices is None:
        batch = Seeds(batch.positions, batch.directions,
                      np.arange(done, done + len(batch), dtype=np.int64))
    return batch


def _keep(counts: np.ndarray, prep: _Prepared, remaining, totals: _Counts) -> np.ndarray:
    """The rows to keep (started and at least the minimum length), in seed order, at most
    `remaining`; the seeds after the last kept one then count as never drawn."""
    n = len(counts)
    started = counts > 0
    short = started & ((counts - 1) * prep.step_size < prep.min_length)
    keep = np.flatnonzero(started & ~short)
    if remaining is not None and len(keep) > remaining:
        keep = keep[:remaining]
        n = int(keep[-1]) + 1 if len(keep) else 0
        started, short = started[:n], short[:n]
    totals.seeds += n
    totals.accepted += len(keep)
    totals.no_start += int((~started).sum())
    totals.too_short += int(short.sum())
    return keep


def _should_stop(totals: _Counts, num_streamlines) -> bool:
    if num_streamlines is None:
        return False
    return (totals.accepted >= num_streamlines
            or totals.seeds >= SEEDS_PER_STREAMLINE * num_streamlines)


class _Gathered:
    """The kept streamlines of a run kept in memory, batch by batch."""

    def __init__(self) -> None:
        self.pieces: list[Tractogram] = []

    def add(self, piece: Tractogram) -> None:
        self.pieces.append(piece)

    def tractogram(self, fod: Volume) -> Tractogram:
        if not self.pieces:
            return Tractogram(np.empty((0, 3), np.float32), [0],
                              streamline_data={"end_status": np.empty((0, 2), np.uint8)},
                              reference=fod.grid)
        lengths = np.concatenate([np.diff(p.offsets) for p in self.pieces])
        return Tractogram(np.concatenate([p.points for p in self.pieces]),
                          np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64),
                          streamline_data={"end_status": np.concatenate(
                              [p.streamline_data["end_status"] for p in self.pieces])},
                          reference=fod.grid)


def _seed_points(seeds):
    if isinstance(seeds, Seeds):
        return seeds.positions
    if isinstance(seeds, MaskSeeds):
        affine = np.asarray(seeds.affine, dtype=np.float64)
        return seeds.voxels.astype(np.float64) @ affine[:3, :3].T + affine[:3, 3]
    return None


def _inside_mask(points: np.ndarray, mask: Volume) -> np.ndarray:
    """Whether each world point's nearest mask voxel (halves away from zero) is >= 0.5."""
    inv = np.linalg.inv(mask.affine)
    coords = points @ inv[:3, :3].T + inv[:3, 3]
    shape = np.asarray(mask.data.shape)
    on = np.all((coords > -0.5) & (coords < shape - 0.5), axis=1)
    lower = np.floor(coords[on])
    vox = (lower + (coords[on] - lower >= 0.5)).astype(np.int64)
    inside = np.zeros(len(points), dtype=bool)
    inside[on] = mask.data[vox[:, 0], vox[:, 1], vox[:, 2]] >= 0.5
    return inside


def _mask_on_grid(mask: Volume, fod: Volume) -> np.ndarray:
    centres = np.indices(fod.data.shape[:3]).reshape(3, -1).T.astype(np.float64)
    world = centres @ fod.affine[:3, :3].T + fod.affine[:3, 3]
    return _inside_mask(world, mask).reshape(fod.data.shape[:3])


def _check_alignment(fod: Volume, seeds, mask: Volume | None) -> None:
    points = _seed_points(seeds)
    if points is None or len(points) == 0:
        return
    inv = np.linalg.inv(fod.affine)
    vox = points @ inv[:3, :3].T + inv[:3, 3]
    on = np.all((vox >= -0.5) & (vox <= np.asarray(fod.data.shape[:3]) - 0.5), axis=1)
    if not on.any():
        raise ValueError("No seed lies on the FOD's grid; are the seeds in the FOD's space?")
    if mask is not None and not _inside_mask(points, mask).any():
        raise ValueError("No seed lies in the mask")
