"""`lar track`: tracks streamlines through an FOD image on the GPU (the command of the reference,
src/larmor/cli/track.py, with its seed sources, options and refusals); writes TCK."""

import dataclasses
import os
import secrets
from enum import Enum
from pathlib import Path
from typing import Optional

import typer

from larmor_amd.cli.fod import SHBasis
from larmor_amd.cli.outputs import check_outputs


class TrackingAlgorithm(str, Enum):
    det = "det"
    prob = "prob"


def _cutoff(value: str):
    return "auto" if value == "auto" else float(value)


def track(
    fod_path: Path = typer.Argument(..., help="FOD image (SH coefficients)."),
    output_path: Path = typer.Argument(..., help="Output tractogram (.tck)."),
    algorithm: TrackingAlgorithm = typer.Option(TrackingAlgorithm.det, "--algorithm"),
    seed_image: Optional[Path] = typer.Option(None, "--seed-image"),
    seed_random_per_voxel: Optional[tuple[Path, int]] = typer.Option(
        None, "--seed-random-per-voxel"),
    seed_file: Optional[Path] = typer.Option(None, "--seed-file"),
    select: Optional[int] = typer.Option(None, "--select"),
    mask_path: Optional[Path] = typer.Option(None, "--mask"),
    unidirectional: bool = typer.Option(False, "--unidirectional"),
    step_size: Optional[float] = typer.Option(None, "--step-size"),
    angle: Optional[float] = typer.Option(None, "--angle"),
    min_length: Optional[float] = typer.Option(None, "--min-length"),
    max_length: Optional[float] = typer.Option(None, "--max-length"),
    cutoff: str = typer.Option("auto", "--cutoff"),
    basis: SHBasis = typer.Option(SHBasis.tournier, "--basis"),
    run_seed: Optional[int] = typer.Option(None, "--seed"),
    json_path: Optional[Path] = typer.Option(None, "--json"),
    output_seeds: Optional[Path] = typer.Option(None, "--output-seeds"),
    force: bool = typer.Option(False, "--force", "-f"),
    jobs: int = typer.Option(-1, "--nt"),
) -> None:
    """Track streamlines through an FOD image."""
    from larmor_amd.io import load_volume
    from larmor_amd.tractography import MaskSeeds, Tracker, read_seeds, seeds_per_voxel

    if output_path.suffix != ".tck":
        raise typer.BadParameter("must be a .tck file", param_hint="OUTPUT")
    if select is not None and select < 1:
        raise typer.BadParameter("must be at least 1", param_hint="--select")
    n_cores = os.cpu_count() or 1
    if jobs == 0 or jobs < -1 or jobs > n_cores:
        raise typer.BadParameter(f"must be -1 (all) or between 1 and {n_cores}",
                                 param_hint="--nt")
    if run_seed is not None and not 0 <= run_seed < 2**64:
        raise typer.BadParameter("must be in [0, 2^64)", param_hint="--seed")
    sources = [seed_image, seed_random_per_voxel, seed_file]
    if sum(s is not None for s in sources) != 1:
        raise typer.BadParameter("give exactly one of --seed-image, --seed-random-per-voxel and "
                                 "--seed-file", param_hint="seeds")
    if seed_random_per_voxel is not None and seed_random_per_voxel[1] < 1:
        raise typer.BadParameter("must be at least 1 seed per voxel",
                                 param_hint="--seed-random-per-voxel")
    if seed_image is not None and select is None:
        raise typer.BadParameter("--seed-image draws seeds without end, so it needs --select",
                                 param_hint="--select")
    tracker = Tracker(algorithm.value, step_size=step_size, max_angle=angle,
                      cutoff=_cutoff(cutoff), min_length=min_length, max_length=max_length,
                      bidirectional=not unidirectional)
    check_outputs([output_path, *[p for p in (json_path, output_seeds) if p is not None]], force)
    seed = run_seed if run_seed is not None else secrets.randbits(64)
    if seed_image is not None:
        seeds = MaskSeeds(load_volume(seed_image))
    elif seed_random_per_voxel is not None:
        seeds = seeds_per_voxel(load_volume(seed_random_per_voxel[0]), seed_random_per_voxel[1],
                                seed)
    else:
        seeds = read_seeds(seed_file)
    fod = load_volume(fod_path)
    mask = None if mask_path is None else load_volume(mask_path)
    run = tracker.track(fod, output_path, seeds, num_streamlines=select, mask=mask,
                        basis=basis.value, seed=seed, output_seeds=output_seeds, n_jobs=jobs)
    inputs = {"fod": str(fod_path), "mask": None if mask_path is None else str(mask_path),
              "seed_image": None if seed_image is None else str(seed_image),
              "seed_file": None if seed_file is None else str(seed_file)}
    run = dataclasses.replace(run, inputs=inputs)
    if json_path is not None:
        run.write_json(json_path)
    typer.echo(f"{run.num_streamlines} streamlines from {run.num_seeds} seeds -> {output_path}")
