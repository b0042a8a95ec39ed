"""Builds liblarmor_sd.so in-tree with hipcc for gfx950.

    python -m larmor_amd.build [--fast]

--fast compiles only the instantiations of the 3-shell, 44/300-direction workload (development).
"""

import argparse
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent
CSRC = ROOT / "csrc"
OUT = ROOT / "liblarmor_sd.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3",
    "--offload-arch=gfx950",
    "-std=c++20",
    "-fPIC",
    # Same rounding as the CPU oracle: no fused multiply-add contraction.
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]
BUILD_DIR = ROOT.parent / "_build" / "larmor_amd"


def _torch_lib_dir() -> str:
    """Directory of the HIP runtime torch ships with (torch/lib)."""
    import importlib.util

    spec = importlib.util.find_spec("torch")
    if spec is None or spec.origin is None:
        raise RuntimeError("torch is needed to build: its HIP runtime is the one to link")
    return str(Path(spec.origin).parent / "lib")


def sources() -> list[Path]:
    """The API units (FP64 and the plain FP32 form) and one unit per (kernel family, column slots)
    instantiation set."""
    return [CSRC / "sd_api.hip", CSRC / "fp32_api.hip", CSRC / "track_api.hip",
            *sorted(CSRC.glob("inst_*.hip"))]


def _compile(src: Path, obj: Path, extra: list[str]) -> tuple[Path, int, str]:
    cmd = [HIPCC, *FLAGS, *extra, "-c", "-o", str(obj), str(src)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    return src, proc.returncode, proc.stderr


def build_library(fast: bool = False, verbose: bool = True, jobs: int | None = None,
                  force: bool = False) -> Path:
    """Compiles the HIP units in parallel and links liblarmor_sd.so; returns its path."""
    from concurrent.futures import ThreadPoolExecutor

    extra = ["-DLSD_FAST_BUILD"] if fast else []
    BUILD_DIR.mkdir(parents=True, exist_ok=True)
    units = sources()
    objs = [BUILD_DIR / (u.stem + (".fast" if fast else "") + ".o") for u in units]
    jobs = jobs or min(len(units), max(1, min(os.cpu_count() or 1, 8)))
    if verbose:
        print(f"hipcc {' '.join(FLAGS + extra)} ({len(units)} units, {jobs} jobs)", flush=True)
    # A unit is rebuilt when its object is older than any source or header (or force).
    newest = max(p.stat().st_mtime for p in [*CSRC.iterdir(), *(ROOT.parent / "include").iterdir(),
                                             Path(__file__)])
    todo = [(u, o) for u, o in zip(units, objs)
            if force or not o.exists() or o.stat().st_mtime < newest]
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        results = list(pool.map(lambda uo: _compile(uo[0], uo[1], extra), todo))
    failed = [(src, err) for src, rc, err in results if rc != 0]
    for src, err in failed:
        sys.stderr.write(f"--- {src.name}\n{err}")
    if failed:
        raise RuntimeError(f"hipcc failed on {len(failed)} unit(s)")
    tmp = OUT.with_suffix(".so.tmp")
    # One HIP runtime per process: link against the HIP runtime torch (the device-memory and
    # stream provider) ships.  It has the SONAME of /opt/rocm's, so once torch has loaded it the
    # loader binds this library to it (larmor_amd._lib imports torch first).
    torch_lib = _torch_lib_dir()
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-o", str(tmp), *map(str, objs),
            "-no-hip-rt", f"-L{torch_lib}", "-lamdhip64"]
    proc = subprocess.run(link, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stderr)
        raise RuntimeError(f"link failed ({proc.returncode})")
    os.replace(tmp, OUT)
    return OUT


def main() -> None:
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--fast", action="store_true")
    args = parser.parse_args()
    print(build_library(fast=args.fast))


if __name__ == "__main__":
    main()
