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
SOURCES = ["sd_core.hip"]


def build_library(fast: bool = False, verbose: bool = True) -> Path:
    """Compiles the HIP sources into liblarmor_sd.so; returns its path."""
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, *(["-DLSD_FAST_BUILD"] if fast else []), "-shared", "-o", str(tmp)]
    cmd += [str(CSRC / s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stderr)
        raise RuntimeError(f"hipcc failed ({proc.returncode})")
    os.replace(tmp, OUT)
    return OUT


def main() -> None:
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--fast", action="store_true")
    args = parser.parse_args()
    print(build_library(fast=args.fast))


if __name__ == "__main__":
    main()
