"""Build libpanoptic_hip.so (gfx950) from csrc/*.hip with hipcc.

    python panoptic-evaluator_amd/build.py [--jobs N]

Objects go to panoptic-evaluator_amd/_build/; the shared library is written
next to the Python package (panoptic_evaluator/libpanoptic_hip.so) so that it
travels with the source tree.  Always recompiles every translation unit.
"""

from __future__ import annotations

import argparse
import concurrent.futures
import os
import pathlib
import shutil
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OBJ = HERE / "_build"
LIB = HERE / "panoptic_evaluator" / "libpanoptic_hip.so"
ARCH = os.environ.get("PANOPTIC_HIP_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # Ratios and sums must round exactly like the CPU reference path.
    "-ffp-contract=off",
    # Keep the compiler from forming packed FP32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    f"-I{HERE.parent / 'include'}",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not pathlib.Path(path).exists():
        raise RuntimeError("hipcc not found; a ROCm toolchain is required to build the HIP kernels")
    return path


def compile_one(source: pathlib.Path) -> pathlib.Path:
    target = OBJ / (source.stem + ".o")
    cmd = [hipcc(), *FLAGS, "-c", str(source), "-o", str(target)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode:
        raise RuntimeError(f"hipcc failed for {source.name}:\n{result.stderr[-6000:]}")
    return target


def build(jobs: int | None = None, verbose: bool = False) -> pathlib.Path:
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    jobs = jobs or min(len(sources), os.cpu_count() or 4, 16)
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        objects = list(pool.map(compile_one, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objects), "-o", str(tmp)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode:
        raise RuntimeError(f"link failed:\n{result.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources")
    return LIB


def main(argv=None) -> int:
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--jobs", type=int, default=None)
    args = parser.parse_args(argv)
    build(args.jobs, verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
