"""Builds libplume_hip.so, the gfx950 kernels behind the C ABI of include/plume_hip.h, with hipcc.

Each .hip source compiles to an object in `_build/` (in parallel), then one shared library lands next to this
file, where `plume_amd._lib` loads it. Flags: gfx950 only; -ffp-contract=off so every float operation rounds
as the reference's float32 CPU path does; packed-FP32 instructions disabled.
"""

import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCES = sorted((HERE / "csrc").glob("*.hip"))
HEADERS = sorted((HERE / "csrc").glob("*.h")) + [HERE.parent / "include" / "plume_hip.h"]
BUILD = HERE.parent / "_build"
TARGET = HERE / "libplume_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
         "-Wno-unused-command-line-argument"]


def _compile(source: Path, verbose: bool) -> Path:
    obj = BUILD / (source.stem + ".o")
    command = [HIPCC, *FLAGS, "-c", str(source), "-o", str(obj)]
    if verbose:
        print(" ".join(command), flush=True)
    result = subprocess.run(command, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"hipcc failed on {source.name}:\n{result.stdout}\n{result.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as pool:
        objects = list(pool.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = TARGET.with_suffix(".so.tmp")
    command = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objects), "-o", str(tmp)]
    if verbose:
        print(" ".join(command), flush=True)
    subprocess.run(command, check=True)
    tmp.replace(TARGET)
    return TARGET


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
