"""Builds the CPU oracle's shared libraries (FP32 and FP64) from `plume_oracle.c` with gcc.

Test infrastructure: the libraries are loaded only by the tests, the smoke check and the benchmark's CPU
baseline leg. No FMA contraction (`-ffp-contract=off`, baseline x86-64 ISA), so the FP32 build performs
the reference kernels' float operations one rounding at a time, in their order.
"""

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCE = HERE / "plume_oracle.c"
TARGETS = {"f32": HERE / "liborc_f32.so", "f64": HERE / "liborc_f64.so"}
FLAGS = ["-O3", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-Wall",
         "-Wno-unused-function"]


def build_native(target: Path, verbose: bool = False) -> Path:
    """The same source as a CPU timing baseline rather than a checker: FP32, `-O3 -march=native -fopenmp`
    with gcc's default contraction (FMA where the host has it), built for the machine it runs on. Its
    results are not bit-identical to the reference's operation order; only its speed is used."""
    tmp = Path(str(target) + ".tmp")
    command = ["gcc", "-O3", "-march=native", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-Wall",
               "-Wno-unused-function", str(SOURCE), "-o", str(tmp), "-lm"]
    if verbose:
        print(" ".join(command), flush=True)
    subprocess.run(command, check=True)
    tmp.replace(target)
    return target


def build(verbose: bool = False) -> dict[str, Path]:
    for precision, target in TARGETS.items():
        defines = ["-DORC_DOUBLE"] if precision == "f64" else []
        tmp = target.with_suffix(".so.tmp")
        command = ["gcc", *FLAGS, *defines, str(SOURCE), "-o", str(tmp), "-lm"]
        if verbose:
            print(" ".join(command), flush=True)
        subprocess.run(command, check=True)
        tmp.replace(target)
    return dict(TARGETS)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
