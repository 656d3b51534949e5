"""TEST INFRASTRUCTURE. Build recipe of the CPU oracle `oracle/libba_oracle.so` (gcc, OpenMP).

-ffp-contract=off keeps every product unfused, as the reference's JavaScript arithmetic is.
The reference itself (TypeScript on a WebGPU runtime that is not vendored) cannot be compiled
here, so there is no `oracle/_ref` build.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCE = HERE / "ba_oracle.c"
LIBRARY = HERE / "libba_oracle.so"


def command(output: Path = LIBRARY) -> list[str]:
    return ["gcc", "-O3", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
            "-D_POSIX_C_SOURCE=199309L", "-Wall", "-Wextra", "-Wno-unused-parameter",
            str(SOURCE), "-o", str(output), "-lm"]


def build(verbose: bool = False) -> Path:
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + result.stdout + result.stderr)
    if result.stderr.strip() and verbose:
        print(result.stderr)
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
