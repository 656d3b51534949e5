"""Build of the HIP engine `typecv_amd/libtypecv_ba.so` for gfx950.

The library is built in-tree so that it travels with the repository snapshot; `hipcc`
cross-compiles without a GPU. Packed FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
LIBRARY = PACKAGE / "libtypecv_ba.so"
SOURCES = [CSRC / "ba_engine.hip"]
DEPENDS = [CSRC / "ba_kernels.hip", CSRC / "ba_factored.hip", CSRC / "ba_onepass.hip",
           CSRC / "ba_dense.hip",
           CSRC / "obs_eval.hip", CSRC / "ba_kernels.h", CSRC / "ba_math.h",
           ROOT / "include" / "typecv_ba.h"]


def hipcc() -> str:
    for candidate in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if candidate and Path(candidate).exists():
            return candidate
    raise RuntimeError("hipcc not found: the HIP engine cannot be built")


def compile_command(output: Path = LIBRARY, extra: list[str] | None = None) -> list[str]:
    return [
        hipcc(),
        "--offload-arch=gfx950",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        # multiply-adds are fused within an expression only, never across statements: a value's
        # bits then depend on its source expression alone, not on what else a kernel computes, so
        # a cross block relinearised in one kernel equals the one another kernel stored
        "-ffp-contract=on",
        "-Wall",
        "-Wno-unused-function",
        "-I", str(ROOT / "include"),
        *(extra or []),
        *[str(s) for s in SOURCES],
        "-o", str(output),
    ]


def build(verbose: bool = False) -> Path:
    """Compile the engine library (always; the call is the build check)."""
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = compile_command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + result.stdout + result.stderr)
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
