"""Build of the Node-API addon `typecv_amd/node/typecv_ba.node` over the engine's C ABI.

Compiled by gcc against the Node-API headers of the installed Node (`/usr/include/node`), linked
to `typecv_amd/libtypecv_ba.so` with an rpath relative to the addon, so that `require()` of the
in-tree file loads the in-tree engine.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
PACKAGE = HERE.parent
ROOT = PACKAGE.parent
ADDON = HERE / "typecv_ba.node"
SOURCE = HERE / "typecv_ba_napi.c"


def node_include() -> Path | None:
    for candidate in ("/usr/include/node", "/usr/local/include/node"):
        if (Path(candidate) / "node_api.h").exists():
            return Path(candidate)
    return None


def build(verbose: bool = False) -> Path | None:
    """Compile the addon (always). Returns None when no Node-API headers are installed."""
    include = node_include()
    if include is None or shutil.which("gcc") is None:
        return None
    if not (PACKAGE / "libtypecv_ba.so").exists():
        raise RuntimeError("build the engine library before the addon")
    tmp = ADDON.with_suffix(".node.tmp")
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall", "-Wextra",
           "-Wno-unused-parameter", "-DNODE_GYP_MODULE_NAME=typecv_ba",
           "-I", str(include), "-I", str(ROOT / "include"), str(SOURCE),
           "-L", str(PACKAGE), "-ltypecv_ba", "-Wl,-rpath,$ORIGIN/..", "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError("gcc failed on the Node-API addon:\n" + result.stdout + result.stderr)
    os.replace(tmp, ADDON)
    return ADDON


if __name__ == "__main__":
    print(build(verbose=True))
