"""Build step for libuncertaintea_hip (gfx950) and the CPU oracle.

`build_library()` compiles every .hip source in `csrc/` with hipcc for gfx950 into
`uncertaintea_amd/libuncertaintea_hip.so`.  Floating-point contraction is disabled
so that kernel arithmetic follows the source order exactly (the CPU restatement
reproduces it bit for bit), and packed-FP32 instructions are turned off.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = CSRC / "_build"
LIB = PKG / "libuncertaintea_hip.so"
REPO = PKG.parent
ORACLE_DIR = REPO / "oracle"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    return "hipcc"


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        msg = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{msg}")


def build_library(verbose: bool = True) -> Path:
    """Compile csrc/*.hip for gfx950 and link libuncertaintea_hip.so."""
    BUILD.mkdir(parents=True, exist_ok=True)
    # largest translation units first, so the longest compiles start at once
    sources = sorted(CSRC.glob("*.hip"), key=lambda p: -p.stat().st_size)
    hipcc = _hipcc()
    objs = []
    cmds = []
    for src in sources:
        obj = BUILD / (src.stem + ".o")
        objs.append(obj)
        cmds.append([hipcc, *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=min(len(cmds), max(1, min(8, os.cpu_count() or 1)))) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-o", str(tmp), *map(str, objs)], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = True) -> Path:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    _run(["make", "-C", str(ORACLE_DIR), "-B"], verbose)
    return ORACLE_DIR / "_build" / "liboracle.so"


if __name__ == "__main__":
    build_library(verbose=True)
    build_oracle(verbose=True)
    sys.exit(0)
