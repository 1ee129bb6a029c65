"""Build of libnq.so, the gfx950 HIP library behind the ctypes boundary.

The library is built in-tree (next to this file) so that it travels with the
repository; ``python -m nqueens.build`` or ``__graft_entry__.build()`` runs it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG.parent / "csrc"
ROOT = PKG.parent.parent
LIB = PKG / "libnq.so"
SOURCES = ["nq_kernels.hip", "nq_mlp.hip", "nq_noalloc.hip", "nq_probe.hip", "nq_host.cpp"]
HEADERS = ["nq_task.h", "nq_dfs.h", "nq_kernels.h", "nq_ctx.h"]
ARCH = os.environ.get("NQ_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libnq)")


def command(out: Path = LIB) -> list[str]:
    # Packed FP32 instructions are disabled for all device code.
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", str(ROOT / "include"), "-o", str(out)] + [str(CSRC / s) for s in SOURCES]


def build(verbose: bool = False) -> Path:
    """Compile libnq.so for gfx950 (always recompiles)."""
    tmp = LIB.with_suffix(".so.tmp")
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"libnq build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
