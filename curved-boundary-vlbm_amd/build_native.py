"""Builds libvlbm.so (HIP, gfx950) in-tree with hipcc.

Usage: python curved-boundary-vlbm_amd/build_native.py
The library is written to curved_lbm/_lib/libvlbm.so next to the package
that loads it.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT_DIR = HERE / "curved_lbm" / "_lib"
INCLUDE = HERE.parent / "include"
SOURCES = ["vlbm3d.hip", "vlbm2d.hip", "geom_host.hip", "nn.hip"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def build(verbose: bool = True) -> Path:
    OUT_DIR.mkdir(parents=True, exist_ok=True)
    out = OUT_DIR / "libvlbm.so"
    tmp = OUT_DIR / "libvlbm.so.tmp"
    cmd = [
        hipcc(),
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics",
        f"-I{INCLUDE}",
        "-o", str(tmp),
    ] + [str(CSRC / s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
