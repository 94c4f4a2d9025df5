"""Build ``libxtbhip.so`` for gfx950 with hipcc (in-tree, next to this file).

    python -m dxtb_amd.build
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SOURCES = ["integrals.hip", "scf.hip", "singlepoint.hip", "large.hip", "forces.hip", "pgrad.hip"]
OUT = PKG / "libxtbhip.so"

FLAGS = [
    "-O3", "-fPIC", "-shared", "-std=c++17", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # mistakes that compile but would misbehave on the device
    "-Werror=infinite-recursion", "-Werror=return-type", "-Werror=uninitialized",
    "-Werror=sometimes-uninitialized",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> Path:
    src = [str(PKG / "csrc" / s) for s in SOURCES]
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), *FLAGS, f"-I{ROOT / 'include'}", *src, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed")
    # show compiler warnings (the host pass always notes the device-only
    # target feature; that one is expected)
    warn = [ln for ln in (res.stdout + res.stderr).splitlines()
            if "warning" in ln and "packed-fp32-ops" not in ln and "hip-link" not in ln]
    if warn:
        sys.stderr.write("\n".join(warn) + "\n")
    tmp.replace(OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
