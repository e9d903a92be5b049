"""Builds libnaer2_hip.so in-tree with hipcc for gfx950.

    python -m naer2_amd.build

-ffp-contract=off keeps every fp32 multiply and add separately rounded (the summation and update order
is part of the results contract); -packed-fp32-ops is disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SRCS = [PKG / "csrc" / "naer2_hip.hip", PKG / "csrc" / "rlif.hip", PKG / "csrc" / "adjoint.hip", PKG / "csrc" / "event.hip"]
OUT = PKG / "libnaer2_hip.so"
ARCH = os.environ.get("NAER2_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]
OBJ_DIR = PKG / "_build"


def _run(cmd: list[str], verbose: bool) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd)}")
    if verbose:
        sys.stderr.write(res.stderr)


def build(verbose: bool = False) -> Path:
    """Compile every source to an object (in parallel), then link the shared library; always rebuilds."""
    from concurrent.futures import ThreadPoolExecutor

    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    OBJ_DIR.mkdir(exist_ok=True)
    objs = [OBJ_DIR / (s.stem + ".o") for s in SRCS]
    with ThreadPoolExecutor(max_workers=len(SRCS)) as pool:
        for f in [pool.submit(_run, [hipcc, *FLAGS, "-c", str(s), "-o", str(o)], verbose) for s, o in zip(SRCS, objs)]:
            f.result()
    tmp = OUT.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *[str(o) for o in objs], "-o", str(tmp)], verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
