"""Build libgdiamond_hip.so in-tree with hipcc for gfx950.

Every source is compiled with -ffp-contract=off (each float operation rounds
once, as the reference's --fmad=false build, CMakeLists.txt:25) and without
packed-FP32 VALU ops. Objects go to gdiamond_amd/_build/, the library to
gdiamond_amd/libgdiamond_hip.so.

Usage: python -m gdiamond_amd.build [--incremental]   (or gdiamond_amd.build.build())
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgdiamond_hip.so")
ARCH = os.environ.get("GD_OFFLOAD_ARCH", "gfx950")

SOURCES = ["gd_capi.hip", "gd_batch.hip", "gd_plain.hip", "gd_zstream.hip", "gd_learned.hip",
           "gd_noalloc.hip", "gd_fused.hip", "gd_host.cpp"]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Wall", "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"),
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found")
    return p


def _stale(obj: str, path: str) -> bool:
    """True when obj is missing or older than its source or a shared header."""
    if not os.path.exists(obj):
        return True
    deps = [path] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".hpp")]
    deps.append(os.path.join(ROOT, "include", "gdiamond", "gd_capi.h"))
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src: str, incremental: bool = False) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if incremental and not _stale(obj, path):
        return obj
    if src.endswith(".hip"):
        cmd = [hipcc(), *COMMON, *DEVICE, "-x", "hip", "-c", path, "-o", obj]
    else:
        # host-only C++: no device code
        cmd = [hipcc(), *COMMON, "-x", "c++", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        filtered = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln
                    and "not a recognized feature" not in ln]
        if any(ln.strip() for ln in filtered):
            sys.stderr.write("\n".join(filtered) + "\n")
    return obj


def build(verbose: bool = True, incremental: bool = False) -> str:
    """Compile every source (only the changed ones with incremental=True) and
    link the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(lambda src: _compile(src, incremental), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
