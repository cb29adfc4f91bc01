"""Build libpfem.so (HIP, gfx950) in-tree: plasticity_amd/libpfem.so.

Each translation unit under csrc/ is compiled with hipcc for gfx950 only, then
linked into one shared library whose exported symbols are the C-ABI of
include/pfem.h.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpfem.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["element.hip", "sparse.hip", "amg.hip", "unet.hip", "build.hip", "peaks.hip", "batch_noalloc.hip", "context.cpp", "driver.cpp", "amg.cpp", "unet.cpp"]
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # never let the compiler form packed FP32 ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _compile(src: str, build_dir: str = BUILD, extra=()) -> str:
    obj = os.path.join(build_dir, src.replace(".", "_") + ".o")
    lang = ["-x", "hip"]
    cmd = [HIPCC, *lang, *FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        msgs = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if msgs:
            sys.stderr.write("\n".join(msgs) + "\n")
    return obj


def build(verbose: bool = False, out: str = OUT, extra=()) -> str:
    """Compile every source and link `out`; extra: additional hipcc flags (objects go to a
    separate build folder)."""
    build_dir = BUILD if not extra else os.path.join(BUILD, "check")
    os.makedirs(build_dir, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda src: _compile(src, build_dir, extra), SOURCES))
    tmp = out + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
