"""Build libdfm.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m dfm.build   (from differentiable-flowmap_amd/), or call build().
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
INCLUDE = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include")
LIB = os.path.join(HERE, "libdfm.so")

SOURCES = ["dfm_flowmap.hip", "dfm_mgpcg.hip", "dfm_init.hip", "dfm_mask.hip", "dfm_force.hip", "dfm_poisson_tiled.hip", "dfm_fields.hip", "dfm_clamp.hip", "dfm_batch.hip", "dfm_errors.cpp"]

# -packed-fp32-ops: keep the compiler from emitting packed FP32 (v_pk_*_f32)
# instructions; ordinary float code is scheduled as scalar VALU ops instead.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: cannot build libdfm.so")


def build(verbose: bool = False) -> str:
    """Compile every source to an object, then link the shared library."""
    objdir = os.path.join(os.path.dirname(HERE), "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    objs = [os.path.join(objdir, src + ".o") for src in SOURCES]
    cmds = [[cc, *HIP_FLAGS, "-I", INCLUDE, "-c", os.path.join(CSRC, src), "-o", obj] for src, obj in zip(SOURCES, objs)]
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd), flush=True)
    # the translation units are independent: compile them side by side
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0")) or (os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(subprocess.run, cmd, check=True) for cmd in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
