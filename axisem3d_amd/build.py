"""Build recipe of the HIP library ``axisem3d_amd/libaxs3d.so`` (gfx950).

Compiles every source under ``axisem3d_amd/csrc`` with hipcc, in-tree, with
packed-FP32 instructions disabled, and links one shared library that
exports the C ABI of ``include/axs3d.h``.
"""
from __future__ import annotations

import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libaxs3d.so")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-result",
]

SOURCES = ["axs3d_kernels.hip", "axs3d_cardinal.hip", "axs3d_fluid.hip",
           "axs3d_adjoint.hip", "axs3d_scratch.hip",
           "axs3d_domain.cpp",
           "axs3d_capi.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("hipcc failed: " + " ".join(cmd))
    return r.stdout


def build(verbose=False, jobs=None):
    """Compiles every source (in parallel, `jobs` hipcc processes at a
    time, default min(8, CPUs)) and links the library."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(BUILD, exist_ok=True)
    t0 = time.time()
    cmds, objs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", path, "-o", obj]
        cmds.append(cmd)
        objs.append(obj)
    if jobs is None:
        jobs = max(1, min(8, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for out in outs:
            if out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp]
         + objs)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f} s")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
