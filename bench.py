"""Benchmark of the stiffness path: GLL-point*mode updates per second.

One step = one Newmark time step of the whole domain (source, element
stiffness, assembly, boundaries, mass, Newmark, receivers), with every
operand already resident in HBM.  The headline value is the FP32-stiffness
+ FP64-accumulation path; the plain FP64 path runs on the same workload and
its throughput and the relative seismogram L2 between the two are reported
beside it.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("GLL-point*mode updates/s, FP32-batched vs FP64 plain at <=1e-4 "
          "seismogram L2")
UNIT = "GLL-point*mode updates/s"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="cartesian8000")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def build_workload(name, nsteps):
    from axisem3d_amd import preloop
    if name == "cartesian8000":
        spec = preloop.cartesian_homogeneous(nsteps=nsteps)
        cfg = {"workload": "cartesian box 80x80 km, 8000 elements, npol 4, "
               "nu=2 (Nr=5), homogeneous isotropic solid, axial moment "
               "tensor, 2 Hz Ricker, 10 surface receivers",
               "elements": spec.n_solid_elements,
               "points": spec.n_solid_points,
               "updates_per_step": spec.n_updates_per_step(),
               "dt": spec.dt}
        return spec, cfg
    raise SystemExit(f"unknown workload {name}")


def stiffness_bytes(spec, real_bytes):
    """Algorithmic HBM bytes of one stiffness launch: displacement read
    (FP64 complex x3), stiffness read+write (Real complex x3) per point
    mode, plus the per-element operands (10 GradQuad rows, lambda, mu,
    connectivity) and the element-mode work list."""
    nupd = spec.n_updates_per_step()
    E = spec.n_solid_elements
    n_em = int(np.sum(spec.elem_nr // 2 + 1))
    per_mode = 48 + 2 * 6 * real_bytes
    per_elem = (10 + 2) * 25 * real_bytes + 25 * 4 + 1
    return nupd * per_mode + E * per_elem + 8 * n_em


def point_bytes(spec, real_bytes):
    """Algorithmic HBM bytes of one fused point launch: stiffness read and
    zeroed (Real complex x3), u, v, a FP64 complex x3 read and written,
    the mode->point index, and the per-point Nr, offset, axis flag and
    inverse mass."""
    nupd = spec.n_updates_per_step()
    U = spec.n_solid_points
    return nupd * (2 * 6 * real_bytes + 6 * 48 + 4) + U * (4 + 4 + 1 + 8)


def time_gpu(spec, prec, steps, warmup):
    from axisem3d_amd import capi
    from axisem3d_amd.solver import GpuSolver
    g = GpuSolver(spec, prec, max_samples=steps + warmup)
    g.run(warmup)
    g.synchronize()
    g.profile(True)
    g.profile_reset()
    t0 = time.perf_counter()
    g.run(steps)
    g.synchronize()
    t1 = time.perf_counter()
    prof = {k: g.profile_read(k) for k in range(capi.NUM_KERNELS)}
    g.profile(False)
    seis = g.seismograms(steps + warmup)
    g.close()
    return (t1 - t0), prof, seis


def cpu_baseline(spec, seconds):
    """The FP64 oracle (a restatement of the reference loop, single core)
    timed on the host on the same workload for about `seconds`."""
    import oracle
    o = oracle.Oracle(spec)
    o.run(2, record=False)          # warm-up
    n, t = 0, 0.0
    while t < seconds:
        k = max(1, n)               # doubling batches
        t0 = time.perf_counter()
        o.run(k, record=False)
        t += time.perf_counter() - t0
        n += k
    nupd = spec.n_updates_per_step()
    return {"value": nupd * n / t, "unit": UNIT, "cores": 1, "kind": "port",
            "sample": f"{n} Newmark steps of the same workload (FP64 "
                      f"oracle, single thread), {t:.1f} s"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is "
              "not supported (every GPU box here has one GPU).",
              file=sys.stderr)
        sys.exit(2)
    from axisem3d_amd import capi
    nsteps = a.steps + a.warmup
    spec, cfg = build_workload(a.workload, nsteps)
    nupd = spec.n_updates_per_step()

    t32, prof32, seis32 = time_gpu(spec, "fp32", a.steps, a.warmup)
    t64, prof64, seis64 = time_gpu(spec, "fp64", a.steps, a.warmup)
    value = nupd * a.steps / t32
    value64 = nupd * a.steps / t64
    l2 = float(np.linalg.norm(seis32 - seis64) /
               max(np.linalg.norm(seis64), 1e-300))

    # dominant kernel of the headline path
    ms = {k: prof32[k][0] for k in prof32}
    dom = max(ms, key=lambda k: ms[k])
    names = {capi.KERNEL_STIFFNESS: "solid stiffness (Fourier path)",
             capi.KERNEL_POINT: "fused axial/mass/Newmark point update",
             capi.KERNEL_SOURCE: "moment source",
             capi.KERNEL_RECORD: "receivers"}
    byts = {capi.KERNEL_STIFFNESS: stiffness_bytes(spec, 4),
            capi.KERNEL_POINT: point_bytes(spec, 4)}
    avg_ms = prof32[dom][0] / max(prof32[dom][1], 1)
    achieved = byts.get(dom, 0) / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0
    kernels = {}
    for k in (capi.KERNEL_STIFFNESS, capi.KERNEL_POINT):
        avg = prof32[k][0] / max(prof32[k][1], 1)
        kernels[names[k]] = {
            "avg_us": avg * 1e3, "launches": prof32[k][1],
            "bytes_per_launch": byts[k],
            "achieved_GBs": byts[k] / (avg * 1e-3) / 1e9 if avg > 0 else 0}

    out = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": t32 / a.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic", "config": cfg,
        "fp64_plain": {"value": value64, "ms_per_step": t64 / a.steps * 1e3},
        "speedup_vs_fp64": value / value64,
        "seismogram_rel_l2_vs_fp64": l2,
        "roofline": {"bound": "hbm", "kernel": names[dom],
                     "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None},
        "kernels": kernels,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(spec, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
