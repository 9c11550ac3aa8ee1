"""Benchmark of the stiffness path: GLL-point*mode updates per second.

One step = one Newmark time step of the whole domain (moment source,
element stiffness with FFTs, 3D stress and attenuation, assembly, axis,
mass, Newmark, receivers) with every operand resident in HBM.  The
headline value is the FP32-stiffness + FP64-accumulation path; the plain
FP64 path runs the same workload and its throughput and the FP32-vs-FP64
relative seismogram L2 are reported beside it.

Default workload: the spherical regional run (1D 6-layer model, +-3% 3D
random field, per-point nu 10..400, 5-SLS CG4 attenuation, ~120k
elements).  `--workload cartesian8000` runs the 8000-element nu = 2
Cartesian case.

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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="regional",
                    choices=["regional", "cartesian8000", "batched",
                             "adjoint"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="both precisions run on, untimed, to this many "
                         "steps from rest; their seismograms give "
                         "seismogram_rel_l2_vs_fp64")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    metavar="MIB",
                    help="only run a plain device copy of MIB MiB three "
                         "times (reference reading for the FETCH_SIZE and "
                         "WRITE_SIZE counters) and exit")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain copy kernel of known size: the HBM counters of one launch
    should read `mib` MiB fetched and `mib` MiB written."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_MiB": mib, "launches": 3}),
          flush=True)


def build_workload(name, nsteps):
    from axisem3d_amd import preloop, synth
    if name == "batched":
        # one GPU's share (6 of 48) of the batched uneven instances:
        # elements ~ U[4k, 40k], nu ~ U[2, 256], seeded two-layer models
        rng = np.random.default_rng(2024)
        specs = [preloop.local_instance(1000 + i,
                                        int(rng.integers(4000, 40001)),
                                        int(rng.integers(2, 257)), nsteps,
                                        monitors=2)
                 for i in range(6)]
        spec = preloop.batch_instances(specs)
        spec._specs = specs
        cfg = {"workload": "batched uneven local instances, one GPU's share "
               "(6 of 48): elements ~ U[4k, 40k], nu ~ U[2, 256] (Nr bounded "
               "by the in-plane resolution), seeded isotropic two-layer "
               "models and moment tensors, per-instance dt, 4 surface "
               "receivers and 2 near-source monitors each",
               "instances": [{"elements": s.n_solid_elements,
                              "nr_max": int(s.elem_nr.max()),
                              "updates_per_step": s.n_updates_per_step(),
                              "dt": s.dt} for s in specs]}
    elif name == "cartesian8000":
        spec = preloop.cartesian_homogeneous(nsteps=nsteps)
        cfg = {"workload": "cartesian box 80x80 km, 8000 elements, npol 4, "
               "nu=2 (Nr=5), homogeneous isotropic solid, axial moment "
               "tensor, 2 Hz Ricker, 10 surface receivers"}
    else:
        spec = preloop.spherical_regional(nsteps=nsteps)
        synth.prepare_3d_streaming(spec)
        preloop.finish_regional(spec, nsteps)
        # four near-source monitors so the short bench window records
        # signal (the surface receivers are reached only after ~100 s)
        r0 = 6371e3 - 50e3
        th = np.array([5e3, 10e3, 20e3, 40e3]) / r0
        keep_e, keep_w, keep_p = (spec.rec_elem, spec.rec_weights,
                                  spec.rec_phi)
        preloop.add_receivers(spec, r0 * np.sin(th), r0 * np.cos(th),
                              np.array([0.3, 1.1, 2.0, 4.0]))
        spec.rec_elem = np.concatenate([keep_e, spec.rec_elem])
        spec.rec_weights = np.concatenate([keep_w, spec.rec_weights])
        spec.rec_phi = np.concatenate([keep_p, spec.rec_phi])
        nr = spec.elem_nr.astype(np.int64)
        cfg = {"workload": "spherical regional sector 30 deg x 1200 km, "
               f"{spec.n_solid_elements} elements, npol 4, 6-layer seeded "
               "1D model, +-3% 3D Gaussian random field (50 km), point nu "
               "10..400, 5-SLS CG4 attenuation, axial moment tensor, "
               "100 surface receivers + 4 near-source monitors",
               "element_nr_range": [int(nr.min()), int(nr.max())],
               "slices": int(nr.sum())}
    cfg.update({"elements": spec.n_solid_elements,
                "points": spec.n_solid_points,
                "updates_per_step": spec.n_updates_per_step(),
                "dt": spec.dt})
    return spec, cfg


def bytes_per_step(spec, real):
    """Algorithmic HBM bytes of one step, by kernel: every array touched
    once, intermediates (the cardinal scratch, the element-local forces'
    read by the point kernel aside) on chip.
    fr: Fourier-path elements, displacement read (FP64 complex x3) and
      element force write (Real complex x3) per point-mode, GradQuad
      operands (10 rows), connectivity, 1D material per point.
    cd_grad: displacement read per point-mode of 3D elements, the 5
      gradient rows, connectivity.
    cd_stress: per slice np x 25 Real material and, with attenuation, per
      attenuation point 2 Real moduli + nsls x 6 FP64 memory variables
      and 6 Real memory stress, the last two read and written.
    cd_quad: element force write per point-mode of 3D elements, the 5
      integral-factor rows.
    point: element forces read, u, v, a read+written (FP64), per-point
      data."""
    nupd = spec.n_updates_per_step()
    E, U = spec.n_solid_elements, spec.n_solid_points
    nr = spec.elem_nr.astype(np.int64)
    kind = getattr(spec, "elem_kind", None)
    is3d = np.zeros(E, bool) if kind is None else (np.asarray(kind) == 1)
    mtype = getattr(spec, "elem_mtype", None)
    npar = np.full(E, 2) if mtype is None else \
        np.array([2, 5, 21])[np.asarray(mtype)]
    # point-modes served by 3D elements, split by the elements' share of
    # each point (the 3D workloads are all-3D, so this is exact there)
    frac3d = float(is3d.mean()) if E else 0.0
    n3d = int(is3d.sum())
    nupd3 = int(round(nupd * frac3d))
    nupd1 = nupd - nupd3
    fr = nupd1 * (48 + 6 * real) + (E - n3d) * (10 * 25 * real + 25 * 4 + 1
                                               + 25 * real * 2)
    grad = nupd3 * 48 + n3d * (5 * 25 * real + 25 * 4 + 1)
    stress = int((nr[is3d] * npar[is3d]).sum()) * 25 * real
    att = getattr(spec, "att", None)
    if att is not None:
        npa = 4 if att["cg4"] else 25
        slices = int(nr[is3d].sum())
        # moduli and memory stress in the working precision, memory
        # variables FP64 (read and written)
        stress += slices * npa * (real * (2 + 2 * 6) +
                                  8 * 2 * 6 * att["nsls"])
    quad = nupd3 * 6 * real + n3d * 5 * 25 * real
    point = nupd * (6 * real + 6 * 48 + 4) + U * (4 + 4 + 4 + 1 + 8 + 16)
    return {"fr": int(fr), "cd_grad": int(grad), "cd_stress": int(stress),
            "cd_quad": int(quad), "point": int(point)}


def source_fingerprint():
    """sha256 (first 12 hex digits) of the native sources and the build
    recipe: what a counter summary in profiles/ was measured on.  The
    repository's history is not needed (a bare source tree has none)."""
    import hashlib
    h = hashlib.sha256()
    base = os.path.join(ROOT, "axisem3d_amd")
    files = [os.path.join("csrc", f) for f in
             sorted(os.listdir(os.path.join(base, "csrc")))
             if f.endswith((".hip", ".hpp", ".cpp", ".h"))] + ["build.py"]
    for f in files:
        h.update(f.encode())
        with open(os.path.join(base, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:12]


def _profile_row(table, kernel, workload, nfields):
    """The fields of `kernel` on `workload` in profiles/<table>, and the
    source fingerprint its header records; None for the fields when the
    table is absent, has no such row, or was measured on other sources
    (the counters then describe kernels that no longer exist)."""
    path = os.path.join(ROOT, "profiles", table)
    if not os.path.exists(path):
        return None, None
    src = None
    for line in open(path):
        f = line.split()
        if line.startswith("#"):
            if "source" in f and f.index("source") + 1 < len(f):
                src = f[f.index("source") + 1]
            continue
        if len(f) >= nfields and f[0] == kernel and f[1] == workload:
            if src != source_fingerprint():
                return None, src
            return f, src
    return None, src


def pmc_valu_busy(kernel, workload):
    """Fraction of the VALU issue slots `kernel` keeps busy on `workload`
    (SQ_INSTS_VALU over the CU cycles, from the committed rocprofv3
    counter summary), or None (no summary, or one of other sources)."""
    f, _ = _profile_row("valu_busy.txt", kernel, workload, 3)
    return float(f[2]) if f else None


def pmc_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by what the counter reads for a copy kernel of known size), or
    None when no summary of these sources covers it."""
    f, _ = _profile_row("hbm_traffic.txt", kernel, workload, 6)
    if not f:
        return None
    fetch_kib, fetch_fac = float(f[2]), float(f[3])
    write_kib, write_fac = float(f[4]), float(f[5])
    return (fetch_kib * fetch_fac + write_kib * write_fac) * 1024.0


def time_gpu(spec, prec, steps, warmup, graph=False, total=0):
    """Times `steps` steps launched kernel by kernel with HIP events around
    every kernel (the per-kernel averages come from these same steps);
    with graph, then the same number of steps as replays of the captured
    step graph, without events.  With total, the run then goes on
    (untimed) to `total` steps from rest, and the seismograms of all of
    them are returned (the accuracy leg)."""
    from axisem3d_amd import capi
    from axisem3d_amd.solver import GpuSolver
    t0 = time.perf_counter()
    g = GpuSolver(spec, prec,
                  max_samples=max(total, 2 * steps + warmup + 1))
    tune = g.autotune_info()
    log(f"{prec}: set-up {time.perf_counter() - t0:.1f} s, "
        f"{g.nmodes} point-modes, autotune {tune}")
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
    log(f"{prec}: {steps} steps in {t1 - t0:.3f} s")
    prof["autotune"] = tune
    if graph:
        g.run(1)    # captures the step graph outside the timed region
        g.synchronize()
        t2 = time.perf_counter()
        g.run(steps)
        g.synchronize()
        prof["graph_s"] = time.perf_counter() - t2
        log(f"{prec}: {steps} graph replays in {prof['graph_s']:.3f} s")
    if total > g.step:
        t3 = time.perf_counter()
        g.run(total - g.step)
        g.synchronize()
        log(f"{prec}: accuracy leg to step {total} in "
            f"{time.perf_counter() - t3:.1f} s")
    seis = g.seismograms(g.step)
    g.close()
    return (t1 - t0), prof, seis


def _fft_weight(nr):
    """Cost weight of an element of Nr slices on the CPU: its transforms
    grow as Nr log2 Nr."""
    nr = np.asarray(nr, dtype=np.float64)
    return nr * np.log2(np.maximum(nr, 2.0))


def host_threads():
    """Threads the host offers this process (OMP_NUM_THREADS when set)."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(spec, seconds, workload):
    """The FP64 oracle (a restatement of the reference loop) timed on the
    host with every host thread (OpenMP over the elements, the forces
    scattered in element order) and with one.  Cartesian: the whole
    workload.  Regional: a stratified sample of 24 blocks of 8 elements
    across the Nr range with their 3D model, scaled to the whole domain by
    sum(Nr log2 Nr), the growth of its transforms."""
    import oracle
    from axisem3d_amd import preloop, synth
    if workload == "cartesian8000":
        sub, scale = spec, 1.0
        what = "the whole workload"
    elif workload == "batched":
        sub = min(spec._specs, key=lambda s: s.n_updates_per_step())
        scale = spec.n_updates_per_step() / sub.n_updates_per_step()
        what = (f"the smallest instance ({sub.n_solid_elements} elements), "
                "scaled to the batch by point-modes")
    else:
        nx = spec._mesh.nx
        nz = spec._mesh.nz
        blocks = []
        for b in range(24):
            ix = min(nx - 8, int(b * (nx - 8) / 23))
            iz = (7 * b + 3) % nz
            blocks.append(iz * nx + ix)
        consts = preloop.att_constants(spec.att["w"], spec.att["y"],
                                       spec._att_in["f_min"],
                                       spec._att_in["f_max"],
                                       spec._att_in["f_ref"])
        wcg4 = preloop.cg4_weights(spec._gq["ifact"])
        elems, props, atts = [], [], []
        for e0 in blocks:
            p, a, _ = synth.generate(spec, e0, e0 + 8, spec._field,
                                     spec._att_in, consts, wcg4)
            elems.extend(range(e0, e0 + 8))
            props.append(p)
            atts.append(a)
        sub = preloop.subdomain(spec, elems, np.concatenate(props),
                                np.concatenate(atts))
        w = _fft_weight(spec.elem_nr)
        scale = float(w.sum()) / float(w[elems].sum())
        what = (f"{len(elems)} of {spec.n_solid_elements} elements (24 "
                "blocks of 8 across the Nr range) with their 3D model and "
                "attenuation, scaled to the whole domain by sum(Nr log2 Nr)")
    nupd = spec.n_updates_per_step()

    def timed(threads):
        oracle.set_threads(threads)
        o = oracle.Oracle(sub)
        rng = np.random.default_rng(0)
        o.displ[:] = rng.standard_normal(o.displ.shape) * 1e-9
        o.run(1, record=False)
        n, t = 0, 0.0
        while t < seconds:
            k = max(1, n)
            t0 = time.perf_counter()
            o.run(k, record=False)
            t += time.perf_counter() - t0
            n += k
        return nupd / (t / n * scale), n, t

    nth = host_threads()
    v_all, n_all, t_all = timed(nth)
    v_one, n_one, t_one = timed(1)
    oracle.set_threads(1)
    return {"value": v_all, "unit": UNIT, "cores": nth, "kind": "port",
            "sample": f"{n_all} Newmark steps of {what}; FP64 oracle, "
                      f"{nth} OpenMP threads, {t_all:.1f} s",
            "cores_1": {"value": v_one, "cores": 1,
                        "sample": f"{n_one} steps, one thread, "
                                  f"{t_one:.1f} s"}}


def accuracy_leg(spec, seis32, seis64):
    """The FP32+FP64 seismograms against the FP64 ones over the whole run
    from rest: relative L2 over all receivers, components and samples,
    and per receiver, with the step of the source-time function's peak
    (the wavelet must have passed the near receivers for the figure to
    mean anything)."""
    n = seis64.shape[0]
    per = np.linalg.norm(seis32 - seis64, axis=(0, 2)) / np.maximum(
        np.linalg.norm(seis64, axis=(0, 2)), 1e-300)
    amp = np.abs(seis64).max(axis=(0, 2))
    live = amp > 1e-6 * amp.max() if amp.max() > 0 else amp > 0
    stf = getattr(spec, "stf", None)
    out = {"steps": int(n),
           "stf_peak_step": (int(np.argmax(np.abs(stf[:n])))
                             if stf is not None else None),
           "rel_l2": float(np.linalg.norm(seis32 - seis64) /
                           max(np.linalg.norm(seis64), 1e-300)),
           "receivers_with_signal": int(live.sum()),
           "receivers": int(len(per)),
           "rel_l2_worst_receiver": float(per[live].max()) if live.any()
           else None}
    return out


def batched_extras(spec, seis32, seis64, steps, warmup, t_batch):
    """Per-instance seismogram L2 of FP32+FP64 against FP64, and the time
    of the same instances run one after the other, each alone (graph
    replays, against the batch's graph replays t_batch)."""
    from axisem3d_amd.solver import GpuSolver
    out = {"seismogram_rel_l2_vs_fp64": [], "alone_ms_per_step": []}
    for i, (_, _, (r0, r1)) in enumerate(spec._instances):
        if seis64 is not None:
            a, b = seis32[:, r0:r1], seis64[:, r0:r1]
            out["seismogram_rel_l2_vs_fp64"].append(
                float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)))
    total = 0.0
    for s in spec._specs:
        g = GpuSolver(s, "fp32", max_samples=steps + warmup)
        g.run(warmup)
        g.synchronize()
        t0 = time.perf_counter()
        g.run(steps)
        g.synchronize()
        dt = time.perf_counter() - t0
        g.close()
        out["alone_ms_per_step"].append(dt / steps * 1e3)
        total += dt
    out["alone_total_ms_per_step"] = total / steps * 1e3
    out["batched_speedup_vs_alone"] = total / t_batch
    log(f"instances alone: {total / steps * 1e3:.3f} ms per step")
    return out


def adjoint_bench(a):
    """The adjoint run on the 8000-element Cartesian workload (1D
    isotropic, Nr 5) with half the elements 3D isotropic: a forward run of
    `steps` steps from rest against axs3d_adjoint_run over the same steps
    (forward with checkpoints, recomputation, adjoint, sensitivities), FP32
    stiffness with FP64 accumulation."""
    from axisem3d_amd import adjoint, preloop
    from axisem3d_amd.solver import GpuSolver
    nsteps = a.steps
    spec = preloop.cartesian_homogeneous(nsteps=nsteps)
    E = spec.n_solid_elements
    preloop.make_3d_isotropic(spec, lambda x, y, z: 0.0 * x,
                              mask3d=(np.arange(E) % 2) == 0)
    nupd = spec.n_updates_per_step()
    model = list(adjoint.model_of(spec)) + list(adjoint.model3d_of(spec))
    sp = adjoint.with_model(spec, *model)
    g = GpuSolver(sp, "fp32", max_samples=nsteps)
    g.run(nsteps)          # warm-up of the forward (graph capture)
    g.synchronize()
    g.close()
    g = GpuSolver(sp, "fp32", max_samples=nsteps)
    t0 = time.perf_counter()
    g.run(nsteps)
    g.synchronize()
    t_fwd = time.perf_counter() - t0
    seis = g.seismograms(nsteps)
    interval = max(1, int(round(np.sqrt(nsteps))))
    t0 = time.perf_counter()
    kern, k3 = g.adjoint_run(seis, interval)
    t_adj = time.perf_counter() - t0
    g.close()
    log(f"adjoint: forward {t_fwd:.3f} s, adjoint run {t_adj:.3f} s")
    out = {
        "metric": "GLL-point*mode updates/s of the adjoint run (forward "
                  "with checkpoints, recomputation, adjoint field, "
                  "sensitivities), per forward step",
        "value": nupd * nsteps / t_adj, "unit": UNIT, "n_gpus": 1,
        "steps": nsteps, "warmup": 0, "ms_per_step": t_adj / nsteps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "cartesian box 80x80 km, 8000 elements, nu=2 "
                   "(Nr 5), isotropic, every other element 3D, axial "
                   "moment tensor, 10 receivers; misfit J = sum(s^2)/2",
                   "checkpoint_interval": interval,
                   "updates_per_step": nupd},
        "forward_ms_per_step": t_fwd / nsteps * 1e3,
        "adjoint_over_forward": t_adj / t_fwd,
        "kernel_norms": {"lambda_mu_rho": float(np.linalg.norm(kern)),
                         "lambda_mu_per_slice": float(np.linalg.norm(k3))},
    }
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is "
              "not supported (every GPU box here has one GPU).",
              file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration > 0:
        copy_calibration(a.copy_calibration)
        return
    if a.workload == "adjoint":
        adjoint_bench(a)
        return
    from axisem3d_amd import capi
    # warm-up, timed steps, one capture step and the graph replays; both
    # precisions then run on to the accuracy leg's step count
    acc = 0 if a.no_fp64 else max(a.accuracy_steps, 2 * a.steps + a.warmup
                                  + 1)
    nsteps = max(2 * a.steps + a.warmup + 1, acc)
    t0 = time.perf_counter()
    spec, cfg = build_workload(a.workload, nsteps)
    log(f"workload built in {time.perf_counter() - t0:.1f} s: "
        f"{cfg['elements']} elements, {cfg['updates_per_step']} "
        "point-modes per step")
    nupd = spec.n_updates_per_step()

    t32, prof32, seis32 = time_gpu(spec, "fp32", a.steps, a.warmup,
                                   graph=True, total=acc)
    value = nupd * a.steps / t32
    out_64 = None
    l2 = None
    accuracy = None
    if not a.no_fp64:
        t64, prof64, seis64 = time_gpu(spec, "fp64", a.steps, a.warmup,
                                       total=acc)
        out_64 = {"value": nupd * a.steps / t64,
                  "ms_per_step": t64 / a.steps * 1e3}
        l2 = float(np.linalg.norm(seis32 - seis64) /
                   max(np.linalg.norm(seis64), 1e-300))
        accuracy = accuracy_leg(spec, seis32, seis64)

    names = {capi.KERNEL_STIFFNESS: "solid stiffness, Fourier path",
             capi.KERNEL_STIFFNESS_3D: "solid stiffness, cardinal path "
                                       "(all three kernels)",
             capi.KERNEL_CD_GRAD: "k_cd_grad",
             capi.KERNEL_CD_STRESS: "k_cd_stress",
             capi.KERNEL_CD_QUAD: "k_cd_quad",
             capi.KERNEL_POINT: "k_point",
             capi.KERNEL_SOURCE: "moment source",
             capi.KERNEL_RECORD: "receivers"}
    b = bytes_per_step(spec, 4)
    # algorithmic bytes per step of each timed group
    byts = {capi.KERNEL_STIFFNESS: b["fr"],
            capi.KERNEL_STIFFNESS_3D: b["cd_grad"] + b["cd_stress"] +
            b["cd_quad"],
            capi.KERNEL_CD_GRAD: b["cd_grad"],
            capi.KERNEL_CD_STRESS: b["cd_stress"],
            capi.KERNEL_CD_QUAD: b["cd_quad"],
            capi.KERNEL_POINT: b["point"]}
    kernels = {}
    for k in byts:
        ms, cnt = prof32[k]
        if cnt == 0:
            continue
        per_step = cnt / a.steps
        kernels[names[k]] = {
            "ms_per_step": ms / a.steps, "launches_per_step": per_step,
            "avg_us": ms / cnt * 1e3,
            "bytes_per_launch": byts[k] / per_step,
            "achieved_GBs": byts[k] / (ms / a.steps * 1e-3) / 1e9}
    tune32 = prof32.pop("autotune")
    tg = prof32.pop("graph_s")
    if out_64 is not None:
        prof64.pop("autotune")
    # dominant single kernel (the cardinal group is three kernels)
    singles = [k for k in byts if k != capi.KERNEL_STIFFNESS_3D and
               prof32[k][1] > 0]
    dom = max(singles, key=lambda k: prof32[k][0])
    dk = kernels[names[dom]]
    achieved = dk["bytes_per_launch"] / (dk["avg_us"] * 1e-6) / 1e9
    traffic = pmc_traffic(names[dom], a.workload)
    step_bytes = sum(v for v in b.values())
    out = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": t32 / a.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic", "config": cfg,
        "fp64_plain": out_64,
        "speedup_vs_fp64": (value / out_64["value"]) if out_64 else None,
        "seismogram_rel_l2_vs_fp64": l2,
        "accuracy_leg": accuracy,
        "roofline": {"bound": "hbm", "kernel": names[dom],
                     "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "valu_busy": pmc_valu_busy(names[dom], a.workload),
                     "counters_source": {
                         "profiles": _profile_row("hbm_traffic.txt", "", "",
                                                  99)[1],
                         "this_tree": source_fingerprint()}},
        "step_roofline": {"algorithmic_bytes_per_step": step_bytes,
                          "achieved_GBs": step_bytes / (t32 / a.steps) / 1e9,
                          "frac": step_bytes / (t32 / a.steps) / 1e9 /
                          HBM_PEAK_GBS},
        "kernels": kernels,
        "graph_replay": {"ms_per_step": tg / a.steps * 1e3,
                         "value": nupd * a.steps / tg},
        "autotune": tune32,
    }
    if a.workload == "batched":
        out["per_instance"] = batched_extras(spec, seis32,
                                             seis64 if out_64 else None,
                                             a.steps, a.warmup, tg)
    if not a.no_cpu:
        t0 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(spec, a.cpu_seconds, a.workload)
        log(f"cpu baseline {time.perf_counter() - t0:.1f} s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
