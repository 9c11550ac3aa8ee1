"""Host side of bench.py (CPU only): the one-GPU refusal, the algorithmic
byte model's parts, and the profile lookups the bench line uses."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_bytes_per_step_counts_each_array_once():
    import bench
    from axisem3d_amd import preloop
    sp = preloop.cartesian_homogeneous(
        nx=4, nz=4, width=4e3, depth=4e3, nsteps=5, src_depth=2e3,
        n_receivers=2, nodal_nr=lambda s, z: np.full(s.shape, 5.0))
    b = bench.bytes_per_step(sp, 4)
    nupd = sp.n_updates_per_step()
    # Fourier path only: no cardinal bytes; the point kernel reads the
    # element forces and reads and writes u, v, a (FP64) per point-mode
    assert b["cd_grad"] == b["cd_stress"] == b["cd_quad"] == 0
    assert b["point"] >= nupd * (6 * 4 + 6 * 48)
    assert b["fr"] > nupd * 48


def test_profile_lookups():
    import bench
    t = bench.pmc_traffic("k_cd_stress", "regional")
    assert t is None or t > 1e9
    v = bench.pmc_valu_busy("k_cd_stress", "regional")
    assert v is None or 0.0 < v <= 1.0
    assert bench.pmc_traffic("no_such_kernel", "regional") is None
