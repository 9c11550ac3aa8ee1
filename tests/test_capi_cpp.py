"""A C++ host of the C ABI with no Python in the loop (tests/cpp: the
Domain seam gpu_domain.hpp and the program perphase_main.cpp).

CPU: the program compiles with -Wall -Wextra -Werror against
include/axs3d.h and links libaxs3d.so.  GPU: it builds a small domain
(Fourier-path and cardinal-path elements, a moment source, receivers)
through the set-up calls, runs the per-phase Newmark loop
(NewmarkTimeScheme::solve order) and its seismograms match the FP64
oracle: FP64 relative L2 1e-12, FP32+FP64 1e-4."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop, spectral

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
PROG = os.path.join(CPP, "axs3d_perphase")
NSTEPS = 150


def test_cpp_host_compiles_and_links():
    r = subprocess.run(["make", "-B", "-C", CPP], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.access(PROG, os.X_OK)


def _spec():
    sp = preloop.cartesian_homogeneous(
        nx=6, nz=7, width=6e3, depth=5.6e3, nsteps=NSTEPS, src_depth=2e3,
        n_receivers=4,
        nodal_nr=lambda s, z: 3.0 + 9.0 * (s / 6e3) + 0.0 * z)
    preloop.make_3d_isotropic(
        sp, lambda x, y, z: 0.02 * np.sin(x / 900.0) * np.cos(z / 700.0),
        mask3d=(np.arange(sp.n_solid_elements) % 2) == 0)
    return sp


def write_domain(sp, path):
    """The arrays of the set-up calls as raw little-endian files plus the
    manifest domain.txt that perphase_main.cpp reads."""
    os.makedirs(path, exist_ok=True)
    t = spectral.tables()
    arrays = {
        "g_gll": (t[4], np.float64), "g_glj": (t[5], np.float64),
        "solid_nr": (sp.solid_nr, np.int32),
        "solid_mass": (sp.solid_mass, np.float64),
        "elem_points": (sp.elem_points, np.int32),
        "elem_axial": (sp.elem_axial, np.uint8),
        "elem_gq": (sp.elem_gq, np.float64),
        "elem_iso": (sp.elem_iso, np.float64),
        "axial_points": (sp.axial_points, np.int32),
        "src_pattern": (sp.src_pattern, np.complex128),
        "stf": (sp.stf, np.float64),
        "rec_elem": (sp.rec_elem, np.int32),
        "rec_weights": (sp.rec_weights, np.float64),
        "rec_phi": (sp.rec_phi, np.float64),
    }
    if getattr(sp, "elem_kind", None) is not None:
        arrays["elem_kind"] = (sp.elem_kind, np.uint8)
        arrays["prop3d"] = (sp.prop3d, np.float64)
    lines = [f"scalar dt {sp.dt!r}", f"scalar src_elem {int(sp.src_elem)}",
             f"scalar src_nu1 {int(np.asarray(sp.src_pattern).shape[0])}"]
    for name, (a, dt) in arrays.items():
        b = np.ascontiguousarray(np.asarray(a), dtype=dt).tobytes()
        with open(os.path.join(path, name + ".bin"), "wb") as f:
            f.write(b)
        lines.append(f"array {name} {len(b)}")
    with open(os.path.join(path, "domain.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
def test_cpp_host_per_phase_matches_oracle(tmp_path, prec, tol):
    assert os.access(PROG, os.X_OK), "build first: make -C tests/cpp"
    sp = _spec()
    write_domain(sp, str(tmp_path / "dom"))
    out = str(tmp_path / "seis.bin")
    r = subprocess.run([PROG, str(tmp_path / "dom"), prec, str(NSTEPS), out],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    nrec = len(sp.rec_elem)
    got = np.fromfile(out, dtype=np.float64).reshape(NSTEPS, nrec, 3)
    ref = oracle.Oracle(sp).run(NSTEPS)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"C++ host {prec}: seismogram L2 vs oracle {err:.3e}")
    assert np.abs(ref).max() > 0
    assert err < tol
