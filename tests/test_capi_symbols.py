"""The C-ABI library loads and exports every symbol include/axs3d.h
declares (no device needed; nothing is computed)."""
import ctypes
import os
import re

from axisem3d_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "axs3d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(axs3d_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_table():
    assert _declared() == sorted(capi.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_errors_without_context():
    lib = capi.load()
    assert lib.axs3d_last_error(None) == b"axs3d || Null context."
    assert lib.axs3d_finalize(None) != 0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    # no fallback: without the HIP library the solver refuses to start
    import pytest
    from axisem3d_amd import preloop
    from axisem3d_amd.solver import GpuSolver
    monkeypatch.setattr(capi, "LIB_PATH", str(tmp_path / "libaxs3d.so"))
    monkeypatch.setattr(capi, "_lib", None)
    spec = preloop.cartesian_homogeneous(nx=2, nz=2, width=2e3, depth=2e3,
                                         nsteps=4, n_receivers=1)
    with pytest.raises(capi.Axs3dError, match="not built"):
        GpuSolver(spec, "fp64")


def test_product_package_does_not_use_the_oracle():
    pkg = os.path.join(ROOT, "axisem3d_amd")
    pat = re.compile(r"^\s*(import\s+oracle|from\s+oracle\b)|oracle/", re.M)
    offenders = []
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hpp", ".hip", ".h")):
                path = os.path.join(dirpath, f)
                if pat.search(open(path, errors="replace").read()):
                    offenders.append(os.path.relpath(path, ROOT))
    assert not offenders, offenders
