"""Edge cases of the C ABI path against the FP64 oracle: the smallest
azimuthal resolutions (Nr = 1 to 4, on the Fourier and the cardinal
path: zero-stage and single-stage FFTs, even-Nr Nyquist rows), a domain
with neither source nor receivers started from a seeded displacement
(free oscillation), and runs of zero steps."""
import copy

import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _small_nr_spec(nsteps, cardinal):
    spec = preloop.cartesian_homogeneous(
        nx=5, nz=5, width=5e3, depth=5e3, nsteps=nsteps, src_depth=2e3,
        n_receivers=3,
        nodal_nr=lambda s, z: 1.0 + 3.0 * np.clip((s - 1.2e3) / 3.8e3, 0, 1)
        + 0.0 * z)
    if cardinal:
        E = spec.n_solid_elements
        preloop.make_media(spec, np.zeros(E, np.uint8), field=None,
                           mask3d=(np.arange(E) % 2) == 0, att=None, seed=4)
    return spec


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
@pytest.mark.parametrize("cardinal", [False, True])
def test_smallest_nr_matches_oracle(prec, tol, cardinal):
    from axisem3d_amd.solver import GpuSolver
    nsteps = 300
    spec = _small_nr_spec(nsteps, cardinal)
    assert set(np.unique(spec.elem_nr)) <= {1, 2, 3, 4}
    assert 1 in set(spec.elem_nr) and len(set(spec.elem_nr)) >= 3
    ref = oracle.Oracle(spec).run(nsteps)
    g = GpuSolver(spec, prec)
    g.run(nsteps)
    got = g.seismograms()
    g.close()
    assert np.abs(ref).max() > 0
    err = _rel(got, ref)
    print(f"Nr 1..4 cardinal={cardinal} {prec}: seismogram L2 {err:.3e}")
    assert err < tol


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
def test_free_oscillation_without_source_or_receivers(prec, tol):
    from axisem3d_amd.solver import GpuSolver
    spec = _small_nr_spec(200, True)
    spec = copy.copy(spec)
    spec.src_elem = -1
    spec.src_pattern = None
    spec.stf = None
    spec.rec_elem = None
    spec.rec_weights = None
    spec.rec_phi = None
    o = oracle.Oracle(spec)
    rng = np.random.default_rng(3)
    u0 = rng.standard_normal((o.nmodes, 3)) * 1e-3
    off = o.offsets
    u0 = u0.astype(complex)
    u0[off[:-1]] = u0[off[:-1]].real
    o.displ[:] = u0
    g = GpuSolver(spec, prec)
    g.set_field("displ", u0)
    g.run(0)
    assert np.array_equal(g.get_field("displ"), u0)
    o.run(150, record=False)
    g.run(150)
    got = g.get_field("displ")
    g.close()
    assert np.abs(o.displ).max() > 0
    err = _rel(got, o.displ)
    print(f"free oscillation {prec}: displacement L2 {err:.3e}")
    assert err < tol


def test_setup_calls_after_finalize_are_refused():
    """The device buffers, the captured step graph and the adjoint caches
    are built once by finalize: a second finalize, or any set-up call
    after it, returns an error instead of leaving them stale, and the
    context keeps running."""
    import ctypes as C
    from axisem3d_amd import capi
    from axisem3d_amd.solver import GpuSolver
    spec = _small_nr_spec(20, False)
    g = GpuSolver(spec, "fp64")
    L = capi.load()
    g.run(5)
    assert L.axs3d_finalize(g.ctx) != 0
    assert b"already finalized" in L.axs3d_last_error(g.ctx)
    pat = np.ascontiguousarray(spec.src_pattern, dtype=np.complex128)
    stf = np.ascontiguousarray(spec.stf, dtype=np.float64)
    assert L.axs3d_add_moment_source(
        g.ctx, int(spec.src_elem), pat.shape[0], pat.ctypes.data, len(stf),
        stf.ctypes.data) != 0
    assert L.axs3d_set_dt(g.ctx, C.c_double(1e-3)) != 0
    mass = np.ones(len(spec.solid_nr))
    nr = np.ascontiguousarray(spec.solid_nr, dtype=np.int32)
    assert L.axs3d_set_solid_points(g.ctx, len(nr), nr.ctypes.data,
                                    mass.ctypes.data) != 0
    # the domain is unchanged: it runs on and matches a fresh run
    g.run(5)
    got = g.seismograms()
    g.close()
    h = GpuSolver(spec, "fp64")
    h.run(10)
    ref = h.seismograms()
    h.close()
    assert np.array_equal(got, ref)
