"""The reference's SEG salt local example (examples/03_salt_body_SEG_local/
input1D: its Exodus mesh, 1900 solid and 100 fluid elements, and its
inparam keys, from the committed fixture) set up by preloop.from_inparam
and run through the C ABI against the FP64 oracle: Clayton-Enquist and
Kosloff boundaries on the right and bottom, the 200 m water layer with its
free surface, the axial source with the inputs' Gaussian STF, Nr as the
inputs give it (1) and with Nr 9 everywhere, 5 receivers below the sea
floor.  Tolerances: seismograms FP64 relative L2 1e-12, FP32+FP64 1e-4."""
import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop
from test_exodus import fixture_inparam, fixture_mesh

pytestmark = pytest.mark.gpu


def _spec(nr, nsteps):
    inp = fixture_inparam()
    inp["nr"] = dict(inp["nr"], constant=nr)
    ex = fixture_mesh()
    z = float(ex.coords[:, 1].max()) - 250.0
    return preloop.from_inparam(
        inp, ex, nsteps, receivers_s=[0.0, 1e3, 2e3, 3.5e3, 4.9e3],
        receivers_z=[z] * 5, receivers_phi=[0.0, 0.5, 1.0, 2.0, 3.0])


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
@pytest.mark.parametrize("nr,nsteps", [(1, 1200), (9, 500)])
def test_salt_example_matches_oracle(nr, nsteps, prec, tol):
    from axisem3d_amd.solver import GpuSolver
    sp = _spec(nr, nsteps)
    oracle.set_threads(8)
    try:
        ref = oracle.Oracle(sp).run(nsteps)
    finally:
        oracle.set_threads(1)
    g = GpuSolver(sp, prec)
    g.run(nsteps)
    got = g.seismograms()
    g.close()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"salt example Nr {nr} {prec}: {nsteps} steps, seismogram L2 "
          f"{err:.3e}")
    assert np.abs(ref).max() > 0
    assert err < tol
