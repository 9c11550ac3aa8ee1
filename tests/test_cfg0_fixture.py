"""The 8000-element Cartesian workload (80 x 80 km, nu = 2, homogeneous
isotropic solid, 2 Hz Ricker moment source, 10 surface receivers) against
the committed 3000-step FP64 oracle seismograms (tests/golden/
cfg0_seismograms.npz, every 10th step, written by make_cfg0_fixture.py).

CPU: the fixture still describes the workload the builder makes (dt,
element count, moment tensor, receiver azimuths) and the oracle
reproduces its first 300 steps.  GPU: all 3000 steps through the C ABI,
FP64 within relative L2 1e-12 and FP32+FP64 within 1e-4 of the fixture."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_cfg0_fixture as fx  # noqa: E402

FIX = np.load(os.path.join(HERE, "golden", "cfg0_seismograms.npz"))


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_fixture_matches_workload():
    sp = fx.spec()
    assert int(FIX["nsteps"]) == fx.NSTEPS and int(FIX["every"]) == fx.EVERY
    assert float(FIX["dt"]) == sp.dt
    assert int(FIX["n_elements"]) == sp.n_solid_elements == 8000
    assert np.array_equal(FIX["rec_phi"], sp.rec_phi)
    assert np.array_equal(FIX["moment_tensor"], sp._mt)
    assert FIX["seismograms"].shape == (fx.NSTEPS // fx.EVERY, 10, 3)


def test_oracle_reproduces_fixture_onset():
    import oracle
    sp = fx.spec()
    oracle.set_threads(os.cpu_count() or 1)
    try:
        seis = oracle.Oracle(sp).run(300)
    finally:
        oracle.set_threads(1)
    ref = FIX["seismograms"][:30]
    assert np.abs(ref).max() > 0
    assert _rel(seis[::fx.EVERY], ref) < 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
def test_gpu_matches_fixture(prec, tol):
    from axisem3d_amd.solver import GpuSolver
    sp = fx.spec()
    g = GpuSolver(sp, prec)
    g.run(fx.NSTEPS)
    got = g.seismograms()[::fx.EVERY]
    g.close()
    ref = FIX["seismograms"]
    err = _rel(got, ref)
    late = _rel(got[150:], ref[150:])
    print(f"cfg0 {prec}: 3000-step seismogram L2 vs oracle fixture "
          f"{err:.3e} (steps 1500-3000: {late:.3e})")
    assert err < tol
