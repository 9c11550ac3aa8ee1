"""Batched independent instances (one domain, each instance with its own
mesh, Fourier order, two-layer model, moment tensor, time step and
receivers) against the FP64 oracle run instance by instance, and against
the same instances run alone on the GPU: the batched launches compute
every instance exactly as a run of its own would, so the seismograms are
bitwise equal; parity with the oracle is to 1e-12 (FP64) and 1e-4
(FP32+FP64) seismogram L2 per instance."""
import numpy as np
import pytest

import oracle
from axisem3d_amd import preloop

pytestmark = pytest.mark.gpu

NSTEPS = 300


def _instances(n=4, seed=12):
    rng = np.random.default_rng(seed)
    return [preloop.local_instance(int(seed * 100 + i),
                                   int(rng.integers(20, 61)),
                                   int(rng.integers(2, 9)), NSTEPS)
            for i in range(n)]


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-4)])
def test_batched_instances_match_oracle_and_single_runs(prec, tol):
    from axisem3d_amd.solver import GpuSolver
    specs = _instances()
    assert len({s.dt for s in specs}) == len(specs)
    batch = preloop.batch_instances(specs)
    g = GpuSolver(batch, prec)
    g.run(NSTEPS)
    seis = g.seismograms()
    g.close()
    for i, (sp, (_, _, (r0, r1))) in enumerate(zip(specs,
                                                   batch._instances)):
        got = seis[:, r0:r1]
        ref = oracle.Oracle(sp).run(NSTEPS)
        assert np.abs(ref).max() > 0
        err = _rel(got, ref)
        print(f"{prec} instance {i}: {sp.n_solid_elements} elements, Nr "
              f"{int(sp.elem_nr.max())}, dt {sp.dt}: seismogram L2 vs "
              f"FP64 oracle {err:.3e}")
        assert err < tol
        alone = GpuSolver(sp, prec)
        alone.run(NSTEPS)
        assert np.array_equal(alone.seismograms(), got)
        alone.close()


def test_batch_rejects_mismatched_sources():
    from axisem3d_amd import capi
    from axisem3d_amd.solver import GpuSolver
    specs = _instances(2)
    batch = preloop.batch_instances(specs)
    e, pat, stf = batch.extra_sources[0]
    batch.extra_sources = [(e, pat, stf[:-1])]
    with pytest.raises(capi.Axs3dError, match="same number of STF steps"):
        GpuSolver(batch, "fp64")
