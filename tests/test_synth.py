"""The C++ streaming generator of the 3D model equals the numpy preloop
(same formulas), and the streaming dt equals the full-evaluation dt."""
import numpy as np
import pytest

from axisem3d_amd import preloop, synth


@pytest.mark.parametrize("cg4", [True, False])
def test_generator_matches_numpy(cg4):
    spec = preloop.spherical_regional(
        ntheta=6, nrad=6, theta_max_deg=2.0, depth=60e3, nu_min=3,
        nu_max=12, nsteps=10, n_receivers=3, cg4=cg4, corr_len=10e3)
    ref = preloop.spherical_regional(
        ntheta=6, nrad=6, theta_max_deg=2.0, depth=60e3, nu_min=3,
        nu_max=12, nsteps=10, n_receivers=3, cg4=cg4, corr_len=10e3)
    ref._build_3d()
    synth.prepare_3d_streaming(spec, chunk=7)
    assert spec.dt == ref.dt
    for k in ("alpha", "beta", "gamma"):
        assert np.allclose(spec.att[k], ref.att[k], rtol=1e-14, atol=0)
    consts = preloop.att_constants(spec.att["w"], spec.att["y"], 0.005, 0.5,
                                   1.0)
    wcg4 = preloop.cg4_weights(spec._gq["ifact"])
    prop, att, _ = synth.generate(spec, 0, spec.n_solid_elements,
                                  spec._field, spec._att_in, consts, wcg4)
    assert np.allclose(prop, ref.prop3d, rtol=1e-12, atol=0)
    assert np.allclose(att, ref.att["d3d"], rtol=1e-12, atol=0)
