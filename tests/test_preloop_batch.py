"""Host side of batched instances: batch_instances concatenates the
instances' points, elements, sources and receivers with the right offsets
and keeps each instance's time step per point (CPU only)."""
import numpy as np
import pytest

from axisem3d_amd import preloop


def _instances():
    return [preloop.local_instance(70 + i, 20 + 7 * i, 2 + i, 50)
            for i in range(3)]


def test_batch_offsets_reproduce_every_instance():
    specs = _instances()
    b = preloop.batch_instances(specs)
    assert len(b._instances) == 3
    srcs = [(b.src_elem, b.src_pattern, b.stf)] + b.extra_sources
    for i, (s, ((p0, p1), (e0, e1), (r0, r1))) in enumerate(
            zip(specs, b._instances)):
        assert p1 - p0 == len(s.solid_nr) and e1 - e0 == s.n_solid_elements
        np.testing.assert_array_equal(b.solid_nr[p0:p1], s.solid_nr)
        np.testing.assert_array_equal(b.solid_mass[p0:p1], s.solid_mass)
        np.testing.assert_array_equal(b.elem_points[e0:e1] - p0,
                                      s.elem_points)
        np.testing.assert_array_equal(b.elem_iso[e0:e1], s.elem_iso)
        np.testing.assert_array_equal(b.elem_gq[e0:e1], s.elem_gq)
        np.testing.assert_array_equal(b.rec_elem[r0:r1] - e0, s.rec_elem)
        np.testing.assert_array_equal(b.rec_weights[r0:r1], s.rec_weights)
        assert np.all(b.point_instance[p0:p1] == i)
        assert b.instance_dt[i] == s.dt
        e, pat, stf = srcs[i]
        assert e - e0 == s.src_elem
        np.testing.assert_array_equal(pat, s.src_pattern)
        np.testing.assert_array_equal(stf, s.stf)
    ax = set(b.axial_points.tolist())
    for s, ((p0, _), _, _) in zip(specs, b._instances):
        assert set((s.axial_points + p0).tolist()) <= ax
    # no element of one instance touches another instance's points
    for (p0, p1), (e0, e1), _ in b._instances:
        assert b.elem_points[e0:e1].min() >= p0
        assert b.elem_points[e0:e1].max() < p1


def test_batch_rejects_unequal_step_counts():
    specs = _instances()
    specs[1] = preloop.local_instance(99, 25, 3, 40)
    with pytest.raises(ValueError, match="same number of steps"):
        preloop.batch_instances(specs)
