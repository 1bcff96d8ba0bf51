"""The oracle against the reference's own code (oracle/_ref) on the whole
plain-method workload: uniform128_cic (128^3 uniform grid, 4096 rays over the
5 x 5 deg narrow field, CIC, to z = 0.5), 2.3 M RK4 steps, bit for bit; and
the narrow-field ray layout against its closed-form area."""
import math

import numpy as np
import pytest

from mpf import synth
import oracle as orc


@pytest.mark.skipif(not orc.ref_available(), reason="oracle/_ref not built")
def test_oracle_equals_reference_on_uniform128_cic():
    w, rec, f64, length, table, in9 = synth.build_workload("uniform128_cic")
    assert in9.shape == (4096, 9) and rec.size == 128 ** 3
    R = orc.RefHarness()
    h = R.tree(rec)
    ref, rsteps = R.integrate(h, w["order"], table, length, 10, 2, w["z_stop"], in9, nthreads=8)
    R.free(h)
    O = orc.Oracle()
    h = O.tree(rec, f64)
    out, steps = O.integrate(h, w["order"], table, length, 10, 2, w["z_stop"], in9, precision=0, nthreads=8)
    O.free(h)
    assert steps == rsteps > 2_000_000
    np.testing.assert_array_equal(out, ref)


def test_narrow_field_geometry():
    """Rays of the regular grid and uniform draws lie in the field
    (|phi| < h, |theta - pi/2| < h in the field's frame), the grid's centre
    ray is the axis direction, and the field's solid angle 4 h sin h (the
    reference's sky portion, miscellaneous.h:332) equals a quadrature of
    the area element sin(theta) dtheta dphi over the field."""
    h = math.radians(5.0)
    F = synth.narrow_field_frame(20.0, 20.0)
    assert np.allclose(F @ F.T, np.eye(3), atol=1e-15)
    assert np.allclose(np.cross(F[0], F[1]), F[2], atol=1e-15)
    for ph, th in (synth.narrow_field_grid(20.0, 20.0, 10.0, 33), synth.narrow_field_uniform(20.0, 20.0, 10.0, 5000, 3)):
        d = synth.launch_photons(ph, th)[:, 6:9] @ F.T
        assert np.all(np.abs(np.arctan2(d[:, 1], d[:, 0])) < h + 1e-12)
        assert np.all(np.abs(np.arccos(d[:, 2]) - math.pi / 2) < h + 1e-12)
    ph, th = synth.narrow_field_grid(20.0, 20.0, 10.0, 33)
    c = synth.launch_photons(ph[[33 * 33 // 2]], th[[33 * 33 // 2]])[0, 6:9]
    assert np.allclose(c, F[0], atol=1e-15)
    t = np.linspace(math.pi / 2 - h, math.pi / 2 + h, 20001)
    area = 2 * h * np.trapezoid(np.sin(t), t) * (180 / math.pi) ** 2
    assert abs(area - synth.narrow_field_area_deg2(10.0)) < 1e-6
    assert abs(synth.narrow_field_area_deg2(10.0) - 99.873) < 1e-3
