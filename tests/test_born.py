"""The Born approximation of the oracle (the homogeneous path of
Integrator::integrate<-1>, integrator.h:505, 573, with the infinitesimal
Jacobian sampled on the octree's data, lensing.h:399-480), pinned against
fixtures from the reference's own code (tests/golden/make_golden.py) and the
compiled reference: the jacobiantype = infinitesimal_born of catalogues.h:670-674
and the lensing_born maps of hmaps.h:705-713.  CPU only."""
import os

import numpy as np
import pytest

import cases
from oracle import Oracle, RefHarness, ref_available

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_born_jacobian_matches_golden():
    g = np.load(os.path.join(GOLDEN, "born.npz"))
    O = Oracle()
    for name in ("amr", "wide"):
        rec, f64 = cases.tree(name)
        assert str(g[f"{name}_digest"]) == cases.records_digest(rec)
        h = O.tree(rec)
        for i, r in enumerate(g[f"{name}_in9"]):
            a = O.jacobian(h, -1, cases.cosmology(), cases.LENGTH, cases.NSTEPS, 0, 0.32, r, [0.15, 0.25, 0.3])
            np.testing.assert_array_equal(a, g[f"{name}_jac"][i], err_msg=f"{name} ray {i}")
        # not the identity: the samples are on the data
        assert np.abs(g[f"{name}_jac"][:, :, 0] - 1).max() > 1e-7
        O.free(h)


@pytest.mark.skipif(not ref_available(), reason="reference harness not built")
def test_born_jacobian_matches_reference_other_stops():
    O, R = Oracle(), RefHarness()
    rec, f64 = cases.tree("wide")
    ho, hr = O.tree(rec), R.tree(rec)
    for stop, val in ((2, 0.3), (1, 5e16), (4, 0.8)):
        for r in cases.rays(n_side=3, aperture=16.0):
            a = O.jacobian(ho, -1, cases.cosmology(), cases.LENGTH, cases.NSTEPS, stop, val, r, [0.1, 0.25])
            b = R.jacobian(hr, -1, cases.cosmology(), cases.LENGTH, cases.NSTEPS, stop, val, r, [0.1, 0.25])
            np.testing.assert_array_equal(a, b)
    O.free(ho)
    R.free(hr)
