"""The oracle's Integrator::propagate (integrator.h:826-1015) pinned: against
fixtures written by the reference's own propagate (tests/golden/
make_golden.py), against the compiled reference on further rays where it is
present, and against the closed form of a homogeneous universe (straight
rays: the angular diameter distance is a chi L, times the chord-to-arc ratio
of the bundle).  CPU only."""
import math
import os

import numpy as np
import pytest

import cases
from mpf import synth
from oracle import Oracle, RefHarness, ref_available

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KINDS = {"redshift": 2, "a": 4, "t": 3, "lambda": 1, "r": 0}


@pytest.fixture(scope="module")
def wide():
    O = Oracle()
    rec, f64 = cases.tree("wide")
    h = O.tree(rec)
    yield O, h, rec
    O.free(h)


def test_propagate_matches_golden(wide):
    O, h, rec = wide
    g = np.load(os.path.join(GOLDEN, "propagate.npz"))
    assert str(g["digest"]) == cases.records_digest(rec)
    tab = cases.cosmology()
    nvalid = 0
    for order in (2, 1):
        for kind in KINDS if order == 2 else ("redshift",):
            for i, r in enumerate(g["in9"]):
                rows, dist = O.propagate(h, order, tab, cases.LENGTH, cases.NSTEPS, r, 6, 0.01, 0.37 * i,
                                         kind=KINDS[kind])
                np.testing.assert_array_equal(rows, g[f"o{order}_{kind}_{i}_rows"], err_msg=f"{order} {kind} {i}")
                np.testing.assert_array_equal(dist, g[f"o{order}_{kind}_{i}_dist"], err_msg=f"{order} {kind} {i}")
                nvalid += len(rows) > 0
    assert nvalid >= 40


@pytest.mark.skipif(not ref_available(), reason="reference harness not built")
def test_propagate_matches_reference(wide):
    O, h, rec = wide
    R = RefHarness()
    hr = R.tree(rec)
    tab = cases.cosmology()
    rng = np.random.default_rng(4)
    in9 = cases.rays(n_side=4, aperture=40.0, phi0=25.0, theta0=15.0)
    for i, r in enumerate(in9):
        count, angle, rot = int(rng.integers(0, 9)), float(rng.uniform(1e-3, 0.05)), float(rng.uniform(0, 6.3))
        kind = list(KINDS)[i % 5]
        amin = 0.0 if i % 3 else 0.5
        a, da = O.propagate(h, 2, tab, cases.LENGTH, cases.NSTEPS, r, count, angle, rot, kind=KINDS[kind], amin=amin)
        b, db = R.propagate(hr, 2, tab, cases.LENGTH, cases.NSTEPS, r, count, angle, rot, interp=kind, amin=amin)
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(da, db)
    R.free(hr)


def test_propagate_rejects_short_bundles():
    # the test trees stored within 0.5 of the observer: every ray ends
    # before EXTENT / 4, so every bundle is rejected
    O = Oracle()
    rec, f64 = cases.tree("amr")
    h = O.tree(rec)
    for r in cases.rays(n_side=2, aperture=10.0):
        rows, dist = O.propagate(h, 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, r, 4, 0.01, 0.0)
        assert len(rows) == 0
    O.free(h)


def test_propagate_homogeneous_closed_form():
    # phi = 0 everywhere (Order -1): straight rays, D_A = a chi L x chord/arc
    lvl, kx, ky, kz = synth.uniform_level_coords(3, half=2.0)
    from mpf import formats
    f = np.zeros((kx.size, 7))
    f[:, 5] = 1.0
    rec, _ = formats.make_records(lvl, kx, ky, kz, f)
    O = Oracle()
    h = O.tree(rec)
    tab = synth.flat_lcdm_table(h=0.7, omega_m=0.3, a_min=0.02, n=2048)
    angle = 0.02
    in9 = np.array([1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.6, 0.48, 0.64])
    for kind in ("redshift", "r", "lambda"):
        rows, dist = O.propagate(h, -1, tab, cases.LENGTH, 64, in9, 8, angle, 0.2, kind=KINDS[kind])
        assert len(rows) > 100
        chi = rows[:, 12]
        a = rows[:, 0]
        expect = a * chi * cases.LENGTH * (2 * math.sin(angle / 2) / angle)
        # away from the start and from the box edge, where the bundle rays end
        # at slightly different parameters and the interpolation extrapolates
        k = (chi > 0.05) & (chi < 0.95 * chi.max())
        np.testing.assert_allclose(dist[k], expect[k], rtol=2e-6, err_msg=kind)
    O.free(h)
