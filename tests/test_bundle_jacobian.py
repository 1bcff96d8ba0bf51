"""The oracle's Lensing::dbetadtheta (lensing.h:140-364, the bundle Jacobian)
pinned: against fixtures written by the reference's own dbetadtheta
(tests/golden/make_golden.py), including the failure entries of rays that
end before a value; against the compiled reference on further directions
where it is present; and against the closed form of a homogeneous universe
(straight rays: A = I / sqrt(1 + tan^2 opening) when the screen is normal to
the direction and dist is the surface radius).  CPU only."""
import math
import os

import numpy as np
import pytest

import cases
from mpf import formats, synth
from oracle import Oracle, RefHarness, ref_available

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KIND = {"r": 0, "lambda": 1, "redshift": 2, "t": 3, "a": 4}
VALUES = {"r": [0.15, 0.3, 0.42], "lambda": [2e16, 5e16], "redshift": [0.1, 0.25, 0.33], "t": [4e16, 1e17],
          "a": [0.95, 0.85, 0.78]}


def test_dbetadtheta_matches_golden():
    g = np.load(os.path.join(GOLDEN, "bundle.npz"))
    O = Oracle()
    nfail = 0
    for name in ("amr", "wide"):
        rec, f64 = cases.tree(name)
        assert str(g[f"{name}_digest"]) == cases.records_digest(rec)
        h = O.tree(rec)
        for kind, vals in VALUES.items():
            for i, (p, t) in enumerate(g[f"{name}_phi_theta"]):
                a = O.dbetadtheta(h, 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, KIND[kind], 0.005,
                                  [0.0, 0.0, 0.0], p, t, vals, g[f"{name}_{kind}_ki"][i], g[f"{name}_{kind}_dist"][i])
                np.testing.assert_array_equal(a, g[f"{name}_{kind}_jac"][i], err_msg=f"{name} {kind} {i}")
                nfail += a[0, 0] == 42
        O.free(h)
    assert nfail > 5   # the fixtures hold failure entries too


@pytest.mark.skipif(not ref_available(), reason="reference harness not built")
def test_dbetadtheta_matches_reference():
    O, R = Oracle(), RefHarness()
    rec, f64 = cases.tree("wide")
    ho, hr = O.tree(rec), R.tree(rec)
    rng = np.random.default_rng(3)
    for kind, vals in VALUES.items():
        for i in range(6):
            p, t = rng.uniform(0.1, 0.6), rng.uniform(0.1, 0.6)
            ki = rng.normal(size=(len(vals), 3))
            dist = rng.uniform(0.1, 0.4, len(vals))
            obs = [0.0, 0.0, 0.0] if i % 2 else rng.uniform(-0.01, 0.01, 3)
            a = O.dbetadtheta(ho, 1 + i % 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, KIND[kind], 0.002, obs,
                              p, t, vals, ki, dist)
            b = R.dbetadtheta(hr, 1 + i % 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, kind, 0.002, obs, p, t,
                              vals, ki, dist)
            np.testing.assert_array_equal(a, b)
    O.free(ho)
    R.free(hr)


def test_dbetadtheta_homogeneous_closed_form():
    lvl, kx, ky, kz = synth.uniform_level_coords(3, half=2.0)
    f = np.zeros((kx.size, 7))
    f[:, 5] = 1.0
    rec, _ = formats.make_records(lvl, kx, ky, kz, f)
    O = Oracle()
    h = O.tree(rec)
    tab = synth.flat_lcdm_table(h=0.7, omega_m=0.3, a_min=0.05, n=2048)
    opening = 0.004
    for p, t in ((0.3, 1.1), (2.0, 0.4)):
        k = np.array([math.cos(p) * math.sin(t), math.sin(p) * math.sin(t), math.cos(t)])
        radii = np.array([0.3, 0.8, 1.5])
        a = O.dbetadtheta(h, -1, tab, cases.LENGTH, 40, 0, opening, [0.0, 0.0, 0.0], p, t, radii,
                          np.tile(k, (3, 1)), radii)
        expect = np.array([1.0, 0.0, 0.0, 1.0]) / math.sqrt(1 + math.tan(opening) ** 2)
        np.testing.assert_allclose(a, np.tile(expect, (3, 1)), atol=1e-9)
    O.free(h)


@pytest.mark.skipif(not ref_available(), reason="reference harness not built")
def test_dbetadtheta_plane_matches_reference():
    """stop_bundle = plane (lensing.h:201, 281-293): the four rays stop on the
    last value along the last ki and are interpolated along each ki; bit for
    bit the reference's, failure entries included."""
    O, R = Oracle(), RefHarness()
    tab = cases.cosmology()
    rng = np.random.default_rng(3)
    fails = 0
    for name, ref in (("amr", [0.1, 0.2, 0.3]), ("wide", [0.2, 0.5, 0.9])):
        rec, _ = cases.tree(name)
        ho, hr = O.tree(rec), R.tree(rec)
        for _ in range(8):
            phi = math.radians(20) + rng.uniform(-0.2, 0.2)
            th = math.radians(20) + rng.uniform(-0.2, 0.2)
            n = np.array([math.cos(phi) * math.sin(th), math.sin(phi) * math.sin(th), math.cos(th)])
            ki = np.tile(n, (3, 1)) + rng.normal(0, 0.01, (3, 3))
            for order in (1, 2):
                a = O.dbetadtheta(ho, order, tab, cases.LENGTH, cases.NSTEPS, 5, 1e-4, [0, 0, 0], phi, th, ref, ki, ref)
                b = R.dbetadtheta(hr, order, tab, cases.LENGTH, cases.NSTEPS, "plane", 1e-4, [0, 0, 0], phi, th, ref,
                                  ki, ref)
                assert np.array_equal(a, b), (name, order)
                fails += a[0, 0] == 42
        O.free(ho)
        R.free(hr)
    assert 0 < fails < 32
