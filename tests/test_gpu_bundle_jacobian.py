"""Lensing::dbetadtheta (the bundle Jacobian, lensing.h:140-364) on the GPU
(mpf_jacobian_bundle) against fixtures written by the reference's own code
and against the oracle's restatement (pinned to it in
tests/test_bundle_jacobian.py): bit-identical entries, failure entries
included, in the exact tiers, with values shared by every direction and with
one value per direction; the split tier against the FP64 tier within
1e-9 + 1e-5 |A - I|."""
import os

import numpy as np
import pytest

import cases
from mpf import capi
from oracle import Oracle

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KIND = {"r": 0, "lambda": 1, "redshift": 2, "t": 3, "a": 4}
VALUES = {"r": [0.15, 0.3, 0.42], "lambda": [2e16, 5e16], "redshift": [0.1, 0.25, 0.33], "t": [4e16, 1e17],
          "a": [0.95, 0.85, 0.78]}


def ctx(name, precision, order=capi.TSC):
    rec, f64 = cases.tree(name)
    c = capi.Context(cases.LENGTH, order=order, precision=precision, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    return c, rec, f64


def test_bundle_jacobian_matches_golden():
    g = np.load(os.path.join(GOLDEN, "bundle.npz"))
    for name in ("amr", "wide"):
        c, rec, _ = ctx(name, capi.PREC_REF)
        for kind, vals in VALUES.items():
            print(f"[case] bundle jacobian fixtures tree={name} kind={kind}", flush=True)
            a = c.jacobian_bundle(g[f"{name}_phi_theta"], kind, vals, g[f"{name}_{kind}_ki"],
                                  g[f"{name}_{kind}_dist"], 0.005)
            np.testing.assert_array_equal(a, g[f"{name}_{kind}_jac"], err_msg=f"{name} {kind}")
        c.close()


@pytest.mark.parametrize("precision", (capi.PREC_REF, capi.PREC_FP64))
def test_bundle_jacobian_per_direction_matches_oracle(precision):
    c, rec, f64 = ctx("wide", precision)
    O = Oracle()
    h = O.tree(rec, f64)
    rng = np.random.default_rng(8)
    n = 40
    pt = np.c_[np.radians(20) + rng.uniform(-0.3, 0.3, n), np.radians(20) + rng.uniform(-0.3, 0.3, n)]
    for kind, lo, hi in (("redshift", 0.05, 0.4), ("r", 0.1, 0.6), ("a", 0.75, 0.97)):
        print(f"[case] bundle jacobian per direction precision={precision} kind={kind}", flush=True)
        ref = rng.uniform(lo, hi, (n, 1))
        ki = rng.normal(size=(n, 1, 3))
        dist = rng.uniform(0.1, 0.5, (n, 1))
        a = c.jacobian_bundle(pt, kind, ref, ki, dist, 0.003)
        for i in range(n):
            b = O.dbetadtheta(h, 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, KIND[kind], 0.003,
                              [0.0, 0.0, 0.0], pt[i, 0], pt[i, 1], ref[i], ki[i], dist[i], precision=precision)
            np.testing.assert_array_equal(a[i], b, err_msg=f"{kind} direction {i}")
    O.free(h)
    c.close()


def test_bundle_jacobian_split_vs_fp64():
    print("[case] bundle jacobian split vs fp64", flush=True)
    g = np.load(os.path.join(GOLDEN, "bundle.npz"))
    cs, _, _ = ctx("wide", capi.PREC_SPLIT)
    cf, _, _ = ctx("wide", capi.PREC_FP64)
    for kind, vals in VALUES.items():
        pt, ki, dist = g["wide_phi_theta"], g[f"wide_{kind}_ki"], g[f"wide_{kind}_dist"]
        a = cs.jacobian_bundle(pt, kind, vals, ki, dist, 0.005)
        b = cf.jacobian_bundle(pt, kind, vals, ki, dist, 0.005)
        ident = np.array([1.0, 0.0, 0.0, 1.0])
        tol = 1e-9 + 1e-5 * np.abs(b - ident)
        assert np.all(np.abs(a - b) <= tol), f"{kind}: max excess {np.max(np.abs(a - b) - tol):.3e}"
    cs.close()
    cf.close()


@pytest.mark.parametrize("precision", (capi.PREC_REF, capi.PREC_FP64, capi.PREC_SPLIT))
def test_bundle_jacobian_plane_vs_oracle(precision):
    """stop_bundle = plane: bit for bit the oracle's (itself the reference's,
    tests/test_bundle_jacobian.py), failure entries included: every third
    direction's last plane is out of reach, so its entries are the
    reference's 42, 42, 1, 0.  A split context runs this path in the FP64
    tier and is compared with the oracle's FP64 mode."""
    print(f"[case] bundle jacobian plane precision={precision}", flush=True)
    import math
    rec, f64 = cases.tree("wide")
    c = capi.Context(cases.LENGTH, order=capi.TSC, precision=precision, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    O = Oracle()
    h = O.tree(rec, f64)
    rng = np.random.default_rng(8)
    n, ref = 12, np.array([0.2, 0.5, 0.9])
    pt = np.c_[math.radians(20) + rng.uniform(-0.2, 0.2, n), math.radians(20) + rng.uniform(-0.2, 0.2, n)]
    dirs = np.c_[np.cos(pt[:, 0]) * np.sin(pt[:, 1]), np.sin(pt[:, 0]) * np.sin(pt[:, 1]), np.cos(pt[:, 1])]
    ki = np.repeat(dirs[:, None, :], 3, axis=1) + rng.normal(0, 0.01, (n, 3, 3))
    refs = np.tile(ref, (n, 1))
    refs[::3, 2] = 5.0          # beyond the data: the rays never reach it
    got = c.jacobian_bundle(pt, "plane", refs, ki, refs, 1e-4)
    oprec = 0 if precision == capi.PREC_REF else 1
    for i in range(n):
        want = O.dbetadtheta(h, 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, 5, 1e-4, [0, 0, 0], pt[i, 0],
                             pt[i, 1], refs[i], ki[i], refs[i], precision=oprec)
        np.testing.assert_array_equal(got[i], want, err_msg=f"direction {i}")
    failed = np.all(got[:, :, 0] == 42, axis=1)
    assert 0 < failed.sum() < n, failed.sum()
    assert np.all(failed[::3])
    # every crossing was placed while integrating (no whole trajectories)
    assert c.bundle_plane_info() == 0
    O.free(h)
    c.close()


@pytest.mark.parametrize("precision", (capi.PREC_REF, capi.PREC_FP64))
def test_bundle_jacobian_plane_whole_trajectory_directions(precision):
    """Directions whose rays' projections on the normal decrease go through
    whole trajectories (upper_bound over them, as the reference), the others
    through the crossings recorded while integrating, in one call: every
    other direction's normal points back to the observer (projections
    decrease from the first step, the plane value behind the observer ends
    the rays at once), and every result equals the oracle's bit for bit."""
    print(f"[case] bundle jacobian plane whole-trajectory directions precision={precision}", flush=True)
    import math
    c, rec, f64 = ctx("wide", precision)
    O = Oracle()
    h = O.tree(rec, f64)
    rng = np.random.default_rng(11)
    n = 16
    pt = np.c_[math.radians(20) + rng.uniform(-0.2, 0.2, n), math.radians(20) + rng.uniform(-0.2, 0.2, n)]
    dirs = np.c_[np.cos(pt[:, 0]) * np.sin(pt[:, 1]), np.sin(pt[:, 0]) * np.sin(pt[:, 1]), np.cos(pt[:, 1])]
    ki = (dirs + rng.normal(0, 0.01, (n, 3)))[:, None, :].copy()
    ref = rng.uniform(0.2, 0.6, (n, 1))
    back = np.arange(n) % 2 == 1
    ki[back] = -ki[back]
    ref[back] = -0.3
    dist = rng.uniform(0.2, 0.6, (n, 1))
    got = c.jacobian_bundle(pt, "plane", ref, ki, dist, 1e-4)
    assert c.bundle_plane_info() == int(back.sum())
    oprec = 0 if precision == capi.PREC_REF else 1
    for i in range(n):
        want = O.dbetadtheta(h, 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, 5, 1e-4, [0, 0, 0], pt[i, 0],
                             pt[i, 1], ref[i], ki[i], dist[i], precision=oprec)
        np.testing.assert_array_equal(got[i], want, err_msg=f"direction {i}")
    O.free(h)
    c.close()


def test_bundle_jacobian_plane_batch_split_context():
    """A catalogue-sized batch of plane bundles (4,096 directions, 16,384
    rays in one launch) from a split context equals the FP64 context's bit for
    bit (both run the FP64 tier), with no direction left to whole
    trajectories; a sample is checked against the oracle."""
    print("[case] bundle jacobian plane batch of 4096 directions", flush=True)
    import math
    cs, rec, f64 = ctx("wide", capi.PREC_SPLIT)
    cf, _, _ = ctx("wide", capi.PREC_FP64)
    rng = np.random.default_rng(12)
    n = 4096
    pt = np.c_[math.radians(20) + rng.uniform(-0.3, 0.3, n), math.radians(20) + rng.uniform(-0.3, 0.3, n)]
    dirs = np.c_[np.cos(pt[:, 0]) * np.sin(pt[:, 1]), np.sin(pt[:, 0]) * np.sin(pt[:, 1]), np.cos(pt[:, 1])]
    ki = (dirs + rng.normal(0, 0.01, (n, 3)))[:, None, :].copy()
    ref = rng.uniform(0.1, 0.7, (n, 1))
    a = cs.jacobian_bundle(pt, "plane", ref, ki, ref, 2e-4)
    assert cs.bundle_plane_info() == 0
    b = cf.jacobian_bundle(pt, "plane", ref, ki, ref, 2e-4)
    np.testing.assert_array_equal(a, b)
    assert np.all(np.isfinite(a)) and np.all(a[:, 0, 0] != 42)
    O = Oracle()
    h = O.tree(rec, f64)
    for i in range(0, n, 512):
        want = O.dbetadtheta(h, 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, 5, 2e-4, [0, 0, 0], pt[i, 0],
                             pt[i, 1], ref[i], ki[i], ref[i], precision=1)
        np.testing.assert_array_equal(a[i], want, err_msg=f"direction {i}")
    O.free(h)
    cs.close()
    cf.close()
