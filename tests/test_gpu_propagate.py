"""Integrator::propagate (integrator.h:826-1015) on the GPU (mpf_propagate)
against the oracle's restatement, which tests/test_propagate.py pins to the
reference's own propagate: bit-identical central trajectories, angular
diameter distances and rejections, for every interpolation parameter, TSC and
CIC, both exact tiers."""
import os

import numpy as np
import pytest

import cases
from mpf import capi
from oracle import Oracle

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KINDS = {"redshift": 2, "a": 4, "t": 3, "lambda": 1, "r": 0}


def ctx(name, order, precision):
    rec, f64 = cases.tree(name)
    c = capi.Context(cases.LENGTH, order=order, precision=precision, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    return c, rec, f64


def test_propagate_matches_golden():
    print("[case] propagate vs reference fixtures", flush=True)
    g = np.load(os.path.join(GOLDEN, "propagate.npz"))
    in9 = g["in9"]
    rot = 0.37 * np.arange(len(in9))
    for order in (capi.TSC, capi.CIC):
        c, rec, _ = ctx("wide", order, capi.PREC_REF)
        for kind in (KINDS if order == capi.TSC else ("redshift",)):
            rows, dist = c.propagate(in9, 6, 0.01, rot, interpolation=kind, capacity=4000)
            for i in range(len(in9)):
                np.testing.assert_array_equal(rows[i], g[f"o{order}_{kind}_{i}_rows"], err_msg=f"{order} {kind} {i}")
                np.testing.assert_array_equal(dist[i], g[f"o{order}_{kind}_{i}_dist"], err_msg=f"{order} {kind} {i}")
        c.close()


@pytest.mark.parametrize("precision", (capi.PREC_REF, capi.PREC_FP64))
def test_propagate_matches_oracle(precision):
    c, rec, f64 = ctx("wide", capi.TSC, precision)
    O = Oracle()
    h = O.tree(rec, f64)
    tab = cases.cosmology()
    in9 = cases.rays(n_side=5, aperture=44.0, phi0=24.0, theta0=18.0)
    rng = np.random.default_rng(12)
    rot = rng.uniform(0, 2 * np.pi, len(in9))
    for count, angle, kind, amin in ((8, 0.004, "redshift", 0.0), (3, 0.05, "r", 0.6), (1, 0.02, "a", 0.0),
                                     (0, 0.01, "t", 0.0), (12, 0.001, "lambda", 0.3)):
        print(f"[case] propagate precision={precision} count={count} kind={kind} amin={amin}", flush=True)
        rows, dist = c.propagate(in9, count, angle, rot, interpolation=kind, amin=amin, capacity=4000)
        nacc = 0
        for i in range(len(in9)):
            ra, da = O.propagate(h, 2, tab, cases.LENGTH, cases.NSTEPS, in9[i], count, angle, rot[i],
                                 kind=KINDS[kind], amin=amin, precision=precision, maxlen=4000)
            np.testing.assert_array_equal(rows[i], ra, err_msg=f"photon {i}")
            np.testing.assert_array_equal(dist[i], da, err_msg=f"photon {i}")
            nacc += len(ra) > 0
        assert nacc > 0 or amin > 0.0   # the amin test may reject every bundle
    O.free(h)
    c.close()


def test_propagate_rejections_and_errors():
    print("[case] propagate rejections and errors", flush=True)
    c, rec, _ = ctx("amr", capi.TSC, capi.PREC_REF)
    rows, dist = c.propagate(cases.rays(n_side=3, aperture=10.0), 4, 0.01, capacity=4000)
    assert all(len(r) == 0 for r in rows)          # rays end before EXTENT / 4
    with pytest.raises(capi.MpfError):
        c.propagate(cases.rays(n_side=1), 4, 0.01, capacity=3)   # more elements than capacity
    c.close()
    cs, _, _ = ctx("wide", capi.TSC, capi.PREC_SPLIT)
    with pytest.raises(capi.MpfError):
        cs.propagate(cases.rays(n_side=1), 4, 0.01)             # whole trajectories: exact tiers only
    cs.close()
