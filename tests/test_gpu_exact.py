"""GPU parity of the exact tiers through the C ABI (libmpf.so).

MPF_PREC_REF must be bit-identical to the reference: checked against the
golden fixtures written from the reference's own code and against the oracle
on more rays.  MPF_PREC_FP64 must be bit-identical to the oracle's FP64 mode
(same operations, FP64 fields)."""
import os

import numpy as np
import pytest

import cases
from mpf import capi
from oracle import Oracle

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def ctx_for(name, order=2, precision=capi.PREC_REF):
    rec, f64 = cases.tree(name)
    c = capi.Context(cases.LENGTH, order=order, precision=precision, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    return c, rec, f64


@pytest.mark.parametrize("name", cases.TREES)
def test_locate_interp_golden(name):
    print(f"[case] locate/interp tree={name}", flush=True)
    c, rec, _ = ctx_for(name)
    g = np.load(os.path.join(GOLDEN, f"tree_{name}.npz"))
    e, lv = c.locate(g["points"])
    np.testing.assert_array_equal(e, g["locate_elem"])
    np.testing.assert_array_equal(lv, g["locate_level"])
    for order in (0, 1, 2, 3):
        np.testing.assert_array_equal(c.interpolate(order, g["points"]), g[f"interp{order}"],
                                      err_msg=f"order {order}")
    c.close()


@pytest.mark.parametrize("name", cases.TREES)
def test_integrate_golden(name):
    g = np.load(os.path.join(GOLDEN, f"tree_{name}.npz"))
    for order in (1, 2):
        c, _, _ = ctx_for(name, order=order)
        for rs in cases.RAYSETS:
            for stop, val in cases.STOPS:
                key = f"rays_{rs}_{stop}_o{order}"
                if key not in g:
                    continue
                print(f"[case] integrate tree={name} {key}", flush=True)
                kit = tuple(cases.DIR) if stop == "plane" else (0.0, 0.0, 0.0)
                out, steps = c.integrate(g[f"in9_{rs}"], stop, val, kit)
                np.testing.assert_array_equal(out, g[key], err_msg=key)
                assert steps == int(np.sum(np.maximum(g[key][:, capi.OUT["nsteps"]], 0)))
        c.close()


@pytest.mark.parametrize("name", ("amr", "coarse"))
@pytest.mark.parametrize("precision", (capi.PREC_REF, capi.PREC_FP64))
def test_integrate_vs_oracle_many_rays(name, precision):
    print(f"[case] integrate-vs-oracle tree={name} precision={precision}", flush=True)
    c, rec, f64 = ctx_for(name, precision=precision)
    O = Oracle()
    h = O.tree(rec, f64)
    in9 = cases.rays(n_side=40, aperture=50.0)
    tab = cases.cosmology()
    for stop, val in (("redshift", 0.35), ("none", 0.0)):
        out, steps = c.integrate(in9, stop, val)
        ref, rsteps = O.integrate(h, 2, tab, cases.LENGTH, cases.NSTEPS, cases.STOP_KIND[stop], val, in9,
                                  precision=precision, nthreads=8)
        np.testing.assert_array_equal(out, ref)
        assert steps == rsteps
    O.free(h)
    c.close()


def test_fp64_interp_vs_oracle():
    print("[case] fp64 interpolation", flush=True)
    c, rec, f64 = ctx_for("amr", precision=capi.PREC_FP64)
    O = Oracle()
    h = O.tree(rec, f64)
    p = cases.points(rec, n=2000, seed=4)
    for order in (0, 1, 2, 3):
        np.testing.assert_array_equal(c.interpolate(order, p), O.interp(h, order, p, precision=1))
    O.free(h)
    c.close()


def test_empty_and_errors():
    print("[case] empty inputs and error paths", flush=True)
    c, rec, _ = ctx_for("uniform")
    out, steps = c.integrate(np.zeros((0, 9)), "redshift", 0.3)
    assert out.shape == (0, capi.NOUT) and steps == 0
    e, lv = c.locate(np.zeros((0, 3)))
    assert e.size == 0
    with pytest.raises(capi.MpfError):
        c.integrate(cases.rays(2), 17, 0.3)
    with pytest.raises(capi.MpfError):
        c.load_fields64(np.zeros((3, 7)))
    bad = rec.copy()
    bad[[1, 2]] = bad[[2, 1]]
    with pytest.raises(capi.MpfError):
        c.load_records(bad)
    c.close()
    c2 = capi.Context(cases.LENGTH)
    with pytest.raises(capi.MpfError):
        c2.integrate(cases.rays(2), "redshift", 0.3)  # no octree loaded
    c2.close()


def test_rays_outside_and_holes():
    """Rays launched outside the stored region or into the hole end with the
    reference's sentinels (single launch element / immediate stop)."""
    print("[case] rays outside the data", flush=True)
    c, rec, f64 = ctx_for("amr")
    O = Oracle()
    h = O.tree(rec, f64)
    in9 = cases.rays(6, aperture=30.0)
    in9[:, 2:5] = [[0.7, 0.0, 0.0]] * 18 + [[0.38 * d for d in cases.DIR]] * 18
    tab = cases.cosmology()
    out, steps = c.integrate(in9, "redshift", 0.3)
    ref, rsteps = O.integrate(h, 2, tab, cases.LENGTH, cases.NSTEPS, 2, 0.3, in9)
    np.testing.assert_array_equal(out, ref)
    O.free(h)
    c.close()
