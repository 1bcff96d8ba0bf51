"""The device root-finder in the REF tier against fixtures written by the
reference's own catalogues.h (tests/golden/catalogue.npz, oracle/_ref,
tests/golden/make_golden.py): every jacobiantype, stop_bundle and plane, an
observer velocity, sources that converge after Jacobian-based updates,
rejected sources, rays that miss their source and failed bundles; the grid
re-run of rejected sources (use_previous_catalogues = 2) and
relCat_with_previous_cat (mpf_catalogue_previous).

Tolerances: status and iteration count exact (and the firstid of rays that
miss their source); observed angles within 1e-12 rad and redshifts within
1e-12 relative (the device's sin, cos, atan2 and acos differ from glibc's in
the last bits); the infinitesimal Jacobian within 1e-9 of |A - 1| (the fused
S1 - S2/D accumulation rounds differently from the reference's
per-distance sum); a bundle Jacobian, a difference of four ray positions over
2 D tan(opening), within 2e-16 / (D tan(opening)): positions that differ in
their last bits (the launch direction above) enter it amplified by that
divisor.  Failure entries exact."""
import os
import sys

import numpy as np
import pytest

import cases
from mpf import capi

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

CAT = np.load(os.path.join(HERE, "golden", "catalogue.npz"))
VOBS = tuple(CAT["vobs"])
_ctx = {}


def ctx(tree):
    if tree not in _ctx:
        rec, _ = cases.tree(tree)
        c = capi.Context(cases.LENGTH, order=capi.TSC, precision=capi.PREC_REF, nsteps=cases.NSTEPS)
        c.load_records(rec)
        c.set_cosmology(cases.cosmology())
        _ctx[tree] = c
    return _ctx[tree]


def newton_kw(jt, bundle):
    kw = dict(vobs=VOBS)
    if jt == "infinitesimal_born":
        kw["jacobian_born"] = True
    if bundle is not None:
        kw["bundle"] = bundle
    return kw


def compare(a, ref, what, ang_tol=1e-12, z_tol=1e-12, jac_tol=1e-9, jac_abs=None):
    """a: device rows (MPF_CAT_*); ref: refc_newton rows of the fixture;
    jac_abs: per-source absolute Jacobian tolerance (bundles)."""
    C = capi.CAT
    outside = (ref[:, 0] == 42) & (ref[:, 1] == 42)
    rejected = ~outside & (ref[:, 3] == 42)
    ok = ~outside & ~rejected
    status = np.where(outside, capi.CAT_OUTSIDE, np.where(rejected, capi.CAT_REJECTED, capi.CAT_OK))
    bad = np.flatnonzero(a[:, C["status"]] != status)
    assert bad.size == 0, f"{what}: status differs on {bad.size} sources {bad[:8]}"
    failed_bundle = outside & (ref[:, 2] == 22)
    it_ok = ~failed_bundle
    bad = np.flatnonzero(a[it_ok, C["iter"]] != ref[it_ok, 3])
    assert bad.size == 0, f"{what}: iterations differ on {bad.size} sources"
    np.testing.assert_array_equal(a[outside, C["sep"]], ref[outside, 2])
    live = ~outside
    for col, rc in (("phi_obs", 0), ("theta_obs", 1)):
        np.testing.assert_allclose(a[live, C[col]], ref[live, rc], rtol=0, atol=ang_tol, err_msg=f"{what} {col}")
    # separations: tiny numbers from a difference of positions, absolute in units of the angles' error
    np.testing.assert_allclose(a[live, C["sep"]], ref[live, 2], rtol=0, atol=ang_tol, err_msg=f"{what} sep")
    for k in range(6):   # z1 of rejected sources too; the others are 0 there
        col = C[f"z{k}"]
        np.testing.assert_allclose(a[live, col], ref[live, 8 + k], rtol=z_tol, atol=1e-300, err_msg=f"{what} z{k}")
    jac = np.flatnonzero(live)
    extra = 0.0 if jac_abs is None else jac_abs[jac]
    for q, (col, one) in enumerate((("a11", 1), ("a12", 0), ("a21", 0), ("a22", 1))):
        x, y = a[jac, C[col]], ref[jac, 4 + q]
        assert np.all(np.abs(x - y) <= 1e-13 + extra + jac_tol * np.abs(y - one)), f"{what} {col}"
    np.testing.assert_array_equal(a[failed_bundle, C["a11"]], ref[failed_bundle, 4])


def bundle_tol(src, bundle):
    if bundle is None:
        return None
    return 2e-16 / (np.linalg.norm(src[:, :3], axis=1) * np.tan(bundle[2]))


@pytest.mark.parametrize("run", make_golden.CAT_RUNS, ids=lambda r: make_golden.cat_name(*r[:4]))
def test_device_newton_matches_reference(run):
    tree, acc, jt, bundle = run[:4]
    name = make_golden.cat_name(tree, acc, jt, bundle)
    print(f"[case] catalogue {name}", flush=True)
    src = CAT[f"src_{name}"]
    a, rays = ctx(tree).catalogue_solve(src, accuracy=acc, **newton_kw(jt, bundle))
    compare(a, CAT[f"ref_{name}"], name, jac_abs=bundle_tol(src, bundle))
    assert rays > 0


@pytest.mark.parametrize("jt,bundle", (("infinitesimal", None), ("bundle", ("redshift", "sachs", 1e-4))))
def test_device_grid_rerun_matches_reference(jt, bundle):
    name = "rerun_" + make_golden.cat_name("strong", 3e-12, jt, bundle)
    print(f"[case] catalogue {name}", flush=True)
    a, rays = ctx("strong").catalogue_solve(CAT[f"src_{name}"], init2=CAT[f"init_{name}"], accuracy=3e-12,
                                            max_iterations=100, use_previous=2, **newton_kw(jt, bundle))
    ref = CAT[f"ref_{name}"]
    compare(a, ref, name, jac_abs=bundle_tol(CAT[f"src_{name}"], bundle))
    # a rejected source of the re-run mode carries no Jacobian
    rej = a[:, capi.CAT["status"]] == capi.CAT_REJECTED
    assert rej.any() and np.all(a[rej, capi.CAT["a11"]:capi.CAT["a22"] + 1] == 0)


@pytest.mark.parametrize("jt,bundle", (("infinitesimal", None), ("infinitesimal_born", None),
                                       ("bundle", ("redshift", "sachs", 1e-4)), ("bundle", ("a", "normal", 1e-4)),
                                       ("bundle", ("plane", "sachs", 1e-4))))
def test_device_previous_catalogue_matches_reference(jt, bundle):
    name = "prev_" + make_golden.cat_name("strong", 0, jt, bundle)
    print(f"[case] catalogue {name}", flush=True)
    prev = CAT["prev18"]
    rows = ctx("strong").catalogue_previous(prev, **newton_kw(jt, bundle))
    ref = CAT[f"ref_{name}"]
    np.testing.assert_array_equal(rows[:, :13], ref[:, :13])
    np.testing.assert_array_equal(rows[:, 17], ref[:, 17])
    # the bundle divides by the crossing's distance, at least 0.1 for these sources
    extra = 0.0 if bundle is None else 2e-16 / (0.1 * np.tan(bundle[2]))
    for q, one in enumerate((1, 0, 0, 1)):
        x, y = rows[:, 13 + q], ref[:, 13 + q]
        assert np.all(np.abs(x - y) <= 1e-13 + extra + 1e-9 * np.abs(y - one)), (name, q)
