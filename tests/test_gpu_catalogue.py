"""Batched catalogue root-finding on the GPU (catalogues.h:410-920) against
the oracle's restatement of iterateNewtonMethod/newtonMethod2d, which is
pinned bit for bit to the reference's own catalogues.h (compiled in
oracle/_ref, tests/test_ref_catalogue.py), and by the closed form phi == 0
(converged at the first iteration on the unlensed direction).

Tolerances: same status and iteration count; observed angles within 1e-12 rad
(the device's sines, cosines, atan2 and acos differ from libm in the last
bits); redshifts within 1e-12 relative; Jacobian within 1e-9 of |A - 1|.
Split tier against the FP64 tier: same status and iterations, angles within
1e-10 rad."""
import numpy as np
import pytest

import cases
from mpf import capi
from oracle import Oracle

pytestmark = pytest.mark.gpu


def sources(n, seed=5, rmin=0.1, rmax=0.35, spread=0.15):
    rng = np.random.default_rng(seed)
    d = rng.uniform(rmin, rmax, n)
    dirs = cases.DIR + rng.normal(scale=spread, size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    src = np.zeros((n, 6))
    src[:, :3] = dirs * d[:, None]
    src[:, 3:] = rng.normal(scale=1e-3, size=(n, 3))
    return src


def ctx(name, precision):
    rec, f64 = cases.tree(name)
    c = capi.Context(cases.LENGTH, order=capi.TSC, precision=precision, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    return c, rec, f64


def compare(a, b, ang_tol, z_tol, jac_tol):
    np.testing.assert_array_equal(a[:, capi.CAT["status"]], b[:, capi.CAT["status"]])
    np.testing.assert_array_equal(a[:, capi.CAT["iter"]], b[:, capi.CAT["iter"]])
    ok = b[:, capi.CAT["status"]] == capi.CAT_OK
    assert ok.sum() > 0.5 * len(b)
    for f in ("phi", "theta"):
        np.testing.assert_array_equal(a[:, capi.CAT[f]], b[:, capi.CAT[f]]) if ang_tol == 0 else \
            np.testing.assert_allclose(a[:, capi.CAT[f]], b[:, capi.CAT[f]], rtol=0, atol=1e-15)
    for f in ("phi_obs", "theta_obs"):
        np.testing.assert_allclose(a[ok, capi.CAT[f]], b[ok, capi.CAT[f]], rtol=0, atol=ang_tol)
    for f in ("z0", "z1", "z4"):
        np.testing.assert_allclose(a[ok, capi.CAT[f]], b[ok, capi.CAT[f]], rtol=z_tol)
    for f, one in (("a11", 1), ("a12", 0), ("a22", 1)):
        x, y = a[ok, capi.CAT[f]], b[ok, capi.CAT[f]]
        assert np.all(np.abs(x - y) <= 1e-13 + jac_tol * np.abs(y - one)), f


@pytest.mark.parametrize("precision", (capi.PREC_REF, capi.PREC_FP64))
def test_catalogue_vs_oracle(precision):
    print(f"[case] catalogue precision={precision}", flush=True)
    c, rec, f64 = ctx("amr", precision)
    O = Oracle()
    h = O.tree(rec, f64)
    src = sources(200)
    for acc in (1e-8, 1e-12):
        a, rays = c.catalogue_solve(src, accuracy=acc)
        b, rrays = O.catalogue(h, 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, src, accuracy=acc,
                               precision=precision)
        compare(a, b, 1e-12, 1e-12, 1e-9)
        assert rays >= rrays
    O.free(h)
    c.close()


def test_catalogue_split_vs_fp64():
    print("[case] catalogue split vs fp64", flush=True)
    src = sources(300, seed=6)
    c64, _, _ = ctx("coarse", capi.PREC_FP64)
    cs, _, _ = ctx("coarse", capi.PREC_SPLIT)
    a, _ = cs.catalogue_solve(src, accuracy=1e-8)
    b, _ = c64.catalogue_solve(src, accuracy=1e-8)
    compare(a, b, 1e-10, 1e-7, 1e-5)
    c64.close()
    cs.close()


def test_catalogue_flrw_first_iteration():
    print("[case] catalogue phi == 0", flush=True)
    from mpf import formats, synth
    lvl, kx, ky, kz = synth.uniform_level_coords(4, half=0.5)
    f = np.zeros((kx.size, 7))
    f[:, 5] = 1.0
    rec, _ = formats.make_records(lvl, kx, ky, kz, f)
    c = capi.Context(cases.LENGTH, order=capi.TSC, precision=capi.PREC_FP64)
    c.load_records(rec)
    c.set_cosmology(cases.cosmology())
    src = sources(50, seed=7, spread=0.05)
    out, rays = c.catalogue_solve(src, accuracy=1e-12)
    assert np.all(out[:, capi.CAT["status"]] == capi.CAT_OK)
    assert np.all(out[:, capi.CAT["iter"]] == 1)
    np.testing.assert_array_equal(out[:, capi.CAT["phi_obs"]], out[:, capi.CAT["phi"]])
    np.testing.assert_array_equal(out[:, capi.CAT["theta_obs"]], out[:, capi.CAT["theta"]])
    np.testing.assert_allclose(out[:, capi.CAT["a11"]], 1.0, atol=1e-15)
    c.close()


@pytest.mark.parametrize("bundle", [("redshift", "sachs", 1e-4), ("radius", "normal", 2e-4)])
def test_catalogue_bundle_vs_oracle(bundle):
    """jacobiantype = bundle: converged sources (their Jacobian is the
    bundle's) and sources run to max_iterations (rounds 5 on update with the
    bundle Jacobian), against the oracle's restatement in the FP64 tier."""
    print(f"[case] catalogue bundle={bundle}", flush=True)
    c, rec, f64 = ctx("amr", capi.PREC_FP64)
    O = Oracle()
    h = O.tree(rec, f64)
    kind = capi.STOP_NAMES[bundle[0]]
    plane = capi.BUNDLE_PLANES[bundle[1]]
    for acc, n in ((1e-8, 120), (1e-300, 40)):
        src = sources(n, seed=13)
        a, rays = c.catalogue_solve(src, accuracy=acc, bundle=bundle)
        b, _ = O.catalogue(h, 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, src, accuracy=acc, precision=1,
                           bundle=(kind, plane, bundle[2]))
        np.testing.assert_array_equal(a[:, capi.CAT["status"]], b[:, capi.CAT["status"]])
        np.testing.assert_array_equal(a[:, capi.CAT["iter"]], b[:, capi.CAT["iter"]])
        np.testing.assert_array_equal(a[:, capi.CAT["nrounds"]], b[:, 17])
        done = b[:, capi.CAT["status"]] != capi.CAT_OUTSIDE
        assert done.sum() > 0.5 * n
        for f in ("phi_obs", "theta_obs"):
            np.testing.assert_allclose(a[done, capi.CAT[f]], b[done, capi.CAT[f]], rtol=0, atol=1e-12)
        for f in ("a11", "a12", "a21", "a22"):
            np.testing.assert_allclose(a[done, capi.CAT[f]], b[done, capi.CAT[f]], rtol=0, atol=1e-9)
        if acc < 1e-100:
            assert np.all(a[done, capi.CAT["status"]] == capi.CAT_REJECTED)
    O.free(h)
    c.close()
