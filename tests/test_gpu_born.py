"""The Born approximation on the GPU (mpf_integrate_surfaces_born[_dev]):
the homogeneous path of Integrator::integrate<-1> with the fused
infinitesimal Jacobian sampled on the octree's data, against the oracle
(pinned to the reference in tests/test_born.py): surface records bit for bit,
the Jacobian within 1e-12 + 1e-9 |A - 1| (the device's S1 - S2/D and its
own trigonometry, as for the other surfaces); the lensing_born maps of
hmaps.h:705-713; the infinitesimal_born columns of the catalogue
(catalogues.h:670-674)."""
import os

import numpy as np
import pytest

import cases
from mpf import capi
from oracle import Oracle

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("valid", "firstid", "f", "x", "y", "z", "dist", "chi", "lambda", "t", "a", "redshift", "s", "isw", "phi",
          "kt", "kx", "ky", "kz")


def ctx(name, precision):
    rec, f64 = cases.tree(name)
    c = capi.Context(cases.LENGTH, order=capi.TSC, precision=precision, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    return c, rec, f64


def jac_close(x, y, rel=1e-9, ab=1e-12):
    for k, one in ((0, 1.0), (1, 0.0), (3, 1.0)):
        tol = ab + rel * np.abs(y[..., k] - one)
        assert np.all(np.abs(x[..., k] - y[..., k]) <= tol), f"entry {k}: {np.max(np.abs(x[..., k] - y[..., k]) - tol)}"


@pytest.mark.parametrize("precision", (capi.PREC_REF, capi.PREC_FP64))
def test_born_surfaces_match_oracle(precision):
    c, rec, f64 = ctx("wide", precision)
    O = Oracle()
    h = O.tree(rec, f64)
    in9 = cases.rays(n_side=6, aperture=30.0)
    vals = [0.15, 0.3, 0.5]
    print(f"[case] born surfaces precision={precision}", flush=True)
    out, surf = c.integrate_surfaces_born(in9, "radius", vals)
    for i, r in enumerate(in9):
        ref = O.surfaces(h, -1, cases.cosmology(), cases.LENGTH, cases.NSTEPS, 0, vals, r, precision=precision)
        for f in FIELDS:
            np.testing.assert_array_equal(surf[i, :, capi.SURF[f]], ref[:, capi.SURF[f]], err_msg=f"ray {i} {f}")
        a = surf[i][:, [capi.SURF["a11"], capi.SURF["a12"], capi.SURF["a12"], capi.SURF["a22"]]]
        b = ref[:, [capi.SURF["a11"], capi.SURF["a12"], capi.SURF["a12"], capi.SURF["a22"]]]
        jac_close(a, b)
        # the path is straight: the direction never changes
        np.testing.assert_allclose(surf[i, :, capi.SURF["kx"]] / surf[i, :, capi.SURF["kz"]], r[6] / r[8], rtol=1e-12)
    O.free(h)
    c.close()


def test_born_jacobian_matches_reference_fixtures():
    """At radius surfaces, the fused Born Jacobian against the reference's own
    dbetadtheta_infinitesimal on its integrate<-1> trajectory."""
    print("[case] born jacobian vs reference fixtures", flush=True)
    g = np.load(os.path.join(GOLDEN, "born.npz"))
    for name in ("amr", "wide"):
        c, rec, _ = ctx(name, capi.PREC_REF)
        vals = [0.15, 0.25, 0.3]
        out, surf = c.integrate_surfaces_born(g[f"{name}_in9"], "radius", vals)
        a = surf[..., [capi.SURF["a11"], capi.SURF["a12"], capi.SURF["a12"], capi.SURF["a22"]]]
        assert np.all(surf[..., capi.SURF["valid"]] == 1.0)
        jac_close(a, g[f"{name}_jac"], rel=1e-8, ab=1e-12)
        c.close()


def test_lensing_born_maps_and_catalogue_born():
    print("[case] lensing_born maps and infinitesimal_born catalogue", flush=True)
    c, rec, f64 = ctx("wide", capi.PREC_SPLIT)
    nside = 32
    npix = 12 * nside * nside
    v = capi.healpix_pix2vec(nside, np.arange(npix))
    pix = np.nonzero(v @ cases.DIR > np.cos(np.radians(20)))[0].astype(np.int64)
    homo = np.array([[1e17, 0.2, 5e16, 0.1, 1 / 1.1], [2e17, 0.45, 1e17, 0.25, 1 / 1.25]])
    dp, din = c.dev_alloc(pix.nbytes), c.dev_alloc(pix.size * 72)
    c.upload(dp, pix)
    c.hmaps_launch_dev(nside, dp, pix.size, din)
    dout = c.dev_alloc(pix.size * capi.NOUT * 8)
    dsb = c.dev_alloc(pix.size * 2 * capi.NSURF * 8)
    c.integrate_surfaces_born_dev(din, pix.size, dout, dsb, "radius", list(homo[:, 1]), sync=True)
    nm = c.hmaps_nmaps(["lensing_born", "steps"])
    with pytest.raises(capi.MpfError):   # steps needs the ray's own surfaces
        c.hmaps_components_dev(nside, dp, None, pix.size, 2, ["lensing_born", "steps"], homo, 1, 10 ** 9,
                               surf_born_ptr=dsb)
    cap = 2 * 4 * npix
    dmaps = c.dev_alloc(cap * 4)
    c.upload(dmaps, np.zeros(cap, dtype=np.float32))
    assert c.hmaps_components_dev(nside, dp, None, pix.size, 2, ["lensing_born"], homo, dmaps, cap,
                                  surf_born_ptr=dsb) == 4
    maps = c.download(dmaps, (2, 4, npix), dtype=np.float32)
    sb = c.download(dsb, (pix.size, 2, capi.NSURF))
    for iz in range(2):
        s = sb[:, iz]
        ok = s[:, capi.SURF["valid"]] == 1.0
        assert ok.all()
        a11, a12, a22 = (s[:, capi.SURF[k]] for k in ("a11", "a12", "a22"))
        exp = np.stack([1 - (a11 + a22) * 0.5, -0.5 * (a11 - a22), -a12, a11 * a22 - a12 * a12]).astype(np.float32)
        np.testing.assert_array_equal(maps[iz][:, pix], exp)
        assert np.abs(maps[iz][0, pix]).max() > 1e-7     # the data's convergence, not zero
    for p in (dp, din, dout, dsb, dmaps):
        c.dev_free(p)
    # catalogue: same Newton solution, Born Jacobian columns
    rng = np.random.default_rng(2)
    n = 300
    d = rng.uniform(0.15, 0.45, n)
    dirs = cases.DIR + rng.normal(scale=0.12, size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    src = np.zeros((n, 6))
    src[:, :3] = dirs * d[:, None]
    plain, _ = c.catalogue_solve(src, accuracy=1e-9)
    born, rays = c.catalogue_solve(src, accuracy=1e-9, jacobian_born=True)
    keep = [capi.CAT[k] for k in ("phi", "theta", "phi_obs", "theta_obs", "sep", "iter", "z0", "z1", "z2", "z3",
                                  "z4", "z5", "status")]
    np.testing.assert_array_equal(born[:, keep], plain[:, keep])
    ok = born[:, capi.CAT["status"]] == capi.CAT_OK
    assert ok.sum() > 0.9 * n
    ang = born[ok][:, [capi.CAT["phi_obs"], capi.CAT["theta_obs"]]]
    in9 = capi.launch_angles(ang)
    _, sb = c.integrate_surfaces_born(in9, "radius", None, ray_value=d[ok])
    for f in ("a11", "a12", "a22"):
        np.testing.assert_allclose(born[ok, capi.CAT[f]], sb[:, 0, capi.SURF[f]], rtol=0, atol=1e-13)
    np.testing.assert_array_equal(born[ok, capi.CAT["a21"]], born[ok, capi.CAT["a12"]])
    c.close()
