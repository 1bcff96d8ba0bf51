"""HEALPix map kernels on the GPU: pixel-centre launches (hmaps.h:599-607),
per-pixel lensing maps (hmaps.h:724-762) and binning of rays by observed
direction with wavefront-reduced fixed-point atomics (deterministic)."""
import numpy as np
import pytest

import cases
from mpf import capi

pytestmark = pytest.mark.gpu


def ctx(name="amr", precision=capi.PREC_SPLIT):
    rec, f64 = cases.tree(name)
    c = capi.Context(cases.LENGTH, order=capi.TSC, precision=precision, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    return c


def test_pixel_launch_matches_host():
    print("[case] pixel launch", flush=True)
    c = ctx()
    nside = 64
    pix = np.arange(0, 12 * nside * nside, 7, dtype=np.int64)
    dp = c.dev_alloc(pix.nbytes)
    c.upload(dp, pix)
    din = c.dev_alloc(pix.size * 72)
    c.hmaps_launch_dev(nside, dp, pix.size, din)
    in9 = c.download(din, (pix.size, 9))
    v = capi.healpix_pix2vec(nside, pix)
    v /= np.linalg.norm(v, axis=1)[:, None]
    np.testing.assert_allclose(in9[:, 6:9], v, rtol=0, atol=2e-15)
    np.testing.assert_array_equal(in9[:, [0, 5]], 1.0)
    np.testing.assert_array_equal(in9[:, 1:5], 0.0)
    c.dev_free(dp)
    c.dev_free(din)
    c.close()


@pytest.mark.parametrize("precision", (capi.PREC_FP64, capi.PREC_SPLIT))
def test_pixel_maps(precision):
    print(f"[case] pixel-centre maps precision={precision}", flush=True)
    c = ctx(precision=precision)
    nside = 32
    npix = 12 * nside * nside
    # pixels whose centres fall in the stored region's cone
    v = capi.healpix_pix2vec(nside, np.arange(npix))
    pix = np.nonzero(v @ cases.DIR > np.cos(np.radians(25)))[0].astype(np.int64)
    zs = [0.1, 0.25]
    dp, din = c.dev_alloc(pix.nbytes), c.dev_alloc(pix.size * 72)
    c.upload(dp, pix)
    c.hmaps_launch_dev(nside, dp, pix.size, din)
    dout = c.dev_alloc(pix.size * capi.NOUT * 8)
    dsurf = c.dev_alloc(pix.size * len(zs) * capi.NSURF * 8)
    c.integrate_surfaces_dev(din, pix.size, dout, dsurf, "redshift", zs, jacobian=True, sync=True)
    dmaps = c.dev_alloc(len(zs) * 4 * npix * 8)
    c.upload(dmaps, np.full(len(zs) * 4 * npix, -7.0))
    c.hmaps_fill_dev(nside, dp, dsurf, pix.size, len(zs), dmaps)
    maps = c.download(dmaps, (len(zs), 4, npix))
    surf = c.download(dsurf, (pix.size, len(zs), capi.NSURF))
    for j in range(len(zs)):
        s = surf[:, j]
        ok = s[:, capi.SURF["valid"]] == 1.0
        a11, a12, a22 = (s[:, capi.SURF[k]] for k in ("a11", "a12", "a22"))
        exp = np.stack([1 - (a11 + a22) * 0.5, -0.5 * (a11 - a22), -a12, a11 * a22 - a12 * a12])
        exp[:, ~ok] = 0.0
        np.testing.assert_array_equal(maps[j][:, pix], exp)
        others = np.ones(npix, bool)
        others[pix] = False
        assert np.all(maps[j][:, others] == -7.0)
        assert ok.mean() > 0.5
        kappa = maps[j, 0, pix[ok]]
        assert np.all(np.isfinite(kappa)) and np.abs(kappa).max() < 0.5
    for p in (dp, din, dout, dsurf, dmaps):
        c.dev_free(p)
    c.close()


def test_binning_deterministic_and_exact():
    print("[case] healpix binning", flush=True)
    c = ctx()
    from mpf import synth
    nside = 64
    npix = 12 * nside * nside
    ph, th = synth.cone_uniform(20.0, 20.0, 30.0, 20000, seed=3)
    in9 = synth.launch_photons(ph, th)
    zs = np.random.default_rng(4).uniform(0.05, 0.3, in9.shape[0])
    n = in9.shape[0]
    din = c.dev_alloc(in9.nbytes)
    c.upload(din, in9)
    drv = c.dev_alloc(zs.nbytes)
    c.upload(drv, zs)
    dout = c.dev_alloc(n * capi.NOUT * 8)
    dsurf = c.dev_alloc(n * capi.NSURF * 8)
    c.integrate_surfaces_dev(din, n, dout, dsurf, "redshift", [0.0], jacobian=True, sync=True, ray_value_ptr=drv)
    surf = c.download(dsurf, (n, capi.NSURF))
    out = c.download(dout, (n, capi.NOUT))
    ok = surf[:, capi.SURF["valid"]] == 1.0
    assert ok.mean() > 0.5
    # each ray stops just past its own redshift and its surface sits at it
    np.testing.assert_allclose(surf[ok, capi.SURF["redshift"]], zs[ok], rtol=1e-12)
    assert np.all(out[ok, capi.OUT["redshift"]] > zs[ok])
    dmaps = c.dev_alloc(4 * npix * 8)
    c.healpix_bin_dev(nside, din, dsurf, n, 1, 0, dmaps)
    m1 = c.download(dmaps, (4, npix))
    c.healpix_bin_dev(nside, din, dsurf, n, 1, 0, dmaps)
    m2 = c.download(dmaps, (4, npix))
    np.testing.assert_array_equal(m1, m2)          # deterministic
    pix = capi.healpix_vec2pix(nside, in9[:, 6:9])
    cnt = np.bincount(pix[ok], minlength=npix)
    np.testing.assert_array_equal(m1[3], cnt)
    a11, a12, a22 = (surf[ok, capi.SURF[k]] for k in ("a11", "a12", "a22"))
    for k, q in enumerate((1 - (a11 + a22) * 0.5, -0.5 * (a11 - a22), -a12)):
        s = np.bincount(pix[ok], weights=q, minlength=npix)
        mean = np.where(cnt > 0, s / np.maximum(cnt, 1), 0.0)
        np.testing.assert_allclose(m1[k], mean, rtol=0, atol=1e-12)
    for p in (din, drv, dout, dsurf, dmaps):
        c.dev_free(p)
    c.close()


COMPONENTS = ["lensing", "dr", "dl", "dt", "da", "dz", "ds", "isw", "dens", "dens_max", "steps", "phi", "deflection"]


def homogeneous_ref_ray(tab):
    """The homogeneous reference photon of hmaps.cpp:167-171 (Order -1)."""
    from test_gpu_trajectory import homogeneous_level5
    rec = homogeneous_level5()
    c = capi.Context(cases.LENGTH, order=capi.HOMOGENEOUS, precision=capi.PREC_REF, nsteps=256)
    c.load_records(rec)
    c.set_cosmology(tab)
    out, trajs = c.trajectory(np.array([[1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0]]), "none", capacity=1 << 20)
    c.close()
    return trajs[0]


def component_run(c, nside, pix, zs, homo, jac=None):
    npix = 12 * nside * nside
    dp, din = c.dev_alloc(pix.nbytes), c.dev_alloc(pix.size * 72)
    c.upload(dp, pix)
    c.hmaps_launch_dev(nside, dp, pix.size, din)
    dout = c.dev_alloc(pix.size * capi.NOUT * 8)
    dsurf = c.dev_alloc(pix.size * len(zs) * capi.NSURF * 8)
    c.integrate_surfaces_dev(din, pix.size, dout, dsurf, "redshift", zs, jacobian=True, sync=True)
    nm = c.hmaps_nmaps(COMPONENTS, bundle=jac is not None)
    cap = len(zs) * nm * npix
    dmaps = c.dev_alloc(cap * 4)
    c.upload(dmaps, np.full(cap, -7.0, dtype=np.float32))
    djac = None
    if jac is not None:
        djac = c.dev_alloc(jac.nbytes)
        c.upload(djac, jac)
    with pytest.raises(capi.MpfError):   # a buffer one map too small is refused
        c.hmaps_components_dev(nside, dp, dsurf, pix.size, len(zs), COMPONENTS, homo, dmaps, cap - npix, djac)
    assert c.hmaps_components_dev(nside, dp, dsurf, pix.size, len(zs), COMPONENTS, homo, dmaps, cap, djac) == nm
    maps = c.download(dmaps, (cap,), dtype=np.float32)
    surf = c.download(dsurf, (pix.size, len(zs), capi.NSURF))
    in9 = c.download(din, (pix.size, 9))
    for p in (dp, din, dout, dsurf, dmaps) + ((djac,) if djac else ()):
        c.dev_free(p)
    return nm, maps.reshape(len(zs), nm, npix), surf, in9


def test_component_maps_homogeneous_closed_forms():
    """phi = 0 everywhere: straight rays on the homogeneous background, so
    convergence, shear, deflection, the perturbations of chi, lambda, t, a,
    z, s, ISW, density and potential vanish, and 1/mu = 1."""
    print("[case] map components, homogeneous universe", flush=True)
    from test_gpu_trajectory import homogeneous_level5
    tab = synth_table()
    rec = homogeneous_level5()
    c = capi.Context(cases.LENGTH, order=capi.TSC, precision=capi.PREC_FP64, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.set_cosmology(tab)
    nside = 16
    pix = np.arange(0, 12 * nside * nside, 5, dtype=np.int64)
    zs = [0.2, 0.6]
    homo, _ = capi.hmaps_homogeneous(homogeneous_ref_ray(tab), zs)
    nm, maps, surf, _ = component_run(c, nside, pix, zs, homo)
    c.close()
    assert nm == 4 + 11 + 2
    for iz in range(len(zs)):
        m = maps[iz][:, pix]
        assert np.all(surf[:, iz, capi.SURF["valid"]] == 1.0)
        k, g1, g2, imu = m[0], m[1], m[2], m[3]
        assert np.abs(k).max() < 1e-9 and np.abs(g1).max() < 1e-9 and np.abs(g2).max() < 1e-9
        np.testing.assert_allclose(imu, 1.0, atol=1e-6)
        for j, name in enumerate(("dr", "dl", "dt", "da", "dz", "ds")):
            assert np.abs(m[4 + j]).max() < 2e-4, f"{name}: {np.abs(m[4 + j]).max():.2e}"
        for j in (10, 11, 12, 14):   # isw, dens, dens_max, phi
            assert np.all(m[j] == 0.0)
        np.testing.assert_array_equal(m[13], surf[:, iz, capi.SURF["firstid"]].astype(np.float32))
        assert np.abs(m[15]).max() < 1e-6 and np.abs(m[16]).max() < 1e-6   # deflection (float)


def synth_table():
    from mpf import synth
    return synth.flat_lcdm_table(h=0.7, omega_m=0.3, a_min=0.05, n=2048)


@pytest.mark.parametrize("bundle", (False, True))
def test_component_maps_layout(bundle):
    """Map layout and formulas of hmaps.h:716-836 on an inhomogeneous tree:
    each map against the surface records (FP32 rounding of the FP64 formula;
    the transcendental ones within 2 FP32 ulps); rejected surfaces zero every
    map of that surface; other pixels untouched."""
    print(f"[case] map components layout bundle={bundle}", flush=True)
    c = ctx(precision=capi.PREC_FP64)
    nside = 32
    npix = 12 * nside * nside
    v = capi.healpix_pix2vec(nside, np.arange(npix))
    pix = np.nonzero(v @ cases.DIR > np.cos(np.radians(25)))[0].astype(np.int64)
    zs = [0.1, 0.25]
    tab = cases.cosmology()
    homo = np.array([[1e17, 0.3, 5e16, z, 1 / (1 + z)] for z in zs])
    jac = None
    if bundle:
        ang = np.c_[np.arctan2(v[pix, 1], v[pix, 0]), np.arccos(v[pix, 2])]
        d = np.tile(np.array([0.2, 0.3]), (pix.size, 1))
        jac = c.jacobian_bundle(ang, "redshift", zs, np.repeat(v[pix][:, None, :], 2, axis=1), d, 0.004)
    nm, maps, surf, in9 = component_run(c, nside, pix, zs, homo, jac)
    c.close()
    assert nm == (5 if bundle else 4) + 11 + 2
    f32 = np.float32
    for iz in range(len(zs)):
        s = surf[:, iz]
        ok = s[:, capi.SURF["valid"]] == 1.0
        if bundle:
            ok &= jac[:, iz, 0] != 42
            a11, a12, a21, a22 = (jac[:, iz, k] for k in range(4))
            w = -np.arctan2(a12 - a21, a11 + a22)
            lens = [1 - (a11 + a22) * 0.5 / np.cos(w), -0.5 * (np.cos(w) * (a11 - a22) + np.sin(w) * (a12 + a21)),
                    -0.5 * (np.cos(w) * (a12 + a21) + np.sin(w) * (a22 - a11)), w, a11 * a22 - a12 * a21]
        else:
            a11, a12, a22 = (s[:, capi.SURF[k]] for k in ("a11", "a12", "a22"))
            lens = [1 - (a11 + a22) * 0.5, -0.5 * (a11 - a22), -a12, a11 * a22 - a12 * a12]
        h = homo[iz]
        g = lambda k: s[:, capi.SURF[k]]
        x, y, z = g("x"), g("y"), g("z")
        r = np.sqrt(x * x + y * y + z * z)
        phi_p = np.arctan2(v[pix, 1], v[pix, 0])
        th_p = np.arccos(v[pix, 2])
        rest = [g("chi") / h[1] - 1, g("lambda") / h[2] - 1, g("t") / h[0] - 1, g("a") / h[4] - 1,
                g("redshift") / h[3] - 1, g("s") / h[1] - 1, g("isw"), g("rho"), g("rhomax"), g("firstid"),
                g("phi"), phi_p - np.arctan2(y, x), th_p - np.arccos(z / r)]
        # FP64 formulas without transcendental functions round identically;
        # cos/sin/atan2/acos may differ in the last FP64 bit between the device
        # and numpy
        exact = list(range(len(lens) + 11)) if not bundle else [4] + list(range(5, 16))
        expect = np.stack(lens + rest).astype(f32)
        expect[:, ~ok] = 0.0
        got = maps[iz][:, pix]
        for j in range(nm):
            if j in exact:
                np.testing.assert_array_equal(got[j], expect[j], err_msg=f"surface {iz} map {j}")
            else:
                ulp = np.spacing(np.abs(expect[j]).astype(f32))
                assert np.all(np.abs(got[j] - expect[j]) <= 2 * ulp + 1e-30), f"surface {iz} map {j}"
        others = np.ones(npix, bool)
        others[pix] = False
        assert np.all(maps[iz][:, others] == -7.0)
        assert ok.mean() > 0.5
