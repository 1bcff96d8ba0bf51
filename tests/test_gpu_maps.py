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
