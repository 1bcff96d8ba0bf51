"""The oracle's catalogue with jacobiantype = bundle (catalogues.h:455-491,
509-523, 666-681), CPU only.  The oracle's bundle branch is pinned bit for
bit to the reference's own catalogues.h (compiled in oracle/_ref,
tests/test_ref_catalogue.py); this file adds what the mathematics fixes,
by closed forms and by a second, independent estimate of the same matrix:
  * phi == 0: every source converges at the first iteration on its unlensed
    direction and the bundle Jacobian is cos(opening) times the identity (the
    chord over the tangent offset of the four rays);
  * the bundle Jacobian of a converged source (a finite difference of four
    rays) agrees with the infinitesimal one (the tidal integral along the
    central ray) to the size of the lensing, 5e-4 here;
  * the Newton path before iteration 5 is the infinitesimal mode's: same
    status, iterations and observed angles, bit for bit."""
import numpy as np

import cases
from mpf import formats, synth
from oracle import Oracle

TAB = cases.cosmology()
ST, IT, A = 16, 5, slice(12, 16)


def sources(n, seed=5, rmin=0.1, rmax=0.35, spread=0.15):
    rng = np.random.default_rng(seed)
    d = rng.uniform(rmin, rmax, n)
    dirs = cases.DIR + rng.normal(scale=spread, size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    src = np.zeros((n, 6))
    src[:, :3] = dirs * d[:, None]
    src[:, 3:] = rng.normal(scale=1e-3, size=(n, 3))
    return src


def test_bundle_catalogue_flrw():
    lvl, kx, ky, kz = synth.uniform_level_coords(4, half=0.5)
    f = np.zeros((kx.size, 7))
    f[:, 5] = 1.0
    rec, _ = formats.make_records(lvl, kx, ky, kz, f)
    O = Oracle()
    h = O.tree(rec)
    src = sources(40, seed=7, spread=0.05)
    for kind, plane in ((2, 0), (0, 1), (4, 0)):
        out, _ = O.catalogue(h, 2, TAB, cases.LENGTH, cases.NSTEPS, src, accuracy=1e-12, precision=1,
                             bundle=(kind, plane, 1e-4))
        assert np.all(out[:, ST] == 0)
        assert np.all(out[:, IT] == 1)
        np.testing.assert_array_equal(out[:, 2:4], out[:, 0:2])
        # straight rays: the four rays leave at +-tan(opening) off the axis, so the finite difference is
        # the chord 2 D sin(opening) over 2 D tan(opening): cos(opening) on the diagonal
        c = np.cos(1e-4)
        np.testing.assert_allclose(out[:, A], np.tile([c, 0.0, 0.0, c], (len(src), 1)), rtol=0, atol=1e-11)
    O.free(h)


def test_bundle_catalogue_against_infinitesimal():
    rec, f64 = cases.tree("amr")
    O = Oracle()
    h = O.tree(rec, f64)
    src = sources(200)
    for acc in (1e-8, 1e-12):
        inf, _ = O.catalogue(h, 2, TAB, cases.LENGTH, cases.NSTEPS, src, accuracy=acc, precision=1)
        for kind, plane in ((2, 0), (0, 1), (3, 0)):
            b, _ = O.catalogue(h, 2, TAB, cases.LENGTH, cases.NSTEPS, src, accuracy=acc, precision=1,
                               bundle=(kind, plane, 1e-4))
            np.testing.assert_array_equal(b[:, ST], inf[:, ST])
            np.testing.assert_array_equal(b[:, IT], inf[:, IT])
            np.testing.assert_array_equal(b[:, 2:4], inf[:, 2:4])
            ok = b[:, ST] == 0
            assert ok.sum() > 100
            assert np.abs(b[ok, A] - inf[ok, A]).max() < 5e-4
            assert np.abs(b[ok, A] - inf[ok, A]).max() > 0      # a different estimate
    O.free(h)


def test_bundle_catalogue_late_iterations():
    """accuracy out of reach: every source that reaches its radius runs all
    max_iterations rounds, rounds 5 on with the bundle Jacobian, and is
    rejected with that Jacobian in its columns."""
    rec, f64 = cases.tree("amr")
    O = Oracle()
    h = O.tree(rec, f64)
    src = sources(60, seed=9)
    b, _ = O.catalogue(h, 2, TAB, cases.LENGTH, cases.NSTEPS, src, accuracy=1e-300, precision=1,
                       bundle=(2, 0, 1e-4))
    inf, _ = O.catalogue(h, 2, TAB, cases.LENGTH, cases.NSTEPS, src, accuracy=1e-300, precision=1)
    rej = b[:, ST] == 2
    assert rej.sum() > 30
    assert np.all(b[rej, IT] == 42) and np.all(b[rej, 17] == 11)
    np.testing.assert_array_equal(b[:, ST], inf[:, ST])
    assert np.abs(b[rej, A] - inf[rej, A]).max() < 5e-4
    # the observed angles follow the bundle updates from round 5 on: close to, not equal to, the infinitesimal ones
    assert np.abs(b[rej, 2:4] - inf[rej, 2:4]).max() < 1e-9
    O.free(h)
