"""HEALPix RING scheme of libmpf (host functions), pinned by the mathematics
of the published scheme: the reference's HEALPix C++ 3.70 is absent, so its
outputs cannot be compared ("parity unpinned" against the library itself).

Checks: pixel centres equal the ring structure written out pixel by pixel
(oracle/healpix.py); every centre maps back to its own pixel; random
directions agree with an independent face/(x, y) route to the RING index;
pixel 0 sits at z = 1 - 1/(3 nside^2), phi = pi/4; equal areas."""
import numpy as np
import pytest

import healpix as hpo
from mpf import capi


@pytest.mark.parametrize("nside", (1, 2, 4, 8, 16, 64))
def test_centres_and_round_trip(nside):
    npix = 12 * nside * nside
    pix = np.arange(npix)
    v = capi.healpix_pix2vec(nside, pix)
    z, phi = hpo.ring_centres(nside)
    np.testing.assert_allclose(v[:, 2], z, rtol=0, atol=2e-16)
    vv = np.stack([np.sqrt(1 - z * z) * np.cos(phi), np.sqrt(1 - z * z) * np.sin(phi), z], axis=1)
    # (the expected sine loses bits near the poles: sqrt(1 - z^2) cancels)
    np.testing.assert_allclose(v, vv, rtol=0, atol=3e-15)
    np.testing.assert_array_equal(capi.healpix_vec2pix(nside, v), pix)


@pytest.mark.parametrize("nside", (1, 4, 32, 256, 2048))
def test_random_directions_vs_face_route(nside):
    rng = np.random.default_rng(nside)
    v = rng.normal(size=(200000, 3))
    a = capi.healpix_vec2pix(nside, v)
    b = hpo.vec2pix_ring(nside, v)
    assert np.mean(a == b) == 1.0, f"{np.sum(a != b)} disagreements"
    assert a.min() >= 0 and a.max() < 12 * nside * nside


def test_pixel_zero_and_equal_area():
    for nside in (1, 2, 16):
        v = capi.healpix_pix2vec(nside, [0])[0]
        assert abs(v[2] - (1 - 1 / (3 * nside * nside))) < 1e-15
        assert abs(np.arctan2(v[1], v[0]) - np.pi / 4) < 1e-15
    nside = 8
    rng = np.random.default_rng(0)
    v = rng.normal(size=(768 * 2000, 3))
    h = np.bincount(capi.healpix_vec2pix(nside, v), minlength=768)
    # multinomial counts around 2000 per pixel
    assert abs(h.mean() - 2000) < 1e-9 and h.std() < 2000 ** 0.5 * 1.25 and h.min() > 1700


def test_invalid_arguments():
    with pytest.raises(capi.MpfError):
        capi.healpix_pix2vec(4, [192])
    with pytest.raises(capi.MpfError):
        capi.healpix_pix2vec(0, [0])
