"""Cones and photon launch (host code of libmpf, no GPU): the product's
mpf_cones_* / mpf_cone_* / mpf_launch_* against fixtures written by the
reference's own headers (HyperSphere::uniform, Cone::inside,
Integrator::launch, tests/golden/make_golden.py), against the compiled
reference where it is present, and against closed forms.  GenerateNarrowCones
and getTargets (miscellaneous.h) are pinned bit for bit to the reference's
own code in tests/test_ref_catalogue.py."""
import math
import os

import numpy as np
import pytest

from mpf import capi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def g():
    return np.load(os.path.join(GOLDEN, "cones.npz"))


def unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


# ------------------------------------------------ spiral and full-sky cones

def test_fullsky_bases_are_the_reference_spiral(g):
    for n in (1, 2, 5, 12, 48, 300):
        c = capi.cones_fullsky(n)
        np.testing.assert_array_equal(c[:, 3:6], g[f"uniform_{n}"])
        assert np.all(c[:, :3] == 0.0)


def test_fullsky_angle_from_smallest_spiral_distance(g):
    # 1.8 asin(d_min / diameter) with d_min the first member of uniform()'s pair
    for n in (1, 2, 5, 12, 48, 300):
        c = capi.cones_fullsky(n)
        assert np.all(c[:, 6] == c[0, 6])
        assert c[0, 6] == 1.8 * math.asin(g[f"uniform_minmax_{n}"][0] / 4.0)
    # brute-force pair distance, independent of the spiral's own loop
    for n in (5, 48):
        p = g[f"uniform_{n}"]
        d = np.linalg.norm(p[:, None] - p[None], axis=-1)
        d[np.arange(n), np.arange(n)] = np.inf
        assert abs(capi.cones_fullsky(n)[0, 6] - 1.8 * math.asin(d.min() / 4.0)) < 1e-15


def test_fullsky_cones_cover_and_partition_the_sky():
    # every direction is inside some cone, and owned by exactly one (the
    # closest axis), for directions away from ties
    rng = np.random.default_rng(2)
    d = unit(rng.normal(size=(20000, 3)))
    for n in (3, 12, 40):
        c = capi.cones_fullsky(n)
        inside = np.stack([capi.cone_inside(c[i], d) for i in range(n)])
        owned = np.stack([capi.cone_select(c, i, d) for i in range(n)])
        assert inside.any(axis=0).all()
        assert np.all(owned.sum(axis=0) == 1)
        # the owner is the cone whose axis makes the smallest angle with d
        cosang = d @ unit(c[:, 3:6]).T
        np.testing.assert_array_equal(owned.argmax(axis=0), cosang.argmax(axis=1))


# ------------------------------------------------------------- Cone::inside

def test_cone_inside_matches_reference(g):
    c = g["launch_cones"]
    for i in range(12):
        np.testing.assert_array_equal(capi.cone_inside(c[i], g["inside_points"]), g["inside"][i].astype(bool))


def test_cone_inside_closed_form():
    cone = np.array([0, 0, 0, 2.0, 0, 0, math.radians(10)])
    pts = np.array([[1.0, 0, 0], [2.0, 0, 0], [2.0000001, 0, 0], [-1.0, 0, 0], [0.0, 0, 0],
                    [1.0, math.tan(math.radians(9.9)), 0], [1.0, 0, math.tan(math.radians(10.1))]])
    # on the axis up to the base, not beyond, not behind, not the vertex; the half-angle bounds it
    np.testing.assert_array_equal(capi.cone_inside(cone, pts), [True, True, False, False, False, True, False])


# ------------------------------------------------------------------- launch

def test_launch_cone_matches_reference_draw_sequence(g):
    c = g["launch_cones"]
    for ic, seed in ((0, 1), (7, 2021), (11, 5)):
        in9, rot = capi.launch_cone(c, ic, 500, seed, micro_radius=1.0 / 3.0)
        np.testing.assert_array_equal(in9, g[f"launch_{ic}_{seed}_in9"])
        np.testing.assert_array_equal(rot, g[f"launch_{ic}_{seed}_rot"])


def test_launch_cone_rays_are_owned_and_unit():
    c = capi.cones_fullsky(12)
    in9, rot = capi.launch_cone(c, 4, 2000, 9)
    assert np.all(in9[:, 0] == 1.0) and np.all(in9[:, 1:5] == 0.0) and np.all(in9[:, 5] == 1.0)
    np.testing.assert_allclose(np.linalg.norm(in9[:, 6:9], axis=1), 1.0, atol=1e-15)
    assert capi.cone_select(c, 4, in9[:, 6:9]).all()
    assert np.all((rot >= 0) & (rot < 2 * math.pi))


def test_launch_angles_towards_bundle_match_reference(g):
    np.testing.assert_array_equal(capi.launch_angles(g["angles_phi_theta"]), g["angles_in9"])
    np.testing.assert_array_equal(capi.launch_towards(g["towards_begin"], g["towards_end"]), g["towards_in9"])
    row = 0
    for k, (cnt, ang, rot, cen) in enumerate(g["bundle_params"]):
        cnt, cen = int(cnt), bool(cen)
        out = capi.launch_bundle(g["angles_in9"][k], cnt, ang, rot, cen)
        np.testing.assert_array_equal(out, g["bundle_out"][row:row + cnt + cen])
        row += cnt + cen


def test_launch_bundle_closed_form():
    # the circle sits at the half-angle from the central photon, equally spaced
    p9 = capi.launch_angles([[0.7, 1.1]])[0]
    b = capi.launch_bundle(p9, 8, 0.05, 0.4, center=True)
    np.testing.assert_array_equal(b[0], p9)
    k0 = p9[6:9]
    ang = np.arccos(np.clip(b[1:, 6:9] @ k0, -1, 1))
    np.testing.assert_allclose(ang, 0.05, atol=1e-12)
    gaps = np.linalg.norm(np.diff(b[1:, 6:9], axis=0), axis=1)
    np.testing.assert_allclose(gaps, gaps[0], rtol=1e-10)


def test_launch_bundle_matches_reference_on_random_directions():
    """Integrator::launch (integrator.h:345-382) compiled from the reference,
    against the library's and the oracle's bundle launch on random photons,
    half-angles and rotations.  The reference build takes sin and cos of one
    variable from one sincos call but keeps the two calls of cos(acos(..)) and
    sin(acos(..)) (and of atan2) separate; glibc's sincos and cos differ in
    the last bit at some arguments (as at 1.8608994686324885), which a few
    thousand directions meet."""
    from oracle import Oracle, RefHarness, ref_available
    if not ref_available():
        pytest.skip("reference harness not built")
    R, O = RefHarness(), Oracle()
    rng = np.random.default_rng(5)
    for _ in range(3000):
        k = rng.normal(size=3)
        k /= np.linalg.norm(k)
        in9 = np.array([1.0, 0.0, 0.0, 0.0, 0.0, 1.0, *k])
        count, angle, rot = int(rng.integers(3, 12)), float(rng.uniform(1e-4, 0.2)), float(rng.uniform(0, 2 * np.pi))
        want = R.launch_bundle(in9, count, angle, rot, True)
        np.testing.assert_array_equal(capi.launch_bundle(in9, count, angle, rot, center=True), want)
        np.testing.assert_array_equal(O.launch_bundle(in9, count, angle, rot), want)


def test_bundle_theta_from_separate_sin_and_cos():
    """The direction that separates sincos from cos in the last bit: the
    bundle's z components come from cos(acos(kz)), not from sincos."""
    kz = -0.28605105861782765
    in9 = np.array([1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.7896345801127843, -0.5428185900959241, kz])
    r = math.sqrt(in9[6] ** 2 + in9[7] ** 2 + in9[8] ** 2)
    th = math.acos(kz / r)
    rot, angle = 3.9788187068425933, 0.01
    b = capi.launch_bundle(in9, 6, angle, rot, center=True)
    sps, cps = math.sin(rot), math.cos(rot)
    rcos, rsin = r * math.cos(angle), r * math.sin(angle)
    step = (math.pi + math.pi) / 6
    for i in range(6):
        x, y = rsin * math.cos(i * step), rsin * math.sin(i * step)
        assert b[i + 1, 8] == math.sin(th) * sps * x + math.sin(th) * cps * y + math.cos(th) * rcos


# ------------------------------------------------------------- narrow cones

NARROW = (16, 20.0, 70.0, 5.0, 5.0)


def test_narrow_cones_in_region_from_reference_spiral():
    from oracle import RefHarness, ref_available
    if not ref_available():
        pytest.skip("reference harness not built")
    n, phi, theta, ty, tz = NARROW
    c, cr, rotm1, size = capi.cones_narrow(n, phi, theta, ty, tz)
    # exactly n spiral points of that size fall in the region, and they are the bases
    xyz, _ = RefHarness().sphere_uniform(size, 2.0)
    ph = np.arctan2(xyz[:, 1], xyz[:, 0])
    th = np.arccos(xyz[:, 2] / 2.0)
    sel = (np.abs(ph) < math.radians(ty)) & (th > math.pi / 2 - math.radians(tz)) & (th < math.pi / 2 + math.radians(tz))
    assert sel.sum() == n
    np.testing.assert_array_equal(c[:, 3:6], xyz[sel])


def test_narrow_cones_rotation_closed_form():
    n, phi, theta, ty, tz = NARROW
    c, cr, rotm1, size = capi.cones_narrow(n, phi, theta, ty, tz)
    p, t = math.radians(phi), math.radians(theta)
    rot = np.array([[math.cos(t) * math.cos(p), math.cos(t) * math.sin(p), -math.sin(t)],
                    [-math.sin(p), math.cos(p), 0.0],
                    [math.cos(p) * math.sin(t), math.sin(p) * math.sin(t), math.cos(t)]])
    np.testing.assert_allclose(rotm1, rot.T, atol=2e-16)
    np.testing.assert_allclose(rotm1 @ rot, np.eye(3), atol=4e-16)
    np.testing.assert_allclose(cr[:, 3:6], c[:, 3:6] @ rotm1.T, atol=1e-15)
    assert np.all(cr[:, 6] == c[:, 6]) and np.all(c[:, 6] >= 0.01)
    # the cones' half-angle is 1.8 asin of the smallest base distance over the diameter
    b = c[:, 3:6]
    d = np.linalg.norm(b[:, None] - b[None], axis=-1) + np.diag(np.full(n, np.inf))
    assert abs(c[0, 6] - max(0.01, 1.8 * math.asin(d.min() / 4.0))) < 1e-15


def test_narrow_cone_info_parsing(tmp_path):
    f = tmp_path / "info_narrow_cone.txt"
    f.write_text("# cone\nphi = 20.7  # degrees\ntheta=70.5\n\nthetay = 5\nthetaz = 4.5\nthetaz = 9\nignored\n")
    # phi goes through std::stoul (integer degrees); adjacent duplicate keys keep the first value
    assert capi.narrow_cone_info(f) == (20.0, 70.5, 5.0, 4.5)
    with pytest.raises(RuntimeError):
        capi.narrow_cone_info(tmp_path / "missing.txt")


def test_narrow_cones_reject_bad_angles():
    with pytest.raises(RuntimeError):
        capi.cones_narrow(4, 0.0, 90.0, 0.0, 5.0)


# ------------------------------------------------ targets and map pixels

def test_targets_by_value_and_pixels_by_index():
    c = capi.cones_fullsky(12)
    rng = np.random.default_rng(8)
    pos = unit(rng.normal(size=(5000, 3))) * rng.uniform(0.05, 1.9, (5000, 1))
    targets = np.concatenate([pos, rng.normal(size=(5000, 5))], axis=1)
    sel = [capi.cone_targets(c, i, targets) for i in range(12)]
    allsel = np.sort(np.concatenate(sel))
    np.testing.assert_array_equal(allsel, np.arange(5000))   # each target in exactly one cone
    for i in range(12):
        np.testing.assert_array_equal(sel[i], np.flatnonzero(capi.cone_select(c, i, pos, by_value=True)))
    # a duplicated cone: by value the copies do not exclude each other, by index they do not either
    # (equal distance is not closer), so both own the same targets
    c2 = np.concatenate([c, c[3:4]])
    np.testing.assert_array_equal(capi.cone_targets(c2, 3, targets), capi.cone_targets(c2, 12, targets))
    # HEALPix pixel centres: owned pixels partition the sphere
    nside = 16
    pix = [capi.hmaps_cone_pixels(nside, c, i) for i in range(12)]
    np.testing.assert_array_equal(np.sort(np.concatenate(pix)), np.arange(12 * nside * nside))
    vec = capi.healpix_pix2vec(nside, np.arange(12 * nside * nside))
    np.testing.assert_array_equal(pix[5], np.flatnonzero(capi.cone_select(c, 5, vec)))
