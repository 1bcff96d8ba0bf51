"""The oracle (C restatement, oracle/mpf_oracle.c) pinned against the
reference: golden fixtures written from the reference's own headers, the
compiled reference itself where it is present, known-answer keys, and closed
forms.  CPU only."""
import math
import os

import numpy as np
import pytest

import cases
from oracle import Oracle, RefHarness, ref_available

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def O():
    return Oracle()


def lohi_int(k):
    return (int(k[1]) << 64) | int(k[0])


# ------------------------------------------------------------------ keys

def test_known_answer_keys(O):
    # printed by the reference headers (simplehyperoctreeindex.h cipher/compute)
    assert lohi_int(O.compute_key(32, [0.0, 0.0, 0.0])[0]) == 0xf8888888888888888888888888888888
    assert lohi_int(O.compute_key(9, [0.3, -1.1, 1.7])[0]) == 0xdcebec9be00000000000000000000000
    assert lohi_int(O.compute_key(2, [0.3, -1.1, 1.7])[0]) == 0xdc000000000000000000000000000000
    assert lohi_int(O.compute_key(1, [0.5, 0.5, 0.5])[0]) == 0xf << 124
    assert lohi_int(O.compute_key(1, [0.5, -0.5, -0.5])[0]) == 0x9 << 124
    assert lohi_int(O.compute_key(1, [-0.5, 0.5, -0.5])[0]) == 0xa << 124
    assert lohi_int(O.compute_key(1, [-0.5, -0.5, 0.5])[0]) == 0xc << 124
    lvl, cen, ext, par = O.key_info(O.compute_key(9, [0.3, -1.1, 1.7]))
    assert lvl[0] == 9
    assert tuple(cen[0]) == (0.30078125, -1.09765625, 1.69921875)
    assert ext[0] == 0.0078125
    assert lohi_int(par[0]) == 0xdcebec9b << 96


def test_keys_match_golden(O):
    g = np.load(os.path.join(GOLDEN, "keys.npz"))
    for i, l in enumerate(g["levels"]):
        np.testing.assert_array_equal(O.compute_key(int(l), g["points"]), g["keys"][i])
    lvl, cen, ext, par = O.key_info(g["keys"][6])
    np.testing.assert_array_equal(lvl, g["info_level"])
    np.testing.assert_array_equal(cen, g["info_center"])
    np.testing.assert_array_equal(ext, g["info_extent"])
    np.testing.assert_array_equal(par, g["info_parent"])


def test_key_closed_forms(O):
    rng = np.random.default_rng(5)
    p = rng.uniform(-1.99, 1.99, (500, 3))
    for l in (3, 8, 15, 21):
        k = O.compute_key(l, p)
        lvl, cen, ext, par = O.key_info(k)
        assert np.all(lvl == l)
        assert np.all(ext == 4.0 / 2 ** l)
        # the cell contains the point, and its centre is -2 + 4(2i+1)/2^(l+1)
        assert np.all(np.abs(cen - p) <= ext[:, None] / 2 + 1e-15)
        idx = np.floor((p + 2.0) * 2 ** l / 4.0)
        np.testing.assert_array_equal(cen, -2.0 + 4.0 * ((2 * idx + 1) / 2.0 ** (l + 1)))
        # parent == key at level l-1; prefix relation
        np.testing.assert_array_equal(par, O.compute_key(l - 1, p))
        # pre-order: ancestors sort before descendants
        ka = [lohi_int(x) for x in O.compute_key(l - 1, p)]
        kb = [lohi_int(x) for x in k]
        assert all(a < b for a, b in zip(ka, kb))


def test_cosmology_interpolate_golden(O):
    g = np.load(os.path.join(GOLDEN, "keys.npz"))
    np.testing.assert_array_equal(O.interpolate(g["interp_x"], g["interp_y"], g["interp_x0"]), g["interp_out"])
    # a linear table is reproduced (and extrapolated) exactly up to rounding
    x = np.linspace(0, 1, 11)
    np.testing.assert_allclose(O.interpolate(x, 3 * x + 1, np.array([-0.5, 0.25, 1.5])), [-0.5, 1.75, 5.5])


# --------------------------------------------------------- octree / interp

@pytest.mark.parametrize("name", cases.TREES)
def test_tree_fixture_digest(name):
    rec, _ = cases.tree(name)
    g = np.load(os.path.join(GOLDEN, f"tree_{name}.npz"))
    assert str(g["digest"]) == cases.records_digest(rec), "test tree generator changed; rewrite fixtures"


@pytest.mark.parametrize("name", cases.TREES)
def test_locate_interp_golden(O, name):
    rec, f64 = cases.tree(name)
    g = np.load(os.path.join(GOLDEN, f"tree_{name}.npz"))
    h = O.tree(rec, f64)
    e, lv = O.locate(h, g["points"])
    np.testing.assert_array_equal(e, g["locate_elem"])
    np.testing.assert_array_equal(lv, g["locate_level"])
    for order in (0, 1, 2, 3):
        np.testing.assert_array_equal(O.interp(h, order, g["points"]), g[f"interp{order}"], err_msg=f"order {order}")
    O.free(h)


@pytest.mark.parametrize("name", cases.TREES)
def test_integrate_golden(O, name):
    rec, f64 = cases.tree(name)
    g = np.load(os.path.join(GOLDEN, f"tree_{name}.npz"))
    h = O.tree(rec, f64)
    tab = cases.cosmology()
    n = 0
    for rs in cases.RAYSETS:
        in9 = g[f"in9_{rs}"]
        for stop, val in cases.STOPS:
            kit = cases.DIR if stop == "plane" else None
            for order in (1, 2):
                key = f"rays_{rs}_{stop}_o{order}"
                if key not in g:
                    continue
                out, steps = O.integrate(h, order, tab, cases.LENGTH, cases.NSTEPS, cases.STOP_KIND[stop], val, in9,
                                         kitarget=kit, nthreads=4)
                np.testing.assert_array_equal(out, g[key], err_msg=key)
                n += 1
    assert n >= 20
    O.free(h)


@pytest.mark.parametrize("name", cases.TREES)
def test_jacobian_golden(O, name):
    rec, f64 = cases.tree(name)
    g = np.load(os.path.join(GOLDEN, f"tree_{name}.npz"))
    h = O.tree(rec, f64)
    tab = cases.cosmology()
    for i, r in enumerate(g["jac_in9"]):
        np.testing.assert_array_equal(O.jacobian(h, 2, tab, cases.LENGTH, cases.NSTEPS, 2, 0.25, r, g["jac_dist"]),
                                      g["jac"][i])
    O.free(h)


def test_closed_form_interpolation(O):
    """A field linear in position is reproduced exactly by CIC and TSC on a
    uniform grid (the weights sum to one and are first-moment exact)."""
    from mpf import formats, synth
    lvl, kx, ky, kz = synth.uniform_level_coords(4, half=0.5)
    cx, cy, cz = formats.cell_centers(lvl, kx, ky, kz)
    f = np.zeros((cx.size, 7))
    f[:, 1] = 0.25 + 2.0 * cx - 1.0 * cy + 0.5 * cz  # phi: linear, exactly representable per cell
    f[:, 5] = 1.0
    rec, order = formats.make_records(lvl, kx, ky, kz, f)
    h = O.tree(rec, f[order])
    rng = np.random.default_rng(3)
    p = rng.uniform(-0.3, 0.3, (400, 3))
    expect = 0.25 + 2.0 * p[:, 0] - 1.0 * p[:, 1] + 0.5 * p[:, 2]
    for order in (1, 2):
        r = O.interp(h, order, p, precision=1)
        np.testing.assert_allclose(r[:, 1], expect, rtol=0, atol=1e-13)
        np.testing.assert_allclose(r[:, 5], 1.0, rtol=0, atol=1e-14)  # partition of unity
    O.free(h)


def test_flrw_rays(O):
    """phi == 0: rays are straight lines from the observer and the redshift
    is a0/a - 1 (integrator.h:598 with phi = 0)."""
    from mpf import formats, synth
    lvl, kx, ky, kz = synth.uniform_level_coords(4, half=0.5)
    f = np.zeros((kx.size, 7))
    f[:, 5] = 1.0
    rec, order = formats.make_records(lvl, kx, ky, kz, f)
    h = O.tree(rec)
    tab = cases.cosmology()
    in9 = cases.rays(4, aperture=20.0)
    out, _ = O.integrate(h, 2, tab, cases.LENGTH, 10, 2, 0.2, in9)
    d = in9[:, 6:9] / np.linalg.norm(in9[:, 6:9], axis=1)[:, None]
    pos = out[:, 2:5]
    np.testing.assert_allclose(np.cross(pos, d), 0.0, atol=1e-14)
    np.testing.assert_allclose(out[:, 9], 1.0 / out[:, 0] - 1.0, rtol=1e-14)
    assert np.all(np.abs(out[:, 25]) < 1e-12)  # null condition error()
    O.free(h)


@pytest.mark.skipif(not ref_available(), reason="reference harness oracle/_ref not built here")
def test_oracle_vs_reference_random():
    """Extra random points and rays against the compiled reference itself."""
    O = Oracle()
    R = RefHarness()
    tab = cases.cosmology()
    rng = np.random.default_rng(77)
    rec, f64 = cases.tree("amr")
    ho, hr = O.tree(rec, f64), R.tree(rec)
    p = rng.uniform(-0.6, 0.6, (4000, 3))
    for order in (0, 1, 2, 3):
        np.testing.assert_array_equal(O.interp(ho, order, p), R.interp(hr, order, p))
    ph = math.radians(20) + rng.uniform(-0.3, 0.3, 48)
    th = math.radians(20) + rng.uniform(-0.3, 0.3, 48)
    from mpf import synth
    in9 = synth.launch_photons(ph, th)
    for kind, val in ((2, 0.35), (-1, 0.0), (0, 0.3)):
        a, sa = O.integrate(ho, 2, tab, cases.LENGTH, 10, kind, val, in9, nthreads=4)
        b, sb = R.integrate(hr, 2, tab, cases.LENGTH, 10, kind, val, in9, nthreads=4)
        np.testing.assert_array_equal(a, b)
        assert sa == sb
