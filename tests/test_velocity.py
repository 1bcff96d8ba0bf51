"""The oracle's velocity field pinned (CPU only).

Integration: the redshift of the velocity-field build takes the Doppler and
ISW terms (integrator.h:537-539, 594-595, 765-766) and the cells interpolate
their velocity with the other fields (gravity2.h).  The oracle is checked
bit for bit against fixtures written by the reference's own integrator
compiled with VELOCITYFIELD (tests/golden/make_golden.py velocity), against
that build directly where oracle/_ref holds it, and against the closed form
of a uniform velocity field in a homogeneous universe.

Deposit (Hmaps::CreateOctreeVelocityWithCIC / WithTSC, hmaps.h:439-541, with
the normalisation and parent fill of hmaps.cpp:389-441): hmaps.h includes
mpi.h and the HDF5 reader, which this image lacks, so the reference's
deposit cannot be compiled here.  The oracle is pinned by closed forms that
do not restate its loops: the weight a cell receives is the textbook
cloud-in-cell tent or triangular-shaped-cloud spline of the particle's
distance to the cell centre in units of the cell size, summed over particles,
for every stored cell of every level; the velocity is the weighted mean of
the particle velocities; a cell without weight carries its parent's
velocity.  Velocities are in units of c, as the reference's are after its
particle reader's conversion (hmaps.h:378-380).
"""
import os

import numpy as np
import pytest

import cases
from oracle import Oracle, RefVelocity, refv_available

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C_LIGHT = 299792458.0


@pytest.fixture(scope="module")
def O():
    return Oracle()


def _with_velocity(O, name, seed=7):
    rec, f64 = cases.tree(name)
    rho, vel = cases.velocity_field(rec, seed)
    h = O.tree(rec)
    O.set_velocity(h, rho, vel, cases.VOBS)
    return rec, h


@pytest.mark.parametrize("name", ["amr", "partial"])
def test_velocity_integration_matches_golden(O, name):
    g = np.load(os.path.join(GOLDEN, "velocity.npz"))
    rec, h = _with_velocity(O, name)
    try:
        assert str(g[f"{name}_digest"]) == cases.records_digest(rec)
        tab = cases.cosmology()
        for order in (0, 1, 2):
            for stop, value in (("redshift", 0.3), ("none", 0.0), ("radius", 0.42)):
                out, _ = O.integrate(h, order, tab, cases.LENGTH, cases.NSTEPS, cases.STOP_KIND[stop], value,
                                     g["in9"])
                ref = g[f"{name}_{order}_{stop}"]
                assert np.array_equal(out, ref), (name, order, stop, np.abs(out - ref).max())
    finally:
        O.free(h)


@pytest.mark.skipif(not refv_available(), reason="oracle/_ref/libmpfrefvel.so not built (no reference sources)")
@pytest.mark.parametrize("name", ["amr", "coarse", "partial"])
def test_velocity_oracle_matches_reference_build(O, name):
    V = RefVelocity()
    rec, _ = cases.tree(name)
    rho, vel = cases.velocity_field(rec, seed=11)
    h = O.tree(rec)
    O.set_velocity(h, rho, vel, (0.4e-3, 1.3e-3, -0.8e-3))
    hr = V.tree(rec, rho, vel)
    try:
        pts = cases.points(rec, 1500, seed=3)
        for order in (0, 1, 2):
            assert np.array_equal(O.interp10(h, order, pts), V.interp10(hr, order, pts).astype(np.float64))
        tab = cases.cosmology()
        in9 = cases.rays(n_side=6, aperture=10.0)
        for order in (-1, 0, 1, 2):
            for stop, value in (("redshift", 0.25), ("a", 0.8), ("none", 0.0)):
                if order == -1 and (stop == "none" or name == "partial"):
                    # homogeneous rays that reach points without a stored cell (past the stored
                    # region, or the partial tree's cut): the reference then dereferences locate()'s
                    # end() for the step length (integrator.h:604, 775), which is undefined
                    continue
                a, sa = O.integrate(h, order, tab, cases.LENGTH, cases.NSTEPS, cases.STOP_KIND[stop], value, in9)
                b, sb = V.integrate(hr, order, tab, cases.LENGTH, cases.NSTEPS, cases.STOP_KIND[stop], value, in9,
                                    vobs=(0.4e-3, 1.3e-3, -0.8e-3))
                assert sa == sb and np.array_equal(a, b), (name, order, stop)
    finally:
        O.free(h)
        V.free(hr)


def test_observer_velocity_closed_form(O):
    """Homogeneous universe (order -1: phi = 0, no ISW, no cell data): only
    the observer's term remains, z = a0/a (1 - u.k0/(c k0_t)) - 1, and k/k_t
    is constant on the straight rays, so the last element's direction gives
    the same u.k/(c k_t)."""
    rec, _ = cases.tree("amr")
    tab = cases.cosmology()
    in9 = cases.rays(n_side=4, aperture=6.0)
    u = np.array([0.7e-3, -0.3e-3, 1.0e-3])
    h = O.tree(rec)
    try:
        base, _ = O.integrate(h, -1, tab, cases.LENGTH, cases.NSTEPS, -1, 0.0, in9)
        O.set_velocity(h, None, np.zeros((rec.size, 3), dtype=np.float32), u)
        moving, _ = O.integrate(h, -1, tab, cases.LENGTH, cases.NSTEPS, -1, 0.0, in9)
        # the velocities change no geodesic
        assert np.array_equal(moving[:, :9], base[:, :9])
        v0n = (moving[:, 6:9] @ u) / (C_LIGHT * moving[:, 5])
        assert np.all(np.abs(v0n) > 1e-4)
        np.testing.assert_allclose(moving[:, 9] + 1, (base[:, 9] + 1) * (1 - v0n), rtol=1e-12, atol=0)
    finally:
        O.free(h)


def test_cell_velocity_closed_form(O):
    """A uniform cell velocity w (FP64 accumulation: TSC weights sum to one)
    adds a0/a w.k/(c k_t) to 1 + z on an unchanged trajectory."""
    rec, _ = cases.tree("amr")
    n = rec.size
    tab = cases.cosmology()
    in9 = cases.rays(n_side=5, aperture=6.0)
    w = np.array([1.0e-3, 0.3e-3, -0.7e-3])
    h = O.tree(rec)
    try:
        O.set_velocity(h, None, np.zeros((n, 3), dtype=np.float32), (0.0, 0.0, 0.0))
        still, _ = O.integrate(h, 2, tab, cases.LENGTH, cases.NSTEPS, -1, 0.0, in9, precision=1)
        O.set_velocity(h, None, np.tile(w.astype(np.float32), (n, 1)), (0.0, 0.0, 0.0))
        moving, _ = O.integrate(h, 2, tab, cases.LENGTH, cases.NSTEPS, -1, 0.0, in9, precision=1)
        assert np.array_equal(np.delete(moving, 9, axis=1), np.delete(still, 9, axis=1))
        # a0: the scale factor of the first trajectory element (every ray starts at the same time)
        a0 = O.trajectory(h, 2, tab, cases.LENGTH, cases.NSTEPS, -1, 0.0, in9[0], precision=1)[0, 0]
        wf = w.astype(np.float32).astype(np.float64)  # the cells hold FP32
        doppler = (moving[:, 6:9] @ wf) / (C_LIGHT * moving[:, 5])
        assert np.all(np.abs(doppler) > 1e-4)
        np.testing.assert_allclose((moving[:, 9] + 1) - (still[:, 9] + 1), a0 / moving[:, 0] * doppler, rtol=1e-9,
                                   atol=0)
    finally:
        O.free(h)


def _cells(O, rec):
    lohi = np.stack([rec["lo"], rec["hi"]], axis=1)
    lvl, cen, ext, par = O.key_info(lohi)
    index = {(int(a), int(b)): i for i, (a, b) in enumerate(lohi)}
    parent = np.array([index.get((int(p[0]), int(p[1])), -1) for p in par])
    return lvl.astype(int), cen, ext, parent


def _tent(d):
    return np.clip(1.0 - np.abs(d), 0.0, None)


def _tsc(d):
    a = np.abs(d)
    return np.where(a < 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


def _expected(O, rec, pos, vel, kernel):
    lvl, cen, ext, parent = _cells(O, rec)
    p = pos.astype(np.float64)
    w = np.zeros(rec.size)
    sv = np.zeros((rec.size, 3))
    for i in range(rec.size):
        kx = kernel((p[:, 0] - cen[i, 0]) / ext[i]) * kernel((p[:, 1] - cen[i, 1]) / ext[i]) * \
            kernel((p[:, 2] - cen[i, 2]) / ext[i])
        w[i] = kx.sum()
        sv[i] = kx @ vel.astype(np.float64)
    v = sv / np.where(w == 0, 1.0, w)[:, None]
    for level in range(lvl.min() + 1, lvl.max() + 1):
        for i in np.nonzero((lvl == level) & (w == 0))[0]:
            v[i] = v[parent[i]]
    return w, sv, v


@pytest.mark.parametrize("scheme,kernel", [(1, _tent), (2, _tsc)])
@pytest.mark.parametrize("name", ["amr", "partial"])
def test_deposit_closed_form(O, name, scheme, kernel):
    rec, _ = cases.tree(name)
    pos, vel = cases.particles(300, seed=scheme)
    h = O.tree(rec)
    try:
        rho, v = O.velocity_deposit(h, rec.size, scheme, pos, vel)
        w, sv, ve = _expected(O, rec, pos, vel, kernel)
        touched = w > 0
        assert touched.sum() > 100
        np.testing.assert_allclose(rho, w, rtol=2e-6, atol=1e-6)
        # the weighted mean of velocities of about 1e-3 c, FP32 sums
        np.testing.assert_allclose(v, ve, rtol=0, atol=1e-8)
        # cells without weight: the parent's velocity, or zero at the coarsest level
        lvl = _cells(O, rec)[0]
        assert np.all(v[(~touched) & (lvl == lvl.min())] == 0)
        assert np.any((~touched) & (lvl > lvl.min()) & np.any(v != 0, axis=1))
    finally:
        O.free(h)


def test_deposit_weights_sum_to_one_per_level(O):
    """A particle whose whole stencil is stored gives each level a total
    weight of 1 (CIC tent and TSC spline are partitions of unity)."""
    rec, _ = cases.tree("uniform")
    h = O.tree(rec)
    try:
        rng = np.random.default_rng(5)
        pos = rng.uniform(-0.3, 0.3, (1, 3)).astype(np.float32)
        vel = np.array([[0.3e-3, 0.7e-3, -1.0e-3]], dtype=np.float32)
        for scheme, cells in ((1, 8), (2, 27)):
            rho, v = O.velocity_deposit(h, rec.size, scheme, pos, vel)
            assert np.count_nonzero(rho) == cells
            assert abs(float(rho.sum(dtype=np.float64)) - 1.0) < 1e-6
            np.testing.assert_allclose(v[rho > 0], np.repeat(vel, cells, axis=0), rtol=1e-6)
    finally:
        O.free(h)


def test_deposit_accumulates_over_files(O):
    """Two particle files deposited one after the other give the sums of the
    union (up to FP32 rounding), normalised once at the end."""
    rec, _ = cases.tree("amr")
    h = O.tree(rec)
    try:
        pa, va = cases.particles(200, seed=11)
        pb, vb = cases.particles(150, seed=12)
        r1, s1 = O.velocity_deposit(h, rec.size, 2, pa, va, finish=False)
        r2, v2 = O.velocity_deposit(h, rec.size, 2, pb, vb, rho=r1, vel3=s1)
        ru, vu = O.velocity_deposit(h, rec.size, 2, np.r_[pa, pb], np.r_[va, vb])
        np.testing.assert_allclose(r2, ru, rtol=2e-6, atol=1e-6)
        np.testing.assert_allclose(v2, vu, rtol=0, atol=1e-8)
        with pytest.raises(ValueError):
            O.velocity_deposit(h, rec.size, 3, pa, va)
    finally:
        O.free(h)
