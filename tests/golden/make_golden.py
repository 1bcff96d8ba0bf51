"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Runs in the container that holds the reference sources: it loads
oracle/_ref/libmpfref.so (the reference headers instantiated by
oracle/ref_harness.cpp, built by oracle/Makefile) and records its outputs on
the deterministic inputs of tests/cases.py.  The fixtures are data only:
inputs are regenerated by tests/cases.py and checked against the stored
digest; outputs are the reference's numbers.

    python tests/golden/make_golden.py [tree ... | cones | propagate | bundle | born | velocity]
    (default: keys, cones, propagate, bundle, born, velocity and every tree)

The velocity fixture comes from oracle/_ref/libmpfrefvel.so, the reference's
integrator compiled with VELOCITYFIELD (oracle/ref_harness_vel.cpp).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import cases  # noqa: E402
from oracle import RefHarness, RefVelocity  # noqa: E402


def main(names):
    R = RefHarness()
    tab = cases.cosmology()
    if names == ["cones"]:
        return cones(R)
    if names == ["propagate"]:
        return propagate(R, tab)
    if names == ["bundle"]:
        return bundle(R, tab)
    if names == ["born"]:
        return born(R, tab)
    if names == ["velocity"]:
        return velocity(tab)
    if names:
        return trees(R, tab, names)
    # keys: compute() at several levels on random and edge points
    rng = np.random.default_rng(11)
    kp = np.concatenate([rng.uniform(-2.2, 2.2, (400, 3)), rng.uniform(-1e-9, 1e-9, (20, 3)),
                         np.array([[0.3, -1.1, 1.7], [0.0, 0.0, 0.0], [-2.0, -2.0, -2.0], [1.9999999999, 0.5, -0.5],
                                   [-2.5, 0.0, 0.0], [2.6, 0.0, 0.0]])])
    levels = np.array([0, 1, 2, 5, 9, 13, 21, 27, 32], dtype=np.uint32)
    keys = np.stack([R.compute_key(int(l), kp) for l in levels])
    info_lvl, info_cen, info_ext, info_par = R.key_info(keys[6])
    xs = np.sort(rng.uniform(0, 10, 50))
    ys = rng.normal(size=50)
    x0 = np.concatenate([rng.uniform(-1, 11, 200), xs[:5], [xs[0], xs[-1]]])
    np.savez_compressed(os.path.join(HERE, "keys.npz"), points=kp, levels=levels, keys=keys, info_level=info_lvl,
                        info_center=info_cen, info_extent=info_ext, info_parent=info_par, interp_x=xs,
                        interp_y=ys, interp_x0=x0, interp_out=R.interpolate(xs, ys, x0))
    cones(R)
    propagate(R, tab)
    bundle(R, tab)
    born(R, tab)
    velocity(tab)
    trees(R, tab, cases.TREES)


PROP_KINDS = ("redshift", "a", "t", "lambda", "r")

VEL_TREES = ("amr", "partial")
VEL_STOPS = (("redshift", 0.3), ("none", 0.0), ("radius", 0.42))


def velocity(tab):
    """Integrations of the velocity-field build: the cell velocities and
    weights of cases.velocity_field and the observer velocity cases.VOBS,
    every interpolation order and three stop criteria."""
    V = RefVelocity()
    out = {}
    in9 = cases.rays(n_side=5, aperture=6.0)
    out["in9"] = in9
    for name in VEL_TREES:
        rec, _ = cases.tree(name)
        rho, vel = cases.velocity_field(rec)
        out[f"{name}_digest"] = np.array(cases.records_digest(rec))
        h = V.tree(rec, rho, vel)
        for order in (0, 1, 2):
            for stop, value in VEL_STOPS:
                o, steps = V.integrate(h, order, tab, cases.LENGTH, cases.NSTEPS, cases.STOP_KIND[stop], value, in9,
                                       vobs=cases.VOBS)
                out[f"{name}_{order}_{stop}"] = o
        V.free(h)
    np.savez_compressed(os.path.join(HERE, "velocity.npz"), **out)


BUNDLE_VALUES = {"r": [0.15, 0.3, 0.42], "lambda": [2e16, 5e16], "redshift": [0.1, 0.25, 0.33],
                 "t": [4e16, 1e17], "a": [0.95, 0.85, 0.78]}


def bundle(R, tab):
    """Lensing::dbetadtheta (the bundle Jacobian) for directions around the
    test cone, every stop parameter, on a tree where some rays end early
    (failure entries) and on one where every ray reaches the values."""
    rng = np.random.default_rng(21)
    out = {}
    for name in ("amr", "wide"):
        rec, f64 = cases.tree(name)
        h = R.tree(rec)
        out[f"{name}_digest"] = np.array(cases.records_digest(rec))
        pt = np.c_[np.radians(20) + rng.uniform(-0.2, 0.2, 10), np.radians(20) + rng.uniform(-0.2, 0.2, 10)]
        out[f"{name}_phi_theta"] = pt
        for kind, vals in BUNDLE_VALUES.items():
            ki = rng.normal(size=(10, len(vals), 3))
            dist = rng.uniform(0.1, 0.4, (10, len(vals)))
            out[f"{name}_{kind}_ki"], out[f"{name}_{kind}_dist"] = ki, dist
            out[f"{name}_{kind}_jac"] = np.stack([
                R.dbetadtheta(h, 2, tab, cases.LENGTH, cases.NSTEPS, kind, 0.005, [0.0, 0.0, 0.0], p, t, vals,
                              ki[i], dist[i]) for i, (p, t) in enumerate(pt)])
        R.free(h)
    np.savez_compressed(os.path.join(HERE, "bundle.npz"), **out)
    print("bundle written")


def born(R, tab):
    """The Born Jacobian of the reference: Integrator::integrate<-1> to a
    radius, then Lensing::dbetadtheta_infinitesimal on the octree's data
    (catalogues.h:670-674)."""
    out = {}
    for name in ("amr", "wide"):
        rec, f64 = cases.tree(name)
        h = R.tree(rec)
        in9 = cases.rays(n_side=5, aperture=24.0)
        out[f"{name}_digest"] = np.array(cases.records_digest(rec))
        out[f"{name}_in9"] = in9
        # stop before the hole of the amr tree at 0.38 (inside a missing region
        # the homogeneous path has no cell level to take its step from)
        out[f"{name}_jac"] = np.stack([R.jacobian(h, -1, tab, cases.LENGTH, cases.NSTEPS, 0, 0.32, r, [0.15, 0.25, 0.3])
                                       for r in in9])
        R.free(h)
    np.savez_compressed(os.path.join(HERE, "born.npz"), **out)
    print("born written")


def propagate(R, tab):
    """Integrator::propagate: central trajectories and angular diameter
    distances of bundles on the tree that reaches past EXTENT / 4."""
    rec, f64 = cases.tree("wide")
    h = R.tree(rec)
    in9 = cases.rays(n_side=3, aperture=30.0)
    out = dict(digest=np.array(cases.records_digest(rec)), in9=in9)
    for order in (2, 1):
        for kind in PROP_KINDS if order == 2 else ("redshift",):
            for i, r in enumerate(in9):
                rows, dist = R.propagate(h, order, tab, cases.LENGTH, cases.NSTEPS, r, 6, 0.01, 0.37 * i,
                                         interp=kind)
                out[f"o{order}_{kind}_{i}_rows"] = rows
                out[f"o{order}_{kind}_{i}_dist"] = dist
    R.free(h)
    np.savez_compressed(os.path.join(HERE, "propagate.npz"), **out)
    print("propagate written")


def cones(R):
    """Spiral points and pair distances of HyperSphere::uniform, and photons
    of Integrator::launch: into a series of cones with the rays.cpp:233-239
    draw sequence, towards angles, towards points, and around a photon."""
    out = {}
    for n in (1, 2, 5, 12, 48, 300):
        out[f"uniform_{n}"], out[f"uniform_minmax_{n}"] = R.sphere_uniform(n, 2.0)
    # 12 cones built from the spiral (vertex at the origin), with their angle
    # from the spiral's smallest pair distance
    xyz, mm = R.sphere_uniform(12, 2.0)
    fs = np.concatenate([np.zeros((12, 3)), xyz, np.full((12, 1), 1.8 * np.arcsin(mm[0] / 4.0))], axis=1)
    out["launch_cones"] = fs
    for ic, seed in ((0, 1), (7, 2021), (11, 5)):
        out[f"launch_{ic}_{seed}_in9"], out[f"launch_{ic}_{seed}_rot"] = R.launch_cone(fs, ic, 1.0 / 3.0, seed, 500)
    rng = np.random.default_rng(17)
    pts = rng.normal(size=(4000, 3)) * rng.uniform(0.1, 3.0, (4000, 1))
    out["inside_points"] = pts
    out["inside"] = np.stack([R.cone_inside(fs[i], pts) for i in range(12)])
    pt = np.c_[rng.uniform(-np.pi, np.pi, 2000), rng.uniform(0, np.pi, 2000)]
    out["angles_phi_theta"] = pt
    out["angles_in9"] = R.launch_angles([0.0, 0.0, 0.0], pt)
    b, e = rng.normal(size=(500, 3)), rng.normal(size=(500, 3))
    out["towards_begin"], out["towards_end"], out["towards_in9"] = b, e, R.launch_towards(b, e)
    bundles = []
    for k in range(200):
        p9 = out["angles_in9"][k]
        cnt, ang, rot = int(rng.integers(1, 17)), rng.uniform(1e-5, 0.3), rng.uniform(0, 2 * np.pi)
        bundles.append((p9, cnt, ang, rot, k % 2 == 0, R.launch_bundle(p9, cnt, ang, rot, k % 2 == 0)))
    out["bundle_params"] = np.array([[c, a, r, ce] for _, c, a, r, ce, _ in bundles])
    out["bundle_out"] = np.concatenate([o for *_, o in bundles])
    np.savez_compressed(os.path.join(HERE, "cones.npz"), **out)
    print("cones written")


def trees(R, tab, names):
    for name in names:
        rec, f64 = cases.tree(name)
        h = R.tree(rec)
        pts = cases.points(rec, n=1500)
        out = dict(digest=np.array(cases.records_digest(rec)), points=pts)
        for order in (0, 1, 2, 3):
            out[f"interp{order}"] = R.interp(h, order, pts)
        e, lv = R.locate(h, pts)
        out["locate_elem"] = e
        out["locate_level"] = lv
        for rs, kw in cases.RAYSETS.items():
            in9 = cases.rays(**kw)
            out[f"in9_{rs}"] = in9
            for stop, val in cases.STOPS:
                kit = cases.DIR if stop == "plane" else None
                for order in (1, 2):
                    if order == 1 and stop not in ("redshift", "none"):
                        continue
                    o, steps = R.integrate(h, order, tab, cases.LENGTH, cases.NSTEPS, cases.STOP_KIND[stop], val,
                                           in9, kitarget=kit)
                    out[f"rays_{rs}_{stop}_o{order}"] = o
        # infinitesimal jacobian of a few rays at two source distances
        in9 = cases.rays(**cases.RAYSETS["narrow"])[::16]
        jac = np.stack([R.jacobian(h, 2, tab, cases.LENGTH, cases.NSTEPS, 2, 0.25, r, np.array([0.2, 0.3]))
                        for r in in9])
        out["jac_in9"] = in9
        out["jac_dist"] = np.array([0.2, 0.3])
        out["jac"] = jac
        R.free(h)
        np.savez_compressed(os.path.join(HERE, f"tree_{name}.npz"), **out)
        print(name, "written")


if __name__ == "__main__":
    main(sys.argv[1:])
