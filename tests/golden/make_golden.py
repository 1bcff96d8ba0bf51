"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Runs in the container that holds the reference sources: it loads
oracle/_ref/libmpfref.so (the reference headers instantiated by
oracle/ref_harness.cpp, built by oracle/Makefile) and records its outputs on
the deterministic inputs of tests/cases.py.  The fixtures are data only:
inputs are regenerated by tests/cases.py and checked against the stored
digest; outputs are the reference's numbers.

    python tests/golden/make_golden.py [tree ... | cones | propagate | bundle | born | velocity |
                                        catalogue | cosmology | targets]
    (default: all of them and every tree)

The velocity fixture comes from oracle/_ref/libmpfrefvel.so, the reference's
integrator compiled with VELOCITYFIELD (oracle/ref_harness_vel.cpp); the
catalogue, cosmology and targets fixtures from oracle/_ref/libmpfrefcat.so
(catalogues.h, input.h and miscellaneous.h, oracle/ref_harness_cat.cpp).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import cases  # noqa: E402
from oracle import RefCatalogue, RefHarness, RefVelocity  # noqa: E402


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
    if names == ["catalogue"]:
        return catalogue(tab)
    if names == ["cosmology"]:
        return cosmology()
    if names == ["targets"]:
        return targets()
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
    catalogue(tab)
    cosmology()
    targets()
    trees(R, tab, cases.TREES)


# catalogue root-finder configurations: (tree, accuracy, jacobiantype, bundle (stop_bundle, plane,
# openingmin) or None, number of sources[, arguments of cases.sources]); the last two bundles are wide
# enough for some of their rays to leave the tree (the failure entries 42, 42, 22)
CAT_VOBS = (3e-4, -2e-4, 1e-4)
CAT_RUNS = (("amr", 1e-8, "infinitesimal", None, 200), ("amr", 1e-12, "infinitesimal", None, 200),
            ("partial", 1e-12, "infinitesimal", None, 200), ("strong", 1e-8, "infinitesimal", None, 200),
            ("strong", 1e-12, "infinitesimal", None, 200), ("strong", 1e-12, "infinitesimal_born", None, 200),
            ("strong", 1e-12, "bundle", ("redshift", "sachs", 1e-4), 80),
            ("strong", 1e-12, "bundle", ("radius", "normal", 2e-4), 80),
            ("strong", 1e-12, "bundle", ("plane", "sachs", 1e-4), 80),
            ("strong", 1e-12, "bundle", ("a", "normal", 1e-4), 80),
            ("strong", 1e-12, "bundle", ("t", "sachs", 1e-4), 80),
            ("strong", 1e-12, "bundle", ("lambda", "normal", 1e-4), 80),
            ("strong", 1e-8, "bundle", ("redshift", "sachs", 0.5), 80, dict(seed=9, rmin=0.3, rmax=0.45, spread=0.25)),
            ("strong", 1e-8, "bundle", ("a", "sachs", 0.8), 80, dict(seed=9, rmin=0.3, rmax=0.45, spread=0.25)))


def cat_name(tree, acc, jt, bundle):
    b = "" if bundle is None else "_" + "_".join(str(v) for v in bundle)
    return f"{tree}_{acc:g}_{jt}{b}"


def catalogue(tab):
    """Catalogues::iterateNewtonMethod on the sources of cases.sources (the
    16 outputs of refc_newton per source), the grid re-run of rejected
    sources (use_previous_catalogues = 2) and relCat_with_previous_cat
    (use_previous_catalogues = 1), with an observer velocity."""
    RC = RefCatalogue()
    out = {"vobs": np.array(CAT_VOBS)}
    handles = {}
    for run in CAT_RUNS:
        tree, acc, jt, bundle, n = run[:5]
        if tree not in handles:
            rec, _ = cases.tree(tree)
            handles[tree] = RC.tree(rec)
            out[f"digest_{tree}"] = np.array(cases.records_digest(rec))
        src = cases.sources(n, **(run[5] if len(run) > 5 else dict(seed=5)))
        kw = dict(jacobiantype=jt)
        if bundle is not None:
            kw.update(stop_bundle=bundle[0], plane=bundle[1], opening=bundle[2])
        name = cat_name(tree, acc, jt, bundle)
        out[f"src_{name}"] = src
        out[f"ref_{name}"] = RC.newton(handles[tree], tab, cases.LENGTH, cases.NSTEPS, src, accuracy=acc,
                                       vobs=CAT_VOBS, **kw)
        print(name, "rejected", int(np.sum(out[f"ref_{name}"][:, 3] == 42)), "failed",
              int(np.sum(out[f"ref_{name}"][:, 0] == 42)))
    # the grid re-run of the first rejected sources of the strong tree, from
    # their last angles (catalogues.cpp:240-257, catalogues.h:826-829)
    h = handles["strong"]
    base = out["ref_" + cat_name("strong", 1e-12, "infinitesimal", None)]
    src = out["src_" + cat_name("strong", 1e-12, "infinitesimal", None)]
    rej = np.flatnonzero(base[:, 3] == 42)[:8]
    for jt, bundle in (("infinitesimal", None), ("bundle", ("redshift", "sachs", 1e-4))):
        kw = dict(jacobiantype=jt)
        if bundle is not None:
            kw.update(stop_bundle=bundle[0], plane=bundle[1], opening=bundle[2])
        name = "rerun_" + cat_name("strong", 3e-12, jt, bundle)
        out[f"src_{name}"] = src[rej]
        out[f"init_{name}"] = base[rej, :2]
        out[f"ref_{name}"] = RC.newton(h, tab, cases.LENGTH, cases.NSTEPS, src[rej], init2=base[rej, :2],
                                       accuracy=3e-12, vobs=CAT_VOBS, use_prev=2, **kw)
        print(name, "converged", int(np.sum(out[f"ref_{name}"][:, 3] != 42)))
    # relCat_with_previous_cat on catalogue rows of converged sources
    conv = out["ref_" + cat_name("strong", 1e-8, "infinitesimal", None)]
    src = out["src_" + cat_name("strong", 1e-8, "infinitesimal", None)]
    ok = np.flatnonzero(conv[:, 3] != 42)[:60]
    d = np.linalg.norm(src[ok, :3], axis=1)
    prev = np.zeros((ok.size, 18))
    prev[:, 0] = ok
    prev[:, 1] = np.arctan2(src[ok, 1], src[ok, 0])
    prev[:, 2] = np.arccos(src[ok, 2] / d)
    prev[:, 3:7] = conv[ok, 0:4]
    prev[:, 7:13] = conv[ok, 8:14]
    prev[:, 13:17] = conv[ok, 4:8]
    prev[:, 17] = 1
    out["prev18"] = prev
    for jt, bundle in (("infinitesimal", None), ("infinitesimal_born", None), ("bundle", ("redshift", "sachs", 1e-4)),
                       ("bundle", ("a", "normal", 1e-4)), ("bundle", ("plane", "sachs", 1e-4))):
        kw = dict(jacobiantype=jt)
        if bundle is not None:
            kw.update(stop_bundle=bundle[0], plane=bundle[1], opening=bundle[2])
        name = "prev_" + cat_name("strong", 0, jt, bundle)
        out[f"ref_{name}"] = RC.relcat_prev(h, tab, cases.LENGTH, cases.NSTEPS, prev, "/tmp/mpf_golden_relprev", **kw)
    # the files relCat writes (catalogue, _err, _reject) for three runs, and
    # those of relCat_with_previous_cat
    import tempfile
    tmp = tempfile.mkdtemp()
    for run in (CAT_RUNS[0], CAT_RUNS[4], CAT_RUNS[13]):
        tree, acc, jt, bundle = run[:4]
        name = cat_name(tree, acc, jt, bundle)
        src = out[f"src_{name}"]
        t8 = np.concatenate([src, np.arange(len(src))[:, None] + 1000.0, np.full((len(src), 1), 3.0)], axis=1)
        kw = dict(jacobiantype=jt)
        if bundle is not None:
            kw.update(stop_bundle=bundle[0], plane=bundle[1], opening=bundle[2])
        base = os.path.join(tmp, name)
        RC.relcat(handles[tree], tab, cases.LENGTH, cases.NSTEPS, t8, base, accuracy=acc, vobs=CAT_VOBS, **kw)
        for suffix in ("", "_err", "_reject"):
            out[f"file_{name}{suffix}"] = np.frombuffer(open(base + suffix + ".txt", "rb").read(), dtype=np.uint8)
    base = os.path.join(tmp, "previous")
    RC.relcat_prev(h, tab, cases.LENGTH, cases.NSTEPS, prev, base)
    for suffix in ("", "_err"):
        out[f"file_previous{suffix}"] = np.frombuffer(open(base + suffix + ".txt", "rb").read(), dtype=np.uint8)
    for hh in handles.values():
        RC.free(hh)
    np.savez_compressed(os.path.join(HERE, "catalogue.npz"), **out)
    print("catalogue written")


COSMO_CASES = ((2000, (700, 1301)), (400, ()), (1200, (10,)))


def cosmology():
    """Input::acquire on RAMSES-style evolution files (the files' text is
    stored with the table), and Input::correct with a perturbed photon."""
    import tempfile
    from mpf import synth
    RC = RefCatalogue()
    out = {}
    for n, bad in COSMO_CASES:
        d = tempfile.mkdtemp()
        evol, par, cell = synth.ramses_evolution_files(d, h=0.7, omega_m=0.3, n=n, bad_rows=bad)
        hol, tab = RC.acquire(evol, par, cell, mpc=3.08568e22)
        key = f"n{n}"
        out[f"{key}_evol"] = np.frombuffer(open(evol, "rb").read(), dtype=np.uint8)
        out[f"{key}_param"] = np.frombuffer(open(par, "rb").read(), dtype=np.uint8)
        out[f"{key}_celldir"] = np.array(os.path.basename(cell))
        out[f"{key}_hol"] = np.array(hol)
        out[f"{key}_table"] = tab
        t, a = tab[0], tab[1]
        tt = np.linspace(t[0], t[min(1200, len(t) - 1)], 600)
        ta = np.interp(tt, t, a) * (1 + 1e-3 * np.sin(tt / tt[-1] * 7))
        out[f"{key}_traj_t"], out[f"{key}_traj_a"] = tt, ta
        out[f"{key}_corrected"] = RC.correct(tab, tt, ta)
    np.savez_compressed(os.path.join(HERE, "cosmology.npz"), **out)
    print("cosmology written")


NARROW_CASES = ((16, 20, 70.0, 5.0, 5.0), (4, 33, 60.0, 10.0, 10.0), (50, 0, 90.0, 20.0, 15.0),
                (1, 200, 45.5, 3.0, 2.0), (9, 20, 70.0, 30.0, 30.0))


def targets():
    """Miscellaneous::GenerateNarrowCones from info_narrow_cone.txt (ASCII
    input), GenerateFullskyCones, and getTargets of random targets for the
    first cones of each series."""
    import tempfile
    RC = RefCatalogue()
    out = {}
    for k, (n, phi, theta, ty, tz) in enumerate(NARROW_CASES):
        d = tempfile.mkdtemp() + "/"
        with open(d + "info_narrow_cone.txt", "w") as fh:
            fh.write(f"phi = {phi}\ntheta = {theta}\nthetay = {ty}\nthetaz = {tz}\n")
        c, cr, rm, tyz = RC.narrow_cones(d, n)
        out[f"narrow{k}_params"] = np.array([n, phi, theta, ty, tz], dtype=np.float64)
        out[f"narrow{k}_cones"], out[f"narrow{k}_rot"], out[f"narrow{k}_rotm1"], out[f"narrow{k}_tyz"] = c, cr, rm, tyz
        rng = np.random.default_rng(n)
        pos = rng.normal(size=(1500, 3))
        pos /= np.linalg.norm(pos, axis=1)[:, None]
        pos *= rng.uniform(0.05, 1.9, (1500, 1))
        t8 = np.concatenate([pos, rng.normal(size=(1500, 3)), np.arange(1500)[:, None], np.ones((1500, 1))], axis=1)
        out[f"narrow{k}_targets"] = t8
        for ic in range(min(n, 3)):
            out[f"narrow{k}_sel{ic}"] = RC.get_targets(t8, c, ic)[:, 6]
    for n in (1, 12, 48):
        out[f"fullsky{n}"] = RC.fullsky_cones(n)
    np.savez_compressed(os.path.join(HERE, "targets.npz"), **out)
    print("targets written")


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
