"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Runs in the container that holds the reference sources: it loads
oracle/_ref/libmpfref.so (the reference headers instantiated by
oracle/ref_harness.cpp, built by oracle/Makefile) and records its outputs on
the deterministic inputs of tests/cases.py.  The fixtures are data only:
inputs are regenerated by tests/cases.py and checked against the stored
digest; outputs are the reference's numbers.

    python tests/golden/make_golden.py [tree ...]   (default: keys and every tree)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import cases  # noqa: E402
from oracle import RefHarness  # noqa: E402


def main(names):
    R = RefHarness()
    tab = cases.cosmology()
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
    trees(R, tab, cases.TREES)


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
