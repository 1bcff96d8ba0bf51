"""Deterministic test inputs shared by the golden-fixture writer and the tests.

Trees are small AMR octrees with analytic fields; rays leave the observer at
the box centre into a cone around (phi, theta) = (20 deg, 20 deg), crossing a
refined patch (level jumps, TSC level drops), a hole (missing neighbours) and
the edge of the stored region.
"""
from __future__ import annotations

import hashlib
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "magrathea-pathfinder_amd"))

from mpf import formats, synth  # noqa: E402

LENGTH = 3000.0 * synth.MPC_SI / 0.7
NSTEPS = 10
DIR = np.array([math.cos(math.radians(20)) * math.sin(math.radians(20)),
                math.sin(math.radians(20)) * math.sin(math.radians(20)),
                math.cos(math.radians(20))])


def cosmology():
    return synth.flat_lcdm_table(h=0.7, omega_m=0.3, a_min=0.2, n=512)


def _patch(center, half):
    c = np.asarray(center)

    def f(level, x, y, z):
        return (np.abs(x - c[0]) < half) & (np.abs(y - c[1]) < half) & (np.abs(z - c[2]) < half)
    return f


def tree(name):
    """(records, fields64) of a named test tree."""
    if name == "uniform":
        return synth.analytic_tree(base_level=6, half=0.5, length=LENGTH, seed=1)
    if name == "amr":
        inner = _patch(0.22 * DIR, 0.09)
        deep = _patch(0.2 * DIR + np.array([0.02, 0.0, 0.0]), 0.04)

        def refine(level, x, y, z):
            m = inner(level, x, y, z)
            if level >= 7:
                m &= deep(level, x, y, z)
            return m
        other = np.array([0.0, 1.0, 0.0]) - DIR[1] * DIR
        other /= np.linalg.norm(other)
        hole_c = 0.38 * DIR + 0.05 * other
        hole = _patch(hole_c, 0.05)
        return synth.analytic_tree(base_level=6, half=0.5, max_level=9, refine=refine, length=LENGTH, seed=2,
                                   holes=hole)
    if name == "partial":
        # the refined patch of "amr" cut by a hole: parents outside the hole
        # keep only some of their children (partial octs, as a cone's edge
        # leaves them in the reference's cone files)
        inner = _patch(0.22 * DIR, 0.09)
        deep = _patch(0.2 * DIR + np.array([0.02, 0.0, 0.0]), 0.04)

        def refine(level, x, y, z):
            m = inner(level, x, y, z)
            if level >= 7:
                m &= deep(level, x, y, z)
            return m
        box = _patch(0.2 * DIR + np.array([0.03, 0.01, -0.01]), 0.012)

        def cut(level, x, y, z):   # cells of levels 8 and 9 only: their level-7 parents stay
            return (level >= 8) & box(level, x, y, z)
        return synth.analytic_tree(base_level=6, half=0.5, max_level=9, refine=refine, length=LENGTH, seed=4,
                                   holes=cut)
    if name == "coarse":
        return synth.analytic_tree(base_level=5, half=0.5, max_level=7, refine=_patch(0.15 * DIR, 0.12),
                                   length=LENGTH, seed=3)
    if name == "strong":
        # the refined patch of "amr" with a potential a hundred times deeper:
        # catalogue sources need the Jacobian-based Newton updates, and at
        # tight accuracies some are rejected
        return synth.analytic_tree(base_level=6, half=0.5, max_level=8, refine=_patch(0.22 * DIR, 0.09),
                                   length=LENGTH, seed=2, amp=1.0e-3)
    if name == "wide":
        # reaches past EXTENT / 4 = 1 from the observer, as the bundles of
        # Integrator::propagate must (integrator.h:907-909)
        return synth.analytic_tree(base_level=5, half=1.25, max_level=7, refine=_patch(0.6 * DIR, 0.15),
                                   length=LENGTH, seed=5)
    raise KeyError(name)


TREES = ("uniform", "amr", "coarse", "partial")


def records_digest(rec):
    return hashlib.sha256(np.ascontiguousarray(rec).tobytes()).hexdigest()


def sources(n, seed=5, rmin=0.1, rmax=0.35, spread=0.15):
    """Catalogue sources (x, y, z, vx, vy, vz) around the cone axis at
    distances in [rmin, rmax) (box units), velocities in simulation units."""
    rng = np.random.default_rng(seed)
    d = rng.uniform(rmin, rmax, n)
    dirs = DIR + rng.normal(scale=spread, size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    src = np.zeros((n, 6))
    src[:, :3] = dirs * d[:, None]
    src[:, 3:] = rng.normal(scale=1e-3, size=(n, 3))
    return src


RAYSETS = {"narrow": dict(n_side=8, aperture=6.0), "wide": dict(n_side=8, aperture=40.0)}


def rays(n_side=8, aperture=6.0, phi0=20.0, theta0=20.0):
    ph, th = synth.cone_grid(phi0, theta0, aperture, n_side)
    return synth.launch_photons(ph, th)


def points(rec, n=3000, seed=0):
    """Query points: uniform in and around the stored region, on cell faces,
    near the refined patch and exactly at the box centre."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(-0.56, 0.56, (n, 3))
    k = n // 5
    p[:k] = np.round(p[:k] * 64) / 64               # on level-8 faces/corners
    p[k:2 * k] = 0.22 * DIR + rng.uniform(-0.1, 0.1, (k, 3))
    p[2 * k:2 * k + 10] = 0.0
    p[2 * k + 10:2 * k + 20] = rng.uniform(-1e-12, 1e-12, (10, 3))
    return p


STOPS = (("redshift", 0.3), ("radius", 0.42), ("none", 0.0), ("lambda", 5e16), ("t", 1.2e17),
         ("a", 0.8), ("plane", 0.3), ("s", 0.33))
STOP_KIND = {"none": -1, "radius": 0, "lambda": 1, "redshift": 2, "t": 3, "a": 4, "plane": 5, "s": 6}


# ---- velocity field (the reference's velocity-field build)
# Velocities are in units of c: the reference converts the particle
# velocities with unit_l 1e-2 / (c unit_t) (hmaps.h:378-380) and its Doppler
# term v.k / (c k_t) is then the line-of-sight velocity over c.
VOBS = (1.2e-3, -0.7e-3, 0.4e-3)   # observer velocity / c


def velocity_field(rec, seed=7):
    """(rho, vel3) per record: weights in [0, 3) with a fifth of the cells at
    zero, Gaussian cell velocities of 1e-3 c per axis (float32)."""
    rng = np.random.default_rng(seed)
    n = rec.size
    rho = rng.uniform(0, 3, n).astype(np.float32)
    rho[rng.random(n) < 0.2] = 0
    vel = rng.normal(0, 1e-3, (n, 3)).astype(np.float32)
    return rho, vel


def particles(n, seed=0, center=None, spread=0.12):
    """(pos3, vel3) float32 particles around the refined patch of the test
    trees: a bulk flow plus a dispersion of 7e-4 c per axis."""
    rng = np.random.default_rng(seed)
    c = 0.22 * DIR if center is None else np.asarray(center)
    pos = (c + rng.uniform(-spread, spread, (n, 3))).astype(np.float32)
    vel = (np.array([5e-4, -2.7e-4, 1.3e-4]) + rng.normal(0, 7e-4, (n, 3))).astype(np.float32)
    return pos, vel
