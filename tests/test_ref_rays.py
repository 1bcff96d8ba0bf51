"""The rays program end to end: the reference's own rays.cpp, built into
oracle/_ref/rays and run as one MPI task on a cone file, a RAMSES evolution
file and a parameter file written here, against the same computation through
the library's host calls and the oracle's restatement of propagate (itself
pinned to the reference's Integrator::propagate in tests/test_propagate.py,
and the device's propagate to it in tests/test_gpu_propagate.py):
  * the cosmology: mpf_cosmology_acquire, the homogeneous reference photon
    (integrate<-1> through the level-5 homogenized octree, rays.cpp:209-213),
    mpf_cosmology_correct;
  * the launch of ntrajectories rays into cone 0 with the seeded draws and
    their bundle rotations (mpf_launch_cone), and each ray's bundle
    (mpf_launch_bundle), for full-sky cones (mpf_cones_fullsky) and for
    narrow cones read from an ASCII info_narrow_cone.txt (typefile 2,
    mpf_cones_narrow: the rays go into the cones turned into the octree's
    frame, rays.cpp:191-197 and 233-236);
  * every ray's central trajectory with its angular diameter distance, and
    the reference ray's, as the program writes them with savemode = 0
    (integrator.h:998-1010): every field of every element, exactly;
  * the statistics of every kind for two interpolation parameters:
    mpf_ray_statistics, written by mpf_ray_statistics_write.
The program clears the octree after each statistic (rays.cpp:358), so that
with statistic = all every statistic after the first meets an empty octree;
each statistic is run in an invocation of its own.
The statistics files are compared line by line: the abscissae and the
reference ray's statistic exactly (17 significant digits round-trip a
double); the mean and the standard deviation within what the order of the
sum allows (the reference adds the rays in the order its threads finish
them): n ulp of the mean, and n ulp of the deviation plus 2 ulp of the mean
(the deviation is the spread of values of the mean's size).  CPU only."""
import math
import os

import numpy as np
import pytest

import cases
from mpf import capi, formats, synth
from oracle import Oracle, rays_available, run_rays

pytestmark = pytest.mark.skipif(not rays_available(), reason="the reference's rays program (oracle/_ref) not built")

NCONES, SEED, MICRO, NTRAJ, NBUNDLE, OPENING, NSTEPS, NCOARSE, NSTAT = 8, 11, 16, 12, 6, 0.01, 10, 6, 40
# the narrow field: phi, theta (degrees; phi is read as an integer) and the half-openings
NARROW = (20, 70.0, 5.0, 5.0)
STATS = ("distance", "distance2", "homogeneous", "inhomogeneous")
INTERPS = (("redshift", capi.STOP_REDSHIFT), ("r", capi.STOP_RADIUS))
EXTENT = 4
MPC = 3.08568e22
# in9 of Integrator::launch from the box centre towards (EXTENT / 2, 0, 0)
TOWARDS_X = np.array([1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0])
# columns of a photon as Output::save writes it (photon.h: index, the core
# a t x y z dt/dl dx/dl dy/dl dz/dl, then the extras in tuple order)
PHOTON_COLUMNS = ("index", "a", "t", "x", "y", "z", "kt", "kx", "ky", "kz", "level", "ah", "rho", "phi", "dphidx",
                  "dphidy", "dphidz", "dphidl", "laplacian", "redshift", "dsdl2", "error", "distance", "isw", "iswold",
                  "chi", "lambda", "dphidt", "s")


def uniform_records(level, a):
    """The cells of one level (no parents), as SimpleHyperOctree::assign(level, 0)
    lays them out, with a = `a` and the other fields zero."""
    side = 1 << level
    c = -2.0 + (np.arange(side) + 0.5) * (4.0 / side)
    xyz = np.stack(np.meshgrid(c, c, c, indexing="ij"), -1).reshape(-1, 3)
    k = capi.prep_keys(np.full(len(xyz), level, dtype=np.uint32), xyz.astype(np.float32))
    rec = np.zeros(len(k), dtype=formats.RECORD_DTYPE)
    rec["lo"], rec["hi"] = k[:, 0], k[:, 1]
    rec["a"] = a
    return rec[np.lexsort((rec["lo"], rec["hi"]))]


def reference_nsteps(n, nref=5):
    """The steps per cell of the homogeneous reference photon (rays.cpp:212)."""
    return EXTENT * 2 ** (int(math.log2(n / 2.0 ** nref + 1.0) + 1.0) + 1)


def name(*parts):
    """Output::name of rays.cpp:272-274 and 441-452, '.' written 'd'."""
    return "_".join(parts).replace(".", "d") + ".txt"


def stat_name(interp, stat, rank=None):
    return name(f"rays_{interp}", "stat", f"{NBUNDLE:05d}", f"{OPENING:8.6f}", interp, stat,
                *(() if rank is None else (f"{rank:05d}",)))


def ray_name(interp, i=None):
    """The reference ray's file (i None) or ray i's (rays.cpp:275, 323)."""
    return name(f"rays_{interp}", f"{NBUNDLE:05d}", f"{OPENING:8.6f}", interp, "00000",
                *(() if i is None else (f"{i:05d}", "00000")))


def read_stat(path):
    lines = open(path).read().splitlines()
    assert lines[0].startswith("# ")
    return int(lines[0][2:]), [ln.split(" ") for ln in lines[1:]]


@pytest.fixture(scope="module", params=("fullsky", "narrow"))
def run(request, tmp_path_factory):
    sky = request.param
    d = tmp_path_factory.mktemp("rays_" + sky)
    evol, par, cell = synth.ramses_evolution_files(str(d), h=0.7, omega_m=0.3, n=1200)
    if sky == "narrow":
        os.makedirs(cell, exist_ok=True)
        phi, theta, ty, tz = NARROW
        with open(os.path.join(cell, "info_narrow_cone.txt"), "w") as fh:
            fh.write(f"phi = {phi}\ntheta = {theta}\nthetay = {ty}\nthetaz = {tz}\n")
    cones = d / "cones"
    out = d / "out"
    cones.mkdir()
    out.mkdir()
    rec, _ = cases.tree("wide")
    rec.tofile(cones / "cone_00000")
    base = dict(conefmt="cone_%05d", ncones=NCONES, conedir=str(cones) + "/", seed=SEED, microcoeff=MICRO,
                ntrajectories=NTRAJ, isfullsky=int(sky == "fullsky"), paramfile=par, evolfile=evol,
                rhoch2=1.87847e-26, mpc=MPC,
                ncoarse=NCOARSE, ray_targets="random", massmin=0, massmax=0, nbundlemin=NBUNDLE, nbundlecnt=1,
                openingmin=OPENING, openingcnt=1, outputdir=str(out) + "/", nsteps=NSTEPS, savemode=0,
                makestat=1, nstat=NSTAT, celldir=cell + "/", typefile=0 if sky == "fullsky" else 2, base="none",
                halos=0)
    for interp, _ in INTERPS:
        for stat in STATS:
            run_rays(dict(base, stop_ray=interp, statistic=stat, outputprefix=f"rays_{interp}"), str(d))
    return d, rec, evol, par, cell, sky


@pytest.fixture(scope="module")
def mine(run):
    """The same computation through the library's host calls and the oracle:
    per interpolation parameter the rays' (central rows, distances) and the
    reference ray's."""
    d, rec, evol, par, cell, sky = run
    # the cosmology of rays.cpp:201-213
    tab, h, om, lb = capi.cosmology_acquire(evol, par, cell + "/", mpc=MPC)
    length = lb * MPC / h
    O = Oracle()
    h5 = O.tree(uniform_records(5, 1.0))
    traj = O.trajectory(h5, -1, tab, length, reference_nsteps(tab.shape[1]), capi.STOP_NONE, 0.0, TOWARDS_X)
    O.free(h5)
    tab = np.array(tab)
    tab[1] = capi.cosmology_correct(tab[0], tab[1], traj[:, capi.OUT["t"]], traj[:, capi.OUT["a"]])
    # the rays of cone 0 and their bundles
    if sky == "fullsky":
        cones = capi.cones_fullsky(NCONES, 2.0)
    else:
        cones = capi.cones_narrow(NCONES, *NARROW)[1]
    in9, rot = capi.launch_cone(cones, 0, NTRAJ, SEED, micro_radius=1.0 / MICRO)
    h = O.tree(rec)
    hh = O.tree(uniform_records(NCOARSE // 2, 0.0))
    res = {}
    for interp, kind in INTERPS:
        runs = [O.propagate(h, 2, tab, length, NSTEPS, in9[i], NBUNDLE, OPENING, rot[i], kind=kind)
                for i in range(NTRAJ)]
        ref = O.propagate(hh, -1, tab, length, EXTENT * NSTEPS * (1 << (NCOARSE - NCOARSE // 2)) * 2, TOWARDS_X,
                          NBUNDLE, OPENING, 0.0, kind=kind)
        res[interp] = (runs, ref)
    O.free(h)
    O.free(hh)
    return res, length


def photon_rows(rows, dist):
    """The oracle's rows in the columns of Output::save."""
    cols = []
    for c in PHOTON_COLUMNS:
        if c == "index":
            cols.append(np.arange(len(rows), dtype=np.float64))
        elif c == "distance":
            cols.append(dist)
        else:
            cols.append(rows[:, capi.OUT[c]])
    return np.stack(cols, 1)


def read_photons(path):
    rows = [ln.split() for ln in open(path).read().splitlines() if ln.strip()]
    assert all(len(r) == len(PHOTON_COLUMNS) for r in rows)
    return np.array([[float(v) for v in r] for r in rows])


def test_trajectories_match_the_reference_rays_program(run, mine):
    """Every ray's central trajectory and distance, and the reference ray's,
    exactly as the program writes them."""
    d = run[0]
    res, _ = mine
    for interp, _ in INTERPS:
        runs, (ref_rows, ref_dist) = res[interp]
        want = read_photons(d / "out" / ray_name(interp))
        got = photon_rows(ref_rows, ref_dist)
        assert got.shape == want.shape and len(want) > 100, interp
        np.testing.assert_array_equal(got, want, err_msg=f"reference ray, {interp}")
        accepted = 0
        for i, (rows, dist) in enumerate(runs):
            path = d / "out" / ray_name(interp, i)
            if not len(rows):
                # a rejected bundle writes nothing (integrator.h:999)
                assert not path.exists(), (interp, i)
                continue
            want = read_photons(path)
            assert want.shape[0] == len(rows), (interp, i)
            np.testing.assert_array_equal(photon_rows(rows, dist), want, err_msg=f"ray {i}, {interp}")
            assert np.any(dist > 0)
            accepted += 1
        assert accepted >= NTRAJ // 2


def test_statistics_files_match_the_reference_rays_program(run, mine):
    d = run[0]
    res, length = mine
    eps = np.finfo(float).eps
    checked = 0
    for interp, _ in INTERPS:
        runs, (ref_rows, ref_dist) = res[interp]
        trajs = [r[0] for r in runs]
        dists = [r[1] for r in runs]
        for stat in STATS:
            rows, used = capi.ray_statistics(trajs, dists, ref_rows, ref_dist, interp, stat, nstat=NSTAT, amin=0.0,
                                             length=length)
            path = d / ("mine_" + stat_name(interp, stat))
            capi.ray_statistics_write(path, rows, used)
            n_mine, got = read_stat(path)
            for rank in (None, 0):
                n_ref, ref = read_stat(d / "out" / stat_name(interp, stat, rank))
                assert n_mine == n_ref == used, (interp, stat, rank)
                assert len(got) == len(ref) > 10, (interp, stat)
                for a, b in zip(got, ref):
                    # the reference ray's redshift and statistic, written as the reference writes them
                    assert a[:2] == b[:2], (interp, stat, a, b)
                    m, s = float(a[2]), float(a[3])
                    rm, rs = float(b[2]), float(b[3])
                    assert abs(m - rm) <= used * eps * abs(rm), (interp, stat, a, b)
                    assert abs(s - rs) <= used * eps * rs + 2 * eps * abs(rm), (interp, stat, a, b)
                checked += 1
    assert checked == 2 * len(STATS) * 2
