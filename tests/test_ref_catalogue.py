"""The catalogue root-finder, the cosmology table and the cone and target
selection against the reference's own catalogues.h, input.h and
miscellaneous.h.

The fixtures (tests/golden/catalogue.npz, cosmology.npz, targets.npz) are the
outputs of oracle/_ref/libmpfrefcat.so, which compiles those headers from the
reference sources (oracle/ref_harness_cat.cpp, tests/golden/make_golden.py):
  * Catalogues::iterateNewtonMethod on 4 trees, 2 accuracies, every
    jacobiantype (infinitesimal, infinitesimal_born, bundle with each
    stop_bundle and both planes, bundles wide enough to fail), an observer
    velocity; sources that converge after Jacobian-based updates, sources
    rejected after 10 iterations, sources whose ray never reaches them;
  * the grid re-run of rejected sources (use_previous_catalogues = 2) and
    relCat_with_previous_cat (use_previous_catalogues = 1);
  * Input::acquire and Input::correct on RAMSES-style evolution files;
  * GenerateNarrowCones, GenerateFullskyCones and getTargets.
The oracle's restatement (oracle/mpf_oracle.c) must match bit for bit, and
so must the product's host code (cosmology table, cones, targets).  Where the
harness is built, fresh inputs are compared live as well.
"""
import os
import sys
import tempfile

import numpy as np
import pytest

import cases
from mpf import capi
from oracle import Oracle, RefCatalogue, refc_available

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

CAT = np.load(os.path.join(HERE, "golden", "catalogue.npz"))
COSMO = np.load(os.path.join(HERE, "golden", "cosmology.npz"))
TARGETS = np.load(os.path.join(HERE, "golden", "targets.npz"))
KIND = {"radius": 0, "lambda": 1, "redshift": 2, "t": 3, "a": 4, "plane": 5}
PLANE = {"sachs": 0, "normal": 1}
# refc_newton columns -> the oracle's MPF_CAT_* columns
PAIRS = ((0, 2), (1, 3), (2, 4), (3, 5), (4, 12), (5, 13), (6, 14), (7, 15), (8, 6), (9, 7), (10, 8), (11, 9),
         (12, 10), (13, 11))
_trees = {}


def _tree(name):
    if name not in _trees:
        rec, f64 = cases.tree(name)
        if name in ("amr", "partial", "strong"):
            assert str(CAT[f"digest_{name}"]) == cases.records_digest(rec)
        O = Oracle()
        _trees[name] = (O, O.tree(rec, f64), rec)
    return _trees[name]


def _oracle_kw(jt, bundle):
    kw = {}
    if bundle is not None:
        kw["bundle"] = (KIND[bundle[0]], PLANE[bundle[1]], bundle[2])
    if jt == "infinitesimal_born":
        kw["born"] = True
    return kw


def assert_matches_reference(ref, b, what):
    """Every output of iterateNewtonMethod bit for bit, except result[1][1]
    of a failed bundle, which the reference leaves unset."""
    for ri, oi in PAIRS:
        m = np.ones(len(ref), dtype=bool) if ri != 3 else ref[:, 2] != 22
        bad = np.flatnonzero(ref[m, ri] != b[m, oi])
        assert bad.size == 0, f"{what}: column {ri} differs on {bad.size} sources, first {bad[:5]}"
    # the routing of relCat (catalogues.h:876-903) as the oracle's status
    status = np.where((ref[:, 0] == 42) & (ref[:, 1] == 42), 1, np.where(ref[:, 3] == 42, 2, 0))
    np.testing.assert_array_equal(b[:, 16], status)


@pytest.mark.parametrize("run", make_golden.CAT_RUNS, ids=lambda r: make_golden.cat_name(*r[:4]))
def test_oracle_newton_matches_reference(run):
    tree, acc, jt, bundle = run[:4]
    name = make_golden.cat_name(tree, acc, jt, bundle)
    O, h, _ = _tree(tree)
    src, ref = CAT[f"src_{name}"], CAT[f"ref_{name}"]
    b, _ = O.catalogue(h, 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, src, accuracy=acc,
                       vobs=tuple(CAT["vobs"]), **_oracle_kw(jt, bundle))
    assert_matches_reference(ref, b, name)


def test_fixture_covers_every_path():
    """The fixtures hold converged sources after 3-5 iterations (so after
    Jacobian-based updates), rejected ones (42 iterations), rays that miss
    the source radius (42, 42) and failed bundles (42, 42, 22)."""
    allref = np.concatenate([CAT["ref_" + make_golden.cat_name(*r[:4])] for r in make_golden.CAT_RUNS])
    it = allref[:, 3][allref[:, 0] != 42]
    assert {3, 4, 42} <= set(it.astype(int).tolist())
    assert np.sum((allref[:, 0] == 42) & (allref[:, 2] == 0)) > 0
    assert np.sum(allref[:, 2] == 22) > 0
    rerun = CAT["ref_rerun_" + make_golden.cat_name("strong", 3e-12, "infinitesimal", None)]
    assert np.sum(rerun[:, 3] == 42) > 0 and np.sum(rerun[:, 3] != 42) > 0


@pytest.mark.parametrize("jt,bundle", (("infinitesimal", None), ("bundle", ("redshift", "sachs", 1e-4))))
def test_oracle_grid_rerun_matches_reference(jt, bundle):
    name = "rerun_" + make_golden.cat_name("strong", 3e-12, jt, bundle)
    O, h, _ = _tree("strong")
    b, _ = O.catalogue(h, 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, CAT[f"src_{name}"],
                       init2=CAT[f"init_{name}"], accuracy=3e-12, vobs=tuple(CAT["vobs"]), use_prev=2,
                       **_oracle_kw(jt, bundle))
    assert_matches_reference(CAT[f"ref_{name}"], b, name)
    assert np.all(b[:, 18] == 0)   # no grid ray fell short of its source


@pytest.mark.parametrize("jt,bundle", (("infinitesimal", None), ("infinitesimal_born", None),
                                       ("bundle", ("redshift", "sachs", 1e-4)), ("bundle", ("a", "normal", 1e-4)),
                                       ("bundle", ("plane", "sachs", 1e-4))))
def test_oracle_previous_catalogue_matches_reference(jt, bundle):
    name = "prev_" + make_golden.cat_name("strong", 0, jt, bundle)
    O, h, _ = _tree("strong")
    prev = CAT["prev18"]
    jac, flag = O.catalogue_prev(h, 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, prev,
                                 jacobiantype={"infinitesimal": 0, "infinitesimal_born": 1, "bundle": 2}[jt],
                                 bundle=None if bundle is None else (KIND[bundle[0]], PLANE[bundle[1]], bundle[2]))
    ref = CAT[f"ref_{name}"]
    np.testing.assert_array_equal(jac, ref[:, 13:17])
    np.testing.assert_array_equal(ref[:, :13], prev[:, :13])   # only the Jacobian columns change
    assert np.all(flag == 0)


@pytest.mark.skipif(not refc_available(), reason="reference catalogue harness not built")
def test_oracle_newton_live_against_harness():
    """Fresh sources (another seed) on the strong tree, infinitesimal and
    bundle, straight against the reference's iterateNewtonMethod."""
    RC = RefCatalogue()
    O, h, rec = _tree("strong")
    hr = RC.tree(rec)
    src = cases.sources(60, seed=23, spread=0.2)
    for acc, jt, bundle in ((1e-12, "infinitesimal", None), (1e-10, "bundle", ("t", "normal", 3e-4))):
        kw = dict(jacobiantype=jt)
        if bundle is not None:
            kw.update(stop_bundle=bundle[0], plane=bundle[1], opening=bundle[2])
        ref = RC.newton(hr, cases.cosmology(), cases.LENGTH, cases.NSTEPS, src, accuracy=acc, vobs=(1e-4, 0, -2e-4),
                        **kw)
        b, _ = O.catalogue(h, 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, src, accuracy=acc,
                           vobs=(1e-4, 0, -2e-4), **_oracle_kw(jt, bundle))
        assert_matches_reference(ref, b, f"live {jt}")
    RC.free(hr)


# ------------------------------------------------------------ cosmology


@pytest.mark.parametrize("n", [c[0] for c in make_golden.COSMO_CASES])
def test_cosmology_acquire_and_correct_match_reference(n, tmp_path):
    key = f"n{n}"
    evol = tmp_path / "ramses_input_00001.dat"
    par = tmp_path / "param_00001.txt"
    evol.write_bytes(COSMO[f"{key}_evol"].tobytes())
    par.write_bytes(COSMO[f"{key}_param"].tobytes())
    cell = str(tmp_path / str(COSMO[f"{key}_celldir"]))
    tab, h, om, lb = capi.cosmology_acquire(str(evol), str(par), cell, mpc=3.08568e22)
    assert (h, om, lb) == tuple(COSMO[f"{key}_hol"])
    np.testing.assert_array_equal(tab, COSMO[f"{key}_table"])
    t, a = tab[0], tab[1]
    out = capi.cosmology_correct(t, a, COSMO[f"{key}_traj_t"], COSMO[f"{key}_traj_a"])
    np.testing.assert_array_equal(out, COSMO[f"{key}_corrected"][1])


# ------------------------------------------------------- cones and targets


@pytest.mark.parametrize("k", range(len(make_golden.NARROW_CASES)))
def test_narrow_cones_and_targets_match_reference(k):
    n, phi, theta, ty, tz = TARGETS[f"narrow{k}_params"]
    n = int(n)
    c, cr, rotm1, _ = capi.cones_narrow(n, phi, theta, ty, tz)
    np.testing.assert_array_equal(c, TARGETS[f"narrow{k}_cones"])
    np.testing.assert_array_equal(cr, TARGETS[f"narrow{k}_rot"])
    np.testing.assert_array_equal(rotm1, TARGETS[f"narrow{k}_rotm1"])
    t8 = TARGETS[f"narrow{k}_targets"]
    for ic in range(min(n, 3)):
        idx = capi.cone_targets(c, ic, t8)
        np.testing.assert_array_equal(t8[idx, 6], TARGETS[f"narrow{k}_sel{ic}"])


@pytest.mark.parametrize("n", (1, 12, 48))
def test_fullsky_cones_match_reference(n):
    np.testing.assert_array_equal(capi.cones_fullsky(n), TARGETS[f"fullsky{n}"])


@pytest.mark.skipif(not refc_available(), reason="reference catalogue harness not built")
def test_narrow_cones_live_against_harness():
    RC = RefCatalogue()
    for n, phi, theta, ty, tz in ((7, 123, 33.3, 12.0, 7.5), (25, 300, 100.0, 8.0, 25.0)):
        d = tempfile.mkdtemp() + "/"
        with open(d + "info_narrow_cone.txt", "w") as fh:
            fh.write(f"phi = {phi}\ntheta = {theta}\nthetay = {ty}\nthetaz = {tz}\n")
        rc, rr, rm, _ = RC.narrow_cones(d, n)
        c, cr, rotm1, _ = capi.cones_narrow(n, float(phi), theta, ty, tz)
        np.testing.assert_array_equal(c, rc)
        np.testing.assert_array_equal(cr, rr)
        np.testing.assert_array_equal(rotm1, rm)


# ------------------------------------------------------- catalogue files


@pytest.mark.parametrize("idx", (0, 4, 13))
def test_catalogue_files_match_reference(idx, tmp_path):
    """mpf_catalogue_write on rows equal to the reference's (the oracle's,
    pinned above) writes relCat's three files byte for byte."""
    tree, acc, jt, bundle = make_golden.CAT_RUNS[idx][:4]
    name = make_golden.cat_name(tree, acc, jt, bundle)
    O, h, _ = _tree(tree)
    src = CAT[f"src_{name}"]
    rows, _ = O.catalogue(h, 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, src, accuracy=acc,
                          vobs=tuple(CAT["vobs"]), **_oracle_kw(jt, bundle))
    idn = np.stack([np.arange(len(src)) + 1000.0, np.full(len(src), 3.0)], axis=1)
    capi.catalogue_write(str(tmp_path / "cat"), rows, idn)
    for suffix in ("", "_err", "_reject"):
        got = (tmp_path / f"cat{suffix}.txt").read_bytes()
        assert got == CAT[f"file_{name}{suffix}"].tobytes(), suffix


def test_previous_catalogue_files_match_reference(tmp_path):
    ref = CAT["ref_prev_" + make_golden.cat_name("strong", 0, "infinitesimal", None)]
    bad = ref.copy()
    bad[::7, 13:15] = 42          # rows routed to the error file
    capi.catalogue_write_previous(str(tmp_path / "prev"), ref)
    assert (tmp_path / "prev.txt").read_bytes() == CAT["file_previous"].tobytes()
    assert (tmp_path / "prev_err.txt").read_bytes() == CAT["file_previous_err"].tobytes()
    capi.catalogue_write_previous(str(tmp_path / "prev2"), bad)
    main = (tmp_path / "prev2.txt").read_text().splitlines()
    err = (tmp_path / "prev2_err.txt").read_text().splitlines()
    assert len(err) == len(range(0, len(ref), 7)) and len(main) + len(err) == len(ref)
