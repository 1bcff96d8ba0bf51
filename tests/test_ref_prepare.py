"""Octree preparation (create_octree's ASCII branch, host code of libmpf, no
GPU) end to end against the reference's own Input::importascii and
Miscellaneous::correctOctree (Input::sistemize, Input::correct, the dphi/dt
conversion), compiled from input.h and miscellaneous.h into
oracle/_ref/libmpfrefcat.so: the records must be byte-identical.

SimpleHyperOctree::update pairs the last cell with the element past the end
of its container (simplehyperoctree.h:1285-1300).  The harness fills the
container's spare storage with copies of the cell of the largest key; the
inputs give every cell its own scale factor (a lightcone) and no duplicate
records, so each update meets a copy of that cell there (same a: no change)
and the comparison covers every byte of every record but its 4 bytes of
padding (which the reference leaves as they come).  The coarsening branch
(correction = 0) leaves moved cells past the end before its update, whose
pairing the reference leaves to whatever is there; it is compared with the
pinned steps in tests/test_prepare.py instead."""
import numpy as np
import pytest

from mpf import capi, synth
from oracle import RefCatalogue, refc_available
from test_prepare import amr_records, by_key, capi_level, write_dat

needs_refc = pytest.mark.skipif(not refc_available(), reason="catalogue harness (oracle/_ref) not built")


def fields(rec):
    """The records' bytes without their padding."""
    r = np.array(rec, copy=True)
    r["pad"] = 0
    return r.tobytes()


def lightcone_a(xyz):
    """A scale factor per cell from its distance to the observer (float32,
    distinct values): a lightcone's cells each have their own."""
    r = np.linalg.norm(xyz.astype(np.float64), axis=1)
    return (0.95 - 0.4 * r).astype(np.float32)


def prepared_records(seed, nan=True):
    """amr_records with partial octs, lightcone scale factors, isolated
    coarse cells of non-normal density (zero, or NaN when nan) and refined
    cells of zero density.  (A "nan" in a cell file stops the reference's
    stream extraction: the import tests write zeros.)"""
    rng = np.random.default_rng(seed)
    rec, lv, xyz = amr_records(rng, a_values=(0.8,), partial=True)
    rec["a"] = lightcone_a(xyz)
    lvl = np.array([capi_level(int(h), int(l)) for h, l in zip(rec["hi"], rec["lo"])])
    coarse = np.nonzero(lvl == lvl.min())[0]
    # isolated coarse cells (no two of them neighbours): every 41st, spaced in key order
    for i in coarse[5::41][:4]:
        rec["rho"][i] = np.float32(np.nan) if (nan and i % 2) else 0.0
    fine = np.nonzero(lvl == lvl.max())[0]
    rec["rho"][fine[3::17][:6]] = 0.0
    return rec, lv, xyz


@needs_refc
@pytest.mark.parametrize("order", ("shuffled", "sorted"))
def test_correct_octree_matches_the_reference(order):
    """correctOctree on a lightcone container with partial octs (completed
    by TSC on the growing container), non-normal coarse densities (means of
    their neighbours) and refined zero densities (the parent's): the
    builder's records byte for byte the reference's."""
    R = RefCatalogue()
    rec, _, _ = prepared_records(21)
    rng = np.random.default_rng(22)
    rec = by_key(rec) if order == "sorted" else rec[rng.permutation(len(rec))]
    keys = (rec["hi"].astype(object) << 64) | rec["lo"].astype(object)
    filler = rec[int(np.argmax(keys))]
    table = np.array(synth.flat_lcdm_table(n=512))
    cfg = dict(h=0.7, omegam=0.3, lboxmpch=1000.0, mpc=capi.MPC_SI)
    P = capi.Prep(correction=1, coarsecorrection=1, **cfg)
    P.load_records(rec)
    P.correct(table)
    got = P.records()
    want = R.prep_correct(rec, filler, 16 * len(rec), table, correction=1, coarsecorrection=1, **cfg)
    assert len(got) > len(rec)          # the partial octs were completed
    assert np.all(np.isfinite(got["rho"])) and np.all(got["rho"] != 0)
    assert np.any(got["dphidt"] != 0)
    assert fields(got) == fields(want)


@needs_refc
@pytest.mark.parametrize("micro", (8.0, 32.0))
def test_import_ascii_matches_the_reference(tmp_path, micro):
    """importascii of one cell file with PreparationASCII's filter (central
    sphere, full-sky cone): the same records in the same order."""
    R = RefCatalogue()
    rec, lv, xyz = prepared_records(23, nan=False)
    path = tmp_path / "cells_00000.dat"
    write_dat(path, (lv - 2).astype(int), xyz.astype(np.float32).astype(float), rec)
    cones = capi.cones_fullsky(4, 2.0)
    levels = set()
    for icone in range(4):
        P = capi.Prep(microcoeff=micro)
        n = P.import_ascii(path, cones[icone])
        got = P.records()
        want = R.prep_import(path, cones[icone], 1.0 / micro)
        assert n == len(want)
        assert fields(got) == fields(want)
        assert 0 < n < len(rec)
        assert np.count_nonzero(got["rho"]) > n // 2
        levels |= set(capi_level(int(h), int(l)) for h, l in zip(got["hi"], got["lo"]))
    assert levels == {5, 6}


@needs_refc
def test_import_ascii_narrow_cone_and_coarseonly_match_the_reference(tmp_path):
    """A narrow cone (positions and forces turned by rotm1 before the
    filter) and coarseonly (refined cells dropped after the selection)."""
    R = RefCatalogue()
    rec, lv, xyz = prepared_records(24, nan=False)
    path = tmp_path / "cells_00000.dat"
    write_dat(path, (lv - 2).astype(int), xyz.astype(np.float32).astype(float), rec)
    th = 0.3
    rot = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]])
    cone = capi.cones_fullsky(2, 2.0)[0]
    P = capi.Prep(isfullsky=0, microcoeff=16.0)
    P.import_ascii(path, cone, rot)
    got = P.records()
    want = R.prep_import(path, cone, 1.0 / 16.0, isfullsky=0, rotm1=rot)
    assert len(want) > 0 and fields(got) == fields(want)
    assert np.count_nonzero(got["rho"]) > len(got) // 2
    P = capi.Prep(coarseonly=1, ncoarse=3, microcoeff=16.0)
    P.import_ascii(path, cone)
    got = P.records()
    want = R.prep_import(path, cone, 1.0 / 16.0, coarseonly=1, ncoarse=3)
    assert len(want) > 0 and fields(got) == fields(want)
    assert set(capi_level(int(h), int(l)) for h, l in zip(got["hi"], got["lo"])) == {5}
