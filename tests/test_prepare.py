"""Octree preparation (create_octree's ASCII branch, host code of libmpf, no
GPU): mpf_prep_* against the reference's own SimpleHyperOctreeIndex,
SimpleHyperOctree and Utility code compiled by oracle/_ref (keys, the cone
and central-sphere filter, update, coarsen, the uncached NGP/CIC/TSC on
containers with an unsorted tail, interpolate2), and the whole correction
(Input::correct and correctOctree, in input.h and miscellaneous.h, which need
HDF5 and MPI here) against expectations built from those pinned calls and
closed forms.

The reference's SimpleHyperOctree::update also reads the element past the
end of its container for the last cell; the harness puts a copy of the last
record there, and the inputs below keep that read harmless (the last cell's
copies, or equal scale factors), so every other byte is compared."""
import math
import os

import numpy as np
import pytest

from mpf import capi, formats, synth
from oracle import RefHarness, ref_available

needs_ref = pytest.mark.skipif(not ref_available(), reason="reference harness (oracle/_ref) not built")


# --------------------------------------------------------------- helpers

def key_words(rec):
    return np.stack([rec["lo"], rec["hi"]], axis=1)


def make_cells(level_r, lo, hi, rng):
    """Cell centres of RAMSES level level_r in [lo, hi)^3 (float32), box units."""
    n = 2 ** level_r
    ids = np.arange(n)
    c = (-0.5 + (ids + 0.5) / n).astype(np.float64)
    c = c[(c >= lo) & (c < hi)]
    g = np.stack(np.meshgrid(c, c, c, indexing="ij"), -1).reshape(-1, 3)
    return g.astype(np.float32)


def records_from(level_oct, xyz32, fields7):
    """Records of cells at octree levels from float positions (the importer's keys)."""
    k = capi.prep_keys(level_oct, xyz32)
    rec = np.zeros(len(k), dtype=formats.RECORD_DTYPE)
    rec["lo"], rec["hi"] = k[:, 0], k[:, 1]
    for i, name in enumerate(formats.FIELDS7):
        rec[name] = fields7[:, i]
    return rec


def amr_records(rng, a_values=(0.8,), partial=False):
    """A small AMR octree in record form: RAMSES level 3 over the central
    box, the cells of one corner refined to level 4 (one oct left partial
    when asked: children 4..7 dropped), one copy per scale factor."""
    coarse = make_cells(3, -0.5, 0.5, rng)
    fine = make_cells(4, -0.5, -0.125, rng)
    lv = np.concatenate([np.full(len(coarse), 5), np.full(len(fine), 6)]).astype(np.uint32)
    xyz = np.concatenate([coarse, fine])
    if partial:
        # drop the children 4..7 (z bit set) of the fine oct at the corner
        k = capi.prep_keys(lv, xyz)
        hi = k[:, 1]
        nib = (hi >> np.uint64(60 - 4 * 5)) & np.uint64(0xF)
        parent = (hi >> np.uint64(64 - 4 * 5))
        corner = parent == parent[lv == 6].min()
        drop = (lv == 6) & corner & ((nib & np.uint64(4)) != 0)
        lv, xyz = lv[~drop], xyz[~drop]
    out = []
    for a in a_values:
        f = np.zeros((len(lv), 7), dtype=np.float32)
        f[:, 0] = rng.lognormal(0.0, 0.5, len(lv))
        f[:, 1] = rng.normal(0.0, 1e-6, len(lv)) * (1.0 + 0.1 * a)
        f[:, 2:5] = rng.normal(0.0, 1e-5, (len(lv), 3))
        f[:, 5] = a
        out.append(records_from(lv, xyz, f))
    return np.concatenate(out), lv, xyz


def by_key(rec):
    return rec[np.lexsort((rec["lo"], rec["hi"]))]


def last_is_max_key(rec):
    """Moves one record of the largest key to the end."""
    last = np.lexsort((rec["lo"], rec["hi"]))[-1]
    return np.concatenate([np.delete(rec, last), rec[last:last + 1]])


# ---------------------------------------------------------- keys, filter

@needs_ref
def test_keys_match_the_reference():
    R = RefHarness()
    rng = np.random.default_rng(1)
    lv = rng.integers(2, 33, 4000).astype(np.uint32)
    xyz = rng.uniform(-1.9, 1.9, (4000, 3)).astype(np.float32)
    # cell centres and cell corners, as the ASCII files hold them
    lvc = np.full(1000, 9, dtype=np.uint32)
    cen = ((rng.integers(0, 512, (1000, 3)) + 0.5) / 128.0 - 2.0).astype(np.float32)
    cor = (rng.integers(0, 512, (1000, 3)) / 128.0 - 2.0).astype(np.float32)
    for L, P in ((lv, xyz), (lvc, cen), (lvc, cor)):
        np.testing.assert_array_equal(capi.prep_keys(L, P), R.prep_keys(L, P))


@needs_ref
def test_filter_matches_the_reference():
    R = RefHarness()
    rng = np.random.default_rng(2)
    lv = rng.integers(3, 12, 20000).astype(np.uint32)
    xyz = rng.uniform(-1.99, 1.99, (20000, 3)).astype(np.float32)
    keys = R.prep_keys(lv, xyz)
    cones = capi.cones_fullsky(6, 2.0)
    _, cr, _, _ = capi.cones_narrow(4, 20.0, 30.0, 10.0, 10.0)
    for c in list(cones) + list(cr):
        for micro in (1.0 / 16, 1.0 / 4):
            a = capi.prep_collide(keys, micro, c)
            b = R.prep_collide(keys, micro, c)
            np.testing.assert_array_equal(a, b)
            assert 0 < a.sum() < len(a)


def test_filter_closed_forms():
    """Inside the central sphere always kept; far behind the cone dropped."""
    cone = np.array([0, 0, 0, 2.0, 0, 0, math.radians(10.0)])
    lv = np.array([8, 8, 8], dtype=np.uint32)
    xyz = np.array([[0.01, 0.01, 0.01], [1.0, 0.0, 0.0], [-1.0, 0.5, 0.5]], dtype=np.float32)
    keep = capi.prep_collide(capi.prep_keys(lv, xyz), 0.1, cone)
    np.testing.assert_array_equal(keep, [True, True, False])


# --------------------------------------------------------- octree calls

@needs_ref
def test_update_matches_the_reference():
    """Duplicates across scale factors (dphidt from the phi difference),
    ties in a, an unsorted container."""
    R = RefHarness()
    rng = np.random.default_rng(3)
    rec, _, _ = amr_records(rng, a_values=(0.7, 0.8, 0.8000001))
    rec = rec[rng.permutation(len(rec))]
    # the last input record is a copy of the largest key (the past-the-end read
    # of the reference then sees one of that cell's copies)
    rec = last_is_max_key(rec)
    P = capi.Prep()
    P.load_records(rec)
    P.update()
    got = P.records()
    exp = R.prep_update(rec)
    assert len(got) == len(exp) < len(rec)
    assert got.tobytes() == exp.tobytes()
    assert np.any(got["dphidt"] != 0)


@needs_ref
def test_update_drops_invalidated_cells():
    R = RefHarness()
    rng = np.random.default_rng(4)
    rec, _, _ = amr_records(rng, a_values=(0.6, 0.9))
    rec = rec[rng.permutation(len(rec))]
    rec[5]["lo"], rec[5]["hi"] = np.uint64(0xFFFFFFFFFFFFFFFF), np.uint64(0x7FFFFFFFFFFFFFFF)
    rec = last_is_max_key(rec)
    P = capi.Prep()
    P.load_records(rec)
    P.update()
    assert P.records().tobytes() == R.prep_update(rec).tobytes()


def capi_level(hi, lo):
    k = (hi << 64) | lo
    m = k & int("88888888888888888888888888888888", 16)
    if m == 0:
        return 0
    tz = (m & -m).bit_length() - 1
    return (128 - tz) // 4 + 1


@needs_ref
@pytest.mark.parametrize("order", (0, 1, 2))
def test_interpolation_matches_the_reference(order):
    """Uncached NGP/CIC/TSC on a sorted AMR container and on the same one
    with children appended at its end (the state during oct completion)."""
    R = RefHarness()
    rng = np.random.default_rng(6 + order)
    rec, _, _ = amr_records(rng)
    rec = by_key(rec)
    # the same a everywhere keeps the reference's past-the-end read harmless
    pts = rng.uniform(-0.55, 0.55, (3000, 3))
    pts = np.concatenate([pts, [[0.0, 0.0, 0.0]], rng.uniform(-0.5, -0.1, (500, 3))])
    for container in (rec, tail_appended(rec, rng)):
        P = capi.Prep()
        P.load_records(container)
        got = P.interpolate(order, pts)
        exp = R.prep_interpolate(container, order, pts)
        assert got.tobytes() == exp.tobytes()
        assert np.count_nonzero(got[:, 1]) > len(pts) // 4


def tail_appended(rec, rng):
    """Children of a few coarse cells appended unsorted at the end."""
    lvl = np.array([capi_level(int(h), int(l)) for h, l in zip(rec["hi"], rec["lo"])])
    coarse = np.nonzero(lvl == 5)[0]
    pick = rng.choice(coarse, 6, replace=False)
    extra = []
    for i in pick:
        k = (int(rec[i]["hi"]) << 64) | int(rec[i]["lo"])
        for s in rng.permutation(8):
            kc = k | ((8 | int(s)) << (124 - 4 * 5))
            e = rec[i].copy()
            e["hi"], e["lo"] = np.uint64(kc >> 64), np.uint64(kc & ((1 << 64) - 1))
            e["phi"] = e["phi"] * np.float32(1.0 + 0.01 * s)
            extra.append(e)
    return np.concatenate([rec, np.array(extra, dtype=rec.dtype)])


@needs_ref
def test_interpolate2_closed_form():
    """The reference's interpolate2 with a float abscissa on the table's
    descending a column gives da/dt = -a^2 H(a) (the correction tests use it
    for the dphi/da -> dphi/dt factor)."""
    R = RefHarness()
    table = synth.flat_lcdm_table(n=512)
    a = np.linspace(0.06, 0.999, 1000).astype(np.float32)
    f = R.prep_interpolate2(a, table[1], table[2])
    assert np.all(np.isfinite(f)) and np.all(f < 0)
    # closed form: the table's da/dt is -a^2 H(a), linear in between
    H0 = 0.7e5 / synth.MPC_SI
    ref = -a.astype(np.float64) ** 2 * H0 * np.sqrt(0.3 / a.astype(np.float64) ** 3 + 0.7)
    assert np.max(np.abs(f / ref - 1)) < 1e-3


# ---------------------------------------------------------- the correction

def sistemize(rec, h, lbox, mpc):
    """Input::sistemize (input.h:382-394), float64 arithmetic, float32 fields."""
    r = np.array(rec, copy=True)
    a = r["a"].astype(np.float64)
    t = (a * a * mpc) / (h * 100 * 1000)
    l = (a * lbox * mpc) / h
    r["phi"] = (r["phi"].astype(np.float64) * ((l * l) / (t * t))).astype(np.float32)
    for c in ("dphidx", "dphidy", "dphidz"):
        r[c] = (r[c].astype(np.float64) * ((l / (t * t)) * (-a))).astype(np.float32)
    r["dphidt"] = 0
    return r


@needs_ref
def test_correct_two_shells_matches_the_pinned_pipeline():
    """Complete octree seen at two scale factors: SI units, the duplicate
    merge (smallest a, dphi/da), the parent fill of cells seen once, and the
    dphi/dt conversion."""
    R = RefHarness()
    rng = np.random.default_rng(11)
    rec, _, _ = amr_records(rng, a_values=(0.8, 0.85))
    # cells of the fine corner seen at one scale factor only: their dphidt
    # comes from the parent
    lvl = np.array([capi_level(int(h), int(l)) for h, l in zip(rec["hi"], rec["lo"])])
    once = np.nonzero((lvl == 6) & (rec["a"] == np.float32(0.85)))[0][:20]
    rec = np.delete(rec, once)
    rec = rec[rng.permutation(len(rec))]
    rec = last_is_max_key(rec)
    table = synth.flat_lcdm_table(n=512)
    cfg = dict(h=0.7, lboxmpch=1000.0, mpc=capi.MPC_SI)
    P = capi.Prep(**cfg)
    P.load_records(rec)
    P.correct(table)
    got = P.records()
    # expected: the pinned steps
    s = sistemize(rec, 0.7, 1000.0, capi.MPC_SI)
    u = R.prep_update(s)
    keys = [(int(h) << 64) | int(l) for h, l in zip(u["hi"], u["lo"])]
    index = {k: i for i, k in enumerate(keys)}
    lv = np.array([capi_level(k >> 64, k & ((1 << 64) - 1)) for k in keys])
    for n in range(lv.min() + 1, lv.max() + 1):
        for i in np.nonzero(lv == n)[0]:
            if not np.isfinite(u[i]["dphidt"]) or u[i]["dphidt"] == 0 or abs(u[i]["dphidt"]) < np.finfo(np.float32).tiny:
                k = keys[i]
                parent = k & ~((1 << (128 - 4 * (n - 1))) - 1)
                u[i]["dphidt"] = u[index[parent]]["dphidt"]
    f = R.prep_interpolate2(u["a"], table[1], table[2])
    u["dphidt"] = (u["dphidt"] * f).astype(np.float32)
    assert got.tobytes() == u.tobytes()
    assert np.all(got["dphidt"] != 0)


@needs_ref
def test_correct_completes_partial_octs():
    """Children 4..7 of one oct missing: all eight children are appended,
    interpolated (TSC) on the container as it grows, then merged with the
    stored ones."""
    R = RefHarness()
    rng = np.random.default_rng(12)
    rec, _, _ = amr_records(rng, a_values=(0.8,), partial=True)
    rec = by_key(rec)
    table = synth.flat_lcdm_table(n=512)
    P = capi.Prep()
    P.load_records(rec)
    P.correct(table)
    got = P.records()
    # expected from the pinned calls: SI units, update, completion with the
    # reference's uncached TSC on the growing container, update, conversion
    s = R.prep_update(sistemize(rec, 0.7, 1000.0, capi.MPC_SI))
    keys = [(int(h) << 64) | int(l) for h, l in zip(s["hi"], s["lo"])]
    lv = [capi_level(k >> 64, k & ((1 << 64) - 1)) for k in keys]
    keyset = set(keys)
    container = s.copy()
    n0 = len(s)
    for i in range(n0):
        k, l = keys[i], lv[i]
        if i + 1 >= n0 or not (keys[i + 1] >> (128 - 4 * l)) == (k >> (128 - 4 * l)) or l == 0:
            continue   # leaf
        children = [k | ((8 | c) << (124 - 4 * l)) for c in range(8)]
        missing = sum(ref_find_missing(R, container[:n0], ch) for ch in children)
        if not missing:
            continue
        for ch in children:
            centre = [cell_center(ch, d) for d in range(3)]
            g = R.prep_interpolate(container, 2, np.array([centre]))[0]
            e = s[i].copy()
            e["hi"], e["lo"] = np.uint64(ch >> 64), np.uint64(ch & ((1 << 64) - 1))
            if np.any(g != 0):
                for j, name in enumerate(formats.FIELDS7):
                    e[name] = g[j]
            container = np.concatenate([container, np.array([e], dtype=s.dtype)])
    assert len(container) == n0 + 8
    u = R.prep_update(container)
    f = R.prep_interpolate2(u["a"], table[1], table[2])
    u["dphidt"] = (u["dphidt"] * f).astype(np.float32)
    assert len(got) == len(keyset) + 4
    assert got.tobytes() == u.tobytes()


@needs_ref
def test_correct_without_correction_coarsens_partial_octs():
    """correction = 0: the cell whose oct is partial is coarsened (its stored
    children dropped), as the reference's coarsen + update do."""
    R = RefHarness()
    rng = np.random.default_rng(18)
    rec, _, _ = amr_records(rng, a_values=(0.8,), partial=True)
    rec = by_key(rec)
    table = synth.flat_lcdm_table(n=512)
    P = capi.Prep(correction=0)
    P.load_records(rec)
    P.correct(table)
    got = P.records()
    s = R.prep_update(sistemize(rec, 0.7, 1000.0, capi.MPC_SI))
    keys = [(int(h) << 64) | int(l) for h, l in zip(s["hi"], s["lo"])]
    lv = [capi_level(k >> 64, k & ((1 << 64) - 1)) for k in keys]
    partial = []
    for i, (k, l) in enumerate(zip(keys, lv)):
        if i + 1 < len(s) and l and (keys[i + 1] >> (128 - 4 * l)) == (k >> (128 - 4 * l)):
            children = [k | ((8 | c) << (124 - 4 * l)) for c in range(8)]
            if sum(ref_find_missing(R, s, ch) for ch in children):
                partial.append(i)
    assert len(partial) == 1
    u = R.prep_coarsen(s, partial)
    f = R.prep_interpolate2(u["a"], table[1], table[2])
    u["dphidt"] = (u["dphidt"] * f).astype(np.float32)
    assert len(got) == len(s) - 4
    assert got.tobytes() == u.tobytes()


def ref_find_missing(R, rec, child):
    found, _ = R.prep_find(rec, np.array([[child & ((1 << 64) - 1), child >> 64]], dtype=np.uint64))
    return int(found[0] == len(rec))


def cell_center(k, axis):
    l = capi_level(k >> 64, k & ((1 << 64) - 1))
    c = 0
    for i in range(l):
        nib = (k >> (124 - 4 * i)) & 0xF
        c = (c << 1) | ((nib >> axis) & 1)
    return -2.0 + 4.0 * ((2 * c + 1) / float(1 << (l + 1)))


def test_correct_coarse_density_and_parent_fill():
    """coarsecorrection: a coarse cell with rho = 0 gets the mean of its
    normal neighbours; correction: a refined cell with rho = NaN gets its
    parent's."""
    rng = np.random.default_rng(13)
    rec, _, _ = amr_records(rng)
    rec = by_key(rec)
    lvl = np.array([capi_level(int(h), int(l)) for h, l in zip(rec["hi"], rec["lo"])])
    coarse = np.nonzero(lvl == 5)[0]
    target = coarse[len(coarse) // 2]
    rec[target]["rho"] = 0.0
    fine = np.nonzero(lvl == 6)[0][3]
    rec[fine]["rho"] = np.nan
    table = synth.flat_lcdm_table(n=512)
    P = capi.Prep(coarsecorrection=1)
    P.load_records(rec)
    P.correct(table)
    got = P.records()
    gk = [(int(h) << 64) | int(l) for h, l in zip(got["hi"], got["lo"])]
    index = {k: i for i, k in enumerate(gk)}
    # neighbours of the target at its level
    kt = (int(rec[target]["hi"]) << 64) | int(rec[target]["lo"])
    c = [0, 0, 0]
    for i in range(5):
        nib = (kt >> (124 - 4 * i)) & 0xF
        for d in range(3):
            c[d] = (c[d] << 1) | ((nib >> d) & 1)
    vals = []
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                if dx == dy == dz == 0:
                    continue
                kk = make_key(5, c[0] + dx, c[1] + dy, c[2] + dz)
                if kk in index and np.isfinite(got[index[kk]]["rho"]) and got[index[kk]]["rho"] != 0:
                    vals.append(np.float32(got[index[kk]]["rho"]))
    acc = np.float32(0)
    for v in vals:
        acc = np.float32(acc + v)
    assert got[index[kt]]["rho"] == np.float32(acc / np.float32(len(vals)))
    kf = (int(rec[fine]["hi"]) << 64) | int(rec[fine]["lo"])
    parent = kf & ~((1 << (128 - 4 * 5)) - 1)
    assert got[index[kf]]["rho"] == got[index[parent]]["rho"]


def make_key(l, x, y, z):
    k = 0
    for i in range(l):
        b = l - 1 - i
        nib = 8 | (((z >> b) & 1) << 2) | (((y >> b) & 1) << 1) | ((x >> b) & 1)
        k |= nib << (124 - 4 * i)
    return k


def test_acorrection_is_refused():
    with pytest.raises(capi.MpfError):
        capi.Prep(acorrection=1)


# ------------------------------------------------------------- ASCII files

def write_dat(path, level_r, xyz, rec):
    with open(path, "w") as f:
        for i in range(len(xyz)):
            r = rec[i]
            v = [float(xyz[i, 0]), float(xyz[i, 1]), float(xyz[i, 2])] + \
                [float(r[c]) for c in ("rho", "phi", "dphidx", "dphidy", "dphidz", "a")]
            f.write(f"{int(level_r[i])} " + " ".join(repr(x) for x in v) + "\n")


@needs_ref
def test_import_ascii_selects_as_the_reference(tmp_path):
    R = RefHarness()
    rng = np.random.default_rng(14)
    rec, lv, xyz = amr_records(rng)
    level_r = (lv - 2).astype(int)
    path = tmp_path / "cells_00000.dat"
    write_dat(path, level_r, xyz.astype(np.float32).astype(float), rec)
    cones = capi.cones_fullsky(4, 2.0)
    for micro in (8.0, 32.0):
        P = capi.Prep(microcoeff=micro)
        n = P.import_ascii(path, cones[1])
        got = P.records()
        keys = R.prep_keys(lv, xyz)
        keep = R.prep_collide(keys, 1.0 / micro, cones[1])
        assert n == keep.sum() == len(got)
        np.testing.assert_array_equal(key_words(got), keys[keep])
        for name in ("rho", "phi", "dphidx", "dphidy", "dphidz", "a"):
            np.testing.assert_array_equal(got[name], rec[name][keep])
        assert np.all(got["dphidt"] == 0)


@needs_ref
def test_import_ascii_narrow_cone_rotation(tmp_path):
    R = RefHarness()
    rng = np.random.default_rng(15)
    rec, lv, xyz = amr_records(rng)
    path = tmp_path / "cells_00001.dat"
    write_dat(path, (lv - 2).astype(int), xyz.astype(float), rec)
    _, cr, rot, _ = capi.cones_narrow(2, 2.0, 20.0, 20.0, 10.0, 10.0)
    P = capi.Prep(isfullsky=0, microcoeff=4.0)
    P.import_ascii(path, cr[0], rot)
    got = P.records()
    x64 = xyz.astype(np.float64)
    rx = np.zeros_like(xyz)
    for r in range(3):
        rx[:, r] = ((x64[:, 0] * rot[r, 0] + x64[:, 1] * rot[r, 1]) + x64[:, 2] * rot[r, 2]).astype(np.float32)
    keys = R.prep_keys(lv, rx)
    keep = R.prep_collide(keys, 0.25, cr[0])
    np.testing.assert_array_equal(key_words(got), keys[keep])
    f64 = np.stack([rec["dphidx"], rec["dphidy"], rec["dphidz"]], 1).astype(np.float64)
    fx = ((f64[:, 0] * rot[0, 0] + f64[:, 1] * rot[0, 1]) + f64[:, 2] * rot[0, 2]).astype(np.float32)
    np.testing.assert_array_equal(got["dphidx"], fx[keep])


def test_import_coarseonly_and_missing_file(tmp_path):
    rng = np.random.default_rng(16)
    rec, lv, xyz = amr_records(rng)
    path = tmp_path / "c.dat"
    write_dat(path, (lv - 2).astype(int), xyz.astype(float), rec)
    cone = capi.cones_fullsky(1, 2.0)[0]
    P = capi.Prep(coarseonly=1, ncoarse=3, microcoeff=1.0)
    P.import_ascii(path, cone)
    got = P.records()
    assert len(got) and all(capi_level(int(h), int(l)) == 5 for h, l in zip(got["hi"], got["lo"]))
    with pytest.raises(capi.MpfError):
        P.import_ascii(tmp_path / "none.dat", cone)


def test_prepare_ascii_writes_loadable_cone_files(tmp_path):
    """The driver: three cell files at two scale factors, two cones, one
    cone file each; the result equals the builder run by hand (the merge
    keeps the smallest a, so the files' order does not matter) and loads as
    an octree."""
    rng = np.random.default_rng(17)
    rec, lv, xyz = amr_records(rng, a_values=(0.8, 0.85))
    n = len(lv)
    celldir = tmp_path / "cells"
    celldir.mkdir()
    parts = np.array_split(np.arange(2 * n), 3)
    for j, p in enumerate(parts):
        write_dat(celldir / f"part_{j}.dat", (np.tile(lv, 2)[p] - 2).astype(int), np.tile(xyz, (2, 1))[p].astype(float),
                  rec[p])
    (celldir / "notes.txt").write_text("not a cell file\n")
    cones = capi.cones_fullsky(2, 2.0)
    table = synth.flat_lcdm_table(n=512)
    out = [tmp_path / "cone_0.bin", tmp_path / "cone_1.bin"]
    cfg = dict(microcoeff=4.0, order=2)
    capi.prepare_ascii(celldir, cones, out, table, seed=42, **cfg)
    for i in range(2):
        P = capi.Prep(**cfg)
        for j in range(3):
            P.import_ascii(celldir / f"part_{j}.dat", cones[i])
        P.correct(table)
        exp = P.records()
        got = formats.read_cone_file(out[i])
        assert got.tobytes() == exp.tobytes()
        ok, lmin, lmax, msg = capi.validate_octree(got)
        assert ok, msg
    # an existing output is refused (Input::save)
    with pytest.raises(capi.MpfError):
        capi.prepare_ascii(celldir, cones, out, table, seed=42, **cfg)


def test_import_malformed_file_follows_the_stream(tmp_path):
    """A token the stream cannot read (here `nan`) stops the reading: that
    value and every later one stay 0, as with std::ifstream >> (input.h:1028)."""
    path = tmp_path / "bad.dat"
    path.write_text("3 -0.4375 -0.4375 -0.4375 1.5 2e-06 1e-05 2e-05 3e-05 0.8\n"
                    "3 0.0625 0.0625 0.0625 nan 4e-06 1e-05 2e-05 3e-05 0.8\n"
                    "3 0.1875 0.1875 0.1875 1.5 2e-06 1e-05 2e-05 3e-05 0.8\n")
    cone = np.array([0, 0, 0, 2.0, 0, 0, math.radians(60.0)])
    P = capi.Prep(microcoeff=0.5)
    assert P.import_ascii(path, cone) == 3
    got = P.records()
    keys = capi.prep_keys(np.array([5, 5, 2], dtype=np.uint32),
                          np.array([[-0.4375] * 3, [0.0625] * 3, [0.0, 0.0, 0.0]], dtype=np.float32))
    np.testing.assert_array_equal(key_words(got), keys)
    assert got[0]["rho"] == np.float32(1.5) and got[0]["dphidz"] == np.float32(3e-05)
    assert got[1]["rho"] == 0 and got[1]["phi"] == 0 and got[1]["a"] == 0
    assert got[2]["rho"] == 0 and got[2]["a"] == 0
