"""C ABI checks that need no GPU: the library loads, exports every entry point
include/mpf.h declares, and its host-side octree validation accepts the
reference's cone-record format and rejects what the reference cannot search."""
import ctypes
import os
import re

import numpy as np
import pytest

import cases
from mpf import capi, formats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "mpf.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mpf_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    lib = capi.load_library()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), f"libmpf.so does not export {s}"
    assert sorted(capi.EXPORTS) == syms


def test_output_layout_matches_header():
    txt = open(os.path.join(ROOT, "include", "mpf.h")).read()
    enum = re.search(r"MPF_OUT_A,(.*?)MPF_NOUT", txt, flags=re.S).group(1)
    names = ["a"] + [n.strip()[len("MPF_OUT_"):].lower() for n in enum.split(",") if n.strip()]
    assert tuple(names) == capi.OUT_FIELDS


@pytest.mark.parametrize("name", cases.TREES)
def test_validate_accepts_test_trees(name):
    rec, _ = cases.tree(name)
    ok, lo, hi, msg = capi.validate_octree(rec)
    assert ok, msg
    lv = [bin(int(h)).count("1") for h in []]
    assert lo <= hi


def test_validate_rejects_bad_inputs():
    rec, _ = cases.tree("amr")
    # unsorted
    bad = rec.copy()
    bad[[3, 10]] = bad[[10, 3]]
    ok, _, _, msg = capi.validate_octree(bad)
    assert not ok and "sorted" in msg
    # duplicate key
    bad = np.concatenate([rec[:5], rec[4:]])
    ok, _, _, msg = capi.validate_octree(bad)
    assert not ok and "sorted" in msg
    # malformed key (a nibble without its marker bit)
    bad = rec.copy()
    bad["hi"][7] &= ~np.uint64(1 << 63)
    ok, _, _, msg = capi.validate_octree(bad)
    assert not ok
    # not parent-closed: drop a parent of refined cells
    lv = np.array([_level(h, l) for h, l in zip(rec["hi"], rec["lo"])])
    first_child = np.nonzero(lv == lv.max())[0][0]
    # the parent precedes its children in key order: find the nearest coarser record before it
    parent = first_child - 1
    while lv[parent] >= lv[first_child]:
        parent -= 1
    bad = np.delete(rec, parent)
    ok, _, _, msg = capi.validate_octree(bad)
    assert not ok and "parent" in msg
    # empty
    ok, _, _, msg = capi.validate_octree(rec[:0])
    assert not ok


def _level(hi, lo):
    k = (int(hi) << 64) | int(lo)
    l = 0
    while l < 32 and (k >> (127 - 4 * l)) & 1:
        l += 1
    return l


def test_validate_levels():
    lvl, kx, ky, kz = __import__("mpf.synth", fromlist=["x"]).uniform_level_coords(3, half=0.5)
    f = np.zeros((kx.size, 7))
    rec, _ = formats.make_records(lvl, kx, ky, kz, f)
    ok, lo, hi, msg = capi.validate_octree(rec)
    assert ok and lo == 5 and hi == 5
