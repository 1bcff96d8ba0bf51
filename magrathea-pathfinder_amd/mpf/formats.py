"""Cone-file record format and 128-bit octree keys.

A cone file is a packed array of 48-byte records, one per octree cell, sorted
by key (datahandler.h:940, input.h:1298-1310; SIZEOFOCTREE=48 in Makefile:43):

    [0, 16)   key: little-endian unsigned 128-bit SimpleHyperOctreeIndex
    [16, 44)  Gravity<float, 3> contents, tuple stored in reverse order:
              dphidt@16 a@20 dphidx@24 dphidy@28 dphidz@32 phi@36 rho@40
    [44, 48)  padding

Keys hold 4 bits per level from the most significant bit; the nibble of level
l is 8 | z<<2 | y<<1 | x with (x, y, z) the l-th bits (from the top) of the
per-axis cell coordinates (simplehyperoctreeindex.h:1167-1189).  Octree
positions span [-2, 2)^3 (EXTENT = 4, Makefile:34).
"""
from __future__ import annotations

import numpy as np

EXTENT = 4.0

RECORD_DTYPE = np.dtype([
    ("lo", "<u8"), ("hi", "<u8"),
    ("dphidt", "<f4"), ("a", "<f4"), ("dphidx", "<f4"), ("dphidy", "<f4"),
    ("dphidz", "<f4"), ("phi", "<f4"), ("rho", "<f4"), ("pad", "<u4"),
])
assert RECORD_DTYPE.itemsize == 48

# Field order of the 7-vectors used by the C ABI and the oracle
# (gravity.h:45-87 tuple order): rho, phi, dphidx, dphidy, dphidz, a, dphidt.
FIELDS7 = ("rho", "phi", "dphidx", "dphidy", "dphidz", "a", "dphidt")


def keys_from_coords(level, kx, ky, kz):
    """Per-axis integer cell coordinates at `level` -> (hi, lo) uint64 key words."""
    level = np.asarray(level, dtype=np.int64)
    kx = np.asarray(kx, dtype=np.uint64)
    ky = np.asarray(ky, dtype=np.uint64)
    kz = np.asarray(kz, dtype=np.uint64)
    level, kx, ky, kz = np.broadcast_arrays(level, kx, ky, kz)
    hi = np.zeros(kx.shape, dtype=np.uint64)
    lo = np.zeros(kx.shape, dtype=np.uint64)
    lmax = int(level.max()) if level.size else 0
    for i in range(lmax):
        active = level > i
        b = (level - 1 - i).clip(min=0).astype(np.uint64)
        nib = (np.uint64(8) | (((kz >> b) & np.uint64(1)) << np.uint64(2))
               | (((ky >> b) & np.uint64(1)) << np.uint64(1)) | ((kx >> b) & np.uint64(1)))
        nib = np.where(active, nib, np.uint64(0))
        if i < 16:
            hi |= nib << np.uint64(60 - 4 * i)
        else:
            lo |= nib << np.uint64(60 - 4 * (i - 16))
    return hi, lo


def cell_centers(level, kx, ky, kz):
    """Centre of cells: -E/2 + E (2k+1) / 2^(l+1) (simplehyperoctreeindex.h:890-893)."""
    level = np.asarray(level, dtype=np.int64)
    den = np.ldexp(1.0, level + 1)
    f = lambda k: -EXTENT / 2 + EXTENT * ((2 * np.asarray(k, dtype=np.float64) + 1) / den)
    return f(kx), f(ky), f(kz)


def make_records(level, kx, ky, kz, fields7):
    """Build a key-sorted record array.  fields7: (n, 7) float array in FIELDS7 order."""
    hi, lo = keys_from_coords(level, kx, ky, kz)
    fields7 = np.asarray(fields7)
    rec = np.zeros(hi.shape[0], dtype=RECORD_DTYPE)
    rec["hi"] = hi
    rec["lo"] = lo
    for j, name in enumerate(FIELDS7):
        rec[name] = fields7[:, j].astype(np.float32)
    order = np.lexsort((lo, hi))
    return rec[order], order


def records_fields7(rec):
    """(n, 7) float32 array in FIELDS7 order from a record array."""
    return np.stack([rec[name] for name in FIELDS7], axis=1)


def key_int(hi, lo):
    return (int(hi) << 64) | int(lo)


def read_cone_file(path):
    """Raw read of a cone file (input.h:1298-1310)."""
    data = np.fromfile(path, dtype=np.uint8)
    if data.size % 48:
        raise ValueError(f"{path}: size {data.size} is not a multiple of 48")
    return data.view(RECORD_DTYPE)


def write_cone_file(path, rec):
    """Raw write of a cone file (input.h:1276-1292)."""
    np.asarray(rec, dtype=RECORD_DTYPE).tofile(path)
