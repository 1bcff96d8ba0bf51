"""The boundary from compiled code on the GPU: tests/capi/mpf_caller.cpp (a
C++ program built against include/mpf.h, linked to libmpf.so) integrates
the golden rays in the REF tier through mpf_integrate, bit for bit equal to
the reference's own output (tests/golden/tree_amr.npz), and solves a
catalogue through mpf_catalogue_solve, equal to the same call made through
the Python binding."""
import os
import subprocess

import numpy as np
import pytest

import cases
from mpf import capi

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
CALLER = os.path.join(HERE, "capi", "_bin", "mpf_caller")


def inputs(d, name="amr"):
    rec, _ = cases.tree(name)
    rec.tofile(os.path.join(d, "records.bin"))
    np.ascontiguousarray(cases.cosmology()[:4], dtype=np.float64).tofile(os.path.join(d, "table.bin"))


def run(*args):
    assert os.path.exists(CALLER), "tests/capi/_bin/mpf_caller is built by __graft_entry__.build()"
    r = subprocess.run([CALLER, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_compiled_caller_rays_ref_tier_match_reference(tmp_path):
    print("[case] compiled caller: mpf_integrate, REF tier, golden rays", flush=True)
    d = str(tmp_path)
    inputs(d)
    g = np.load(os.path.join(HERE, "golden", "tree_amr.npz"))
    for rs in ("narrow", "wide"):
        g[f"in9_{rs}"].astype(np.float64).tofile(os.path.join(d, "in9.bin"))
        out = run("rays", d, capi.TSC, capi.PREC_REF, cases.NSTEPS, repr(cases.LENGTH), 0.3)
        res = np.fromfile(os.path.join(d, "out.bin"), dtype=np.float64).reshape(-1, capi.NOUT)
        ref = g[f"rays_{rs}_redshift_o2"]
        np.testing.assert_array_equal(res, ref)
        assert int(out.split()[1]) == int(np.sum(np.maximum(ref[:, capi.OUT["nsteps"]], 0)))


def test_compiled_caller_catalogue_matches_binding(tmp_path):
    print("[case] compiled caller: mpf_catalogue_solve, REF tier", flush=True)
    d = str(tmp_path)
    inputs(d)
    src = cases.sources(200, seed=6)
    src.tofile(os.path.join(d, "src6.bin"))
    run("catalogue", d, capi.TSC, capi.PREC_REF, cases.NSTEPS, repr(cases.LENGTH), 1e-8)
    rows = np.fromfile(os.path.join(d, "cat.bin"), dtype=np.float64).reshape(-1, capi.NCAT)
    rec, _ = cases.tree("amr")
    c = capi.Context(cases.LENGTH, order=capi.TSC, precision=capi.PREC_REF, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.set_cosmology(cases.cosmology())
    ref, _ = c.catalogue_solve(src, accuracy=1e-8, max_iterations=10, h=0.7, mpc=3.08568e22)
    c.close()
    assert np.sum(ref[:, capi.CAT["status"]] == capi.CAT_OK) > 100
    np.testing.assert_array_equal(rows, ref)


def test_device_memory_reports_the_device():
    """mpf_device_memory: free <= total, the total an MI355X's (about 288 GiB),
    and an allocation shows in the free bytes."""
    print("[case] mpf_device_memory", flush=True)
    free0, total = capi.device_memory(0)
    assert 0 < free0 <= total and total > 250e9
    c = capi.Context(cases.LENGTH)
    p = c.dev_alloc(8 << 30)
    free1, _ = capi.device_memory(0)
    c.dev_free(p)
    c.close()
    assert free0 - free1 >= (8 << 30) - (64 << 20)
