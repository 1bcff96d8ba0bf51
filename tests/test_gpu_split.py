"""GPU checks of the split-precision tier (FP32 perturbation, FP64 background)
against the FP64 tier on the same rays.

Tolerance (stated by the benchmark contract): relative end-redshift error
<= 1e-7 against the FP64 tier; end positions within 1e-9 of the distance
travelled; identical step counts except on rays whose last step crosses the
stop surface within the precision difference (allowed on at most 2% of rays).
"""
import numpy as np
import pytest

import cases
from mpf import capi

pytestmark = pytest.mark.gpu
Z_TOL = 1e-7
POS_TOL = 1e-9


def run(name, precision, in9, stop, val, variant=None):
    rec, f64 = cases.tree(name)
    c = capi.Context(cases.LENGTH, order=capi.TSC, precision=precision, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    if variant is not None and precision == capi.PREC_SPLIT:
        c.set_launch_config(waves_per_simd=variant)
    out, steps = c.integrate(in9, stop, val)
    c.close()
    return out, steps


@pytest.mark.parametrize("variant", (1, 2, 3, 4))
@pytest.mark.parametrize("name", cases.TREES)
def test_split_vs_fp64(name, variant):
    print(f"[case] split-vs-fp64 tree={name} variant={variant}", flush=True)
    in9 = cases.rays(n_side=32, aperture=50.0)
    for stop, val in (("redshift", 0.35), ("none", 0.0), ("radius", 0.3)):
        a, sa = run(name, capi.PREC_SPLIT, in9, stop, val, variant)
        b, sb = run(name, capi.PREC_FP64, in9, stop, val)
        na, nb = a[:, capi.OUT["nsteps"]], b[:, capi.OUT["nsteps"]]
        same = na == nb
        assert np.mean(~same) <= 0.02, f"{stop}: step counts differ on {np.sum(~same)} rays"
        ok = same & (nb > 0)
        assert ok.sum() > 0.5 * len(in9)
        rz = np.abs(a[ok, capi.OUT["redshift"]] - b[ok, capi.OUT["redshift"]]) / np.abs(b[ok, capi.OUT["redshift"]])
        assert rz.max() <= Z_TOL, f"{stop}: max relative redshift error {rz.max():.3e}"
        dp = np.linalg.norm(a[ok, 2:5] - b[ok, 2:5], axis=1) / np.maximum(b[ok, capi.OUT["s"]], 1e-30)
        assert dp.max() <= POS_TOL, f"{stop}: max position error {dp.max():.3e}"
        for f in ("a", "t", "lambda", "chi", "s"):
            r = np.abs(a[ok, capi.OUT[f]] - b[ok, capi.OUT[f]]) / np.maximum(np.abs(b[ok, capi.OUT[f]]), 1e-300)
            assert r.max() <= 1e-8, f"{stop}: {f} relative error {r.max():.3e}"
        # rays that end with no data (holes, edge) end the same way
        np.testing.assert_array_equal(na < 0, nb < 0)
        # the finest cell of the last point is found at the same level
        # (a cell boundary within the position tolerance aside)
        lv = a[ok, capi.OUT["level"]] == b[ok, capi.OUT["level"]]
        assert np.mean(~lv) <= 0.01, f"{stop}: level differs on {np.sum(~lv)} rays"


def test_split_requires_tsc():
    print("[case] split tier with CIC is refused", flush=True)
    rec, f64 = cases.tree("uniform")
    c = capi.Context(cases.LENGTH, order=capi.CIC, precision=capi.PREC_SPLIT)
    c.load_records(rec)
    c.set_cosmology(cases.cosmology())
    with pytest.raises(capi.MpfError):
        c.integrate(cases.rays(2), "redshift", 0.3)
    c.close()


def test_split_device_resident_repeat():
    """Repeated device-resident runs give identical results and step counts."""
    print("[case] split device-resident repeat", flush=True)
    rec, f64 = cases.tree("amr")
    c = capi.Context(cases.LENGTH, order=capi.TSC, precision=capi.PREC_SPLIT)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    in9 = cases.rays(n_side=40, aperture=40.0)
    d_in = c.dev_alloc(in9.nbytes)
    c.upload(d_in, in9)
    d_out = c.dev_alloc(len(in9) * capi.NOUT * 8)
    outs = []
    for _ in range(3):
        c.integrate_dev(d_in, len(in9), d_out, "redshift", 0.35, sync=True)
        steps, ms = c.last_run()
        outs.append((c.download(d_out, (len(in9), capi.NOUT)), steps))
    c.dev_free(d_in)
    c.dev_free(d_out)
    c.close()
    for o, s in outs[1:]:
        np.testing.assert_array_equal(o, outs[0][0])
        assert s == outs[0][1]
