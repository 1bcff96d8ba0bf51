"""Constant-parameter surfaces (hmaps.h:615-660) and the infinitesimal lensing
Jacobian (lensing.h:399-480) on the GPU, against the oracle's restatement of
Hmaps::FillMap on its stored trajectory (whose trajectory and Jacobian are
pinned to the reference's own code, tests/test_oracle.py).

Exact tiers: surface index, interpolation factor and every interpolated
quantity bit-identical; the Jacobian within 1e-12 absolute + 1e-9 of |A - 1|
(the device uses the position's own sines and cosines and S1 - S2/D instead
of libm trigonometry and a per-distance sum).  Split tier against the FP64
tier: relative redshift <= 1e-7, positions <= 1e-9, Jacobian within
1e-10 absolute + 1e-5 of |A - 1|, rho and rhomax within 1e-5 relative (FP32
polynomial blocks of rho)."""
import numpy as np
import pytest

import cases
from mpf import capi
from oracle import Oracle

pytestmark = pytest.mark.gpu
EXACT_FIELDS = ("valid", "firstid", "f", "x", "y", "z", "dist", "chi", "lambda", "t", "a", "redshift", "s", "isw",
                "rho", "phi", "rhomax", "kt", "kx", "ky", "kz")
SURFS = {"redshift": [0.08, 0.16, 0.25, 0.33], "radius": [0.1, 0.2, 0.3, 0.38], "a": [0.95, 0.85, 0.78],
         "lambda": [2e16, 4e16, 6e16], "t": [3e16, 6e16, 1e17], "s": [0.12, 0.25, 0.36]}


def ctx(name, precision):
    rec, f64 = cases.tree(name)
    c = capi.Context(cases.LENGTH, order=capi.TSC, precision=precision, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    return c, rec, f64


def jac_close(a, b, rel, ab):
    for f in ("a11", "a12", "a22"):
        x = a[..., capi.SURF[f]]
        y = b[..., capi.SURF[f]]
        one = 1.0 if f != "a12" else 0.0
        tol = ab + rel * np.abs(y - one)
        assert np.all(np.abs(x - y) <= tol), f"{f}: max excess {np.max(np.abs(x - y) - tol):.3e}"


@pytest.mark.parametrize("precision", (capi.PREC_REF, capi.PREC_FP64))
@pytest.mark.parametrize("name", ("amr", "coarse"))
def test_surfaces_exact_vs_oracle(name, precision):
    c, rec, f64 = ctx(name, precision)
    O = Oracle()
    h = O.tree(rec, f64)
    tab = cases.cosmology()
    in9 = cases.rays(n_side=6, aperture=40.0)
    for kind, vals in SURFS.items():
        print(f"[case] surfaces tree={name} precision={precision} kind={kind}", flush=True)
        out, surf = c.integrate_surfaces(in9, kind, vals, jacobian=True)
        ref = np.stack([O.surfaces(h, 2, tab, cases.LENGTH, cases.NSTEPS, capi.STOP_NAMES[kind], vals, r,
                                   precision=precision) for r in in9])
        valid = ref[..., 0] == 1.0
        np.testing.assert_array_equal(surf[..., 0], ref[..., 0], err_msg=f"{kind} valid")
        np.testing.assert_array_equal(surf[..., 1], ref[..., 1], err_msg=f"{kind} firstid")
        for fld in EXACT_FIELDS[2:]:
            k = capi.SURF[fld]
            np.testing.assert_array_equal(surf[valid][:, k], ref[valid][:, k], err_msg=f"{kind} {fld}")
        jac_close(surf[valid], ref[valid], 1e-9, 1e-12)
        assert valid.sum() > 0.5 * valid.size
    O.free(h)
    c.close()


@pytest.mark.parametrize("tidal", (1, 0))
@pytest.mark.parametrize("name", ("amr", "uniform"))
def test_surfaces_split_vs_fp64(name, tidal):
    """Split against FP64, with the tidal blocks (the Jacobian's samples
    from one block per qualifying cell) and with every sample looked up."""
    print(f"[case] surfaces split-vs-fp64 tree={name} tidal={tidal}", flush=True)
    c64, _, _ = ctx(name, capi.PREC_FP64)
    cs, _, _ = ctx(name, capi.PREC_SPLIT)
    cs.split_tidal_blocks(tidal)
    in9 = cases.rays(n_side=24, aperture=40.0)
    vals = SURFS["redshift"]
    o64, s64 = c64.integrate_surfaces(in9, "redshift", vals)
    os_, ss = cs.integrate_surfaces(in9, "redshift", vals)
    nt = cs.split_tidal_blocks()
    print(f"  tidal blocks: {nt}", flush=True)
    assert (nt > 0) == bool(tidal)
    same = (s64[..., 0] == ss[..., 0]) & (s64[..., 1] == ss[..., 1])
    assert np.mean(~same) <= 0.02
    ok = same & (s64[..., 0] == 1.0)
    assert ok.sum() > 0.5 * ok.size
    a, b = ss[ok], s64[ok]
    rz = np.abs(a[:, capi.SURF["redshift"]] - b[:, capi.SURF["redshift"]]) / np.abs(b[:, capi.SURF["redshift"]])
    assert rz.max() <= 1e-7
    dp = np.abs(a[:, capi.SURF["dist"]] - b[:, capi.SURF["dist"]]) / b[:, capi.SURF["dist"]]
    assert dp.max() <= 1e-9
    jac_close(a, b, 1e-5, 1e-10)
    # densities from the split tier's FP32 rho blocks: the value at the
    # crossing and the running maximum before it
    for f in ("rho", "rhomax"):
        x, y = a[:, capi.SURF[f]], b[:, capi.SURF[f]]
        assert np.all(np.isfinite(x)), f
        assert np.max(np.abs(x - y) / np.abs(y)) <= 1e-5, f
    c64.close()
    cs.close()


def test_surfaces_flrw_identity():
    """phi == 0: the Jacobian is the identity and surfaces sit at the background distances."""
    print("[case] surfaces FLRW identity", flush=True)
    from mpf import formats, synth
    lvl, kx, ky, kz = synth.uniform_level_coords(4, half=0.5)
    f = np.zeros((kx.size, 7))
    f[:, 5] = 1.0
    rec, _ = formats.make_records(lvl, kx, ky, kz, f)
    for prec in (capi.PREC_FP64, capi.PREC_SPLIT):
        c = capi.Context(cases.LENGTH, order=capi.TSC, precision=prec)
        c.load_records(rec)
        c.set_cosmology(cases.cosmology())
        out, surf = c.integrate_surfaces(cases.rays(4, aperture=20.0), "redshift", [0.1, 0.2])
        assert np.all(surf[..., 0] == 1.0)
        np.testing.assert_allclose(surf[..., capi.SURF["a11"]], 1.0, atol=1e-15)
        np.testing.assert_allclose(surf[..., capi.SURF["a22"]], 1.0, atol=1e-15)
        np.testing.assert_allclose(surf[..., capi.SURF["a12"]], 0.0, atol=1e-15)
        # all rays reach each surface at the same distance
        d = surf[..., capi.SURF["dist"]]
        np.testing.assert_allclose(d, np.broadcast_to(d[:1], d.shape), rtol=1e-12)
        c.close()
