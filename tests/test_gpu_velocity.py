"""GPU parity of the velocity field through the C ABI (libmpf.so).

Integration with a velocity field (the Doppler and ISW terms of the
reference's velocity-field build, integrator.h:537-539, 594-595): bit for bit
equal to the oracle in both exact tiers, which is itself pinned against the
reference's integrator compiled with VELOCITYFIELD (tests/test_velocity.py).

Deposit, normalisation and parent fill (hmaps.h:439-541, hmaps.cpp:389-441):
the device sums the weights in FP64 with atomics and rounds once to FP32,
the reference (and the oracle) sum in FP32 one particle after the other, so
the two agree to FP32 rounding of sums of a few hundred terms: weights to
rtol 2e-6 + atol 1e-6, velocities (about 1e-3 c) to atol 1e-8."""
import numpy as np
import pytest

import cases
from mpf import capi
from oracle import Oracle

pytestmark = pytest.mark.gpu

STOPS = (("redshift", 0.3), ("none", 0.0), ("radius", 0.42), ("a", 0.8))


def ctx_for(name, order, precision, rho=None, vel=None, vobs=cases.VOBS):
    rec, f64 = cases.tree(name)
    c = capi.Context(cases.LENGTH, order=order, precision=precision, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    if vel is not None:
        c.velocity_set(vel, rho)
        c.set_observer_velocity(vobs)
    return c, rec, f64


@pytest.mark.parametrize("precision", (capi.PREC_REF, capi.PREC_FP64))
@pytest.mark.parametrize("name", ("amr", "partial"))
def test_velocity_integration_vs_oracle(name, precision):
    rec, f64 = cases.tree(name)
    rho, vel = cases.velocity_field(rec)
    O = Oracle()
    h = O.tree(rec, f64)
    O.set_velocity(h, rho, vel, cases.VOBS)
    tab = cases.cosmology()
    in9 = cases.rays(n_side=8, aperture=12.0)
    for order in (0, 1, 2):
        c, _, _ = ctx_for(name, order, precision, rho, vel)
        for stop, value in STOPS:
            print(f"[case] velocity integrate tree={name} precision={precision} order={order} stop={stop}",
                  flush=True)
            out, steps = c.integrate(in9, stop, value)
            ref, rsteps = O.integrate(h, order, tab, cases.LENGTH, cases.NSTEPS, cases.STOP_KIND[stop], value, in9,
                                      precision=precision)
            np.testing.assert_array_equal(out, ref, err_msg=f"{name} order {order} {stop}")
            assert steps == rsteps
        c.close()
    O.free(h)


def test_velocity_changes_redshift_only():
    """The velocities change no trajectory: every column but the redshift
    equals the run without a field; the redshift differs."""
    print("[case] velocity redshift only", flush=True)
    rec, _ = cases.tree("amr")
    rho, vel = cases.velocity_field(rec)
    in9 = cases.rays(n_side=6, aperture=12.0)
    c, _, _ = ctx_for("amr", capi.TSC, capi.PREC_FP64)
    c.velocity_set(vel)                 # rho kept
    c.set_observer_velocity(cases.VOBS)
    with_v, _ = c.integrate(in9, "none")
    c.velocity_clear()
    without, _ = c.integrate(in9, "none")
    k = capi.OUT["redshift"]
    assert np.array_equal(np.delete(with_v, k, axis=1), np.delete(without, k, axis=1))
    assert np.all(with_v[:, k] != without[:, k])
    c.close()


@pytest.mark.parametrize("name", ("amr", "coarse"))
def test_velocity_surfaces_vs_oracle(name):
    """Redshift surfaces (hmaps' inhomogeneous redshift maps) with the Doppler
    term: bit for bit equal to the oracle in the reference tier."""
    rec, f64 = cases.tree(name)
    rho, vel = cases.velocity_field(rec)
    c, _, _ = ctx_for(name, capi.TSC, capi.PREC_REF, rho, vel)
    O = Oracle()
    h = O.tree(rec, f64)
    O.set_velocity(h, rho, vel, cases.VOBS)
    tab = cases.cosmology()
    in9 = cases.rays(n_side=6, aperture=12.0)
    vals = [0.1, 0.2, 0.3]
    print(f"[case] velocity surfaces tree={name}", flush=True)
    out, surf = c.integrate_surfaces(in9, "redshift", vals, jacobian=True)
    ref = np.stack([O.surfaces(h, 2, tab, cases.LENGTH, cases.NSTEPS, capi.STOP_NAMES["redshift"], vals, r)
                    for r in in9])
    valid = ref[..., 0] == 1.0
    assert valid.sum() > 0.5 * valid.size
    for fld in ("valid", "firstid", "f", "x", "y", "z", "chi", "t", "a", "redshift", "isw", "rho", "phi"):
        kk = capi.SURF[fld]
        np.testing.assert_array_equal(surf[valid][:, kk], ref[valid][:, kk], err_msg=fld)
    O.free(h)
    c.close()


@pytest.mark.parametrize("scheme", ("cic", "tsc"))
@pytest.mark.parametrize("name", ("amr", "partial"))
def test_velocity_deposit_vs_oracle(name, scheme):
    print(f"[case] velocity deposit tree={name} scheme={scheme}", flush=True)
    rec, f64 = cases.tree(name)
    pa, va = cases.particles(400, seed=21)
    pb, vb = cases.particles(300, seed=22, spread=0.2)
    c, _, _ = ctx_for(name, capi.TSC, capi.PREC_REF)
    c.velocity_field([(pa, va), (pb, vb)], scheme)
    rho, vel = c.velocity_get()
    O = Oracle()
    h = O.tree(rec)
    sid = 1 if scheme == "cic" else 2
    r1, s1 = O.velocity_deposit(h, rec.size, sid, pa, va, finish=False)
    r2, v2 = O.velocity_deposit(h, rec.size, sid, pb, vb, rho=r1, vel3=s1)
    O.free(h)
    assert np.count_nonzero(r2) > 200
    np.testing.assert_array_equal(rho == 0, r2 == 0)
    np.testing.assert_allclose(rho, r2, rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(vel, v2, rtol=0, atol=1e-8)
    # the finished field drives the integration: bit for bit the oracle's on the same field
    O = Oracle()
    h = O.tree(rec)
    O.set_velocity(h, rho, vel, cases.VOBS)
    c.set_observer_velocity(cases.VOBS)
    in9 = cases.rays(n_side=6, aperture=12.0)
    out, steps = c.integrate(in9, "redshift", 0.3)
    ref, rsteps = O.integrate(h, 2, cases.cosmology(), cases.LENGTH, cases.NSTEPS, 2, 0.3, in9)
    np.testing.assert_array_equal(out, ref)
    assert steps == rsteps
    O.free(h)
    c.close()


def test_velocity_empty_and_refusals():
    print("[case] velocity refusals", flush=True)
    rec, _ = cases.tree("amr")
    c, _, _ = ctx_for("amr", capi.TSC, capi.PREC_REF)
    with pytest.raises(capi.MpfError):
        c.velocity_get()                       # no field yet
    with pytest.raises(ValueError):
        c.velocity_field([], "ngp")
    # no particles: every weight and velocity zero (hmaps.cpp:391, 423-441)
    c.velocity_field([], "cic")
    rho, vel = c.velocity_get()
    assert not rho.any() and not vel.any()
    # the split tier does not carry the field
    c.set_precision(capi.PREC_SPLIT)
    with pytest.raises(capi.MpfError):
        c.integrate(cases.rays(n_side=2), "redshift", 0.3)
    c.velocity_clear()
    c.set_precision(capi.PREC_REF)
    out, _ = c.integrate(cases.rays(n_side=2), "redshift", 0.3)
    assert np.all(out[:, capi.OUT["nsteps"]] > 0)
    c.close()
