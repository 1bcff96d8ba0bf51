"""mpf_integrate_dev_static, the allocation-free variant of
mpf_integrate_dev, through the C ABI: the same outputs, bit for bit, as the
entry point it varies, on two calls in a row (the second on a smaller batch,
so the workspace array holds the first call's contents when it starts), in
the REF, FP64 and split tiers, with the ray order on and off; and the limit
on the batch size is an error, not a fault."""
import numpy as np
import pytest

import cases
from mpf import capi, synth

pytestmark = pytest.mark.gpu


def _ctx(precision):
    rec, f64 = cases.tree("amr")
    c = capi.Context(cases.LENGTH, order=capi.TSC, precision=precision, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    return c


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    ph = np.radians(20.0 + rng.uniform(-12, 12, n))
    th = np.radians(20.0 + rng.uniform(-12, 12, n))
    return synth.launch_photons(ph, th)


@pytest.mark.parametrize("precision", (capi.PREC_REF, capi.PREC_FP64, capi.PREC_SPLIT))
@pytest.mark.parametrize("sort_bucket", (0, 4096))
def test_static_matches_dev(precision, sort_bucket):
    print(f"[case] static integrate precision={precision} sort_bucket={sort_bucket}", flush=True)
    c = _ctx(precision)
    c.set_launch_config(refill=64 if sort_bucket else 1, sort_bucket=sort_bucket)
    nmax = 20000
    pin = c.dev_alloc(9 * 8 * nmax)
    pa = c.dev_alloc(capi.NOUT * 8 * nmax)
    pb = c.dev_alloc(capi.NOUT * 8 * nmax)
    try:
        for n, seed, stop, value in ((nmax, 1, "redshift", 0.3), (7001, 2, "radius", 0.4)):
            in9 = _rays(n, seed)
            c.upload(pin, in9)
            c.integrate_dev(pin, n, pa, stop, value, sync=True)   # builds the split blocks on first use
            ref_steps = c.last_run()[0]
            steps = c.integrate_dev_static(pin, n, pb, stop, value)
            a = c.download(pa, (n, capi.NOUT))
            b = c.download(pb, (n, capi.NOUT))
            assert steps == ref_steps > 0
            np.testing.assert_array_equal(b, a)
    finally:
        for p in (pin, pa, pb):
            c.dev_free(p)
        c.close()


def test_static_rejects_oversized_batch():
    print("[case] static integrate size limit", flush=True)
    c = _ctx(capi.PREC_REF)
    n = capi.STATIC_MAX_RAYS + 1
    pin = c.dev_alloc(9 * 8 * 2)
    pout = c.dev_alloc(capi.NOUT * 8 * 2)
    try:
        with pytest.raises(capi.MpfError, match="at most"):
            c.integrate_dev_static(pin, n, pout, "redshift", 0.3)
    finally:
        c.dev_free(pin)
        c.dev_free(pout)
        c.close()


def test_static_at_the_limit():
    print("[case] static integrate at MPF_STATIC_MAX_RAYS", flush=True)
    c = _ctx(capi.PREC_SPLIT)
    c.set_launch_config(refill=64, sort_bucket=65536)
    n = capi.STATIC_MAX_RAYS
    in9 = _rays(n, 3)
    pin = c.dev_alloc(9 * 8 * n)
    pa = c.dev_alloc(capi.NOUT * 8 * n)
    pb = c.dev_alloc(capi.NOUT * 8 * n)
    try:
        c.upload(pin, in9)
        c.integrate_dev(pin, n, pa, "redshift", 0.2, sync=True)
        c.integrate_dev_static(pin, n, pb, "redshift", 0.2)
        np.testing.assert_array_equal(c.download(pb, (n, capi.NOUT)), c.download(pa, (n, capi.NOUT)))
    finally:
        for p in (pin, pa, pb):
            c.dev_free(p)
        c.close()
