"""Statistics of the rays program (rays.cpp:300-456) on the host
(mpf_ray_statistics), over central trajectories and distances from the
oracle's propagate (pinned to the reference's own propagate in
tests/test_propagate.py).  rays.cpp computes its statistics inside main(),
which no harness can call: the reduction is pinned by re-interpolating each
curve with the reference's own Utility::interpolate (the compiled reference
harness) and by closed forms
(identical rays: zero spread; the redshift against itself: the identity).
CPU only."""
import numpy as np
import pytest

import cases
from mpf import capi
from oracle import Oracle, RefHarness, ref_available


@pytest.fixture(scope="module")
def bundles():
    O = Oracle()
    rec, f64 = cases.tree("wide")
    h = O.tree(rec)
    tab = cases.cosmology()
    in9 = cases.rays(n_side=4, aperture=30.0)
    out = [O.propagate(h, 2, tab, cases.LENGTH, cases.NSTEPS, r, 6, 0.01, 0.2 * i) for i, r in enumerate(in9)]
    ref = O.propagate(h, 2, tab, cases.LENGTH, cases.NSTEPS, in9[5], 6, 0.01, 0.0)
    O.free(h)
    return [o[0] for o in out], [o[1] for o in out], ref


def test_statistics_reinterpolation_matches_reference_interpolate(bundles):
    if not ref_available():
        pytest.skip("reference harness not built")
    R = RefHarness()
    trajs, dists, (rr, rd) = bundles
    inv = 1.0 / cases.LENGTH
    out, used = capi.ray_statistics(trajs, dists, rr, rd, "redshift", "distance", nstat=50, length=cases.LENGTH)
    assert used == sum(len(t) > 0 for t in trajs) and used > 5
    mod = max(1, len(rr) // 50)
    refx = rr[::mod, 9]
    np.testing.assert_array_equal(out[:, 0], rr[::mod, 9])
    np.testing.assert_array_equal(out[:, 1], rd[::mod] * inv)
    ys = np.stack([R.interpolate(t[:, 9], d * inv, refx) for t, d in zip(trajs, dists) if len(t)])
    mean = np.zeros(len(refx))
    for y in ys:
        mean += y
    mean /= len(ys)
    np.testing.assert_allclose(out[:, 2], mean, rtol=1e-14)
    np.testing.assert_allclose(out[:, 3], ys.std(axis=0, ddof=1), rtol=1e-10)


def test_statistics_closed_forms(bundles):
    trajs, dists, (rr, rd) = bundles
    t0 = [t for t in trajs if len(t)][0]
    d0 = [d for t, d in zip(trajs, dists) if len(t)][0]
    # the same ray five times: mean = that ray, spread 0 (to the rounding of the mean)
    out, used = capi.ray_statistics([t0] * 5, [d0] * 5, rr, rd, "lambda", "distance2", nstat=30,
                                    length=cases.LENGTH)
    one, _ = capi.ray_statistics([t0] * 2, [d0] * 2, rr, rd, "lambda", "distance2", nstat=30, length=cases.LENGTH)
    assert used == 5
    np.testing.assert_allclose(out[:, 2], one[:, 2], rtol=1e-15)
    assert np.all(out[:, 3] <= 1e-15 * np.abs(out[:, 2]))
    # the redshift against itself (inhomogeneous statistic over the redshift abscissa): the identity
    out, _ = capi.ray_statistics(trajs, dists, rr, rd, "redshift", "inhomogeneous", nstat=40, length=cases.LENGTH)
    inside = (out[:, 0] > 0.01) & (out[:, 0] < 0.9 * min(t[-1, 9] for t in trajs if len(t)))
    np.testing.assert_allclose(out[inside, 2], out[inside, 0], rtol=1e-12)
    # amin drops the reference elements beyond it
    out_a, _ = capi.ray_statistics(trajs, dists, rr, rd, "redshift", "distance", nstat=10 ** 6, amin=0.7,
                                   length=cases.LENGTH)
    assert len(out_a) == np.sum(rr[:, 0] >= 0.7)
