"""Launch configuration of the persistent integrators (ray order, refill
granularity, grid size, occupancy variant) and its autotuner on the GPU.

The configuration only changes which lanes integrate which rays and when:
every output must be bit-identical to the default configuration's, for the
exact and split tiers, with and without surfaces, with uniform and per-ray
stop values (the path-length sort reads the latter)."""
import numpy as np
import pytest

import cases
from mpf import capi, synth

pytestmark = pytest.mark.gpu

CONFIGS = (
    dict(refill=1, sort_bucket=0, blocks_per_cu=0, waves_per_simd=0),
    dict(refill=64, sort_bucket=0, blocks_per_cu=2, waves_per_simd=0),
    dict(refill=16, sort_bucket=7, blocks_per_cu=4, waves_per_simd=1),
    dict(refill=32, sort_bucket=300, blocks_per_cu=16, waves_per_simd=2),
    dict(refill=1, sort_bucket=1 << 40, blocks_per_cu=8, waves_per_simd=3),
)


def ctx(name, precision):
    rec, f64 = cases.tree(name)
    c = capi.Context(cases.LENGTH, order=capi.TSC, precision=precision, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    return c


def random_rays(n, seed):
    ph, th = synth.cone_uniform(20.0, 20.0, 40.0, n, seed)
    zs = np.random.default_rng(seed + 1).uniform(0.05, 0.4, n)
    return synth.launch_photons(ph, th), zs


@pytest.mark.parametrize("precision", (capi.PREC_REF, capi.PREC_SPLIT))
@pytest.mark.parametrize("name", ("amr", "partial"))
def test_launch_configs_give_identical_results(name, precision):
    c = ctx(name, precision)
    in9, zs = random_rays(3000, 5)
    base = base_s = None
    for cfg in CONFIGS:
        if precision != capi.PREC_SPLIT and cfg["waves_per_simd"] > 0:
            continue
        print(f"[case] launch config tree={name} precision={precision} {cfg}", flush=True)
        c.set_launch_config(**cfg)
        c.set_launch_config(surfaces=True, **dict(cfg, waves_per_simd=min(cfg["waves_per_simd"], 2)))
        out, steps = c.integrate(in9, "redshift", 0.35)
        o2, surf = c.integrate_surfaces(in9, "redshift", None, jacobian=True, ray_value=zs)
        if base is None:
            base, base_s, base_o2 = (out, steps), surf, o2
            continue
        np.testing.assert_array_equal(out, base[0])
        assert steps == base[1]
        np.testing.assert_array_equal(o2, base_o2)
        np.testing.assert_array_equal(surf, base_s)
    c.close()


def test_autotune_returns_valid_config_and_same_results():
    print("[case] autotune split tier, plain and surfaces", flush=True)
    c = ctx("amr", capi.PREC_SPLIT)
    in9, zs = random_rays(20000, 9)
    n = len(in9)
    d_in, d_out = c.dev_alloc(in9.nbytes), c.dev_alloc(n * capi.NOUT * 8)
    d_rv, d_surf = c.dev_alloc(zs.nbytes), c.dev_alloc(n * capi.NSURF * 8)
    c.upload(d_in, in9)
    c.upload(d_rv, zs)
    c.integrate_dev(d_in, n, d_out, "redshift", 0.35, sync=True)
    ref = c.download(d_out, (n, capi.NOUT))
    best, trials = c.autotune_dev(d_in, n, d_out, "redshift", 0.35)
    assert len(trials) >= 8
    assert 1 <= best["refill"] <= 64 and best["blocks_per_cu"] in (4, 8, 16) and best["waves_per_simd"] in (1, 2, 3)
    assert c.launch_config() == best
    assert np.all(trials[:, 5] == trials[0, 5])   # every trial integrates the same steps
    np.testing.assert_array_equal(c.download(d_out, (n, capi.NOUT)), ref)
    c.integrate_dev(d_in, n, d_out, "redshift", 0.35, sync=True)
    np.testing.assert_array_equal(c.download(d_out, (n, capi.NOUT)), ref)
    # surfaces with per-ray source redshifts
    c.integrate_surfaces_dev(d_in, n, d_out, d_surf, "redshift", [0.0], jacobian=True, sync=True, ray_value_ptr=d_rv)
    sref = c.download(d_surf, (n, capi.NSURF))
    best_s, trials_s = c.autotune_dev(d_in, n, d_out, surf_ptr=d_surf, kind="redshift", values=[0.0],
                                      ray_value_ptr=d_rv)
    assert best_s["waves_per_simd"] in (1, 2) and len(trials_s) >= 8
    assert c.launch_config(surfaces=True) == best_s
    np.testing.assert_array_equal(c.download(d_surf, (n, capi.NSURF)), sref)
    for p in (d_in, d_out, d_rv, d_surf):
        c.dev_free(p)
    c.close()


def test_launch_config_rejects_bad_values():
    print("[case] launch config validation", flush=True)
    c = ctx("uniform", capi.PREC_REF)
    for bad in (dict(refill=0), dict(refill=65), dict(blocks_per_cu=33), dict(waves_per_simd=5)):
        with pytest.raises(capi.MpfError):
            c.set_launch_config(**bad)
    c.close()
