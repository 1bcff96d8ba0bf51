"""The benchmark workloads themselves under test, through the C ABI.

uniform128_cic (128^3 uniform grid, 4096 rays, CIC, to z = 0.5) in full: the
REF tier bit-identical to the reference's own code (oracle/_ref) on every
ray, the FP64 tier bit-identical to the oracle's FP64 mode.

amr256_tsc (37.5 M cells, 1024 x 1024 rays over the 10 x 10 deg narrow field,
TSC, to z = 2): a 16,384-ray stride in the REF tier bit-identical to the
reference; all rays in the split tier against the FP64 tier, relative
end-redshift and end-position error <= 1e-7 (the split tier's bound).

The uneven Jacobian batch (4,194,304 rays uniform over the field, per-ray
source redshift in [0.1, 2], infinitesimal Jacobian there): convergence
kappa = 1 - (a11 + a22)/2 and shear gamma1 = -(a11 - a22)/2, gamma2 = -a12 of
the split tier against the FP64 tier, |error| <= 1e-7 (absolute; their rms
is about 1e-3 here), stated in the output; two split runs of the whole batch
bit-identical to each other (bit patterns), and their rho and rhomax within
1e-5 relative of the FP64 tier's.

The catalogue (1,048,576 sources uniform in comoving volume in the field, z
in [0.1, 1.5], Newton to 1e-8 rad, split tier): every source inside the
field converges or is rejected after max_iterations as the reference would
(at most 0.1 %), the Born-seeded and plain runs agree within 20 x the
accuracy; 256 of the sources in the REF tier against the oracle (status and
iterations exact, angles within 1e-12 rad)."""
import os
import time

import numpy as np
import pytest

import cases
from mpf import capi

pytestmark = pytest.mark.gpu
capi.load_library()          # the HIP runtime before torch (the generators use torch on the CPU)
from mpf import proposal, synth  # noqa: E402
import oracle as orc             # noqa: E402

_W = {}


def workload(name):
    if name not in _W:
        t0 = time.time()
        _W[name] = synth.build_workload(name)
        print(f"[workload] {name} built in {time.time() - t0:.1f}s", flush=True)
    return _W[name]


def context(name, precision):
    w, rec, f64, length, table, _ = workload(name)
    c = capi.Context(length, order=w["order"], precision=precision, nsteps=10)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(table)
    return c


def kappa_gamma(surf):
    """kappa, gamma1, gamma2 of per-ray surface records (hmaps.h:759-762)."""
    a11, a12, a22 = (surf[:, capi.SURF[k]] for k in ("a11", "a12", "a22"))
    return np.stack([1 - (a11 + a22) / 2, -(a11 - a22) / 2, -a12], 1)


@pytest.mark.skipif(not orc.ref_available(), reason="oracle/_ref not built")
def test_uniform128_cic_full_ref_tier_vs_reference():
    print("[case] uniform128_cic: all 4096 rays, REF tier vs the reference", flush=True)
    w, rec, f64, length, table, in9 = workload("uniform128_cic")
    c = context("uniform128_cic", capi.PREC_REF)
    out, steps = c.integrate(in9, "redshift", w["z_stop"])
    c.close()
    R = orc.RefHarness()
    h = R.tree(rec)
    ref, rsteps = R.integrate(h, w["order"], table, length, 10, 2, w["z_stop"], in9, nthreads=16)
    R.free(h)
    print(f"  {steps} RK4 steps (reference {rsteps})", flush=True)
    assert steps == rsteps and steps > 2_000_000
    np.testing.assert_array_equal(out, ref)


def test_uniform128_cic_full_fp64_tier_vs_oracle():
    print("[case] uniform128_cic: all 4096 rays, FP64 tier vs the oracle's FP64 mode", flush=True)
    w, rec, f64, length, table, in9 = workload("uniform128_cic")
    c = context("uniform128_cic", capi.PREC_FP64)
    out, steps = c.integrate(in9, "redshift", w["z_stop"])
    c.close()
    O = orc.Oracle()
    h = O.tree(rec, f64)
    ref, rsteps = O.integrate(h, w["order"], table, length, 10, 2, w["z_stop"], in9, precision=1, nthreads=16)
    O.free(h)
    assert steps == rsteps
    np.testing.assert_array_equal(out, ref)


@pytest.mark.skipif(not orc.ref_available(), reason="oracle/_ref not built")
def test_amr256_tsc_stride_ref_tier_vs_reference():
    print("[case] amr256_tsc: every 64th ray (16384), REF tier vs the reference", flush=True)
    w, rec, f64, length, table, in9 = workload("amr256_tsc")
    sub = np.ascontiguousarray(in9[::64])
    assert sub.shape[0] == 16384
    c = context("amr256_tsc", capi.PREC_REF)
    out, steps = c.integrate(sub, "redshift", w["z_stop"])
    c.close()
    R = orc.RefHarness()
    t0 = time.time()
    h = R.tree(rec)
    ref, rsteps = R.integrate(h, w["order"], table, length, 10, 2, w["z_stop"], sub, nthreads=16)
    R.free(h)
    print(f"  {steps} RK4 steps (reference {rsteps}, {time.time() - t0:.1f}s on the CPU)", flush=True)
    assert steps == rsteps
    np.testing.assert_array_equal(out, ref)


def test_amr256_tsc_full_split_vs_fp64():
    print("[case] amr256_tsc: all 1048576 rays, split tier vs FP64 tier", flush=True)
    w, rec, f64, length, table, in9 = workload("amr256_tsc")
    c = context("amr256_tsc", capi.PREC_SPLIT)
    out_s, steps_s = c.integrate(in9, "redshift", w["z_stop"])
    c.set_precision(capi.PREC_FP64)
    out_d, steps_d = c.integrate(in9, "redshift", w["z_stop"])
    c.close()
    ns = capi.OUT["nsteps"]
    ok = (out_d[:, ns] >= 0) & (out_s[:, ns] >= 0)
    assert ok.sum() == in9.shape[0], "every ray of the field starts inside the octree"
    z_d, z_s = out_d[:, capi.OUT["redshift"]], out_s[:, capi.OUT["redshift"]]
    rel_z = np.abs(z_s - z_d) / np.abs(z_d)
    rel_x = np.linalg.norm(out_s[:, 2:5] - out_d[:, 2:5], axis=1) / np.linalg.norm(out_d[:, 2:5], axis=1)
    mism = int(np.sum(out_s[:, ns] != out_d[:, ns]))
    print(f"  steps split {steps_s} fp64 {steps_d}; rel end-redshift error max {rel_z.max():.3e} median "
          f"{np.median(rel_z):.3e}; rel end-position error max {rel_x.max():.3e}; {mism} rays differ in step count",
          flush=True)
    assert rel_z.max() <= 1e-7 and rel_x.max() <= 1e-7
    assert mism <= in9.shape[0] // 10000


def test_jacobian_batch_split_vs_fp64_and_repeatable():
    n = 4194304
    print(f"[case] Jacobian batch: {n} rays, per-ray z_s, split vs FP64 kappa/gamma and split repeatability",
          flush=True)
    w, rec, f64, length, table, _ = workload("amr256_tsc")
    ph, th = synth.narrow_field_uniform(w["phi0"], w["theta0"], w["aperture"], n, 7)
    in9 = synth.launch_photons(ph, th)
    zs = np.random.default_rng(8).uniform(0.1, 2.0, n)
    c = context("amr256_tsc", capi.PREC_SPLIT)
    out1, surf1 = c.integrate_surfaces(in9, "redshift", [0.0], jacobian=True, ray_value=zs)
    out2, surf2 = c.integrate_surfaces(in9, "redshift", [0.0], jacobian=True, ray_value=zs)
    print(f"  tidal blocks: {c.split_tidal_blocks()}", flush=True)
    # bit patterns (a NaN entry would compare unequal as a float)
    dout = np.flatnonzero(np.any(out1.view(np.uint64) != out2.view(np.uint64), axis=0))
    dsurf = np.flatnonzero(np.any(surf1[:, 0, :].view(np.uint64) != surf2[:, 0, :].view(np.uint64), axis=0))
    same = dout.size == 0 and dsurf.size == 0
    if not same:
        bad = np.any(surf1[:, 0, :].view(np.uint64) != surf2[:, 0, :].view(np.uint64), axis=1)
        print(f"  repeat differs: out columns {dout.tolist()}, surface columns {dsurf.tolist()}, "
              f"{int(bad.sum())} rays", flush=True)
    c.set_precision(capi.PREC_FP64)
    out_d, surf_d = c.integrate_surfaces(in9, "redshift", [0.0], jacobian=True, ray_value=zs)
    c.close()
    s1, sd = surf1[:, 0, :], surf_d[:, 0, :]
    valid = (s1[:, capi.SURF["valid"]] != 0) & (sd[:, capi.SURF["valid"]] != 0)
    nv = int(valid.sum())
    assert nv >= n - n // 10000
    kg_s, kg_d = kappa_gamma(s1[valid]), kappa_gamma(sd[valid])
    err = np.abs(kg_s - kg_d).max(0)
    rms = np.sqrt(np.mean(kg_d ** 2, 0))
    rel_z = np.abs(s1[valid, capi.SURF["redshift"]] - sd[valid, capi.SURF["redshift"]]) / sd[valid, capi.SURF["redshift"]]
    rr = max(float(np.max(np.abs(s1[valid, capi.SURF[f]] - sd[valid, capi.SURF[f]]) / np.abs(sd[valid, capi.SURF[f]])))
             for f in ("rho", "rhomax"))
    print(f"  rho/rhomax max relative error {rr:.3e}", flush=True)
    assert rr <= 1e-5
    print(f"  {nv} valid rays; max |error| kappa {err[0]:.3e} gamma1 {err[1]:.3e} gamma2 {err[2]:.3e} "
          f"(rms kappa {rms[0]:.3e}, gamma1 {rms[1]:.3e}, gamma2 {rms[2]:.3e}); redshift rel {rel_z.max():.3e}; "
          f"repeat bit-identical {same}", flush=True)
    assert same, "two split runs of the same batch differ"
    assert err.max() <= 1e-7
    assert rel_z.max() <= 1e-7


def test_catalogue_1m_sources():
    n = 1048576
    acc = 1e-8
    print(f"[case] catalogue: {n} sources, plain vs Born-seeded Newton to {acc} rad (split tier)", flush=True)
    w, rec, f64, length, table, _ = workload("amr256_tsc")
    box = w["box_mpch"]
    dmin = synth.comoving_distance_of_z(0.1) * 0.7 / synth.MPC_SI / box
    dmax = synth.comoving_distance_of_z(1.5) * 0.7 / synth.MPC_SI / box
    src = proposal.field_sources(n, w["phi0"], w["theta0"], w["aperture"], dmin, dmax, 11)
    c = context("amr256_tsc", capi.PREC_SPLIT)
    plain, rays_p = c.catalogue_solve(src, accuracy=acc)
    init = c.catalogue_seed(src, use_network=False)
    seeded, rays_s = c.catalogue_solve(src, init2=init, accuracy=acc, jacobian_all_rounds=True)
    c.close()
    st, it = capi.CAT["status"], capi.CAT["iter"]
    ok_p, ok_s = plain[:, st] == capi.CAT_OK, seeded[:, st] == capi.CAT_OK
    out_p = int(np.sum(plain[:, st] == capi.CAT_OUTSIDE))
    print(f"  plain: {ok_p.sum()} converged, {int(np.sum(plain[:, st] == capi.CAT_REJECTED))} rejected, {out_p} "
          f"outside, mean iterations {plain[ok_p, it].mean():.3f}, {rays_p} rays; seeded: {ok_s.sum()} converged, "
          f"mean iterations {seeded[ok_s, it].mean():.3f}, {rays_s} rays", flush=True)
    assert out_p == 0, "sources inside the field start inside the octree"
    assert ok_p.sum() >= n - n // 1000 and ok_s.sum() >= n - n // 1000
    both = ok_p & ok_s
    d = np.abs(plain[both][:, [capi.CAT["phi_obs"], capi.CAT["theta_obs"]]] -
               seeded[both][:, [capi.CAT["phi_obs"], capi.CAT["theta_obs"]]]).max()
    print(f"  max |plain - seeded| observed angle {d:.3e} rad", flush=True)
    assert d <= 20 * acc
    assert seeded[ok_s, it].mean() < plain[ok_p, it].mean()


def test_catalogue_subset_ref_tier_vs_oracle():
    print("[case] catalogue: 256 sources of the field, REF tier vs the oracle", flush=True)
    w, rec, f64, length, table, _ = workload("amr256_tsc")
    box = w["box_mpch"]
    dmin = synth.comoving_distance_of_z(0.1) * 0.7 / synth.MPC_SI / box
    dmax = synth.comoving_distance_of_z(1.5) * 0.7 / synth.MPC_SI / box
    src = proposal.field_sources(256, w["phi0"], w["theta0"], w["aperture"], dmin, dmax, 12)
    c = context("amr256_tsc", capi.PREC_REF)
    dev, _ = c.catalogue_solve(src, accuracy=1e-8)
    c.close()
    O = orc.Oracle()
    h = O.tree(rec, f64)
    ref, _ = O.catalogue(h, w["order"], table, length, 10, src, accuracy=1e-8, precision=0, nthreads=16)
    O.free(h)
    C = capi.CAT
    np.testing.assert_array_equal(dev[:, C["status"]], ref[:, C["status"]])
    np.testing.assert_array_equal(dev[:, C["iter"]], ref[:, C["iter"]])
    ok = dev[:, C["status"]] != capi.CAT_OUTSIDE
    for col in ("phi_obs", "theta_obs"):
        np.testing.assert_allclose(dev[ok, C[col]], ref[ok, C[col]], rtol=0, atol=1e-12, err_msg=col)
    print(f"  {int(np.sum(dev[:, C['status']] == capi.CAT_OK))} converged, iterations "
          f"{dev[:, C['iter']].mean():.2f} on average", flush=True)
