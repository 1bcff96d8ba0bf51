"""Benchmark: ray-RK4-steps/s of the photon-propagation path on one MI355X.

Workload "amr256_tsc": a level-8 base grid (256^3) refined
to level 11 wherever the lognormal overdensity exceeds 8, FP64 fields; 1024 x
1024 rays on a regular grid over a 10 x 10 degree narrow field (the
reference's narrow-cone shape, 99.9 deg^2), TSC interpolation, integrated to
redshift 2.  One bench step = one pass of every ray to its stop (a launch of
the split-precision integrator over the whole batch), inputs resident in HBM.
The plain FP64 tier runs the same rays for the accuracy check (relative
end-redshift error <= 1e-7) and its own throughput; the reference's CPU path
(or the oracle when the reference is absent) is timed on a bounded subset of
the rays.  The roofline is the VALU issue rate (the kernels are VALU-bound;
their VALU instruction counts and HBM traffic come from the rocprofv3 counter
summaries under profiles/).  Further records on the same octree: the uneven
Jacobian/maps batch, the catalogue root-finder and, on its own 128^3 grid,
the plain-method workload in full with the reference's CPU path on every ray.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload amr256_tsc]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "magrathea-pathfinder_amd"))

HBM_PEAK = 8.0e12       # MI355X HBM3E, bytes/s (AMD public spec)
# VALU issue peak: 256 CUs x 2.4 GHz x one wave64 VALU instruction per CU per
# cycle (4 SIMDs, each retiring a full-rate wave64 FP32 or FP64 instruction
# every 4 cycles on 16 lanes).  x 64 lanes x 2 flop (FMA) = 78.6 TFLOP/s, the
# published FP64 vector peak.  Packed FP32 is not generated (build flags).
VALU_PEAK = 256 * 2.4e9
WORKLOAD_TEXT = {
    "amr256_tsc": "amr256_tsc: level-8 base (256^3, 8000 Mpc/h box) refined to level 11 where the lognormal "
                  "overdensity > 8 (37.5 M cells), 1024x1024 rays on a regular grid over a 10x10 deg narrow field "
                  "(99.9 deg^2) around (20, 20) deg, TSC, to z=2",
    "uniform128_cic": "uniform128_cic: uniform 128^3 grid (3000 Mpc/h box), 64x64 rays on a regular grid over a "
                      "5x5 deg narrow field (25.0 deg^2) around (20, 20) deg, CIC, to z=0.5",
    "cone600m_tsc": "cone600m_tsc: level-9 base (512^3, 10000 Mpc/h box) refined to level 13 where the lognormal "
                    "(sigma 2) overdensity > 8, 4096x4096 rays on a regular grid over a 20x20 deg narrow field "
                    "(399 deg^2) around (20, 20) deg, TSC, to z=3 with the infinitesimal Jacobian, nside-4096 maps",
}
B_STEP_SPLIT_TSC = 540  # 27 cells x 5 fields x 4 bytes per RK4 step (equivalent gather, not a roofline)
BF16_PEAK = 2.5e15      # MI355X dense BF16 matrix-core peak, flop/s (AMD public spec)
MLP_FLOP_PER_SAMPLE = 2 * (10 * 128 + 128 * 128 + 128 * 2)   # 35,840: the 10-128-128-2 network's forward
# threads of the CPU baseline: one GPU's share of the host's cores (the
# reference's Utility::parallelize, utility.h:55, starts hardware_concurrency()
# threads; here the host is shared by the node's GPUs)
CPU_SHARE = int(os.environ.get("MPF_CPU_THREADS", "16"))
PROFILES = {"split": "integrate_split_tsc_amr256_1m_rays.txt", "fp64": "integrate_exact_fp64_tsc_amr256_1m_rays.txt",
            "ref": "integrate_exact_ref_tsc_amr256_1m_rays.txt",
            "jacobian": "integrate_split_jacobian_amr256_4m_rays.txt",
            "uniform128_cic": "integrate_exact_ref_cic_uniform128_4k_rays.txt"}


def profiled(name, key):
    """A 'key: value' figure of a committed rocprofv3 summary under
    profiles/ (counter passes of the same kernel and workload), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return float(line.split(":")[1].split()[0])
    except (OSError, ValueError):
        pass
    return None


def valu_roofline(profile, steps, kms, gather_bytes=None):
    """VALU roofline of a ray kernel: VALU wave-instructions per ray-step from
    the SQ_INSTS_VALU pass of the committed profile (same kernel, same
    workload) x this launch's ray-steps / its HIP-event kernel time, against
    VALU_PEAK; HBM traffic (FETCH_SIZE/WRITE_SIZE passes) beside it."""
    insts = profiled(profile, "valu_insts_per_launch")
    psteps = profiled(profile, "ray_steps_per_launch")
    traffic = profiled(profile, "traffic_bytes_per_launch")
    busy = profiled(profile, "valu_busy")
    per_step = insts / psteps if insts and psteps else None
    achieved = per_step * steps / (kms / 1e3) if per_step else None
    r = {"bound": "valu", "achieved": achieved / 1e9 if achieved else None, "peak": VALU_PEAK / 1e9,
         "unit": "G VALU wave-instr/s", "frac": achieved / VALU_PEAK if achieved else None,
         "traffic": traffic, "kernel_ms": kms, "valu_insts_per_ray_step": per_step,
         "valu_busy_profiled": busy,
         "hbm_GB_per_s": traffic / (kms / 1e3) / 1e9 if traffic else None,
         "hbm_frac": traffic / (kms / 1e3) / HBM_PEAK if traffic else None,
         "profile": "profiles/" + profile,
         "note": "achieved = SQ_INSTS_VALU per ray-step (profile) x ray-steps of this launch / its HIP-event time; "
                 "peak = 256 CU x 2.4 GHz x 1 wave64 VALU instr per CU-cycle (x 64 lanes x 2 = the 78.6 TFLOP/s "
                 "FP64 vector peak); traffic = FETCH_SIZE x 2 + WRITE_SIZE x 1 (copy-kernel calibration), HBM bytes "
                 "per launch"}
    if gather_bytes:
        r["equivalent_gather_GB_per_s"] = steps * gather_bytes / (kms / 1e3) / 1e9
        r["equivalent_gather_note"] = (f"{gather_bytes} B per step: a stencil gather the kernel does not make "
                                       "(it evaluates precomputed blocks held in L1/L2); not a roofline")
    return r


def host_cpu():
    """Host CPU description: model, logical CPUs, CPUs this process may use."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def autotune_summary(trials, seconds):
    """Compact record of the autotuner's trials: (refill, sort_bucket,
    blocks_per_cu, waves_per_simd, kernel ms) per run."""
    return {"seconds": seconds, "trials": [[int(r[0]), int(r[1]), int(r[2]), int(r[3]), round(float(r[4]), 2)]
                                           for r in trials]}


FIXED_LAUNCH = None   # launch configuration given on the command line (no autotuning)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_workload(name):
    from mpf import synth
    return synth.build_workload(name, log=log)


def cpu_baseline(rec, table, length, in9, w, nthreads, target_s=15.0):
    """Reference CPU path on a bounded subset of the rays (the reference's own
    headers in oracle/_ref when present, else the oracle restatement)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    kind = "reference" if orc.ref_available() else "port"
    lib = orc.RefHarness() if kind == "reference" else orc.Oracle()
    h = lib.tree(rec)
    order = w["order"]
    # calibrate on a small sample, then run a sample sized to ~target_s
    stride = max(1, in9.shape[0] // 256)
    sample = in9[::stride][:256]
    t0 = time.time()
    kw = dict(nthreads=nthreads)
    _, steps = lib.integrate(h, order, table, length, 10, 2, w["z_stop"], sample, **kw)
    dt = time.time() - t0
    rate = steps / max(dt, 1e-9)
    per_ray = steps / sample.shape[0]
    nray = int(min(in9.shape[0], max(256, target_s * rate / max(per_ray, 1.0))))
    stride = max(1, in9.shape[0] // nray)
    sample = in9[::stride][:nray]
    t0 = time.time()
    _, steps = lib.integrate(h, order, table, length, 10, 2, w["z_stop"], sample, **kw)
    dt = time.time() - t0
    lib.free(h)
    return dict({"value": steps / dt, "unit": "ray-RK4-steps/s", "cores": nthreads, "kind": kind,
                 "sample": f"{sample.shape[0]} of the {in9.shape[0]} rays (every {stride}th), {steps} RK4 steps, "
                           f"{dt:.1f}s, {nthreads} threads"}, **host_cpu())


def uniform_cic(steps, warmup, nthreads, cpu=True):
    """The plain-method workload in full (uniform128_cic: 128^3 uniform grid,
    4096 rays, CIC, to z = 0.5) in the REF tier (bit-identical to the
    reference) and the FP64 tier: kernel throughput over `steps` passes, and
    the reference's CPU path (oracle/_ref, or the oracle) on every ray, whose
    output the REF tier must equal bit for bit."""
    from mpf import capi
    w, rec, f64, length, table, in9 = build_workload("uniform128_cic")
    ctx = capi.Context(length, order=w["order"], precision=capi.PREC_REF, nsteps=10, device=0)
    ctx.load_records(rec)
    ctx.load_fields64(f64)
    ctx.set_cosmology(table)
    n = in9.shape[0]
    d_in, d_out = ctx.dev_alloc(in9.nbytes), ctx.dev_alloc(n * capi.NOUT * 8)
    ctx.upload(d_in, in9)
    stop = ("redshift", w["z_stop"])
    rec_out = {"workload": WORKLOAD_TEXT["uniform128_cic"], "rays": n}
    outs = {}
    for name, prec in (("ref", capi.PREC_REF), ("fp64", capi.PREC_FP64)):
        ctx.set_precision(prec)
        for _ in range(warmup):
            ctx.integrate_dev(d_in, n, d_out, *stop, sync=True)
        kms = []
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.integrate_dev(d_in, n, d_out, *stop, sync=False)
            st, _ = ctx.last_run()
            kms.append(ctx.last_kernel_ms())
        ctx.synchronize()
        wall = time.perf_counter() - t0
        outs[name] = ctx.download(d_out, (n, capi.NOUT))
        k = float(np.mean(kms))
        rec_out[name] = {"value": st * steps / wall, "unit": "ray-RK4-steps/s", "ms_per_step": wall / steps * 1e3,
                         "kernel_ms": k, "kernel_steps_per_s": st / (k / 1e3), "rk4_steps_per_pass": int(st)}
        if name == "ref":
            rec_out[name]["roofline"] = valu_roofline(PROFILES["uniform128_cic"], st, k)
    ctx.dev_free(d_in)
    ctx.dev_free(d_out)
    ctx.close()
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        kind = "reference" if orc.ref_available() else "port"
        lib = orc.RefHarness() if kind == "reference" else orc.Oracle()
        h = lib.tree(rec)
        t0 = time.time()
        ref, rsteps = lib.integrate(h, w["order"], table, length, 10, 2, w["z_stop"], in9, nthreads=nthreads)
        dt = time.time() - t0
        lib.free(h)
        rec_out["cpu_baseline"] = dict({"value": rsteps / dt, "unit": "ray-RK4-steps/s", "cores": nthreads,
                                        "kind": kind, "sample": f"all {n} rays, {rsteps} RK4 steps, {dt:.2f}s, "
                                                                f"{nthreads} threads"}, **host_cpu())
        rec_out["ref_tier_bit_identical_to_cpu"] = bool(np.array_equal(outs["ref"], ref)) if kind == "reference" \
            else None
        rec_out["steps_match_cpu"] = bool(rsteps == rec_out["ref"]["rk4_steps_per_pass"])
    z, zr = outs["fp64"][:, capi.OUT["redshift"]], outs["ref"][:, capi.OUT["redshift"]]
    rec_out["fp64_vs_ref_rel_end_redshift_max"] = float(np.max(np.abs(z - zr) / np.abs(zr)))
    return rec_out


def production_cone(nside_map=4096, ncat=4194304, acc=1e-8, seed=17):
    """The single-cone production workload (cone600m_tsc): a level-9 base
    (512^3) refined to level 13, about 600 M cells with FP64 fields, built by
    torch on the GPU and handed to libmpf; 4096 x 4096 rays over the 20 x 20
    deg narrow field to z = 3 with the infinitesimal Jacobian (split tier),
    kappa/gamma1/gamma2 binned into nside-4096 maps; ncat catalogue sources in
    the field, Newton to acc from the network's first guess (trained on 4 other
    cones of the same octree), beside the Born estimate alone.  Device memory
    is reported at each stage (hipMemGetInfo through torch)."""
    import subprocess
    import threading

    from mpf import capi, proposal, synth
    rec_out = {"workload": WORKLOAD_TEXT["cone600m_tsc"]}
    mem = {}
    peak = [0]

    def guard(limit=220e9):
        # the host memory of this process (the tree is built on the GPU, but
        # its records, fields and the library's re-layout pass through here)
        while True:
            try:
                rss = next(int(l.split()[1]) * 1024 for l in open("/proc/self/status") if l.startswith("VmRSS"))
            except (OSError, StopIteration):
                return
            peak[0] = max(peak[0], rss)
            if rss > limit:
                print(f"[bench] host memory {rss / 1e9:.0f} GB above {limit / 1e9:.0f} GB: stopping", flush=True)
                os._exit(3)
            time.sleep(1.0)
    threading.Thread(target=guard, daemon=True).start()

    def used(tag):
        free, total = capi.device_memory()
        mem[tag] = round((total - free) / 1e9, 2)
        log(f"device memory in use after {tag}: {mem[tag]} GB of {total / 1e9:.1f}")

    def tree_from_child(name):
        # torch's HIP runtime and libmpf's cannot both drive the device in one
        # process: a child builds the tree with torch on the GPU and streams
        # the records and FP64 fields back through a pipe
        p = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--emit-tree", name],
                             stdout=subprocess.PIPE, stderr=None)
        head = json.loads(p.stdout.readline())
        rec = np.empty(head["cells"], dtype=synth.formats.RECORD_DTYPE)
        f64 = np.empty((head["cells"], 7))
        for arr in (rec, f64):
            view = memoryview(arr.reshape(-1).view(np.uint8))
            got = 0
            while got < view.nbytes:
                k = p.stdout.readinto(view[got:])
                if not k:
                    raise RuntimeError(f"tree generator ended early (exit status {p.wait()})")
                got += k
        if p.wait() != 0:
            raise RuntimeError(f"tree generator failed ({p.returncode})")
        w = dict(synth.WORKLOADS[name])
        table = synth.flat_lcdm_table(h=0.7, omega_m=0.3, a_min=0.05, n=4096)
        ph, th = synth.narrow_field_grid(w["phi0"], w["theta0"], w["aperture"], w["nside"])
        return w, rec, f64, head["length"], table, synth.launch_photons(ph, th)

    t0 = time.time()
    w, rec, f64, length, table, in9 = tree_from_child("cone600m_tsc")
    rec_out["generate_s"] = time.time() - t0
    used("tree generation (in a child process)")
    ctx = capi.Context(length, order=2, precision=capi.PREC_SPLIT, nsteps=10, device=0)
    t0 = time.time()
    ctx.load_records(rec)
    ctx.load_fields64(f64)
    ctx.set_cosmology(table)
    rec_out["load_s"] = time.time() - t0
    ncells, lmin, lmax = ctx.octree_info()
    rec_out.update(cells=int(ncells), levels=[int(lmin), int(lmax)])
    del rec, f64
    used("octree load")
    n = in9.shape[0]
    npix = 12 * nside_map * nside_map
    d_in, d_out = ctx.dev_alloc(in9.nbytes), ctx.dev_alloc(n * capi.NOUT * 8)
    d_surf, d_maps = ctx.dev_alloc(n * capi.NSURF * 8), ctx.dev_alloc(4 * npix * 8)
    ctx.upload(d_in, in9)
    used("ray and map buffers")
    ctx.set_launch_config(surfaces=True, refill=64, sort_bucket=65536, blocks_per_cu=8, waves_per_simd=2)
    runs = []
    for rep in range(2):   # the first pass also builds the stencil and rho blocks
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.integrate_surfaces_dev(d_in, n, d_out, d_surf, "redshift", [w["z_stop"]], jacobian=True, sync=True)
        t1 = time.perf_counter()
        steps, ms = ctx.last_run()
        runs.append((t1 - t0, steps, ctx.last_kernel_ms()))
        if rep == 0:
            used("first pass (stencil blocks built)")
    wall, steps, kms = runs[-1]
    t0 = time.perf_counter()
    ctx.healpix_bin_dev(nside_map, d_in, d_surf, n, 1, 0, d_maps)
    ctx.synchronize()
    t_bin = time.perf_counter() - t0
    hits = ctx.download(d_maps + 3 * npix * 8, (npix,))
    surf = ctx.download(d_surf, (n, capi.NSURF))
    valid = surf[:, capi.SURF["valid"]] != 0
    for p in (d_in, d_out, d_surf, d_maps):
        ctx.dev_free(p)
    rec_out["rays"] = {"rays": int(n), "rk4_steps": int(steps), "first_pass_s": runs[0][0], "wall_s": wall,
                       "kernel_ms": kms, "value": steps / wall, "kernel_steps_per_s": steps / (kms / 1e3),
                       "unit": "ray-RK4-steps/s", "rays_valid": int(valid.sum()), "bin_ms": t_bin * 1e3,
                       "pixels_hit": int(np.count_nonzero(hits)), "tidal_blocks": ctx.split_tidal_blocks(),
                       "kappa_rms": float(np.sqrt(np.mean((1 - (surf[valid, capi.SURF["a11"]] +
                                                                 surf[valid, capi.SURF["a22"]]) / 2) ** 2)))}
    del surf, hits
    # catalogue: sources in the field, Born-seeded Newton
    box = w["box_mpch"]
    dmin = synth.comoving_distance_of_z(0.1) * 0.7 / synth.MPC_SI / box
    dmax = synth.comoving_distance_of_z(w["z_stop"]) * 0.7 / synth.MPC_SI / box
    src = proposal.field_sources(ncat, w["phi0"], w["theta0"], w["aperture"], dmin, dmax, seed)
    # the sources are 4x sparser in angle than the ray grid: larger buckets keep
    # a wavefront's rays close in direction (seeded solve: 1.63 s with buckets
    # of 65,536, 1.25 s with 262,144, 1.22 s with 1,048,576, 1.36 s unbucketed)
    ctx.set_launch_config(surfaces=True, refill=64, sort_bucket=1048576, blocks_per_cu=8, waves_per_simd=2)
    # the first guess's network, trained on 4 other cones of this octree
    fit, first_guess = train_first_guess(ctx, w, dmin, dmax, acc, 1048576, 262144, 12, seed)
    d_src, d_init, d_cat = ctx.dev_alloc(src.nbytes), ctx.dev_alloc(ncat * 16), ctx.dev_alloc(ncat * capi.NCAT * 8)
    ctx.upload(d_src, src)
    cruns = {}
    for name in ("born", "seeded"):
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.catalogue_seed_dev(d_src, ncat, d_init, use_network=fit["use_network"] and name == "seeded")
        crays = ctx.catalogue_solve_dev(d_src, d_init, ncat, d_cat, accuracy=acc, jacobian_all_rounds=2)
        ctx.synchronize()
        cwall = time.perf_counter() - t0
        cat = ctx.download(d_cat, (ncat, capi.NCAT))
        ok = cat[:, capi.CAT["status"]] == capi.CAT_OK
        cruns[name] = {"converged": int(ok.sum()), "wall_s": cwall, "sources_per_s": float(ok.sum() / cwall),
                       "rays": int(crays), "mean_iterations": float(cat[ok, capi.CAT["iter"]].mean())}
    for p in (d_src, d_init, d_cat):
        ctx.dev_free(p)
    rec_out["catalogue"] = dict(cruns["seeded"], sources=int(ncat), born=cruns["born"], first_guess=first_guess,
                                note="seeded: the Born estimate plus the network's correction (when it beat Born on "
                                     "the held-out cone); born: the Born estimate alone")
    used("catalogue")
    rec_out["device_memory_GB"] = mem
    rec_out["host_peak_rss_GB"] = round(peak[0] / 1e9, 1)
    ctx.close()
    return rec_out


def jacobian_maps(ctx, w, length, table, n=4194304, nside=2048, zmin=0.1, zmax=2.0, seed=7):
    """Uneven batch with the lensing Jacobian and HEALPix maps: n rays uniform
    over the cone, each stopped at its own source redshift z_s ~ U[zmin, zmax]
    with the infinitesimal Jacobian there, then kappa, gamma1, gamma2 binned
    into RING maps of the given nside (split tier)."""
    from mpf import capi, synth
    ph, th = synth.narrow_field_uniform(w["phi0"], w["theta0"], w["aperture"], n, seed)
    in9 = synth.launch_photons(ph, th)
    zs = np.random.default_rng(seed + 1).uniform(zmin, zmax, n)
    npix = 12 * nside * nside
    d_in, d_rv = ctx.dev_alloc(in9.nbytes), ctx.dev_alloc(zs.nbytes)
    d_out = ctx.dev_alloc(n * capi.NOUT * 8)
    d_surf = ctx.dev_alloc(n * capi.NSURF * 8)
    d_maps = ctx.dev_alloc(4 * npix * 8)
    ctx.upload(d_in, in9)
    ctx.upload(d_rv, zs)
    res = []
    # launch configuration tuned on this batch (ray order by source redshift
    # and direction, refill granularity, grid, occupancy)
    tuned, trials, t_tune = ctx.launch_config(surfaces=True), np.zeros((0, 6)), 0.0
    if FIXED_LAUNCH is None:
        t0 = time.perf_counter()
        tuned, trials = ctx.autotune_dev(d_in, n, d_out, surf_ptr=d_surf, kind="redshift", values=[0.0],
                                         ray_value_ptr=d_rv)
        t_tune = time.perf_counter() - t0
        log(f"jacobian/maps autotune ({t_tune:.1f}s): {tuned}; trials ms {np.round(trials[:, 4], 1).tolist()}")
    for rep in range(2):   # the second pass is reported (the first also allocates the map accumulator)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.integrate_surfaces_dev(d_in, n, d_out, d_surf, "redshift", [0.0], jacobian=True, sync=True,
                                   ray_value_ptr=d_rv)
        steps, ms = ctx.last_run()
        kms = ctx.last_kernel_ms()
        t1 = time.perf_counter()
        ctx.healpix_bin_dev(nside, d_in, d_surf, n, 1, 0, d_maps)
        ctx.synchronize()
        t2 = time.perf_counter()
        res.append((steps, ms, kms, (t1 - t0) * 1e3, (t2 - t1) * 1e3))
    steps, ms, kms, wall_int, wall_bin = res[-1]
    surf = ctx.download(d_surf, (n, capi.NSURF))
    maps = ctx.download(d_maps, (4, npix))
    for p in (d_in, d_rv, d_out, d_surf, d_maps):
        ctx.dev_free(p)
    valid = surf[:, capi.SURF["valid"]] != 0
    return {"workload": f"{n} rays uniform over the {w['aperture']}x{w['aperture']} deg narrow field "
                        f"({synth.narrow_field_area_deg2(w['aperture']):.1f} deg^2), per-ray z_s ~ U[{zmin}, {zmax}], "
                        f"infinitesimal Jacobian at z_s, kappa/gamma1/gamma2 binned into nside {nside} RING maps",
            "value": steps / (wall_int / 1e3 + wall_bin / 1e3), "unit": "ray-RK4-steps/s (integration + maps)",
            "kernel_steps_per_s": steps / (kms / 1e3), "rk4_steps": int(steps), "integrate_ms": ms,
            "kernel_ms": kms,
            "integrate_wall_ms": wall_int, "bin_wall_ms": wall_bin, "rays_valid": int(valid.sum()),
            "pixels_hit": int(np.count_nonzero(maps[3])), "kappa_rms": float(np.sqrt(np.mean(
                maps[0][maps[3] > 0] ** 2))),
            "launch_config": tuned, "autotune": autotune_summary(trials, t_tune),
            "tidal_blocks": ctx.split_tidal_blocks(),
            "roofline": valu_roofline(PROFILES["jacobian"], steps, kms)}


def velocity_field(ctx, rec, n=1 << 24, seed=13):
    """Velocity field of the velocity-field build: n particles inside random
    cells of the octree (denser where it refines) deposited with CIC and with
    TSC at every stored level, then normalised and filled from the parents.
    Particles resident in HBM; the deposit is the timed part."""
    from mpf import synth
    pos, vel = synth.particles_in_cells(rec, n, seed)
    ncells, lmin, lmax = ctx.octree_info()
    nlev = lmax - lmin + 1
    d_pos, d_vel = ctx.dev_alloc(pos.nbytes), ctx.dev_alloc(vel.nbytes)
    ctx.upload(d_pos, pos)
    ctx.upload(d_vel, vel)
    out = {"workload": f"{n} particles in random cells of the octree ({ncells} cells, {nlev} levels), "
                       f"deposit at every level, normalisation and parent fill"}
    for scheme, cells in (("cic", 8), ("tsc", 27)):
        res = []
        for rep in range(2):   # the second pass is reported (the first also allocates)
            ctx.synchronize()
            ctx.velocity_begin()
            t0 = time.perf_counter()
            ctx.velocity_deposit_dev(d_pos, d_vel, n, scheme)
            t1 = time.perf_counter()
            ctx.velocity_finish()
            t2 = time.perf_counter()
            res.append(((t1 - t0) * 1e3, (t2 - t1) * 1e3))
        dep_ms, fin_ms = res[-1]
        b_pl = 24 + cells * 4 * 8     # particle read + 4 FP64 atomic adds per cell
        out[scheme] = {"value": n / (dep_ms / 1e3), "unit": "particles/s (all levels)",
                       "particle_levels_per_s": n * nlev / (dep_ms / 1e3), "deposit_ms": dep_ms,
                       "finish_ms": fin_ms, "bytes_per_particle_level": b_pl,
                       "achieved_GB_per_s": n * nlev * b_pl / (dep_ms / 1e3) / 1e9}
    rho, _ = ctx.velocity_get()
    out["cells_with_weight"] = int(np.count_nonzero(rho))
    ctx.velocity_clear()
    ctx.dev_free(d_pos)
    ctx.dev_free(d_vel)
    return out


TRAIN_AXES = ((80.0, 50.0), (150.0, 70.0), (230.0, 40.0), (300.0, 110.0), (120.0, 130.0))


def train_first_guess(ctx, w, dmin, dmax, acc, ntrain, nval, epochs, seed):
    """The network of the first guess trained on the root-finder's own pairs
    from 4 cones of the same octree (ntrain sources each, same field shape as
    the workload, other axes), early-stopped and gated on a 5th (nval)."""
    from mpf import proposal
    t0 = time.perf_counter()
    pairs = []
    for k, (p, t) in enumerate(TRAIN_AXES):
        s = proposal.field_sources(ntrain if k < 4 else nval, p, t, w["aperture"], dmin, dmax, seed + 100 + k)
        fe, ta, _, _ = proposal.training_pairs(ctx, s, accuracy=acc)
        pairs.append((fe, ta))
    t_pairs = time.perf_counter() - t0
    t0 = time.perf_counter()
    fit = proposal.train(ctx, np.concatenate([p[0] for p in pairs[:4]]), np.concatenate([p[1] for p in pairs[:4]]),
                         pairs[4][0], pairs[4][1], epochs=epochs, batch=1024, lr=1e-3, seed=seed, check_every=1)
    t_train = time.perf_counter() - t0
    return fit, {"training_pairs": int(sum(len(p[0]) for p in pairs[:4])), "cones": 4,
                 "epochs_kept": fit["epoch"], "held_out_rms_arcsec_born": fit["val_rms_born"],
                 "held_out_rms_arcsec_network": fit["val_rms_network"], "use_network": bool(fit["use_network"]),
                 "pairs_s": t_pairs, "train_s": t_train}


def catalogue(ctx, w, length, table, n=1048576, zmin=0.1, zmax=1.5, acc=1e-8, ntrain=1048576, nval=262144,
              epochs=12, seed=11):
    """Catalogue root-finding: n sources uniform in comoving volume in the
    narrow field of the ray workload (same axis, same area) at z in [zmin, zmax],
    Newton to acc (rad) with the infinitesimal Jacobian, split tier: plain
    (unlensed first guess) against the first guess from the Born features
    and the network's correction when it beats the Born estimate alone on a
    held-out cone.  The network is trained on the root-finder's own pairs
    from 4 other cones of the same octree (ntrain sources each), early-stopped
    and gated on a 5th (nval sources)."""
    from mpf import capi, proposal, synth
    box = w["box_mpch"]
    dmin = synth.comoving_distance_of_z(zmin) * 0.7 / synth.MPC_SI / box
    dmax = synth.comoving_distance_of_z(zmax) * 0.7 / synth.MPC_SI / box
    ap = w["aperture"]
    src = proposal.field_sources(n, w["phi0"], w["theta0"], ap, dmin, dmax, seed)
    fit, first_guess = train_first_guess(ctx, w, dmin, dmax, acc, ntrain, nval, epochs, seed)
    d_src, d_init = ctx.dev_alloc(src.nbytes), ctx.dev_alloc(n * 16)
    d_out = ctx.dev_alloc(n * capi.NCAT * 8)
    ctx.upload(d_src, src)
    st, it = capi.CAT["status"], capi.CAT["iter"]
    runs = {}
    # the Jacobian with every round's rays but the first: a seeded source
    # rarely converges at its first ray, and most converge at the second
    jac_rounds = 2
    for name in ("plain", "born", "seeded"):
        ctx.synchronize()
        t0 = time.perf_counter()
        if name == "plain":
            rays = ctx.catalogue_solve_dev(d_src, 0, n, d_out, accuracy=acc)
        else:
            ctx.catalogue_seed_dev(d_src, n, d_init, use_network=fit["use_network"] and name == "seeded")
            rays = ctx.catalogue_solve_dev(d_src, d_init, n, d_out, accuracy=acc, jacobian_all_rounds=jac_rounds)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        out = ctx.download(d_out, (n, capi.NCAT))
        ok = out[:, st] == capi.CAT_OK
        runs[name] = {"sources_per_s": float(ok.sum() / wall), "wall_s": wall, "converged": int(ok.sum()),
                      "rays": int(rays), "mean_iterations": float(out[ok, it].mean()),
                      "rejected": int(np.sum(out[:, st] == capi.CAT_REJECTED))}
        runs[name + "_angles"] = out[:, [capi.CAT["phi_obs"], capi.CAT["theta_obs"]]].copy()
    ok_both = np.all(np.isfinite(runs["plain_angles"]), 1)
    dmax_ang = float(np.max(np.abs(runs["plain_angles"][ok_both] - runs["seeded_angles"][ok_both])))
    # the network's throughput on every source (features from the split tier's stencil blocks)
    d_feat, d_pred = ctx.dev_alloc(n * capi.MLP_NIN * 4), ctx.dev_alloc(n * capi.MLP_NOUT * 4)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.catalogue_features_dev(d_src, n, d_feat)
    t_feat = time.perf_counter() - t0
    t_inf, k_inf = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        ctx.mlp_infer_dev(d_feat, n, d_pred)
        t_inf.append(time.perf_counter() - t0)
        k_inf.append(ctx.mlp_last_ms())
    flop = n * MLP_FLOP_PER_SAMPLE
    kms = min(k_inf)
    for p in (d_src, d_init, d_out, d_feat, d_pred):
        ctx.dev_free(p)
    network = {"features_s": t_feat, "inference_call_s": min(t_inf), "inference_kernel_ms": kms,
               "inference_sources_per_s": n / (kms / 1e3),
               "flop_per_sample": MLP_FLOP_PER_SAMPLE, "achieved_TFLOP_s": flop / (kms / 1e3) / 1e12,
               "peak_TFLOP_s": BF16_PEAK / 1e12, "mfma_frac": flop / (kms / 1e3) / BF16_PEAK,
               "note": "one persistent kernel (weights in LDS, normalisation fused, no allocation): its HIP-event "
                       "time; flop = the three matrix products; the 256 GELU (erf) per sample run on the VALU"}
    return {"workload": f"{n} sources uniform in comoving volume, z in [{zmin}, {zmax}], {ap}x{ap} deg narrow field "
                        f"({synth.narrow_field_area_deg2(ap):.1f} deg^2) of the ray workload, Newton to {acc} rad, "
                        f"infinitesimal Jacobian, split tier",
            "metric": "catalogue sources/s", "value": runs["seeded"]["sources_per_s"], "unit": "sources/s",
            "plain": runs["plain"], "born": runs["born"], "seeded": runs["seeded"],
            "speedup_seeded_over_plain": runs["seeded"]["sources_per_s"] / runs["plain"]["sources_per_s"],
            "iterations_saved_by_network": runs["born"]["mean_iterations"] - runs["seeded"]["mean_iterations"],
            "max_angle_difference_rad": dmax_ang,
            "first_guess": first_guess, "network": network,
            "runs_note": "plain: unlensed start; born: the Born estimate of the features alone; seeded: the Born "
                         "estimate plus the network's correction (when it beat Born on the held-out cone)"}


def emit_tree(name):
    """Child of production_cone: the named workload's tree built by torch on
    the GPU, written to stdout as one JSON line (cells, length) and the raw
    records and FP64 fields."""
    import torch
    from mpf import synth
    w = dict(synth.WORKLOADS[name])
    table = synth.flat_lcdm_table(h=0.7, omega_m=0.3, a_min=0.05, n=4096)
    import contextlib
    dev = "cuda" if torch.cuda.is_available() else "cpu"
    with contextlib.redirect_stdout(sys.stderr):   # the generator's progress lines; stdout carries the tree
        rec, f64, length = synth.lightcone_tree(w["base_input_level"], w["box_mpch"], w["sigma"], w["threshold"],
                                                w["max_input_level"], w["seed"], noise=w.get("noise", 0.0),
                                                decay=w.get("decay", 0.0), table=table, device=dev, verbose=True)
    f64 = np.ascontiguousarray(f64)
    status = {l.split(":")[0]: l.split(":")[1].strip() for l in open("/proc/self/status") if l.startswith("Vm")}
    avail = next((l.split(":")[1].strip() for l in open("/proc/meminfo") if l.startswith("MemAvailable")), "?")
    print(f"[bench] tree generator: VmRSS {status.get('VmRSS')}, VmHWM {status.get('VmHWM')}, host MemAvailable "
          f"{avail}", file=sys.stderr, flush=True)
    fd = sys.stdout.fileno()
    os.write(fd, (json.dumps({"cells": int(rec.size), "length": float(length)}) + "\n").encode())
    for arr in (rec.view(np.uint8), f64.reshape(-1).view(np.uint8)):
        mv = memoryview(arr)
        for i in range(0, mv.nbytes, 1 << 28):
            chunk = mv[i:i + (1 << 28)]
            done = 0
            while done < chunk.nbytes:
                done += os.write(fd, chunk[done:])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="amr256_tsc", choices=("amr256_tsc", "uniform128_cic"))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-tier comparison")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the Jacobian/maps and catalogue workloads on the same octree")
    ap.add_argument("--only", choices=("jacobian_maps", "catalogue", "velocity", "cone600m"), default=None,
                    help="run one of the extra workloads alone and print its record (profiling runs)")
    ap.add_argument("--launch-config", default=None, metavar="R,B,C,W",
                    help="use this launch configuration (refill, sort bucket, blocks per CU, waves per SIMD) for "
                         "every workload instead of the autotuner's (profiling runs)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch configuration (profiling runs)")
    ap.add_argument("--tier", choices=("split", "fp64", "ref"), default="split",
                    help="precision tier of the timed passes (profiling runs of the exact tiers)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="only run the plain copy kernel over MB megabytes (HBM counter calibration)")
    ap.add_argument("--emit-tree", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.emit_tree:
        emit_tree(args.emit_tree)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards by rays or "
              f"cones with no exchange until the map reduction; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    from mpf import capi
    # libmpf's HIP runtime (ROCm) is loaded before torch, which the workload
    # generator uses on the CPU only
    capi.load_library()
    if args.calibrate_copy:
        ctx = capi.Context(1.0, device=0)
        nb = args.calibrate_copy << 20
        src, dst = ctx.dev_alloc(nb), ctx.dev_alloc(nb)
        ms = [ctx.dev_copy(dst, src, nb) for _ in range(3)]
        ctx.dev_free(src)
        ctx.dev_free(dst)
        ctx.close()
        print(json.dumps({"copy_bytes": nb, "kernel_ms": ms, "GB_per_s_read_plus_write": 2 * nb / (min(ms) / 1e3) / 1e9}))
        return

    if args.only == "cone600m":
        print(json.dumps({"cone600m_tsc": production_cone()}), flush=True)
        return
    w, rec, f64, length, table, in9 = build_workload(args.workload)
    order = w["order"]
    tier = {"split": capi.PREC_SPLIT, "fp64": capi.PREC_FP64, "ref": capi.PREC_REF}[args.tier]
    if order != 2 and tier == capi.PREC_SPLIT:
        tier = capi.PREC_FP64
    ctx = capi.Context(length, order=order, precision=tier, nsteps=10, device=0)
    t0 = time.time()
    ctx.load_records(rec)
    ctx.load_fields64(f64)
    ctx.set_cosmology(table)
    ncells, lmin, lmax = ctx.octree_info()
    log(f"octree loaded in {time.time() - t0:.1f}s: {ncells} cells, levels {lmin}..{lmax}")
    fixed = None
    if args.launch_config:
        r, b, cpu, wv = (int(v) for v in args.launch_config.split(","))
        fixed = dict(refill=r, sort_bucket=b, blocks_per_cu=cpu, waves_per_simd=wv)
        ctx.set_launch_config(**fixed)
        ctx.set_launch_config(surfaces=True, **dict(fixed, waves_per_simd=min(wv, 2)))
    global FIXED_LAUNCH
    FIXED_LAUNCH = fixed
    if args.only == "velocity":
        print(json.dumps({"velocity_field": velocity_field(ctx, rec)}), flush=True)
        ctx.close()
        return
    if args.only:
        fn = jacobian_maps if args.only == "jacobian_maps" else catalogue
        print(json.dumps({args.only: fn(ctx, w, length, table)}), flush=True)
        ctx.close()
        return
    # inputs and outputs resident in HBM (libmpf device buffers)
    n = in9.shape[0]
    d_in = ctx.dev_alloc(in9.nbytes)
    ctx.upload(d_in, in9)
    d_out = ctx.dev_alloc(n * capi.NOUT * 8)
    stop = ("redshift", w["z_stop"])
    for i in range(args.warmup):
        t0 = time.time()
        ctx.integrate_dev(d_in, n, d_out, *stop, sync=True)
        steps, ms = ctx.last_run()
        log(f"warmup {i}: {steps} steps, kernel {ms:.1f} ms, wall {time.time() - t0:.2f}s")
    # launch configuration tuned on this batch (untimed, like the warmup)
    tuned, trials, t_tune = ctx.launch_config(), np.zeros((0, 6)), 0.0
    if FIXED_LAUNCH is None and not args.no_autotune:
        t0 = time.perf_counter()
        tuned, trials = ctx.autotune_dev(d_in, n, d_out, *stop)
        t_tune = time.perf_counter() - t0
        log(f"autotune ({t_tune:.1f}s): {tuned}; trials ms {np.round(trials[:, 4], 2).tolist()}")
    ctx.synchronize()
    kernel_ms, run_ms = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctx.integrate_dev(d_in, n, d_out, *stop, sync=False)
        steps, ms = ctx.last_run()   # synchronizes the context stream
        run_ms.append(ms)
        kernel_ms.append(ctx.last_kernel_ms())
    ctx.synchronize()
    wall = time.perf_counter() - t0
    total_steps = steps * args.steps
    value = total_steps / wall
    kms = float(np.mean(kernel_ms))
    kernel_rate = steps / (kms / 1e3)
    log(f"{args.tier}: {steps} steps per pass, {value / 1e9:.3f} G steps/s wall, {kernel_rate / 1e9:.3f} G steps/s kernel")
    out_split = ctx.download(d_out, (n, capi.NOUT))
    tier_text = {"split": ("split FP32-perturbation/FP64-background", "fp32-perturbation/fp64-background"),
                 "fp64": ("exact FP64 tier", "f64"), "ref": ("exact REF tier, bit-identical to the reference", "f32")}
    if tier != {"split": capi.PREC_SPLIT, "fp64": capi.PREC_FP64, "ref": capi.PREC_REF}[args.tier]:
        args.tier = "fp64"
    scheme = "TSC" if order == 2 else "CIC"
    result = {
        "metric": f"ray-RK4-steps/s ({tier_text[args.tier][0]}, {scheme})",
        "value": value, "unit": "ray-RK4-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": tier_text[args.tier][1], "data": "synthetic (seeded GRF lightcone octree)",
        "config": {"workload": WORKLOAD_TEXT[args.workload],
                   "cells": int(ncells), "levels": [int(lmin), int(lmax)], "rays": int(n),
                   "rk4_steps_per_pass": int(steps)},
        "launch_config": tuned, "autotune": autotune_summary(trials, t_tune),
    }
    prof = PROFILES[args.tier] if args.workload == "amr256_tsc" else PROFILES["uniform128_cic"]
    result["roofline"] = valu_roofline(prof, steps, kms, B_STEP_SPLIT_TSC if args.tier == "split" else None)
    result["roofline"]["integration_ms"] = float(np.mean(run_ms))
    if not args.no_fp64 and args.tier == "split":
        ctx.set_precision(capi.PREC_FP64)
        t0 = time.time()
        ctx.integrate_dev(d_in, n, d_out, *stop, sync=True)
        s64, ms64 = ctx.last_run()
        k64 = ctx.last_kernel_ms()
        out64 = ctx.download(d_out, (n, capi.NOUT))
        log(f"fp64 tier: {s64} steps, kernel {k64:.1f} ms ({s64 / (k64 / 1e3) / 1e9:.3f} G steps/s)")
        ok = (out64[:, capi.OUT["nsteps"]] >= 0) & (out_split[:, capi.OUT["nsteps"]] >= 0)
        z64 = out64[ok, capi.OUT["redshift"]]
        zs = out_split[ok, capi.OUT["redshift"]]
        rel_z = np.abs(zs - z64) / np.abs(z64)
        dpos = np.linalg.norm(out_split[ok, 2:5] - out64[ok, 2:5], axis=1) / np.linalg.norm(out64[ok, 2:5], axis=1)
        rf = valu_roofline(PROFILES["fp64"], s64, k64) if order == 2 else None
        result["fp64_tier"] = {"value": s64 / (k64 / 1e3), "unit": "ray-RK4-steps/s", "kernel_ms": k64,
                               "integration_ms": ms64,
                               "steps": int(s64), "speedup_split_over_fp64": kernel_rate / (s64 / (k64 / 1e3)),
                               "rel_end_redshift_err_max": float(rel_z.max()),
                               "rel_end_redshift_err_median": float(np.median(rel_z)),
                               "rel_end_position_err_max": float(dpos.max()),
                               "step_count_mismatch_rays": int(np.sum(out64[:, capi.OUT["nsteps"]] !=
                                                                      out_split[:, capi.OUT["nsteps"]])),
                               "roofline": rf}
        ctx.set_precision(capi.PREC_SPLIT)
    if not args.no_extra and order == 2:
        t0 = time.time()
        result["jacobian_maps"] = jacobian_maps(ctx, w, length, table)
        log(f"jacobian/maps: {result['jacobian_maps']} ({time.time() - t0:.1f}s)")
        t0 = time.time()
        result["catalogue"] = catalogue(ctx, w, length, table)
        log(f"catalogue: {result['catalogue']} ({time.time() - t0:.1f}s)")
        t0 = time.time()
        result["uniform128_cic"] = uniform_cic(max(args.steps, 5), max(args.warmup, 1),
                                               min(CPU_SHARE, host_cpu()["usable_cpus"]), cpu=not args.no_cpu)
        log(f"uniform128_cic: {result['uniform128_cic']} ({time.time() - t0:.1f}s)")
    if not args.no_cpu:
        nth = min(CPU_SHARE, host_cpu()["usable_cpus"])
        t0 = time.time()
        result["cpu_baseline"] = cpu_baseline(rec, table, length, in9, w, nth)
        log(f"cpu baseline: {result['cpu_baseline']} ({time.time() - t0:.1f}s)")
    ctx.dev_free(d_in)
    ctx.dev_free(d_out)
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
