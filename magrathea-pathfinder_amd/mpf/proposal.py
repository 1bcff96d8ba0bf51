"""Learned first guess of the catalogue root-finder (host side).

The root-finder (Catalogues::iterateNewtonMethod, catalogues.h:757-800)
starts every source from its unlensed direction (catalogues.h:832-836).
Here a small network proposes the observed direction instead; it is trained
on pairs that the root-finder itself produced on other cones of the same
octree: the features of a source (a Born estimate of its deflection from the
octree along the line of sight, and what else the network may use) and the
deflection the root-finder found, less the Born estimate.  The first guess
is the unlensed direction + the Born estimate + the network's correction,
and the Newton iterations still take every source to the requested
accuracy.  All compute goes through libmpf
(features, training, inference, seeding and the root-finder on the device).
"""
from __future__ import annotations

import math

import numpy as np

from . import capi

ARCSEC = 206264.80624709636


def cone_sources(n, axis_phi_deg, axis_theta_deg, aperture_deg, dmin, dmax, seed, vscale=1e-3):
    """n sources uniform in comoving volume inside a circular cone of
    half-angle aperture_deg around (axis_phi, axis_theta), at comoving
    distances [dmin, dmax] (box units).  Returns n x 6 (x, y, z, vx, vy, vz)."""
    rng = np.random.default_rng(seed)
    r = np.cbrt(dmin ** 3 + rng.uniform(0, 1, n) * (dmax ** 3 - dmin ** 3))
    ca = rng.uniform(math.cos(math.radians(aperture_deg)), 1.0, n)
    sa = np.sqrt(1 - ca * ca)
    az = rng.uniform(0, 2 * math.pi, n)
    local = np.stack([sa * np.cos(az), sa * np.sin(az), ca], 1)   # around +z
    ph, th = math.radians(axis_phi_deg), math.radians(axis_theta_deg)
    # rotation taking +z to the axis: R = Rz(phi) Ry(theta)
    ry = np.array([[math.cos(th), 0, math.sin(th)], [0, 1, 0], [-math.sin(th), 0, math.cos(th)]])
    rz = np.array([[math.cos(ph), -math.sin(ph), 0], [math.sin(ph), math.cos(ph), 0], [0, 0, 1]])
    d = local @ (rz @ ry).T
    src = np.zeros((n, 6))
    src[:, :3] = d * r[:, None]
    src[:, 3:] = rng.normal(scale=vscale, size=(n, 3))
    return src


def field_sources(n, axis_phi_deg, axis_theta_deg, aperture_deg, dmin, dmax, seed, vscale=1e-3):
    """n sources uniform in comoving volume inside the narrow field
    (aperture x aperture, the reference's narrow-cone shape) around
    (axis_phi, axis_theta), at comoving distances [dmin, dmax] (box units):
    the same sky area as synth.narrow_field_grid.  Returns n x 6."""
    from . import synth
    rng = np.random.default_rng(seed)
    r = np.cbrt(dmin ** 3 + rng.uniform(0, 1, n) * (dmax ** 3 - dmin ** 3))
    ph, th = synth.narrow_field_uniform(axis_phi_deg, axis_theta_deg, aperture_deg, n, seed + 1)
    src = np.zeros((n, 6))
    src[:, :3] = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1) * r[:, None]
    src[:, 3:] = rng.normal(scale=vscale, size=(n, 3))
    return src


def deflections(out):
    """observed - unlensed direction (arcsec along e_phi and e_theta) of
    root-finder records (MPF_NCAT)."""
    dphi = out[:, capi.CAT["phi_obs"]] - out[:, capi.CAT["phi"]]
    dphi = (dphi + math.pi) % (2 * math.pi) - math.pi
    dth = out[:, capi.CAT["theta_obs"]] - out[:, capi.CAT["theta"]]
    return np.stack([dphi * np.sin(out[:, capi.CAT["theta"]]), dth], 1) * ARCSEC


def targets(out, feat):
    """What the network learns: the deflection left after the Born estimate
    carried by the features (columns 1, 2), arcsec."""
    return (deflections(out) - feat[:, 1:3]).astype(np.float32)


def born_guess(src, feat):
    """First guesses from the Born estimate alone (no network)."""
    src = np.asarray(src, dtype=np.float64).reshape(-1, 6)
    feat = np.asarray(feat, dtype=np.float64)
    D = np.linalg.norm(src[:, :3], axis=1)
    theta = np.arccos(src[:, 2] / D)
    return np.stack([np.arctan2(src[:, 1], src[:, 0]) + feat[:, 1] / (ARCSEC * np.sin(theta)),
                     theta + feat[:, 2] / ARCSEC], 1)


def training_pairs(ctx, src, **newton):
    """(features, targets, records, rays) from the root-finder on src: only
    converged sources give pairs."""
    out, rays = ctx.catalogue_solve(src, **newton)
    ok = out[:, capi.CAT["status"]] == capi.CAT_OK
    feat = ctx.catalogue_features(src[ok])
    return feat, targets(out[ok], feat), out, rays


# feature pairs that are vectors in the screen basis (e_phi, e_theta)
VECTOR_PAIRS = ((1, 2), (3, 4), (5, 6), (7, 8))


def turned(feat, target, rng):
    """Every sample's screen vectors (the feature pairs and the target) turned
    by one random angle, after a reflection e_theta -> -e_theta for half of
    the samples.  The lensing field has no preferred direction on the screen,
    so the turned pairs are as likely as the originals; a network trained on
    them cannot tell where in the training cones a source lies from the
    directions of its deflection, only how the target follows the features."""
    n = len(feat)
    psi = rng.uniform(0.0, 2.0 * math.pi, n).astype(np.float32)
    c, s = np.cos(psi), np.sin(psi)
    flip = np.where(rng.uniform(size=n) < 0.5, -1.0, 1.0).astype(np.float32)
    f = np.array(feat, dtype=np.float32)
    t = np.array(target, dtype=np.float32)

    def turn(x, y):
        y = y * flip
        return c * x - s * y, s * x + c * y
    for a, b in VECTOR_PAIRS:
        f[:, a], f[:, b] = turn(f[:, a], f[:, b])
    t[:, 0], t[:, 1] = turn(t[:, 0], t[:, 1])
    return f, t


def train(ctx, feat, target, val_feat, val_target, epochs=100, batch=256, lr=1e-3, seed=0, check_every=5,
          turn=True):
    """Fresh network trained on (feat, target), early-stopped on a held-out
    set (another cone): the state with the lowest held-out residual is kept.
    With turn, every epoch sees the pairs turned and reflected at random
    (turned()).  The network is worth using only if that residual is below
    the one of the Born estimate alone (the held-out targets' rms).  Returns
    a dict: loss per epoch, held-out rms without and with the network, the
    epoch kept and use_network."""
    ctx.mlp_init(seed)
    val_target = np.asarray(val_target, dtype=np.float64)
    rms0 = float(np.sqrt(np.mean(val_target ** 2)))
    rng = np.random.default_rng(seed)
    losses = []
    done = 0
    best = (math.inf, 0, None)
    while done < epochs:
        k = min(check_every, epochs - done)
        if turn:
            for e in range(k):
                f, t = turned(feat, target, rng)
                losses += list(ctx.mlp_train(f, t, epochs=1, batch=batch, lr=lr, seed=seed + done + e))
        else:
            losses += list(ctx.mlp_train(feat, target, epochs=k, batch=batch, lr=lr, seed=seed + done))
        done += k
        r = float(np.sqrt(np.mean((ctx.mlp_infer(val_feat) - val_target) ** 2)))
        if r < best[0]:
            best = (r, done, ctx.mlp_state())
    if best[2] is not None:
        ctx.mlp_state(best[2])
    return dict(loss=np.array(losses), val_rms_born=rms0, val_rms_network=best[0], epoch=best[1],
                use_network=best[0] < rms0)


C2 = 299792458.0 ** 2


def point_features(ctx, src, length):
    """The alternative feature set (n x 10): the unlensed angles phi, theta
    (rad), the comoving distance D (box units), the transverse gradient of
    phi (along e_phi, e_theta) at chi = D/4, D/2, 3D/4 in arcsec
    (2 D L grad_perp(phi) / c^2, L the box length in metres) and phi / c^2 at
    D/2 (units of 1e-5): raw samples of the octree on the straight line of
    sight (TSC through mpf_interpolate), no Born sum.  The network then learns
    the whole deflection (targets: deflections())."""
    src = np.asarray(src, dtype=np.float64).reshape(-1, 6)
    x = src[:, :3]
    D = np.linalg.norm(x, axis=1)
    th = np.arccos(x[:, 2] / D)
    ph = np.arctan2(x[:, 1], x[:, 0])
    ep = np.stack([-np.sin(ph), np.cos(ph), np.zeros_like(ph)], 1)
    et = np.stack([np.cos(ph) * np.cos(th), np.sin(ph) * np.cos(th), -np.sin(th)], 1)
    n = x / D[:, None]
    f = np.zeros((len(src), capi.MLP_NIN), dtype=np.float32)
    f[:, 0], f[:, 1], f[:, 2] = ph, th, D
    prec = ctx.precision
    if prec == capi.PREC_SPLIT:          # interpolation is an exact-tier call
        ctx.set_precision(capi.PREC_FP64)
    try:
        g3 = [ctx.interpolate(2, n * (frac * D)[:, None]) for frac in (0.25, 0.5, 0.75)]
    finally:
        if prec != ctx.precision:
            ctx.set_precision(prec)
    for k, frac in enumerate((0.25, 0.5, 0.75)):
        g = g3[k]
        grad = g[:, 2:5]
        s = 2.0 * D * length / C2 * ARCSEC
        f[:, 3 + 2 * k] = s * np.sum(grad * ep, 1)
        f[:, 4 + 2 * k] = s * np.sum(grad * et, 1)
        if frac == 0.5:
            f[:, 9] = g[:, 1] / C2 * 1.0e5
    return f


def unlensed(src):
    src = np.asarray(src, dtype=np.float64).reshape(-1, 6)
    D = np.linalg.norm(src[:, :3], axis=1)
    return np.arctan2(src[:, 1], src[:, 0]), np.arccos(src[:, 2] / D)


def guess_from_deflection(src, defl):
    """First guess = unlensed direction + a deflection (arcsec along e_phi, e_theta)."""
    ph, th = unlensed(src)
    defl = np.asarray(defl, dtype=np.float64)
    return np.stack([ph + defl[:, 0] / (ARCSEC * np.sin(th)), th + defl[:, 1] / ARCSEC], 1)


def compare_feature_sets(ctx, length, train_src, val_src, test_src, accuracy=1e-8, epochs=20, batch=1024,
                         lr=1e-3, seed=0):
    """The two feature sets of the learned first guess on the same
    root-finder pairs: the Born-residual set (catalogue_features; the network
    learns what the Born estimate leaves) and the raw point-sample set
    (point_features; the network learns the whole deflection).  Trained on
    the cones of train_src, early-stopped on val_src, judged on test_src:
    rms of the first guess's error there (arcsec) and the Newton iterations
    and rays from each first guess.  Returns a dict."""
    def pairs(src):
        out, _ = ctx.catalogue_solve(src, accuracy=accuracy)
        ok = out[:, capi.CAT["status"]] == capi.CAT_OK
        fb = ctx.catalogue_features(src[ok])
        fp = point_features(ctx, src[ok], length)
        d = deflections(out[ok]).astype(np.float32)
        return src[ok], fb, fp, d, out[ok]
    tr = [pairs(s) for s in train_src]
    va = pairs(val_src)
    te = pairs(test_src)
    fb = np.concatenate([t[1] for t in tr])
    fp = np.concatenate([t[2] for t in tr])
    d = np.concatenate([t[3] for t in tr])
    res = {"training_pairs": int(len(d))}
    # Born-residual network: targets = deflection - Born estimate
    fit_b = train(ctx, fb, (d - fb[:, 1:3]).astype(np.float32), va[1], va[3] - va[1][:, 1:3], epochs=epochs,
                  batch=batch, lr=lr, seed=seed)
    state_b = ctx.mlp_state()
    pred_b = ctx.mlp_infer(te[1])
    # absolute angles among the point features: nothing to turn
    fit_p = train(ctx, fp, d, va[2], va[3], epochs=epochs, batch=batch, lr=lr, seed=seed, turn=False)
    pred_p = ctx.mlp_infer(te[2])
    ctx.mlp_state(state_b)
    born = te[1][:, 1:3].astype(np.float64)
    truth = te[3].astype(np.float64)
    guesses = {"plain": np.zeros_like(truth), "born": born, "born+network": born + pred_b,
               "points network": pred_p.astype(np.float64)}
    src_t = te[0]
    for name, defl in guesses.items():
        rms = float(np.sqrt(np.mean((defl - truth) ** 2)))
        out, rays = ctx.catalogue_solve(src_t, init2=guess_from_deflection(src_t, defl), accuracy=accuracy,
                                        jacobian_all_rounds=True)
        ok = out[:, capi.CAT["status"]] == capi.CAT_OK
        res[name] = {"rms_arcsec": rms, "mean_iterations": float(out[ok, capi.CAT["iter"]].mean()),
                     "rays": int(rays), "converged": int(ok.sum()), "sources": int(len(src_t))}
    res["held_out"] = {"born_rms": fit_b["val_rms_born"], "born+network_rms": fit_b["val_rms_network"],
                       "points_network_rms": fit_p["val_rms_network"], "deflection_rms": fit_p["val_rms_born"]}
    res["loss_born_net"] = [float(fit_b["loss"][0]), float(fit_b["loss"][-1])]
    res["loss_points_net"] = [float(fit_p["loss"][0]), float(fit_p["loss"][-1])]
    return res


def solve_seeded(ctx, src, use_network=True, **newton):
    """Root-finder started from the first guesses: (records, rays)."""
    init = ctx.catalogue_seed(src, use_network)
    return ctx.catalogue_solve(src, init, **newton)
