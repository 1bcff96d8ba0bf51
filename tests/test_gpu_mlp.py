"""Learned first guess of the catalogue root-finder on the GPU, against the
CPU oracle (oracle/mlp_oracle.py, pinned by finite differences and closed
forms): the matrix-core operand layout, the forward pass, the gradient, one
Adam step, training, the features, and the gain on the root-finder.

Tolerances (BF16 operands, FP32 accumulation on the device; the oracle rounds
to BF16 at the same points and sums in FP64, so a value near a BF16 rounding
boundary may round the other way):
  * one 16x16x16 product: 1e-6 of sum |a||b| (exact products, FP32 sums);
  * outputs: 1e-2 of their rms; gradient: 2e-2 of its norm, per layer;
  * Adam step: 2e-6 absolute at lr = 1e-3 (a sign error moves a weight 2e-3).
"""
import math

import numpy as np
import pytest

import cases
import mlp_oracle as mlp
from mpf import capi, proposal, synth
from oracle import Oracle

pytestmark = pytest.mark.gpu


def context(precision=capi.PREC_FP64):
    return capi.Context(cases.LENGTH, order=capi.TSC, precision=precision, nsteps=cases.NSTEPS)


def random_state(seed):
    rng = np.random.default_rng(seed)
    W1 = np.zeros((mlp.KIN, mlp.NH))
    W1[:mlp.NIN] = rng.uniform(-1, 1, (mlp.NIN, mlp.NH)) * math.sqrt(6 / (mlp.NIN + mlp.NH))
    W2 = rng.uniform(-1, 1, (mlp.NH, mlp.NH)) * math.sqrt(6 / (2 * mlp.NH))
    W3 = np.zeros((mlp.NH, mlp.KOUT))
    W3[:, :mlp.NOUT] = rng.uniform(-1, 1, (mlp.NH, mlp.NOUT)) * math.sqrt(6 / (mlp.NH + mlp.NOUT))
    b1, b2 = rng.normal(scale=0.1, size=mlp.NH), rng.normal(scale=0.1, size=mlp.NH)
    b3 = np.zeros(mlp.KOUT)
    b3[:mlp.NOUT] = rng.normal(scale=0.1, size=mlp.NOUT)
    norm = np.concatenate([rng.normal(size=mlp.NIN), rng.uniform(0.5, 2, mlp.NIN), [1.5, -0.5], [2.0, 0.3]])
    return np.concatenate([mlp.pack(W1, b1, W2, b2, W3, b3), norm]).astype(np.float32)


def data(seed, n):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, mlp.NIN)).astype(np.float32), rng.normal(size=(n, mlp.NOUT)).astype(np.float32)


def test_mfma_operand_layout():
    print("[case] mfma probe", flush=True)
    c = context()
    rng = np.random.default_rng(1)
    a = mlp.bf16(rng.normal(size=(16, 16)).astype(np.float32))
    b = mlp.bf16(rng.normal(size=(16, 16)).astype(np.float32))
    out = c.mfma_probe(a, b)
    bound = np.abs(a) @ np.abs(b)
    assert np.all(np.abs(out - a @ b) <= 1e-6 * bound)
    assert np.max(np.abs(out - a.T @ b)) > 0.1   # a transposed operand would be caught
    c.close()


def test_mfma32_operand_layout():
    """gfx950's 16x16x32 product with the kernels' layout gives a b; a B
    operand whose slots pair with other K than A's does not (and a
    transposed operand would not either)."""
    print("[case] mfma 16x16x32 probe", flush=True)
    c = context()
    rng = np.random.default_rng(2)
    a = mlp.bf16(rng.normal(size=(16, 32)).astype(np.float32))
    b = mlp.bf16(rng.normal(size=(32, 16)).astype(np.float32))
    bound = np.abs(a) @ np.abs(b)
    out = c.mfma_probe32(a, b, 0)
    assert np.all(np.abs(out - a @ b) <= 1e-6 * bound)
    assert np.max(np.abs(out - (a @ b).T)) > 0.1
    bad = c.mfma_probe32(a, b, 1)
    assert np.max(np.abs(bad - a @ b)) > 0.1
    c.close()


def test_inference_kernel_time_and_large_batch():
    """The persistent inference kernel on 1 M samples: equal to the oracle's
    forward pass (1e-2 of the outputs' rms, as above) on a sample of rows,
    the same outputs from a second call, its HIP-event duration reported."""
    print("[case] mlp inference, 1048576 samples", flush=True)
    c = context()
    c.mlp_init(0)
    state = random_state(21)
    c.mlp_state(state)
    x, _ = data(22, 1048576)
    y = c.mlp_infer(x)
    ms = c.mlp_last_ms()
    y2 = c.mlp_infer(x)
    np.testing.assert_array_equal(y, y2)
    rows = np.random.default_rng(23).choice(len(x), 4096, replace=False)
    yo = mlp.predict(state, x[rows], mixed=True)
    assert np.max(np.abs(y[rows] - yo)) <= 1e-2 * np.sqrt(np.mean(yo ** 2))
    flop = len(x) * 2 * (10 * 128 + 128 * 128 + 128 * 2)
    print(f"  inference kernel {ms:.4f} ms: {flop / (ms / 1e3) / 1e12:.1f} TFLOP/s", flush=True)
    assert 0 < ms < 50
    c.close()


def test_init_state():
    print("[case] mlp init", flush=True)
    c = context()
    c.mlp_init(3)
    s = c.mlp_state().astype(np.float64)
    W1, b1, W2, b2, W3, b3 = mlp.unpack(s[:mlp.NPARAM])
    assert np.all(W1[mlp.NIN:] == 0) and np.all(W3 == 0)
    assert np.all(b1 == 0) and np.all(b2 == 0) and np.all(b3 == 0)
    assert np.max(np.abs(W1)) <= math.sqrt(6 / (mlp.NIN + mlp.NH)) and np.std(W1[:mlp.NIN]) > 0.05
    assert np.max(np.abs(W2)) <= math.sqrt(6 / (2 * mlp.NH))
    norm = s[mlp.NPARAM:]
    np.testing.assert_array_equal(norm, np.concatenate([np.zeros(10), np.ones(10), np.zeros(2), np.ones(2)]))
    c.close()


def test_forward_vs_oracle():
    print("[case] mlp forward", flush=True)
    c = context()
    c.mlp_init(0)
    state = random_state(5)
    c.mlp_state(state)
    np.testing.assert_array_equal(c.mlp_state(), state)
    for n in (1, 63, 1000):
        x, _ = data(6 + n, n)
        y = c.mlp_infer(x)
        yo = mlp.predict(state, x, mixed=True)
        scale = np.sqrt(np.mean(yo ** 2))
        assert np.max(np.abs(y - yo)) <= 1e-2 * scale, n
    c.close()


def test_gradient_vs_oracle():
    print("[case] mlp gradient", flush=True)
    c = context()
    c.mlp_init(0)
    state = random_state(7)
    c.mlp_state(state)
    x, t = data(8, 777)   # ragged: not a multiple of 16, 64 or 256
    g, loss = c.mlp_gradient(x, t)
    xn, tn = mlp.normalise(state, x, t)
    lo, go = mlp.loss_and_grad(state[:mlp.NPARAM].astype(np.float64), xn, tn, mixed=True)
    assert abs(loss - lo) <= 1e-3 * lo
    gb, ob = mlp.unpack(g), mlp.unpack(go)
    for k, (u, v) in enumerate(zip(gb, ob)):
        assert np.linalg.norm(u - v) <= 2e-2 * np.linalg.norm(v), k
    W1, _, _, _, W3, b3 = gb
    assert np.all(W1[mlp.NIN:] == 0) and np.all(W3[:, mlp.NOUT:] == 0) and np.all(b3[mlp.NOUT:] == 0)
    c.close()


def test_adam_first_step():
    print("[case] mlp adam", flush=True)
    c = context()
    c.mlp_init(0)
    state = random_state(9)
    c.mlp_state(state)
    x, t = data(10, 500)
    g, _ = c.mlp_gradient(x, t)
    c.mlp_train(x, t, epochs=1, batch=500, lr=1e-3, seed=1)
    p1 = c.mlp_state()[:mlp.NPARAM].astype(np.float64)
    want, _, _ = mlp.adam_step(state[:mlp.NPARAM].astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, 1e-3)
    assert np.max(np.abs(p1 - want)) <= 2e-6
    c.close()


def test_training_fits_a_smooth_function():
    print("[case] mlp training", flush=True)
    rng = np.random.default_rng(11)

    def f(x):
        return np.stack([np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2], np.tanh(x[:, 3]) - 0.3 * x[:, 4] ** 2], 1)
    x = rng.normal(size=(8192, mlp.NIN)).astype(np.float32)
    xt = rng.normal(size=(2048, mlp.NIN)).astype(np.float32)
    c = context()
    c.mlp_init(2)
    loss = c.mlp_train(x, f(x).astype(np.float32), epochs=40, batch=256, lr=2e-3, seed=3)
    print("[loss]", loss[0], loss[-1], flush=True)
    assert loss[-1] < 0.05 * loss[0]
    y = c.mlp_infer(xt)
    err = np.mean((y - f(xt)) ** 2, 0) / np.var(f(xt), 0)
    assert np.all(err < 0.05), err
    state = c.mlp_state()
    assert np.max(np.abs(mlp.predict(state, xt, mixed=True) - y)) <= 1e-2 * np.sqrt(np.mean(y ** 2))
    c.close()


@pytest.mark.parametrize("precision", (capi.PREC_FP64, capi.PREC_SPLIT))
def test_features_vs_oracle(precision):
    """Exact tier: FP64 TSC with a cache per line, as the oracle's cached
    tsc(); split tier: FP32 stencil blocks (FP32 coefficients: 1e-5).
    Features 3, 4 (5, 6) are differences of two Born sums, each as accurate
    as features 1, 2 (7, 8): their error is measured against that scale."""
    print(f"[case] catalogue features precision={precision}", flush=True)
    rec, f64 = cases.tree("amr")
    c = context(precision)
    c.load_records(rec)
    c.load_fields64(f64)
    src = proposal.cone_sources(300, 20.0, 20.0, 8.0, 0.05, 0.45, seed=12)
    feat = c.catalogue_features(src)
    O = Oracle()
    h = O.tree(rec, f64)
    fo = mlp.features(lambda pts: O.interp(h, 3, pts, precision=1), src, cases.LENGTH, synth.C_LIGHT ** 2)
    O.free(h)
    scale = np.max(np.abs(fo), 0) + 1e-30
    scale[3:5] = scale[1:3]
    scale[5:7] = scale[7:9]
    tol = 1e-6 if precision == capi.PREC_FP64 else 1e-5
    err = np.max(np.abs(feat - fo) / scale, 0)
    assert np.all(err <= tol), err
    assert np.max(np.abs(fo[:, 1:9])) > 0   # the tree's potential is not flat along these lines
    c.close()


@pytest.mark.parametrize("precision", (capi.PREC_FP64, capi.PREC_SPLIT))
def test_features_large_batch_stable(precision):
    """A batch of several workgroups per compute unit: two calls agree bit for
    bit, and with the same sources computed in two halves."""
    print(f"[case] catalogue features, 131072 sources, precision={precision}", flush=True)
    rec, f64 = cases.tree("amr")
    c = context(precision)
    c.load_records(rec)
    c.load_fields64(f64)
    src = proposal.cone_sources(131072, 20.0, 20.0, 8.0, 0.05, 0.45, seed=13)
    a = c.catalogue_features(src)
    b = c.catalogue_features(src)
    h = np.concatenate([c.catalogue_features(src[:65536]), c.catalogue_features(src[65536:])])
    c.close()
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a, h)


def test_jacobian_all_rounds_same_results():
    print("[case] catalogue jacobian_all_rounds", flush=True)
    rec, f64 = cases.tree("amr")
    c = context(capi.PREC_FP64)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    src = proposal.cone_sources(200, 20.0, 20.0, 8.0, 0.05, 0.45, seed=13)
    a, ra = c.catalogue_solve(src, accuracy=1e-12)
    b, rb = c.catalogue_solve(src, accuracy=1e-12, jacobian_all_rounds=True)
    d, rd = c.catalogue_solve(src, accuracy=1e-12, jacobian_all_rounds=2)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a, d)
    ok = a[:, capi.CAT["status"]] == capi.CAT_OK
    early = np.sum(ok & (a[:, capi.CAT["iter"]] <= 3))
    assert ra - rb == early
    # schedule 2: one more ray for the sources that converged at their first ray
    assert rd - rb == np.sum(ok & (a[:, capi.CAT["iter"]] <= 1))
    c.close()


def lightcone():
    return synth.lightcone_tree(6, 3000.0, 1.0, 4.0, 8, 7, device="cpu", noise=1.0, decay=0.75)


@pytest.mark.parametrize("precision", (capi.PREC_FP64, capi.PREC_SPLIT))
def test_proposal_reduces_newton_rounds(precision):
    """Trained on the root-finder's own pairs from 3 cones and early-stopped
    on a 4th, the first guess cuts the rays a 5th cone needs at
    cat_accuracy = 1e-8 rad; every source still converges, to the same root
    within the accuracy.  The network's correction is used only when it beats
    the Born estimate alone on the held-out cone; both are reported."""
    print(f"[case] proposal gain precision={precision}", flush=True)
    rec, f64, length = lightcone()
    c = capi.Context(length, order=capi.TSC, precision=precision, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    acc = 1e-8
    axes = ((70.0, 40.0), (140.0, 60.0), (220.0, 35.0), (300.0, 110.0))
    pairs = []
    for k, (p, t) in enumerate(axes):
        src = proposal.cone_sources(1500, p, t, 6.0, 0.08, 0.45, seed=100 + k)
        fe, ta, _, _ = proposal.training_pairs(c, src, accuracy=acc)
        pairs.append((fe, ta))
    feat = np.concatenate([p[0] for p in pairs[:3]])
    targ = np.concatenate([p[1] for p in pairs[:3]])
    fit = proposal.train(c, feat, targ, pairs[3][0], pairs[3][1], epochs=60, batch=256, lr=1e-3, seed=4)
    print("[train] pairs", len(feat), "loss", fit["loss"][0], fit["loss"][-1], "held-out rms born",
          fit["val_rms_born"], "network", fit["val_rms_network"], "epoch", fit["epoch"], "use", fit["use_network"],
          flush=True)
    src = proposal.cone_sources(1500, 20.0, 20.0, 6.0, 0.08, 0.45, seed=200)
    fe = c.catalogue_features(src)
    plain, rays_plain = c.catalogue_solve(src, accuracy=acc)
    born, rays_born = proposal.solve_seeded(c, src, use_network=False, accuracy=acc, jacobian_all_rounds=True)
    seeded, rays_seeded = proposal.solve_seeded(c, src, use_network=fit["use_network"], accuracy=acc,
                                                jacobian_all_rounds=True)
    np.testing.assert_allclose(c.catalogue_seed(src, False), proposal.born_guess(src, fe), rtol=0, atol=1e-15)
    st, it = capi.CAT["status"], capi.CAT["iter"]
    ok = (plain[:, st] == capi.CAT_OK) & (seeded[:, st] == capi.CAT_OK)
    assert ok.mean() > 0.98
    resid = proposal.targets(plain[ok], fe[ok])
    pred = c.mlp_infer(fe[ok])
    print("[gain] test residual rms", np.sqrt(np.mean(resid ** 2, 0)), "after network",
          np.sqrt(np.mean((resid - pred) ** 2, 0)), flush=True)
    print("[gain] mean iterations plain", plain[ok, it].mean(), "born", born[ok, it].mean(), "seeded",
          seeded[ok, it].mean(), "rays", rays_plain, rays_born, rays_seeded, flush=True)
    # the network against the Born estimate alone, with the same Newton finish
    net, rays_net = proposal.solve_seeded(c, src, use_network=True, accuracy=acc, jacobian_all_rounds=True)
    okn = ok & (net[:, st] == capi.CAT_OK)
    gain = born[okn, it].mean() - net[okn, it].mean()
    print(f"[gain] network - Born: {net[okn, it].mean():.3f} - {born[okn, it].mean():.3f} iterations "
          f"({gain:+.3f} saved by the network), rays {rays_net} vs {rays_born}; held-out rms network "
          f"{fit['val_rms_network']:.4f} vs Born {fit['val_rms_born']:.4f} arcsec, use_network {fit['use_network']}",
          flush=True)
    for f in ("phi_obs", "theta_obs"):
        assert np.max(np.abs(plain[ok, capi.CAT[f]] - seeded[ok, capi.CAT[f]])) < 20 * acc
    assert seeded[ok, it].mean() < plain[ok, it].mean() - 0.5
    assert rays_seeded < 0.6 * rays_plain
    # the gate: the first guess in use costs at most noise against Born alone,
    # and the network is in use only where it beat Born on the held-out cone
    assert seeded[ok, it].mean() <= born[ok, it].mean() + 0.05
    if fit["use_network"]:
        assert fit["val_rms_network"] < fit["val_rms_born"]
    c.close()


def test_feature_sets_on_held_out_cone():
    """The two feature sets of the learned first guess on the same pairs
    (4 training cones of 16,384 sources, a validation cone, a test cone):
    both networks learn (the loss falls), every first guess still converges
    to the accuracy, and the comparison is printed: rms of each first guess's
    error and the Newton iterations it leaves."""
    print("[case] feature sets: Born-residual vs point samples", flush=True)
    rec, f64, length = lightcone()
    c = capi.Context(length, order=capi.TSC, precision=capi.PREC_SPLIT, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.load_fields64(f64)
    c.set_cosmology(cases.cosmology())
    axes = ((70.0, 40.0), (140.0, 60.0), (220.0, 35.0), (300.0, 110.0))
    train_src = [proposal.cone_sources(16384, p, t, 6.0, 0.08, 0.45, seed=300 + k) for k, (p, t) in enumerate(axes)]
    val = proposal.cone_sources(4096, 250.0, 80.0, 6.0, 0.08, 0.45, seed=310)
    test = proposal.cone_sources(4096, 20.0, 20.0, 6.0, 0.08, 0.45, seed=311)
    r = proposal.compare_feature_sets(c, length, train_src, val, test, accuracy=1e-8, epochs=20, batch=1024,
                                      lr=1e-3, seed=5)
    for k, v in r.items():
        print(f"[compare] {k}: {v}", flush=True)
    assert r["loss_born_net"][1] < r["loss_born_net"][0] and r["loss_points_net"][1] < r["loss_points_net"][0]
    for name in ("plain", "born", "born+network", "points network"):
        assert r[name]["converged"] >= 0.98 * r[name]["sources"], name
    assert r["born"]["mean_iterations"] < r["plain"]["mean_iterations"]
    c.close()
