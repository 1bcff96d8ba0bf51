"""Pins the CPU oracle of the learned first guess (no reference counterpart):
finite-difference gradients of the FP64 mode, the BF16 rounding rule on
known values, the mixed-precision mode against the FP64 mode, Adam's closed
form at its first step, and the features in closed-form cases."""
import math

import numpy as np
import pytest

import mlp_oracle as mlp


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    W1 = np.zeros((mlp.KIN, mlp.NH))
    W1[:mlp.NIN] = rng.uniform(-1, 1, (mlp.NIN, mlp.NH)) * math.sqrt(6 / (mlp.NIN + mlp.NH)) * scale
    W2 = rng.uniform(-1, 1, (mlp.NH, mlp.NH)) * math.sqrt(6 / (2 * mlp.NH)) * scale
    W3 = np.zeros((mlp.NH, mlp.KOUT))
    W3[:, :mlp.NOUT] = rng.uniform(-1, 1, (mlp.NH, mlp.NOUT)) * math.sqrt(6 / (mlp.NH + mlp.NOUT)) * scale
    b1 = rng.normal(scale=0.1, size=mlp.NH)
    b2 = rng.normal(scale=0.1, size=mlp.NH)
    b3 = np.zeros(mlp.KOUT)
    b3[:mlp.NOUT] = rng.normal(scale=0.1, size=mlp.NOUT)
    return mlp.pack(W1, b1, W2, b2, W3, b3)


def batch(seed, n=32):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, mlp.NIN)), rng.normal(size=(n, mlp.NOUT))


def fd_check(p, x, t, idx, h=1e-6):
    _, g = mlp.loss_and_grad(p, x, t, mixed=False)
    worst = 0.0
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        num = (mlp.loss_and_grad(pp, x, t, mixed=False)[0] - mlp.loss_and_grad(pm, x, t, mixed=False)[0]) / (2 * h)
        worst = max(worst, abs(num - g[i]) / max(abs(num), 1e-6))
    return worst


def sample_indices(seed):
    rng = np.random.default_rng(seed)
    blocks = [(mlp.OFF_W1, mlp.NIN * mlp.NH), (mlp.OFF_B1, mlp.NH), (mlp.OFF_W2, mlp.NH * mlp.NH),
              (mlp.OFF_B2, mlp.NH), (mlp.OFF_W3, mlp.NH * mlp.KOUT), (mlp.OFF_B3, mlp.NOUT)]
    idx = []
    for off, size in blocks:
        idx += list(off + rng.choice(size, size=min(size, 12), replace=False))
    # W3 entries of the real outputs
    idx += [mlp.OFF_W3 + k * mlp.KOUT + o for k in (0, 77) for o in range(mlp.NOUT)]
    return idx


def test_gradient_finite_differences():
    p = random_params(1)
    x, t = batch(2)
    assert fd_check(p, x, t, sample_indices(3)) < 1e-6


def test_gradient_of_padding_is_zero():
    p = random_params(4)
    x, t = batch(5)
    for mixed in (False, True):
        _, g = mlp.loss_and_grad(p, x, t, mixed=mixed)
        W1, b1, W2, b2, W3, b3 = mlp.unpack(g)
        assert np.all(W1[mlp.NIN:] == 0)
        assert np.all(W3[:, mlp.NOUT:] == 0)
        assert np.all(b3[mlp.NOUT:] == 0)


def test_planted_sign_error_is_caught(monkeypatch):
    """A wrong sign in GELU' (planted in the oracle) fails the check."""
    p = random_params(6)
    x, t = batch(7)
    good = mlp.gelu_d
    monkeypatch.setattr(mlp, "gelu_d", lambda z: good(z) - 2 * z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi))
    assert fd_check(p, x, t, sample_indices(8)) > 1e-2


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1 + 2 ** -8, 1 + 2 ** -7 + 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -3.0, 0.0, 2.0 ** -130,
                  3.140625, 65504.0], dtype=np.float32)
    want = np.array([1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -3.0, 0.0, 2.0 ** -130, 3.140625, 65536.0])
    np.testing.assert_array_equal(mlp.bf16(v), want)


def test_mixed_precision_close_to_fp64():
    p = random_params(9)
    x, t = batch(10, n=256)
    y64, _ = mlp.forward(p, x, mixed=False)
    y16, _ = mlp.forward(p, x, mixed=True)
    assert np.max(np.abs(y16 - y64)) < 3e-2 * np.sqrt(np.mean(y64 ** 2))
    _, g64 = mlp.loss_and_grad(p, x, t, mixed=False)
    _, g16 = mlp.loss_and_grad(p, x, t, mixed=True)
    assert np.linalg.norm(g16 - g64) < 3e-2 * np.linalg.norm(g64)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(11)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    g[:5] = 0.0
    p1, m, v = mlp.adam_step(p, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    # m_hat = g, v_hat = g^2: p - lr g / (|g| + eps)
    np.testing.assert_allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    np.testing.assert_array_equal(p1[:5], p[:5])


def test_predict_uses_the_normalisation():
    p = random_params(12)
    state = np.zeros(mlp.NPARAM + mlp.NNORM)
    state[:mlp.NPARAM] = p
    mean_x, std_x = np.linspace(-1, 1, mlp.NIN), np.linspace(0.5, 2, mlp.NIN)
    mean_t, std_t = np.array([3.0, -2.0]), np.array([0.1, 10.0])
    state[mlp.NPARAM:] = np.concatenate([mean_x, std_x, mean_t, std_t])
    x, _ = batch(13)
    y, _ = mlp.forward(p, (x - mean_x) / std_x, mixed=False)
    np.testing.assert_allclose(mlp.predict(state, x, mixed=False), y * std_t + mean_t, rtol=1e-14, atol=1e-14)


def test_features_closed_forms():
    """phi == 0 gives zero deflection features; a uniform gradient g gives the
    Born sum sum_k (1 - (k + 1/2)/N) / N = 1/2 of 2 g_perp L D / c^2 (N
    samples) and 1/8 of it over the far half; the second line, towards
    theta + f2 (a gradient along z deflects along e_theta only), sees
    g_perp = -sin(theta + f2) g, so f4 = 1/2 2 g L D / c^2 (sin(theta) -
    sin(theta + f2)) and f6 = 1/8 of the same."""
    src = np.array([[0.1, 0.2, 0.3, 0, 0, 0], [0.3, -0.1, 0.2, 0, 0, 0]])
    L, c2 = 1.0e25, 8.98755178736818e16
    f0 = mlp.features(lambda pts: np.zeros((len(pts), 7)), src, L, c2)
    D = np.linalg.norm(src[:, :3], axis=1)
    np.testing.assert_allclose(f0[:, 0], D, rtol=1e-15)
    assert np.all(f0[:, 1:] == 0)
    g = 1e-6

    def uniform_z(pts):
        d = np.zeros((len(pts), 7))
        d[:, 4] = g
        return d
    f = mlp.features(uniform_z, src, L, c2)
    theta = np.arccos(src[:, 2] / D)
    # grad along +z: e_theta component -sin(theta) g, e_phi component 0
    full = 2 * (-np.sin(theta) * g) * L * D / c2 * mlp.ARCSEC
    np.testing.assert_allclose(f[:, [1, 3, 5, 7]], 0, atol=1e-20)
    np.testing.assert_allclose(f[:, 2], 0.5 * full, rtol=1e-12)
    theta1 = theta + f[:, 2] / mlp.ARCSEC
    second = 0.5 * 2 * g * L * D / c2 * mlp.ARCSEC * (np.sin(theta) - np.sin(theta1))
    assert np.all(np.abs(second) > 0)
    np.testing.assert_allclose(f[:, 4], second, rtol=1e-6)
    np.testing.assert_allclose(f[:, 6], 0.25 * second, rtol=1e-6)
    np.testing.assert_allclose(f[:, 8], 0.125 * full, rtol=1e-12)
    assert np.all(f[:, 9] == 0)
