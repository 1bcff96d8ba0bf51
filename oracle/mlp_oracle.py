"""CPU oracle of the learned first guess of the catalogue root-finder.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
network (forward, loss, gradient, Adam step) and of the source features.
The network has no counterpart in the reference; what pins this restatement:
  * the gradient of the FP64 mode against central finite differences,
  * the mixed-precision mode (BF16 rounding at the points where the device
    rounds) against the FP64 mode, within the BF16 resolution,
  * the features' TSC samples come from the oracle's cached interpolation,
    itself pinned against the reference's tsc() (simplehyperoctree.h:1554-1644),
    and the Born sums against closed forms (uniform gradient).

Network: y = gelu(gelu(x W1 + b1) W2 + b2) W3 + b3, loss = mean over samples
and outputs of (y - t)^2, parameters packed as in include/mpf.h (W1 16 x 128
with rows 10..15 zero, W3 128 x 16 with columns 2..15 zero).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

NIN, NOUT, NH, KIN, KOUT = 10, 2, 128, 16, 16
OFF_W1 = 0
OFF_B1 = OFF_W1 + KIN * NH
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH * KOUT
NPARAM = OFF_B3 + KOUT
NNORM = 24
ARCSEC = 206264.80624709636
FEAT_SAMPLES = 256


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(KIN, NH), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH, NH),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(NH, KOUT), p[OFF_B3:NPARAM])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def _pad_x(x):
    x = np.asarray(x, dtype=np.float64).reshape(-1, NIN)
    xp = np.zeros((x.shape[0], KIN))
    xp[:, :NIN] = x
    return xp


def forward(p, x, mixed=True):
    """Outputs (n x NOUT) and the intermediate tensors.  mixed: round the
    input, the weights and the hidden activations to BF16 as the device does
    (FP32 biases, sums in FP64 here)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if mixed else (lambda a: np.asarray(a, dtype=np.float64))
    xp = r(_pad_x(x))
    W1r, W2r, W3r = r(W1), r(W2), r(W3)
    z1 = xp @ W1r + b1
    h1 = r(gelu(z1))
    z2 = h1 @ W2r + b2
    h2 = r(gelu(z2))
    y = h2 @ W3r + b3
    return y[:, :NOUT], dict(x=xp, z1=z1, h1=h1, z2=z2, h2=h2, W1=W1r, W2=W2r, W3=W3r, yfull=y)


def loss_and_grad(p, x, t, mixed=True):
    """Mean squared error over samples and outputs and its gradient
    (NPARAM).  mixed: the device's rounding points (BF16 dL/dZ of every
    layer, BF16 operands)."""
    y, c = forward(p, x, mixed)
    t = np.asarray(t, dtype=np.float64).reshape(-1, NOUT)
    n = y.shape[0]
    r = bf16 if mixed else (lambda a: np.asarray(a, dtype=np.float64))
    e = y - t
    loss = float(np.sum(e * e) / (n * NOUT))
    d3 = np.zeros((n, KOUT))
    d3[:, :NOUT] = 2.0 * e / (n * NOUT)
    d3 = r(d3)
    d2 = r((d3 @ c["W3"].T) * gelu_d(c["z2"]))
    d1 = r((d2 @ c["W2"].T) * gelu_d(c["z1"]))
    g = pack(c["x"].T @ d1, d1.sum(0), c["h1"].T @ d2, d2.sum(0), c["h2"].T @ d3, d3.sum(0))
    return loss, g


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Adam (Kingma & Ba 2015, algorithm 1), step counted from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def normalise(state, x, t=None):
    s = np.asarray(state, dtype=np.float64)[NPARAM:]
    xn = (np.asarray(x, dtype=np.float64) - s[:NIN]) / s[NIN:2 * NIN]
    if t is None:
        return xn
    tn = (np.asarray(t, dtype=np.float64) - s[2 * NIN:2 * NIN + NOUT]) / s[2 * NIN + NOUT:]
    return xn, tn


def predict(state, x, mixed=True):
    """Denormalised prediction of a full state (parameters + normalisation)."""
    s = np.asarray(state, dtype=np.float64)
    y, _ = forward(s[:NPARAM], normalise(s, x), mixed)
    return y * s[NPARAM + 2 * NIN + NOUT:] + s[NPARAM + 2 * NIN:NPARAM + 2 * NIN + NOUT]


def _line_basis(phi, theta):
    cp, sp, ct, st = np.cos(phi), np.sin(phi), np.cos(theta), np.sin(theta)
    nv = np.stack([st * cp, st * sp, ct], -1)
    ep = np.stack([-sp, cp, np.zeros_like(sp)], -1)
    et = np.stack([cp * ct, sp * ct, -st], -1)
    return nv, ep, et


def features(sample_line, src6, length, c2):
    """Source features (include/mpf.h).  sample_line(points (FEAT_SAMPLES x 3),
    ordered outward along one line of sight) -> (FEAT_SAMPLES x 7) fields
    (rho, phi, dx, dy, dz, a, dt): the device samples each line with one
    stencil cache carried outward.  Two lines per source: to the source, and
    towards the source's angles plus the first line's Born deflection."""
    src = np.asarray(src6, dtype=np.float64).reshape(-1, 6)
    n = src.shape[0]
    D = np.linalg.norm(src[:, :3], axis=1)
    theta = np.arccos(src[:, 2] / D)
    phi = np.arctan2(src[:, 1], src[:, 0])
    g2 = 2.0 * length / c2 * ARCSEC
    k = np.arange(FEAT_SAMPLES)
    far = k >= FEAT_SAMPLES // 2

    def line(ph, th, Di):
        nv, ep, et = _line_basis(ph, th)
        chi = Di * (k + 0.5) / FEAT_SAMPLES
        d = sample_line(nv[None, :] * chi[:, None])
        wgt = (Di - chi) / Di * (Di / FEAT_SAMPLES) * g2
        wp = wgt * (d[:, 2:5] @ ep)
        wt = wgt * (d[:, 2:5] @ et)
        return wp, wt, d

    f = np.zeros((n, NIN))
    for i in range(n):
        wp, wt, d = line(phi[i], theta[i], D[i])
        b = np.array([wp.sum(), wt.sum()])
        bf = np.array([wp[far].sum(), wt[far].sum()])
        f[i, 1:3] = b
        f[i, 7:9] = bf
        f[i, 9] = d[FEAT_SAMPLES // 2 - 1, 1] / c2 * 1.0e5
        wp1, wt1, _ = line(phi[i] + b[0] / (ARCSEC * np.sin(theta[i])), theta[i] + b[1] / ARCSEC, D[i])
        f[i, 3:5] = wp1.sum() - b[0], wt1.sum() - b[1]
        f[i, 5:7] = wp1[far].sum() - bf[0], wt1[far].sum() - bf[1]
    f[:, 0] = D
    return f
