"""Seeded synthetic inputs for the photon-propagation path.

Everything here writes inputs in the reference's own formats (48-byte cone
records, a (t, a, da/dt, d2a/dt2) cosmology table, launch photons), so the same
inputs can be fed to the reference code, the oracle and the HIP library.

Units follow the reference: time in seconds, octree positions in box units
(the simulation box of length `length` metres spans [-0.5, 0.5) around the
observer at the origin), phi in m^2/s^2, dphid{x,y,z} in m/s^2 (gradient per
metre), dphidt in m^2/s^3.
"""
from __future__ import annotations

import math

import numpy as np

from . import formats

C_LIGHT = 299792458.0
MPC_SI = 3.08568e22  # jobs/step_00002_rays.sh:57


# --------------------------------------------------------------- cosmology

def flat_lcdm_table(h=0.7, omega_m=0.3, a_min=0.05, n=2048, mpc=MPC_SI):
    """Cosmology table (t, a, da/dt, d2a/dt2) of a flat LCDM universe.

    t is the conformal lookback time from the observer (seconds, increasing),
    so da/dt = -a^2 H(a) < 0, the convention of the reference's corrected
    table (input.h:1091-1233, 1330-1378).  Rows are uniform in a; t is
    integrated with composite Simpson on a fine grid.  Returns a (4, n) array.
    """
    H0 = h * 1.0e5 / mpc
    ol = 1.0 - omega_m
    H = lambda a: H0 * np.sqrt(omega_m / a ** 3 + ol)
    a = np.linspace(1.0, a_min, n)
    # tau(a) = int_a^1 da' / (a'^2 H(a')), refined Simpson between rows
    sub = 16
    tau = np.zeros(n)
    for i in range(1, n):
        x = np.linspace(a[i], a[i - 1], 2 * sub + 1)
        f = 1.0 / (x ** 2 * H(x))
        dx = (a[i - 1] - a[i]) / (2 * sub)
        tau[i] = tau[i - 1] + dx / 3.0 * (f[0] + f[-1] + 4 * f[1:-1:2].sum() + 2 * f[2:-1:2].sum())
    dadt = -a ** 2 * H(a)
    # d/dtau(-a^2 H) = -(2 a H + a^2 dH/da) * da/dtau
    dHda = H0 * 0.5 / np.sqrt(omega_m / a ** 3 + ol) * (-3.0 * omega_m / a ** 4)
    d2 = -(2 * a * H(a) + a ** 2 * dHda) * dadt
    return np.stack([tau, a, dadt, d2])


def comoving_distance_of_z(z, h=0.7, omega_m=0.3, mpc=MPC_SI, n=4096):
    """Comoving distance (metres) to redshift z in flat LCDM."""
    H0 = h * 1.0e5 / mpc
    zz = np.linspace(0.0, z, n)
    f = C_LIGHT / (H0 * np.sqrt(omega_m * (1 + zz) ** 3 + 1 - omega_m))
    return float(np.trapezoid(f, zz))


def a_of_chi_table(table):
    """(chi_metres, a) along the background null ray: dchi/dtau = c."""
    return C_LIGHT * table[0], table[1]


# ------------------------------------------------------------------- rays

def launch_photons(phi, theta, observer=(0.0, 0.0, 0.0)):
    """Launch photons like Integrator::launch(x, y, z, phi, theta)
    (integrator.h:304-329): a=1, t=0, kt=1, k=(cos phi sin th, sin phi sin th, cos th).
    Returns (n, 9) float64: a, t, x, y, z, kt, kx, ky, kz."""
    phi = np.asarray(phi, dtype=np.float64).ravel()
    theta = np.asarray(theta, dtype=np.float64).ravel()
    out = np.zeros((phi.size, 9))
    out[:, 0] = 1.0
    out[:, 2:5] = observer
    out[:, 5] = 1.0
    out[:, 6] = np.cos(phi) * np.sin(theta)
    out[:, 7] = np.sin(phi) * np.sin(theta)
    out[:, 8] = np.cos(theta)
    return out


def cone_grid(phi0_deg, theta0_deg, aperture_deg, nside):
    """Regular nside x nside grid of (phi, theta) over an aperture centred on (phi0, theta0)."""
    half = math.radians(aperture_deg) / 2
    u = (np.arange(nside) + 0.5) / nside * 2 * half - half
    pp, tt = np.meshgrid(math.radians(phi0_deg) + u, math.radians(theta0_deg) + u, indexing="ij")
    return pp.ravel(), tt.ravel()


def cone_uniform(phi0_deg, theta0_deg, aperture_deg, n, seed):
    """n directions uniform (in solid angle) over the square aperture."""
    rng = np.random.default_rng(seed)
    half = math.radians(aperture_deg) / 2
    t0 = math.radians(theta0_deg)
    cmin, cmax = math.cos(t0 + half), math.cos(t0 - half)
    ct = rng.uniform(cmin, cmax, n)
    ph = math.radians(phi0_deg) + rng.uniform(-half, half, n)
    return ph, np.arccos(ct)


# ------------------------------------------------------------------ fields

def grf_potential(n, seed, slope=-3.0, rms=1.0e-5):
    """Periodic Gaussian random field on an n^3 grid with P(k) ~ k^slope, zero
    mean, given rms (in units of c^2 when used for phi/c^2)."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((n, n, n))
    fk = np.fft.rfftn(white)
    kx = np.fft.fftfreq(n)[:, None, None]
    ky = np.fft.fftfreq(n)[None, :, None]
    kz = np.fft.rfftfreq(n)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = 1.0
    amp = k2 ** (slope / 4.0)
    amp[0, 0, 0] = 0.0
    field = np.fft.irfftn(fk * amp, s=(n, n, n))
    field -= field.mean()
    field *= rms / field.std()
    return field


def central_gradient(f, dx):
    """Second-order periodic central differences along each axis."""
    return tuple((np.roll(f, -1, axis=ax) - np.roll(f, 1, axis=ax)) / (2 * dx) for ax in range(3))


def lognormal(g, sigma):
    """rho/rho_mean = exp(sigma g' - sigma^2/2) with g' the field normalised to unit variance."""
    gs = g / g.std()
    return np.exp(sigma * gs - 0.5 * sigma * sigma)


# ------------------------------------------------------------------- trees

def uniform_level_coords(level_input, half=0.5):
    """Octree coordinates of a uniform input-level grid spanning [-half, half)^3.
    Octree level = input level + log2(EXTENT) (input.h:1044)."""
    lvl = level_input + 2
    ncell = 1 << lvl
    size = formats.EXTENT / ncell
    lo = int(round((-half + 2.0) / size))
    hi = int(round((half + 2.0) / size))
    ax = np.arange(lo, hi, dtype=np.uint64)
    kx, ky, kz = np.meshgrid(ax, ax, ax, indexing="ij")
    return lvl, kx.ravel(), ky.ravel(), kz.ravel()


def analytic_fields(x, y, z, length, amp=1.0e-5, seed=0):
    """Smooth analytic potential (a few plane waves) with exact gradients.
    Returns (n, 7) float64 in FIELDS7 order (rho, phi, dphidx, dphidy, dphidz, a, dphidt)."""
    rng = np.random.default_rng(seed)
    c2 = C_LIGHT ** 2
    phi = np.zeros_like(x)
    gx = np.zeros_like(x)
    gy = np.zeros_like(x)
    gz = np.zeros_like(x)
    for _ in range(4):
        k = rng.uniform(2.0, 9.0, 3) * rng.choice([-1, 1], 3)
        p = rng.uniform(0, 2 * np.pi)
        ph = k[0] * x + k[1] * y + k[2] * z + p
        s = np.sin(ph)
        c = np.cos(ph)
        phi += amp * c2 * s / 4
        gx += amp * c2 * k[0] * c / 4 / length
        gy += amp * c2 * k[1] * c / 4 / length
        gz += amp * c2 * k[2] * c / 4 / length
    r = np.sqrt(x * x + y * y + z * z)
    a = 1.0 / (1.0 + 0.8 * r)
    rho = 1.0 + 0.5 * np.sin(3 * x) * np.cos(2 * y)
    dphidt = -1.0e-18 * phi
    return np.stack([rho, phi, gx, gy, gz, a, dphidt], axis=1)


def amr_coords(base_level, half, refine, max_level):
    """Cells of a tree with every base cell in [-half, half)^3 at `base_level`
    (octree levels), where each cell for which refine(level, cx, cy, cz) is
    true is split into its 8 children, recursively up to max_level.  Parents
    are kept.  Returns (level, kx, ky, kz) arrays."""
    ncell = 1 << base_level
    size = formats.EXTENT / ncell
    lo = int(round((-half + 2.0) / size))
    hi = int(round((half + 2.0) / size))
    ax = np.arange(lo, hi, dtype=np.uint64)
    kx, ky, kz = [a.ravel() for a in np.meshgrid(ax, ax, ax, indexing="ij")]
    levels = [np.full(kx.size, base_level, dtype=np.int64)]
    xs, ys, zs = [kx], [ky], [kz]
    cur = (kx, ky, kz)
    lvl = base_level
    while lvl < max_level and cur[0].size:
        cx, cy, cz = formats.cell_centers(lvl, *cur)
        m = refine(lvl, cx, cy, cz)
        px, py, pz = (c[m] for c in cur)
        if px.size == 0:
            break
        ch = []
        for s in range(8):
            ch.append((2 * px + np.uint64(s & 1), 2 * py + np.uint64((s >> 1) & 1), 2 * pz + np.uint64((s >> 2) & 1)))
        cur = tuple(np.concatenate([c[d] for c in ch]) for d in range(3))
        lvl += 1
        levels.append(np.full(cur[0].size, lvl, dtype=np.int64))
        xs.append(cur[0]); ys.append(cur[1]); zs.append(cur[2])
    return np.concatenate(levels), np.concatenate(xs), np.concatenate(ys), np.concatenate(zs)


def analytic_tree(base_level=6, half=0.5, max_level=None, refine=None, length=1.0e25, seed=0,
                  holes=None):
    """A small AMR tree with analytic fields, for tests.  holes: optional
    callable (level, cx, cy, cz) -> bool mask of cells to drop (and with them
    every descendant, so the tree stays parent-closed).  Returns (records,
    fields64 in record order)."""
    if max_level is None:
        max_level = base_level
    if refine is None:
        refine = lambda l, x, y, z: np.zeros(x.shape, dtype=bool)
    level, kx, ky, kz = amr_coords(base_level, half, refine, max_level)
    cx, cy, cz = formats.cell_centers(level, kx, ky, kz)
    if holes is not None:
        keep = ~holes(level, cx, cy, cz)
        level, kx, ky, kz, cx, cy, cz = (v[keep] for v in (level, kx, ky, kz, cx, cy, cz))
    f = analytic_fields(cx, cy, cz, length, seed=seed)
    rec, order = formats.make_records(level, kx, ky, kz, f)
    return rec, f[order]
