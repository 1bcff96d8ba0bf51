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


def ramses_evolution_files(directory, h=0.7, omega_m=0.3, n=2000, a_min=1e-3, box_mpch=2625.5, bad_rows=()):
    """Writes a RAMSES-style evolution file (rows: a, H/H0 scaled as read by
    the reference, tau H0, t_lookback H0, t_proper H0), a parameter file and
    returns (evolution path, parameter path, cell directory name) for a flat
    LCDM model.  Columns 1 and 4 are what the reference reads: da/deta / H0 =
    a^2 E(a) and eta H0 (conformal time from today, negative in the past);
    rows listed in bad_rows repeat the previous a (the reference drops them)."""
    import os
    a = np.linspace(a_min, 1.0, n)
    E = np.sqrt(omega_m / a ** 3 + 1.0 - omega_m)
    # eta(a) H0 = -int_a^1 da' / (a'^2 E(a')), on a fine grid
    fine = np.linspace(a_min, 1.0, 200 * n)
    Ef = np.sqrt(omega_m / fine ** 3 + 1.0 - omega_m)
    f = 1.0 / (fine ** 2 * Ef)
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(fine))])
    eta = np.interp(a, fine, cum - cum[-1])
    g = 1.0 / (fine * Ef)
    cumt = np.concatenate([[0.0], np.cumsum(0.5 * (g[1:] + g[:-1]) * np.diff(fine))])
    lookback = np.interp(a, fine, cumt[-1] - cumt)
    rows = np.stack([a, a * a * E, eta, lookback, eta], 1)
    for i in bad_rows:
        rows[i, 0] = rows[i - 1, 0]
    os.makedirs(directory, exist_ok=True)
    evol = os.path.join(directory, "ramses_input_00001.dat")
    with open(evol, "w") as fh:
        for r in rows:
            fh.write(" ".join(f"{v:.17e}" for v in r) + "\n")
    par = os.path.join(directory, "param_00001.txt")
    with open(par, "w") as fh:
        fh.write(f"h={h:.8f}   \nOmega_m={omega_m:.8f}\nOmega_b=0.043557099\nn_s=0.96300000\nw=-1.0000000\n")
    cell = os.path.join(directory, "boxlen" + f"{box_mpch:g}".replace(".", "-") + "_n1024_lcdmw7")
    return evol, par, cell


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


def narrow_field_frame(phi0_deg, theta0_deg):
    """Orthonormal frame (rows e_x, e_y, e_z) of a narrow field whose local
    +x axis points at (phi0, theta0): e_y along e_phi, e_z along -e_theta."""
    p, t = math.radians(phi0_deg), math.radians(theta0_deg)
    return np.array([[math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)],
                     [-math.sin(p), math.cos(p), 0.0],
                     [-math.cos(t) * math.cos(p), -math.cos(t) * math.sin(p), math.sin(t)]])


def _field_to_world(phi0_deg, theta0_deg, lphi, ltheta):
    local = np.stack([np.sin(ltheta) * np.cos(lphi), np.sin(ltheta) * np.sin(lphi), np.cos(ltheta)], 1)
    d = local @ narrow_field_frame(phi0_deg, theta0_deg)
    return np.arctan2(d[:, 1], d[:, 0]), np.arccos(np.clip(d[:, 2], -1.0, 1.0))


def narrow_field_area_deg2(aperture_deg):
    """Solid angle (deg^2) of a narrow field with semi-angles
    thetay = thetaz = aperture / 2: 4 thetay sin(thetaz), the sky portion of
    Miscellaneous::GenerateNarrowCones (miscellaneous.h:332)."""
    h = math.radians(aperture_deg) / 2
    return 4 * h * math.sin(h) * (180 / math.pi) ** 2


def narrow_field_grid(phi0_deg, theta0_deg, aperture_deg, nside):
    """Regular nside x nside grid over the narrow field of the reference's
    narrow cones (miscellaneous.h:338-343: |phi| < thetay and
    |theta - pi/2| < thetaz in the field's own frame, around its +x axis),
    here with thetay = thetaz = aperture / 2 and the axis at (phi0, theta0):
    an aperture x aperture field of narrow_field_area_deg2(aperture) deg^2
    (99.9 for 10 deg).  Cell-centred local angles; returns world (phi, theta)."""
    half = math.radians(aperture_deg) / 2
    u = (np.arange(nside) + 0.5) / nside * 2 * half - half
    lp, lt = np.meshgrid(u, math.pi / 2 + u, indexing="ij")
    return _field_to_world(phi0_deg, theta0_deg, lp.ravel(), lt.ravel())


def narrow_field_uniform(phi0_deg, theta0_deg, aperture_deg, n, seed):
    """n directions uniform in solid angle over the narrow field of
    narrow_field_grid: local phi uniform, local cos(theta) uniform."""
    rng = np.random.default_rng(seed)
    half = math.radians(aperture_deg) / 2
    lp = rng.uniform(-half, half, n)
    lt = np.arccos(rng.uniform(-math.sin(half), math.sin(half), n))
    return _field_to_world(phi0_deg, theta0_deg, lp, lt)


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


def _parents_kept(level, kx, ky, kz, keep):
    """Mask of the cells whose every stored ancestor is kept (a dropped cell
    takes its descendants with it, so the tree stays parent-closed); cells
    whose centre is outside a hole may keep only some of their children."""
    ok = keep.copy()
    lmin = int(level.min())

    def code(l, x, y, z):
        return (np.asarray(x, dtype=np.uint64) << np.uint64(42)) | (np.asarray(y, dtype=np.uint64) << np.uint64(21)) | \
            np.asarray(z, dtype=np.uint64)
    for l in range(lmin + 1, int(level.max()) + 1):
        par = level == l - 1
        kept_par = np.sort(code(l - 1, kx[par & ok], ky[par & ok], kz[par & ok]))
        cur = np.nonzero(level == l)[0]
        c = code(l - 1, kx[cur] >> np.uint64(1), ky[cur] >> np.uint64(1), kz[cur] >> np.uint64(1))
        pos = np.searchsorted(kept_par, c)
        found = (pos < kept_par.size) & (kept_par[np.minimum(pos, kept_par.size - 1)] == c)
        ok[cur] &= found
    return ok


def analytic_tree(base_level=6, half=0.5, max_level=None, refine=None, length=1.0e25, seed=0,
                  holes=None, amp=1.0e-5):
    """A small AMR tree with analytic fields, for tests (amp: amplitude of
    phi / c^2).  holes: optional
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
        keep &= _parents_kept(level, kx, ky, kz, keep)
        level, kx, ky, kz, cx, cy, cz = (v[keep] for v in (level, kx, ky, kz, cx, cy, cz))
    f = analytic_fields(cx, cy, cz, length, amp=amp, seed=seed)
    rec, order = formats.make_records(level, kx, ky, kz, f)
    return rec, f[order]


# ------------------------------------------------------------ particles

def record_cells(rec):
    """(level, kx, ky, kz) of cone records from their 128-bit keys."""
    level = np.zeros(rec.size, dtype=np.int64)
    k = [np.zeros(rec.size, dtype=np.uint64) for _ in range(3)]
    for i in range(32):
        word = rec["hi"] if i < 16 else rec["lo"]
        nib = (word >> np.uint64(60 - 4 * (i % 16))) & np.uint64(0xF)
        present = (nib & np.uint64(8)) != 0
        level += present
        for d in range(3):
            bit = (nib >> np.uint64(d)) & np.uint64(1)
            k[d] = np.where(present, (k[d] << np.uint64(1)) | bit, k[d])
    return level, k[0], k[1], k[2]


def particles_in_cells(rec, n, seed, bulk=(5e-4, -2.7e-4, 1.3e-4), sigma=7e-4):
    """n particles (float32 positions in box units and velocities in units
    of c) placed uniformly inside randomly chosen cells of a cone octree, so
    that they crowd where the tree refines, with a bulk flow plus a Gaussian
    dispersion per axis."""
    rng = np.random.default_rng(seed)
    level, kx, ky, kz = record_cells(rec)
    pick = rng.integers(0, rec.size, n)
    size = 4.0 / np.exp2(level[pick].astype(np.float64))
    pos = np.empty((n, 3), dtype=np.float32)
    for d, kd in enumerate((kx, ky, kz)):
        pos[:, d] = -2.0 + (kd[pick].astype(np.float64) + rng.uniform(0.0, 1.0, n)) * size
    vel = (np.asarray(bulk) + rng.normal(0.0, sigma, (n, 3))).astype(np.float32)
    return pos, vel


# ----------------------------------------------------- benchmark workloads

def _part1by3(v):
    """Spread the low 16 bits of v to every 4th bit (bit b -> bit 4b)."""
    v = v.astype(np.uint64) & np.uint64(0xFFFF)
    v = (v | (v << np.uint64(24))) & np.uint64(0x000000FF000000FF)
    v = (v | (v << np.uint64(12))) & np.uint64(0x000F000F000F000F)
    v = (v | (v << np.uint64(6))) & np.uint64(0x0303030303030303)
    v = (v | (v << np.uint64(3))) & np.uint64(0x1111111111111111)
    return v


def keys_hi_fast(level, kx, ky, kz):
    """High key word for levels <= 16 (the low word is then zero)."""
    level = np.asarray(level, dtype=np.uint64)
    code = _part1by3(kx) | (_part1by3(ky) << np.uint64(1)) | (_part1by3(kz) << np.uint64(2))
    marks = np.uint64(0x8888888888888888) & ((np.uint64(1) << (np.uint64(4) * level)) - np.uint64(1))
    code = code | marks
    return code << (np.uint64(64) - np.uint64(4) * level)


def _trilinear(grid, fx, fy, fz):
    """Periodic trilinear interpolation of grid (n^3) at fractional grid
    coordinates (cell centres at integer + 0.5)."""
    n = grid.shape[0]
    x = fx - 0.5
    y = fy - 0.5
    z = fz - 0.5
    x0 = np.floor(x).astype(np.int64)
    y0 = np.floor(y).astype(np.int64)
    z0 = np.floor(z).astype(np.int64)
    tx = x - x0
    ty = y - y0
    tz = z - z0
    out = np.zeros(fx.shape)
    for dx in (0, 1):
        wx = tx if dx else 1 - tx
        ix = (x0 + dx) % n
        for dy in (0, 1):
            wy = ty if dy else 1 - ty
            iy = (y0 + dy) % n
            for dz in (0, 1):
                wz = tz if dz else 1 - tz
                iz = (z0 + dz) % n
                out += wx * wy * wz * grid[ix, iy, iz]
    return out


def _hash_uniform(kx, ky, kz, lvl, seed):
    """Deterministic uniform noise in [-sqrt(3), sqrt(3)) (unit variance) from
    cell coordinates: the same on every device."""
    import torch
    h = (kx * 73856093) ^ (ky * 19349663) ^ (kz * 83492791) ^ (lvl * 2654435761 + seed * 97)
    h = h & 0xFFFFFFFF
    h = (h ^ (h >> 16)) * 0x45d9f3b & 0xFFFFFFFF
    h = (h ^ (h >> 16)) * 0x45d9f3b & 0xFFFFFFFF
    h = h ^ (h >> 16)
    u = h.to(torch.float64) / 4294967296.0
    return (2.0 * u - 1.0) * math.sqrt(3.0)


def _trilinear_t(grid, fx, fy, fz):
    """Periodic trilinear interpolation (torch) at fractional grid coordinates
    (cell centres at integer + 0.5)."""
    import torch
    n = grid.shape[0]
    x = fx - 0.5
    y = fy - 0.5
    z = fz - 0.5
    x0 = torch.floor(x)
    y0 = torch.floor(y)
    z0 = torch.floor(z)
    tx, ty, tz = x - x0, y - y0, z - z0
    x0 = x0.long(); y0 = y0.long(); z0 = z0.long()
    flat = grid.reshape(-1)
    out = torch.zeros_like(fx)
    for dx in (0, 1):
        wx = tx if dx else 1 - tx
        ix = (x0 + dx) % n
        for dy in (0, 1):
            wxy = wx * (ty if dy else 1 - ty)
            iy = (y0 + dy) % n
            for dz in (0, 1):
                w = wxy * (tz if dz else 1 - tz)
                iz = (z0 + dz) % n
                out += w * flat[(ix * n + iy) * n + iz]
    return out


def _part1by3_t(v):
    v = v & 0xFFFF
    v = (v | (v << 24)) & 0x000000FF000000FF
    v = (v | (v << 12)) & 0x000F000F000F000F
    v = (v | (v << 6)) & 0x0303030303030303
    v = (v | (v << 3)) & 0x1111111111111111
    return v


def lightcone_tree(base_input_level, box_mpch, sigma, threshold, max_input_level, seed, h=0.7, omega_m=0.3,
                   rms=1.0e-5, noise=0.0, decay=0.0, table=None, device=None, verbose=False):
    """Synthetic lightcone octree of the benchmark configurations.

    phi: Gaussian random field with P(k) ~ k^-3 on the base grid, rms phi/c^2 =
    rms; grad phi by 2nd-order periodic central differences; density lognormal
    from the same field, rho = exp(sigma g - sigma^2/2) with g the unit-variance
    field; a(cell) = a(chi = |x| L) along the background null ray.  Base cells
    with rho - 1 > threshold are split, and each child again while its own
    overdensity (trilinear log-density plus `noise` x unit-variance noise, less
    `decay` per level below the base) exceeds the threshold, down to max_input_level.  Parents are kept.
    Refined cells take phi and grad phi trilinearly interpolated from the base
    grid.  Runs on `device` (torch; default: the GPU when present).
    Returns (records, fields64 in record order, length in metres)."""
    import time

    import torch
    t0 = time.time()
    if device is None:
        device = "cuda" if torch.cuda.is_available() else "cpu"
    c2 = C_LIGHT ** 2
    mpc = MPC_SI
    length = box_mpch * mpc / h
    n = 1 << base_input_level
    g = torch.from_numpy(grf_potential(n, seed, rms=1.0)).to(device)
    phi_grid = g * (rms * c2)
    dxm = length / n
    gx, gy, gz = ((torch.roll(phi_grid, -1, ax) - torch.roll(phi_grid, 1, ax)) / (2 * dxm) for ax in range(3))
    lg = sigma * g - 0.5 * sigma * sigma   # log density
    if table is None:
        table = flat_lcdm_table(h=h, omega_m=omega_m, a_min=0.05, n=4096)
    chi_t = torch.from_numpy(np.ascontiguousarray(C_LIGHT * table[0])).to(device)
    a_t = torch.from_numpy(np.ascontiguousarray(table[1])).to(device)
    H0 = h * 1.0e5 / mpc
    lbase = base_input_level + 2
    off = (1 << lbase) * 3 // 8   # the base grid spans [-0.5, 0.5) of the octree's [-2, 2)
    levels, kxs, kys, kzs, fields = [], [], [], [], []

    def emit(lvl, kx, ky, kz, phi, fx, fy, fz, logrho):
        den = float(1 << (lvl + 1))
        cx = -2.0 + 4.0 * ((2 * kx.double() + 1) / den)
        cy = -2.0 + 4.0 * ((2 * ky.double() + 1) / den)
        cz = -2.0 + 4.0 * ((2 * kz.double() + 1) / den)
        r = torch.sqrt(cx * cx + cy * cy + cz * cz) * length
        idx = torch.searchsorted(chi_t, r).clamp(1, chi_t.numel() - 1)
        w = (r - chi_t[idx - 1]) / (chi_t[idx] - chi_t[idx - 1])
        a = a_t[idx - 1] + w * (a_t[idx] - a_t[idx - 1])
        f = torch.stack([torch.exp(logrho), phi, fx, fy, fz, a, -(1.0 - a) * H0 * phi], dim=1)
        levels.append(torch.full_like(kx, lvl))
        kxs.append(kx); kys.append(ky); kzs.append(kz); fields.append(f)

    ax = torch.arange(n, device=device, dtype=torch.int64) + off
    bx, by, bz = (v.reshape(-1) for v in torch.meshgrid(ax, ax, ax, indexing="ij"))
    emit(lbase, bx, by, bz, phi_grid.reshape(-1), gx.reshape(-1), gy.reshape(-1), gz.reshape(-1), lg.reshape(-1))
    sel = torch.nonzero(torch.exp(lg.reshape(-1)) - 1.0 > threshold).squeeze(1)
    px, py, pz = bx[sel], by[sel], bz[sel]
    for il in range(base_input_level + 1, max_input_level + 1):
        lvl = il + 2
        s = torch.arange(8, device=device)
        kx = (2 * px[:, None] + (s & 1)[None, :]).reshape(-1)
        ky = (2 * py[:, None] + ((s >> 1) & 1)[None, :]).reshape(-1)
        kz = (2 * pz[:, None] + ((s >> 2) & 1)[None, :]).reshape(-1)
        scale = float(1 << (lvl - lbase))
        fx = (kx.double() + 0.5) / scale - off
        fy = (ky.double() + 0.5) / scale - off
        fz = (kz.double() + 0.5) / scale - off
        phi = _trilinear_t(phi_grid, fx, fy, fz)
        gxi = _trilinear_t(gx, fx, fy, fz)
        gyi = _trilinear_t(gy, fx, fy, fz)
        gzi = _trilinear_t(gz, fx, fy, fz)
        lr = _trilinear_t(lg, fx, fy, fz) + noise * _hash_uniform(kx, ky, kz, lvl, seed)
        lsel = lr - decay * (il - base_input_level)
        del fx, fy, fz
        emit(lvl, kx, ky, kz, phi, gxi, gyi, gzi, lr)
        sel = torch.nonzero(torch.exp(lsel) - 1.0 > threshold).squeeze(1)
        px, py, pz = kx[sel], ky[sel], kz[sel]
        if verbose:
            print(f"  level {il}: {kx.numel()} cells, {sel.numel()} refined ({time.time() - t0:.1f}s)", flush=True)
    level = torch.cat(levels)
    kx = torch.cat(kxs); ky = torch.cat(kys); kz = torch.cat(kzs)
    f = torch.cat(fields)
    del levels, kxs, kys, kzs, fields
    if int(level.max()) > 16:
        raise ValueError("lightcone_tree builds trees up to octree level 16")
    code = _part1by3_t(kx) | (_part1by3_t(ky) << 1) | (_part1by3_t(kz) << 2)
    # key = (interleaved bits | marker bit of every nibble) << (64 - 4 level)
    allm = torch.tensor(np.uint64(0x8888888888888888).view(np.int64), device=device)
    nib = 4 * level
    mask = torch.where(level >= 16, torch.full_like(level, -1), (torch.ones_like(level) << nib.clamp(max=63)) - 1)
    key = torch.bitwise_left_shift(code | (allm & mask), (64 - nib).clamp(min=0, max=63))
    # unsigned order of keys whose top bit is set: flip it and sort as signed
    key_signed = key ^ torch.tensor(-0x8000000000000000, device=device, dtype=torch.int64)
    order = torch.argsort(key_signed, stable=True)
    key = key[order]
    f = f[order]
    rec = np.zeros(key.numel(), dtype=formats.RECORD_DTYPE)
    rec["hi"] = key.cpu().numpy().view(np.uint64)
    f64 = f.cpu().numpy()
    for j, name in enumerate(formats.FIELDS7):
        rec[name] = f64[:, j].astype(np.float32)
    if verbose:
        print(f"  tree: {rec.size} cells in {time.time() - t0:.1f}s", flush=True)
    return rec, f64, length


WORKLOADS = {
    # uniform level-7 grid, CIC, 64x64 rays over a 5x5 deg narrow field to z = 0.5
    "uniform128_cic": dict(base_input_level=7, box_mpch=3000.0, sigma=1.0, threshold=1e30, max_input_level=7, seed=42,
                 order=1, nside=64, aperture=5.0, z_stop=0.5, phi0=20.0, theta0=20.0),
    # level-8 base refined to level 11 where the overdensity exceeds 8, TSC,
    # 1024x1024 rays over a 10x10 deg narrow field to z = 2
    "amr256_tsc": dict(base_input_level=8, box_mpch=8000.0, sigma=1.5, threshold=8.0, max_input_level=11, seed=42,
                 noise=1.0, decay=0.75, order=2, nside=1024, aperture=10.0, z_stop=2.0, phi0=20.0, theta0=20.0),
    # level-9 base (512^3) refined to level 13 where the sigma-2 overdensity
    # exceeds 8 (decay 0.67 per level below the base: about 600 M cells),
    # 4096x4096 rays over a 20x20 deg narrow field to z = 3
    "cone600m_tsc": dict(base_input_level=9, box_mpch=10000.0, sigma=2.0, threshold=8.0, max_input_level=13,
                         seed=42, noise=1.0, decay=0.67, order=2, nside=4096, aperture=20.0, z_stop=3.0,
                         phi0=20.0, theta0=20.0),
}


def build_workload(name, log=None, device="cpu"):
    """The named workload of WORKLOADS: (parameters, octree records, FP64
    fields, box length in metres, cosmology table, launch records n x 9 of
    the rays on the regular grid over the narrow field).  device: where
    torch generates the tree (the records come back to the host)."""
    import time
    w = dict(WORKLOADS[name])
    t0 = time.time()
    table = flat_lcdm_table(h=0.7, omega_m=0.3, a_min=0.05, n=4096)
    rec, f64, length = lightcone_tree(w["base_input_level"], w["box_mpch"], w["sigma"], w["threshold"],
                                      w["max_input_level"], w["seed"], noise=w.get("noise", 0.0),
                                      decay=w.get("decay", 0.0), table=table, device=device,
                                      verbose=log is not None and device != "cpu")
    ph, th = narrow_field_grid(w["phi0"], w["theta0"], w["aperture"], w["nside"])
    in9 = launch_photons(ph, th)
    if log:
        log(f"workload {name}: {rec.size} cells, {in9.shape[0]} rays, built in {time.time() - t0:.1f}s")
    return w, rec, f64, length, table, in9

