"""HEALPix RING scheme in numpy, for pinning the library's implementation.

TEST INFRASTRUCTURE ONLY.  The reference links HEALPix C++ 3.70 / chealpix
(hmaps.h:58-60, 599; hmaps.cpp:185), which is absent here: "parity
unpinned" against that library.  This restatement follows a route different
from the library's (libmpf's) one: a point is first placed on a base face
with in-face coordinates (x, y), and the RING index is assembled from the
face's ring and phi offsets (Gorski et al. 2005, section 4 and the face
numbering of their figure 4), instead of the direct edge-line arithmetic.
Pixel centres come from the ring structure written out pixel by pixel.
"""
import numpy as np

JRLL = np.array([2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4])
JPLL = np.array([1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7])


def ring_centres(nside):
    """(z, phi) of every RING pixel, ring by ring: 4i pixels on cap ring i,
    4 nside on equatorial rings, phi offset by half a pixel on odd rings."""
    z = []
    phi = []
    for i in range(1, 4 * nside):
        if i < nside:
            npr = 4 * i
            zi = 1.0 - i * i / (3.0 * nside * nside)
            ph = (np.arange(npr) + 0.5) * (np.pi / 2) / i
        elif i <= 3 * nside:
            npr = 4 * nside
            zi = (2 * nside - i) * 2.0 / (3.0 * nside)
            shift = 0.5 if ((i - nside) % 2 == 0) else 1.0
            ph = (np.arange(npr) + 1 - shift) * np.pi / (2 * nside)
        else:
            k = 4 * nside - i
            npr = 4 * k
            zi = -(1.0 - k * k / (3.0 * nside * nside))
            ph = (np.arange(npr) + 0.5) * (np.pi / 2) / k
        z.append(np.full(npr, zi))
        phi.append(ph)
    return np.concatenate(z), np.concatenate(phi)


def zphi2xyf(nside, z, phi):
    """Base face and in-face coordinates of directions (nside a power of 2)."""
    za = np.abs(z)
    tt = np.mod(phi, 2 * np.pi) * (2 / np.pi)
    tt = np.where(tt >= 4.0, 0.0, tt)
    ix = np.zeros(z.shape, dtype=np.int64)
    iy = np.zeros(z.shape, dtype=np.int64)
    face = np.zeros(z.shape, dtype=np.int64)
    eq = za <= 2.0 / 3.0
    t1 = nside * (0.5 + tt[eq])
    t2 = nside * (z[eq] * 0.75)
    jp = (t1 - t2).astype(np.int64)
    jm = (t1 + t2).astype(np.int64)
    ifp = jp // nside
    ifm = jm // nside
    face[eq] = np.where(ifp == ifm, ifp | 4, np.where(ifp < ifm, ifp, ifm + 8))
    ix[eq] = jm & (nside - 1)
    iy[eq] = nside - (jp & (nside - 1)) - 1
    pc = ~eq
    ntt = np.minimum(3, tt[pc].astype(np.int64))
    tp = tt[pc] - ntt
    tmp = nside * np.sqrt(3 * (1 - za[pc]))
    jp = np.minimum(nside - 1, (tp * tmp).astype(np.int64))
    jm = np.minimum(nside - 1, ((1 - tp) * tmp).astype(np.int64))
    north = z[pc] >= 0
    face[pc] = np.where(north, ntt, ntt + 8)
    ix[pc] = np.where(north, nside - jm - 1, jp)
    iy[pc] = np.where(north, nside - jp - 1, jm)
    return ix, iy, face


def xyf2ring(nside, ix, iy, face):
    nl4 = 4 * nside
    jr = JRLL[face] * nside - ix - iy - 1
    npix = 12 * nside * nside
    ncap = 2 * nside * (nside - 1)
    nr = np.where(jr < nside, jr, np.where(jr > 3 * nside, nl4 - jr, nside))
    n_before = np.where(jr < nside, 2 * jr * (jr - 1),
                        np.where(jr > 3 * nside, npix - 2 * (nl4 - jr + 1) * (nl4 - jr), ncap + (jr - nside) * nl4))
    kshift = np.where((jr < nside) | (jr > 3 * nside), 0, (jr - nside) & 1)
    jp = (JPLL[face] * nr + ix - iy + 1 + kshift) // 2
    jp = np.where(jp > nl4, jp - nl4, jp)
    jp = np.where(jp < 1, jp + nl4, jp)
    return n_before + jp - 1


def vec2pix_ring(nside, v):
    v = np.asarray(v, dtype=np.float64).reshape(-1, 3)
    r = np.linalg.norm(v, axis=1)
    ix, iy, f = zphi2xyf(nside, v[:, 2] / r, np.arctan2(v[:, 1], v[:, 0]))
    return xyf2ring(nside, ix, iy, f)
