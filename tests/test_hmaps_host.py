"""Host parts of the hmaps program (no GPU): the HEALPix FITS writer (read
back with a minimal FITS reader written here from the FITS standard: 2880-
byte blocks of 80-character cards, big-endian IEEE data; the chealpix library
the reference links is absent, so the file layout is pinned by the format
only) and the homogeneous quantities at the map redshifts (hmaps.cpp:297-325)
against closed forms on the homogeneous reference photon of the oracle."""
import math

import numpy as np
import pytest

import cases
from mpf import capi, formats, synth
from oracle import Oracle


def read_fits_table(path):
    raw = open(path, "rb").read()
    hdus, pos = [], 0
    while pos < len(raw):
        cards = {}
        while True:
            block = raw[pos:pos + 2880]
            pos += 2880
            done = False
            for k in range(36):
                c = block[80 * k:80 * k + 80].decode("ascii")
                key = c[:8].strip()
                if key == "END":
                    done = True
                    break
                if c[8:10] == "= ":
                    v = c[10:].split("/")[0].strip()
                    cards[key] = v.strip("'").strip() if v.startswith("'") else v
            if done:
                break
        nbytes = 0
        if int(cards.get("NAXIS", 0)) > 0:
            nbytes = int(cards["NAXIS1"]) * int(cards["NAXIS2"])
        data = raw[pos:pos + nbytes]
        pos += (nbytes + 2879) // 2880 * 2880
        hdus.append((cards, data))
    return hdus


def test_fits_round_trip(tmp_path):
    nside = 8
    npix = 12 * nside * nside
    m = np.linspace(-1, 1, npix).astype(np.float32)
    m[3] = 0.0
    m[5] = np.nan
    m[7] = np.float32(1e-40)   # subnormal
    capi.healpix_write_fits(tmp_path / "m.fits", m, nside, coordsys="G", clean=True)
    hdus = read_fits_table(tmp_path / "m.fits")
    assert len(hdus) == 2
    assert hdus[0][0]["SIMPLE"] == "T" and hdus[0][0]["NAXIS"] == "0"
    h, data = hdus[1]
    assert h["XTENSION"] == "BINTABLE" and h["TFORM1"] == "1E" and h["PIXTYPE"] == "HEALPIX"
    assert h["ORDERING"] == "RING" and int(h["NSIDE"]) == nside and int(h["NAXIS2"]) == npix
    assert int(h["FIRSTPIX"]) == 0 and int(h["LASTPIX"]) == npix - 1 and h["COORDSYS"] == "G"
    v = np.frombuffer(data, dtype=">f4")
    expect = m.copy()
    expect[[3, 5, 7]] = np.float32(-1.6375e30)   # not normal -> Healpix_undef
    np.testing.assert_array_equal(v, expect)
    assert (tmp_path / "m.fits").stat().st_size % 2880 == 0
    capi.healpix_write_fits(tmp_path / "n.fits", m, nside, nest=True, clean=False)
    h2, d2 = read_fits_table(tmp_path / "n.fits")[1]
    assert h2["ORDERING"] == "NESTED"
    np.testing.assert_array_equal(np.frombuffer(d2, dtype=">f4")[[3, 7]], m[[3, 7]])
    assert np.isnan(np.frombuffer(d2, dtype=">f4")[5])


def test_homogeneous_quantities_closed_form():
    # the homogeneous reference photon (Order -1) of a flat LCDM table
    lvl, kx, ky, kz = synth.uniform_level_coords(3, half=2.0)
    f = np.zeros((kx.size, 7))
    f[:, 5] = 1.0
    rec, _ = formats.make_records(lvl, kx, ky, kz, f)
    O = Oracle()
    h = O.tree(rec)
    tab = synth.flat_lcdm_table(h=0.7, omega_m=0.3, a_min=0.05, n=2048)
    in9 = np.array([1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0])
    ref = O.trajectory(h, -1, tab, cases.LENGTH, 256, -1, 0.0, in9, maxlen=1 << 20)
    O.free(h)
    z = np.array([0.1, 0.5, 1.0, 1.5])
    homo, iref = capi.hmaps_homogeneous(ref, z, "redshift")
    np.testing.assert_array_equal(homo[:, 3], z)
    np.testing.assert_array_equal(homo[:, 4], 1.0 / (1.0 + z))
    np.testing.assert_array_equal(iref, z)
    chi_m = np.array([synth.comoving_distance_of_z(zz) for zz in z])   # metres
    np.testing.assert_allclose(homo[:, 1] * cases.LENGTH, chi_m, rtol=1e-5)
    # conformal time of a null ray: c t = chi
    np.testing.assert_allclose(homo[:, 0] * synth.C_LIGHT, homo[:, 1] * cases.LENGTH, rtol=1e-6)
    assert np.all(np.diff(homo[:, 2]) > 0)
    for name, col in (("t", 0), ("r", 1), ("lambda", 2), ("a", 4), ("s", 1)):
        np.testing.assert_array_equal(capi.hmaps_homogeneous(ref, z, name)[1], homo[:, col])
    with pytest.raises(capi.MpfError):
        capi.hmaps_homogeneous(ref, [1e3], "redshift")
