"""The cosmology table of the reference, restated on the host
(Input::acquire input.h:1091-1233, Input::correct input.h:1330-1378).

The tables are bit-identical to the reference's own acquire/correct
(input.h compiled in oracle/_ref, tests/test_ref_catalogue.py); this file
adds what the mathematics fixes: a flat LCDM evolution file gives back the
conformal lookback time, da/dtau and d2a/dtau2 of the model; the parsed h,
Omega_m and box length; dropped rows; and correct() is the identity when the
homogeneous photon reproduces the table, and the interpolant of the photon
inside its time range otherwise.  The reference photon itself (integrate<-1>,
integrator.h) is pinned bit for bit against the reference on the GPU."""
import numpy as np
import pytest

from mpf import capi, synth
from oracle import Oracle

MPC = 3.08568e22
H, OM = 0.7, 0.3


def model(a):
    H0 = H * 1.0e5 / MPC
    E = np.sqrt(OM / a ** 3 + 1 - OM)
    fine = np.linspace(a.min() * 0.999, 1.0, 400000)
    Ef = np.sqrt(OM / fine ** 3 + 1 - OM)
    f = 1.0 / (fine ** 2 * Ef * H0)
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(fine))])
    tau = np.interp(a, fine, cum[-1] - cum)           # conformal lookback time (s)
    dadtau = -a * a * E * H0
    # d2a/dtau2 = d(a^2 E H0)/da * a^2 E H0
    dEda = -1.5 * OM / a ** 4 / E
    d2 = (2 * a * E + a * a * dEda) * H0 * a * a * E * H0
    return tau, dadtau, d2


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    return synth.ramses_evolution_files(str(tmp_path_factory.mktemp("cosmo")), h=H, omega_m=OM, n=2000,
                                        bad_rows=(700, 1301))


def test_parameters_and_shape(files):
    tab, h, om, lb = capi.cosmology_acquire(*files, mpc=MPC)
    assert (h, om, lb) == (0.7, 0.3, 2625.5)
    t, a = tab[0], tab[1]
    assert tab.shape == (4, 1998)            # the two repeated-a rows are dropped
    assert t[0] == 0.0 and a[0] == 1.0       # today first, t from the observer
    assert np.all(np.diff(t) > 0) and np.all(np.diff(a) < 0)


def test_table_matches_flat_lcdm(files):
    tab, *_ = capi.cosmology_acquire(*files, mpc=MPC)
    t, a, dadt, d2 = tab
    tau, dadtau, d2m = model(a)
    sel = a > 0.05
    # Euler integration of dtau = da / (da/dtau) on 2000 rows and the kernel smoother
    assert np.max(np.abs(t[sel] - tau[sel]) / tau[sel].max()) < 2e-3
    assert np.max(np.abs(dadt[sel] / dadtau[sel] - 1)) < 2e-3
    inner = (a > 0.1) & (a < 0.95)
    assert np.max(np.abs(d2[inner] / d2m[inner] - 1)) < 3e-2


def test_short_file_uses_the_whole_window(tmp_path):
    """Under 500 rows the smoother's window is the whole table (size/500 = 0)."""
    files = synth.ramses_evolution_files(str(tmp_path), h=H, omega_m=OM, n=400)
    tab, *_ = capi.cosmology_acquire(*files, mpc=MPC)
    t, a, dadt, _ = tab
    tau, dadtau, _ = model(a)
    sel = a > 0.1
    # 400 rows: the first-order steps of t (da / (da/dt) per row) cost ~1 %
    assert np.max(np.abs(dadt[sel] / dadtau[sel] - 1)) < 2.5e-2
    assert np.max(np.abs(t[sel] - tau[sel]) / tau[sel].max()) < 1e-2


def test_missing_inputs(tmp_path):
    evol, par, cell = synth.ramses_evolution_files(str(tmp_path), n=600)
    with pytest.raises(capi.MpfError):
        capi.cosmology_acquire(evol, str(tmp_path / "nope.txt"), cell)
    with pytest.raises(capi.MpfError):
        capi.cosmology_acquire(evol, par, "cells_without_length")
    with pytest.raises(capi.MpfError):
        capi.cosmology_acquire(str(tmp_path / "missing.dat"), par, cell)


def test_correct_identity_and_interpolant(files):
    tab, *_ = capi.cosmology_acquire(*files, mpc=MPC)
    t, a = tab[0], tab[1]
    # the photon reproduces the table at its own times: nothing changes
    np.testing.assert_array_equal(capi.cosmology_correct(t, a, t, a), a)
    # a shorter photon, with a(t) perturbed: inside its range a is its
    # interpolant (Utility::interpolate), outside smoothed toward it
    m = 600
    tt = np.linspace(t[0], t[1200], m)
    ta = np.interp(tt, t, a) * (1 + 1e-3 * np.sin(tt / tt[-1] * 7))
    out = capi.cosmology_correct(t, a, tt, ta)
    inside = (t >= tt[0]) & (t <= tt[-1])
    O = Oracle()
    np.testing.assert_array_equal(out[inside], O.interpolate(tt, ta, t[inside]))
    far = t > tt[-1] * 1.5
    np.testing.assert_allclose(out[far], a[far], rtol=2e-3)
    assert np.all(np.isfinite(out))
