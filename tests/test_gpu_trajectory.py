"""Whole trajectories (the reference's Evolution<Photon>) from the exact
tiers, against the reference's own Integrator::integrate (oracle/_ref, via the
oracle when the reference build is absent): bit-identical elements, for TSC
on an AMR tree and for the homogeneous order (Order -1, integrator.h:505,
573) of the cosmology's reference photon; and the cosmology preparation of
rays.cpp:200-215 (acquire, that photon on a homogeneous level-5 octree,
correct) against its parts."""
import math

import numpy as np
import pytest

import cases
from mpf import capi, formats, synth
from oracle import Oracle, RefHarness, ref_available

pytestmark = pytest.mark.gpu
MPC = 3.08568e22


def checker():
    return RefHarness() if ref_available() else Oracle()


def homogeneous_level5():
    lvl, kx, ky, kz = synth.uniform_level_coords(3, half=2.0)   # input level 3 -> octree level 5, whole box
    f = np.zeros((kx.size, 7))
    f[:, 5] = 1.0                                              # homogenize: a = 1
    rec, _ = formats.make_records(lvl, kx, ky, kz, f)
    return rec


@pytest.mark.parametrize("precision", (capi.PREC_REF, capi.PREC_FP64))
def test_tsc_trajectories(precision):
    print(f"[case] trajectories tsc precision={precision}", flush=True)
    rec, f64 = cases.tree("amr")
    c = capi.Context(cases.LENGTH, order=capi.TSC, precision=precision, nsteps=cases.NSTEPS)
    c.load_records(rec)
    c.load_fields64(f64)
    tab = cases.cosmology()
    c.set_cosmology(tab)
    in9 = cases.rays(n_side=4, aperture=6.0)
    out, trajs = c.trajectory(in9, "redshift", 0.3, capacity=4000)
    R = Oracle()
    h = R.tree(rec, f64)
    for i in range(len(in9)):
        ref = R.trajectory(h, 2, tab, cases.LENGTH, cases.NSTEPS, 2, 0.3, in9[i], precision=precision, maxlen=4000)
        np.testing.assert_array_equal(trajs[i], ref, err_msg=f"ray {i}")
        np.testing.assert_array_equal(out[i, :27], trajs[i][-1, :27])
    R.free(h)
    if precision == capi.PREC_REF and ref_available():
        H = RefHarness()
        hh = H.tree(rec)
        for i in (0, 5, 11):
            ref = H.trajectory(hh, 2, tab, cases.LENGTH, cases.NSTEPS, 2, 0.3, in9[i], maxlen=4000)
            np.testing.assert_array_equal(trajs[i], ref, err_msg=f"ray {i} vs reference")
        H.free(hh)
    c.close()


def test_homogeneous_reference_photon():
    print("[case] trajectories homogeneous", flush=True)
    rec = homogeneous_level5()
    tab = cases.cosmology()
    nsteps = int(4 * 2 ** (int(math.log2(tab.shape[1] / 32 + 1) + 1) + 1))
    c = capi.Context(cases.LENGTH, order=capi.HOMOGENEOUS, precision=capi.PREC_REF, nsteps=nsteps)
    c.load_records(rec)
    c.set_cosmology(tab)
    in9 = np.array([[1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0]])
    out, trajs = c.trajectory(in9, "none", capacity=200000)
    R = checker()
    h = R.tree(rec)
    ref = R.trajectory(h, -1, tab, cases.LENGTH, nsteps, -1, 0.0, in9[0], maxlen=200000)
    R.free(h)
    assert len(ref) > 1000
    np.testing.assert_array_equal(trajs[0], ref)
    # phi = 0: the redshift is a0/a - 1 along the whole photon
    np.testing.assert_allclose(trajs[0][:, capi.OUT["redshift"]], 1 / trajs[0][:, capi.OUT["a"]] - 1, rtol=1e-12)
    c.close()
    with pytest.raises(capi.MpfError):
        cs = capi.Context(cases.LENGTH, order=capi.HOMOGENEOUS, precision=capi.PREC_SPLIT, nsteps=nsteps)
        cs.load_records(rec)
        cs.set_cosmology(tab)
        cs.integrate(in9, "none")


def test_cosmology_prepare(tmp_path):
    """prepare = acquire, then the homogeneous photon (pinned above against
    the reference), then correct with that photon's (t, a)."""
    print("[case] cosmology prepare", flush=True)
    files = synth.ramses_evolution_files(str(tmp_path), n=1000)
    tab, h, om, lb = capi.cosmology_acquire(*files, mpc=MPC)
    prep, h2, om2, lb2 = capi.cosmology_prepare(*files, mpc=MPC)
    assert (h2, om2, lb2) == (h, om, lb)
    length = lb * MPC / h
    nsteps = int(4 * 2 ** (int(math.log2(tab.shape[1] / 32 + 1) + 1) + 1))
    R = checker()
    hh = R.tree(homogeneous_level5())
    in9 = np.array([1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0])
    ref = R.trajectory(hh, -1, tab, length, nsteps, -1, 0.0, in9, maxlen=1 << 20)
    R.free(hh)
    a = capi.cosmology_correct(tab[0], tab[1], ref[:, capi.OUT["t"]], ref[:, capi.OUT["a"]])
    np.testing.assert_array_equal(prep[0], tab[0])
    np.testing.assert_array_equal(prep[1], a)
    np.testing.assert_array_equal(prep[2], tab[2])
    np.testing.assert_array_equal(prep[3], tab[3])
    assert np.any(prep[1] != tab[1])      # the correction did something
