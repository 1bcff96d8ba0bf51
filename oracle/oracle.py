"""ctypes bindings of the oracle libraries.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.

  Oracle      -- oracle/libmpforacle.so, the C restatement (always built)
  RefHarness  -- oracle/_ref/libmpfref.so, the reference's own headers
                 instantiated by oracle/ref_harness.cpp (built only where the
                 reference sources exist; travels as a built library)
  RefVelocity -- oracle/_ref/libmpfrefvel.so, the same integrator compiled
                 with VELOCITYFIELD (oracle/ref_harness_vel.cpp)
  RefCatalogue -- oracle/_ref/libmpfrefcat.so, catalogues.h, input.h and
                 miscellaneous.h instantiated by oracle/ref_harness_cat.cpp
                 (needs the HDF5 library it was linked against)
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "libmpforacle.so")
REF_LIB = os.path.join(HERE, "_ref", "libmpfref.so")
REFV_LIB = os.path.join(HERE, "_ref", "libmpfrefvel.so")
REFC_LIB = os.path.join(HERE, "_ref", "libmpfrefcat.so")
HDF5_LIB = "/opt/conda/lib/libhdf5.so.103"
NREC = 28

_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64


def _p(a):
    return _vp(a.ctypes.data) if a is not None else None


class _Base:
    prefix = ""

    def __init__(self, path):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = ctypes.CDLL(path)
        pf = self.prefix
        L = self.lib
        getattr(L, pf + "octree_new").restype = _vp
        getattr(L, pf + "octree_new").argtypes = [_vp, _u64]
        getattr(L, pf + "octree_free").argtypes = [_vp]
        getattr(L, pf + "integrate").restype = _u64
        getattr(L, pf + "trajectory").restype = _u64
        getattr(L, pf + "jacobian_infinitesimal").restype = _u64
        getattr(L, pf + "propagate").restype = _u64
        if pf == "mpfo_":
            L.mpfo_surfaces.restype = _u64
            L.mpfo_catalogue.restype = _u64

    def tree(self, rec):
        rec = np.ascontiguousarray(rec)
        h = getattr(self.lib, self.prefix + "octree_new")(_p(rec), _u64(rec.size))
        return h

    def free(self, h):
        getattr(self.lib, self.prefix + "octree_free")(_vp(h))

    def compute_key(self, level, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((xyz.shape[0], 2), dtype=np.uint64)
        getattr(self.lib, self.prefix + "compute_key")(ctypes.c_uint(level), _p(xyz), _u64(xyz.shape[0]), _p(out))
        return out

    def key_info(self, lohi):
        lohi = np.ascontiguousarray(lohi, dtype=np.uint64).reshape(-1, 2)
        m = lohi.shape[0]
        lvl = np.zeros(m, dtype=np.uint32)
        cen = np.zeros((m, 3))
        ext = np.zeros(m)
        par = np.zeros((m, 2), dtype=np.uint64)
        getattr(self.lib, self.prefix + "key_info")(_p(lohi), _u64(m), _p(lvl), _p(cen), _p(ext), _p(par))
        return lvl, cen, ext, par

    def locate(self, h, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        n = xyz.shape[0]
        e = np.zeros(n, dtype=np.int64)
        lv = np.zeros(n, dtype=np.uint32)
        getattr(self.lib, self.prefix + "locate")(_vp(h), _p(xyz), _u64(n), _p(e), _p(lv))
        return e, lv


class Oracle(_Base):
    """The C restatement (oracle/mpf_oracle.c)."""
    prefix = "mpfo_"

    def __init__(self, path=ORACLE_LIB):
        super().__init__(path)
        self.lib.mpfo_octree_set_fields64.argtypes = [_vp, _vp, _u64]

    def tree(self, rec, fields64=None):
        h = super().tree(rec)
        if fields64 is not None:
            f = np.ascontiguousarray(fields64, dtype=np.float64)
            self.lib.mpfo_octree_set_fields64(_vp(h), _p(f), _u64(f.shape[0]))
        return h

    def interp(self, h, order, xyz, precision=0):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((xyz.shape[0], 7))
        self.lib.mpfo_interp(_vp(h), int(order), int(precision), _p(xyz), _u64(xyz.shape[0]), _p(out))
        return out

    # ---- velocity field (velocity-field build)
    def set_velocity(self, h, rho, vel3, vobs=(0.0, 0.0, 0.0)):
        """Replaces rho (if given) and sets the per-record velocity; vel3 None removes it."""
        r = None if rho is None else np.ascontiguousarray(rho, dtype=np.float32)
        v = None if vel3 is None else np.ascontiguousarray(vel3, dtype=np.float32).reshape(-1, 3)
        vo = np.ascontiguousarray(vobs, dtype=np.float64)
        n = (r.size if r is not None else v.shape[0])
        self.lib.mpfo_octree_set_velocity.argtypes = [_vp, _vp, _vp, _u64, _vp]
        rc = self.lib.mpfo_octree_set_velocity(_vp(h), _p(r), _p(v), _u64(n), _p(vo))
        if rc != 0:
            raise ValueError("count differs from the octree's")

    def velocity_deposit(self, h, n, scheme, pos, vel, rho=None, vel3=None, finish=True):
        """Deposit of one particle set into record-order sums (rho, vel3);
        with rho None the sums start from zero.  Returns (rho, vel3)."""
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 3)
        acc = rho is not None
        r = np.array(rho, dtype=np.float32) if acc else np.zeros(n, dtype=np.float32)
        v = np.array(vel3, dtype=np.float32).reshape(-1, 3) if acc else np.zeros((n, 3), dtype=np.float32)
        rc = self.lib.mpfo_velocity_deposit(_vp(h), int(scheme), _p(pos), _p(vel), _u64(pos.shape[0]), int(acc),
                                            int(bool(finish)), _p(r), _p(v))
        if rc != 0:
            raise ValueError("scheme must be 1 (CIC) or 2 (TSC)")
        return r, v

    def interp10(self, h, order, xyz, precision=0):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((xyz.shape[0], 10))
        self.lib.mpfo_interp10(_vp(h), int(order), int(precision), _p(xyz), _u64(xyz.shape[0]), _p(out))
        return out

    def integrate(self, h, order, table, length, nsteps, stop_kind, stop_value, in9, precision=0, nthreads=1,
                  kitarget=None):
        tab = np.ascontiguousarray(table, dtype=np.float64)
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(-1, 9)
        out = np.zeros((in9.shape[0], NREC))
        kit = None if kitarget is None else np.ascontiguousarray(kitarget, dtype=np.float64)
        steps = self.lib.mpfo_integrate(_vp(h), int(order), int(precision), _p(tab), _u64(tab.shape[1]),
                                        ctypes.c_double(length), ctypes.c_uint(nsteps), int(stop_kind),
                                        ctypes.c_double(stop_value), _p(kit), _p(in9), _u64(in9.shape[0]), _p(out),
                                        int(nthreads))
        return out, steps

    def trajectory(self, h, order, table, length, nsteps, stop_kind, stop_value, in9, precision=0, maxlen=100000,
                   kitarget=None):
        tab = np.ascontiguousarray(table, dtype=np.float64)
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(9)
        rows = np.zeros((maxlen, NREC))
        kit = None if kitarget is None else np.ascontiguousarray(kitarget, dtype=np.float64)
        n = self.lib.mpfo_trajectory(_vp(h), int(order), int(precision), _p(tab), _u64(tab.shape[1]),
                                     ctypes.c_double(length), ctypes.c_uint(nsteps), int(stop_kind),
                                     ctypes.c_double(stop_value), _p(kit), _p(in9), _p(rows), _u64(maxlen))
        return rows[:min(n, maxlen)]

    def propagate(self, h, order, table, length, nsteps, in9, count, angle, rotation, kind=2, amin=0.0,
                  precision=0, maxlen=100000):
        """Integrator::propagate: (central rows, distance per element)."""
        tab = np.ascontiguousarray(table, dtype=np.float64)
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(9)
        rows, dist = np.zeros((maxlen, NREC)), np.zeros(maxlen)
        n = self.lib.mpfo_propagate(_vp(h), int(order), int(precision), _p(tab), _u64(tab.shape[1]),
                                    ctypes.c_double(length), ctypes.c_uint(nsteps), _p(in9), ctypes.c_uint(count),
                                    ctypes.c_double(angle), ctypes.c_double(rotation), int(kind),
                                    ctypes.c_double(amin), _p(rows), _p(dist), _u64(maxlen))
        return rows[:min(n, maxlen)], dist[:min(n, maxlen)]

    def dbetadtheta(self, h, order, table, length, nsteps, kind, opening, observer, phi, theta, ref, ki, dist,
                    precision=0):
        """Lensing::dbetadtheta: nd x (a11, a12, a21, a22)."""
        tab = np.ascontiguousarray(table, dtype=np.float64)
        ref = np.ascontiguousarray(ref, dtype=np.float64).ravel()
        ki = np.ascontiguousarray(ki, dtype=np.float64).reshape(-1, 3)
        dist = np.ascontiguousarray(dist, dtype=np.float64).ravel()
        obs = np.ascontiguousarray(observer, dtype=np.float64)
        out = np.zeros((ref.size, 4))
        self.lib.mpfo_dbetadtheta(_vp(h), int(order), int(precision), _p(tab), _u64(tab.shape[1]),
                                  ctypes.c_double(length), ctypes.c_uint(nsteps), int(kind), ctypes.c_double(opening),
                                  _p(obs), ctypes.c_double(phi), ctypes.c_double(theta), _p(ref), _p(ki), _p(dist),
                                  _u64(ref.size), _p(out))
        return out

    def launch_bundle(self, in9, count, angle, rotation):
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(9)
        out = np.zeros((count + 1, 9))
        self.lib.mpfo_launch_bundle(_p(in9), ctypes.c_uint(count), ctypes.c_double(angle),
                                    ctypes.c_double(rotation), _p(out))
        return out

    def jacobian(self, h, order, table, length, nsteps, stop_kind, stop_value, in9, dist, precision=0,
                 kitarget=None):
        tab = np.ascontiguousarray(table, dtype=np.float64)
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(9)
        dist = np.ascontiguousarray(dist, dtype=np.float64).ravel()
        out = np.zeros((dist.size, 4))
        kit = None if kitarget is None else np.ascontiguousarray(kitarget, dtype=np.float64)
        self.lib.mpfo_jacobian_infinitesimal(_vp(h), int(order), int(precision), _p(tab), _u64(tab.shape[1]),
                                             ctypes.c_double(length), ctypes.c_uint(nsteps), int(stop_kind),
                                             ctypes.c_double(stop_value), _p(kit), _p(in9), _p(dist),
                                             _u64(dist.size), _p(out))
        return out

    def surfaces(self, h, order, table, length, nsteps, kind, vals, in9, jacobian=True, precision=0):
        """Per-surface records (nsurf x 32) of one ray, hmaps.h:615-660 + lensing.h:399-480."""
        tab = np.ascontiguousarray(table, dtype=np.float64)
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(9)
        vals = np.ascontiguousarray(vals, dtype=np.float64).ravel()
        out = np.zeros((vals.size, 32))
        self.lib.mpfo_surfaces(_vp(h), int(order), int(precision), _p(tab), _u64(tab.shape[1]),
                               ctypes.c_double(length), ctypes.c_uint(nsteps), int(kind), _p(vals),
                               ctypes.c_uint32(vals.size), int(bool(jacobian)), _p(in9), _p(out))
        return out

    def catalogue(self, h, order, table, length, nsteps, src6, init2=None, accuracy=1e-8, max_iterations=None,
                  vobs=(0.0, 0.0, 0.0), hubble=0.7, mpc=3.08568e22, precision=0, nthreads=8, bundle=None,
                  born=False, use_prev=0):
        """Newton root-finding of catalogues.h:410-800; returns (n x 20 records, rays integrated).
        bundle: None (jacobiantype = infinitesimal, or infinitesimal_born with born) or (stop_bundle
        kind, plane 0 sachs / 1 normal, opening) for jacobiantype = bundle.  use_prev 2: the grid
        re-run (max_iterations 100 by default, else 10).  Column 18: grid rays short of the source."""
        tab = np.ascontiguousarray(table, dtype=np.float64)
        src6 = np.ascontiguousarray(src6, dtype=np.float64).reshape(-1, 6)
        i2 = None if init2 is None else np.ascontiguousarray(init2, dtype=np.float64).reshape(-1, 2)
        vo = np.ascontiguousarray(vobs, dtype=np.float64)
        out = np.zeros((src6.shape[0], 20))
        kind, plane, opening = bundle if bundle is not None else (-1, 0, 0.0)
        jt = 2 if bundle is not None else (1 if born else 0)
        if max_iterations is None:
            max_iterations = 100 if use_prev else 10
        self.lib.mpfo_catalogue_ex.restype = _u64
        rays = self.lib.mpfo_catalogue_ex(_vp(h), int(order), int(precision), _p(tab), _u64(tab.shape[1]),
                                          ctypes.c_double(length), ctypes.c_uint(nsteps), _p(src6), _p(i2),
                                          _u64(src6.shape[0]), ctypes.c_double(accuracy),
                                          ctypes.c_uint(max_iterations), _p(vo), ctypes.c_double(hubble),
                                          ctypes.c_double(mpc), int(jt), int(kind), int(plane),
                                          ctypes.c_double(opening), int(use_prev), _p(out), int(nthreads))
        return out, rays

    def catalogue_prev(self, h, order, table, length, nsteps, prev18, precision=0, jacobiantype=0, bundle=None):
        """relCat_with_previous_cat: (n x 4 new Jacobian columns, n flags)."""
        tab = np.ascontiguousarray(table, dtype=np.float64)
        pv = np.ascontiguousarray(prev18, dtype=np.float64).reshape(-1, 18)
        n = pv.shape[0]
        jac = np.zeros((n, 4))
        flag = np.zeros(n, dtype=np.int32)
        kind, plane, opening = bundle if bundle is not None else (-1, 0, 0.0)
        self.lib.mpfo_catalogue_prev(_vp(h), int(order), int(precision), _p(tab), _u64(tab.shape[1]),
                                     ctypes.c_double(length), ctypes.c_uint(nsteps), _p(pv), _u64(n),
                                     int(jacobiantype), int(kind), int(plane), ctypes.c_double(opening), _p(jac),
                                     _p(flag))
        return jac, flag

    def interpolate(self, x, y, x0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        out = np.zeros(x0.size)
        self.lib.mpfo_interpolate(_p(x), _p(y), _u64(x.size), _p(x0), _u64(x0.size), _p(out))
        return out


class RefHarness(_Base):
    """The reference's own headers, compiled (oracle/_ref)."""
    prefix = "ref_"

    def __init__(self, path=REF_LIB):
        super().__init__(path)
        self.lib.ref_prep_update.restype = _u64
        self.lib.ref_prep_coarsen.restype = _u64

    # ---- octree preparation (create_octree)
    def prep_keys(self, level, xyz):
        level = np.ascontiguousarray(level, dtype=np.uint32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        out = np.zeros((xyz.shape[0], 2), dtype=np.uint64)
        self.lib.ref_prep_keys(_p(level), _p(xyz), _u64(xyz.shape[0]), _p(out))
        return out

    def prep_collide(self, key2, micro_radius, cone7):
        key2 = np.ascontiguousarray(key2, dtype=np.uint64).reshape(-1, 2)
        c7 = np.ascontiguousarray(cone7, dtype=np.float64).reshape(7)
        out = np.zeros(key2.shape[0], dtype=np.uint8)
        self.lib.ref_prep_collide(_p(key2), _u64(key2.shape[0]), ctypes.c_double(micro_radius), _p(c7), _p(out))
        return out.astype(bool)

    def prep_update(self, rec):
        rec = np.array(rec, copy=True)
        n = self.lib.ref_prep_update(_p(rec), _u64(rec.size))
        return rec[:n]

    def prep_coarsen(self, rec, idx):
        rec = np.array(rec, copy=True)
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        n = self.lib.ref_prep_coarsen(_p(rec), _u64(rec.size), _p(idx), _u64(idx.size))
        return rec[:n]

    def prep_find(self, rec, key2):
        rec = np.ascontiguousarray(rec)
        key2 = np.ascontiguousarray(key2, dtype=np.uint64).reshape(-1, 2)
        found = np.zeros(key2.shape[0], dtype=np.uint64)
        leaf = np.zeros(rec.size, dtype=np.uint8)
        self.lib.ref_prep_find(_p(rec), _u64(rec.size), _p(key2), _u64(key2.shape[0]), _p(found), _p(leaf))
        return found, leaf.astype(bool)

    def prep_interpolate(self, rec, order, xyz):
        rec = np.ascontiguousarray(rec)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((xyz.shape[0], 7), dtype=np.float32)
        self.lib.ref_prep_interpolate(_p(rec), _u64(rec.size), int(order), _p(xyz), _u64(xyz.shape[0]), _p(out))
        return out

    def prep_interpolate2(self, x0, x, y):
        x0 = np.ascontiguousarray(x0, dtype=np.float32)
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(x0.size, dtype=np.float32)
        self.lib.ref_prep_interpolate2(_p(x0), _u64(x0.size), _p(x), _p(y), _u64(x.size), _p(out))
        return out

    def interp(self, h, order, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((xyz.shape[0], 7), dtype=np.float32)
        self.lib.ref_interp(_vp(h), int(order), _p(xyz), _u64(xyz.shape[0]), _p(out))
        return out.astype(np.float64)

    def integrate(self, h, order, table, length, nsteps, stop_kind, stop_value, in9, nthreads=1, kitarget=None):
        tab = np.ascontiguousarray(table, dtype=np.float64)
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(-1, 9)
        out = np.zeros((in9.shape[0], NREC))
        kit = None if kitarget is None else np.ascontiguousarray(kitarget, dtype=np.float64)
        steps = self.lib.ref_integrate(_vp(h), int(order), _p(tab), _u64(tab.shape[1]), ctypes.c_double(length),
                                       ctypes.c_uint(nsteps), int(stop_kind), ctypes.c_double(stop_value), _p(kit),
                                       _p(in9), _u64(in9.shape[0]), _p(out), int(nthreads))
        return out, steps

    def trajectory(self, h, order, table, length, nsteps, stop_kind, stop_value, in9, maxlen=100000, kitarget=None):
        tab = np.ascontiguousarray(table, dtype=np.float64)
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(9)
        rows = np.zeros((maxlen, NREC))
        kit = None if kitarget is None else np.ascontiguousarray(kitarget, dtype=np.float64)
        n = self.lib.ref_trajectory(_vp(h), int(order), _p(tab), _u64(tab.shape[1]), ctypes.c_double(length),
                                    ctypes.c_uint(nsteps), int(stop_kind), ctypes.c_double(stop_value), _p(kit),
                                    _p(in9), _p(rows), _u64(maxlen))
        return rows[:min(n, maxlen)]

    def propagate(self, h, order, table, length, nsteps, in9, count, angle, rotation, interp="redshift", amin=0.0,
                  maxlen=100000):
        tab = np.ascontiguousarray(table, dtype=np.float64)
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(9)
        rows, dist = np.zeros((maxlen, NREC)), np.zeros(maxlen)
        n = self.lib.ref_propagate(_vp(h), int(order), _p(tab), _u64(tab.shape[1]), ctypes.c_double(length),
                                   ctypes.c_uint(nsteps), _p(in9), ctypes.c_uint(count), ctypes.c_double(angle),
                                   ctypes.c_double(rotation), interp.encode(), ctypes.c_double(amin), _p(rows),
                                   _p(dist), _u64(maxlen))
        return rows[:min(n, maxlen)], dist[:min(n, maxlen)]

    def dbetadtheta(self, h, order, table, length, nsteps, stop, opening, observer, phi, theta, ref, ki, dist):
        tab = np.ascontiguousarray(table, dtype=np.float64)
        ref = np.ascontiguousarray(ref, dtype=np.float64).ravel()
        ki = np.ascontiguousarray(ki, dtype=np.float64).reshape(-1, 3)
        dist = np.ascontiguousarray(dist, dtype=np.float64).ravel()
        obs = np.ascontiguousarray(observer, dtype=np.float64)
        out = np.zeros((ref.size, 4))
        self.lib.ref_dbetadtheta(_vp(h), int(order), _p(tab), _u64(tab.shape[1]), ctypes.c_double(length),
                                 ctypes.c_uint(nsteps), stop.encode(), ctypes.c_double(opening), _p(obs),
                                 ctypes.c_double(phi), ctypes.c_double(theta), _p(ref), _p(ki), _p(dist),
                                 _u64(ref.size), _p(out))
        return out

    def jacobian(self, h, order, table, length, nsteps, stop_kind, stop_value, in9, dist, kitarget=None):
        tab = np.ascontiguousarray(table, dtype=np.float64)
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(9)
        dist = np.ascontiguousarray(dist, dtype=np.float64).ravel()
        out = np.zeros((dist.size, 4))
        kit = None if kitarget is None else np.ascontiguousarray(kitarget, dtype=np.float64)
        self.lib.ref_jacobian_infinitesimal(_vp(h), int(order), _p(tab), _u64(tab.shape[1]), ctypes.c_double(length),
                                            ctypes.c_uint(nsteps), int(stop_kind), ctypes.c_double(stop_value),
                                            _p(kit), _p(in9), _p(dist), _u64(dist.size), _p(out))
        return out

    def interpolate(self, x, y, x0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        out = np.zeros(x0.size)
        self.lib.ref_interpolate(_p(x), _p(y), _u64(x.size), _p(x0), _u64(x0.size), _p(out))
        return out

    # cones and launch (cone.h, abstracthypersphere.h, integrator.h:124-382, rays.cpp:233-239)
    def sphere_uniform(self, n, radius):
        xyz = np.zeros((n, 3))
        mm = np.zeros(2)
        self.lib.ref_sphere_uniform(_u64(n), ctypes.c_double(radius), _p(xyz), _p(mm))
        return xyz, mm

    def cone_inside(self, cone7, xyz):
        cone7 = np.ascontiguousarray(cone7, dtype=np.float64).ravel()
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(xyz.shape[0], dtype=np.uint8)
        self.lib.ref_cone_inside(_p(cone7), _p(xyz), _u64(xyz.shape[0]), _p(out))
        return out.astype(bool)

    def launch_cone(self, cones7, icone, micro_radius, seed, n):
        cones7 = np.ascontiguousarray(cones7, dtype=np.float64).reshape(-1, 7)
        in9, rot = np.zeros((n, 9)), np.zeros(n)
        self.lib.ref_launch_cone(_p(cones7), ctypes.c_uint32(cones7.shape[0]), ctypes.c_uint32(icone),
                                 ctypes.c_double(micro_radius), ctypes.c_uint32(seed), _u64(n), _p(in9), _p(rot))
        return in9, rot

    def launch_angles(self, observer, phi_theta):
        obs = np.ascontiguousarray(observer, dtype=np.float64)
        pt = np.ascontiguousarray(phi_theta, dtype=np.float64).reshape(-1, 2)
        out = np.zeros((pt.shape[0], 9))
        self.lib.ref_launch_angles(_p(obs), _p(pt), _u64(pt.shape[0]), _p(out))
        return out

    def launch_towards(self, begin, end):
        b = np.ascontiguousarray(begin, dtype=np.float64).reshape(-1, 3)
        e = np.ascontiguousarray(end, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((b.shape[0], 9))
        self.lib.ref_launch_towards(_p(b), _p(e), _u64(b.shape[0]), _p(out))
        return out

    def launch_bundle(self, in9, count, angle, rotation, center):
        in9 = np.ascontiguousarray(in9, dtype=np.float64).ravel()
        out = np.zeros((count + bool(center), 9))
        self.lib.ref_launch_bundle(_p(in9), ctypes.c_uint32(count), ctypes.c_double(angle),
                                   ctypes.c_double(rotation), int(bool(center)), _p(out))
        return out

    def inv3(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(9)
        r = np.zeros(9)
        self.lib.ref_inv3(_p(a), _p(r))
        return r.reshape(3, 3)

    def constants(self):
        self.lib.ref_deg.restype = ctypes.c_double
        self.lib.ref_pi.restype = ctypes.c_double
        return {"deg": self.lib.ref_deg(), "pi": self.lib.ref_pi()}


class RefVelocity:
    """The reference's integrator compiled with VELOCITYFIELD (oracle/_ref)."""

    def __init__(self, path=REFV_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = ctypes.CDLL(path)
        L.refv_octree_new.restype = _vp
        L.refv_octree_new.argtypes = [_vp, _u64, _vp, _vp]
        L.refv_octree_free.argtypes = [_vp]
        L.refv_integrate.restype = _u64

    def tree(self, rec, rho=None, vel3=None):
        rec = np.ascontiguousarray(rec)
        r = None if rho is None else np.ascontiguousarray(rho, dtype=np.float32)
        v = None if vel3 is None else np.ascontiguousarray(vel3, dtype=np.float32).reshape(-1, 3)
        return self.lib.refv_octree_new(_p(rec), _u64(rec.size), _p(r), _p(v))

    def free(self, h):
        self.lib.refv_octree_free(_vp(h))

    def interp10(self, h, order, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.zeros((xyz.shape[0], 10), dtype=np.float32)
        self.lib.refv_interp(_vp(h), int(order), _p(xyz), _u64(xyz.shape[0]), _p(out))
        return out

    def integrate(self, h, order, table, length, nsteps, stop_kind, stop_value, in9, vobs=(0.0, 0.0, 0.0),
                  kitarget=None):
        tab = np.ascontiguousarray(table, dtype=np.float64)
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(-1, 9)
        out = np.zeros((in9.shape[0], NREC))
        kit = None if kitarget is None else np.ascontiguousarray(kitarget, dtype=np.float64)
        vo = np.ascontiguousarray(vobs, dtype=np.float64)
        steps = self.lib.refv_integrate(_vp(h), int(order), _p(tab), _u64(tab.shape[1]), ctypes.c_double(length),
                                        ctypes.c_uint(nsteps), int(stop_kind), ctypes.c_double(stop_value), _p(kit),
                                        _p(vo), _p(in9), _u64(in9.shape[0]), _p(out))
        return out, steps


JACTYPES = {"infinitesimal": 0, "infinitesimal_born": 1, "bundle": 2}
# one element of the reference's octree container: the 128-bit key, then
# Gravity<float> in reverse tuple order (gravity.h:45-87), 48 bytes
OCTREE_RECORD = np.dtype([("lo", "<u8"), ("hi", "<u8"), ("dphidt", "<f4"), ("a", "<f4"), ("dphidx", "<f4"),
                          ("dphidy", "<f4"), ("dphidz", "<f4"), ("phi", "<f4"), ("rho", "<f4"), ("pad", "<u4")])
PLANES = {"sachs": 0, "normal": 1}


class RefCatalogue:
    """The reference's catalogue root-finder, relCat, Input::acquire/correct,
    the cone generators and getTargets (oracle/_ref/libmpfrefcat.so)."""

    def __init__(self, path=REFC_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        # the library names libhdf5.so.103; map it first from where it was
        # linked, so the loader finds it without a search path
        ctypes.CDLL(HDF5_LIB, mode=ctypes.RTLD_GLOBAL)
        L = self.lib = ctypes.CDLL(path)
        L.refc_octree_new.restype = _vp
        L.refc_octree_new.argtypes = [_vp, _u64]
        L.refc_octree_free.argtypes = [_vp]
        L.refc_get_targets.restype = _u64

    def tree(self, rec):
        rec = np.ascontiguousarray(rec)
        return self.lib.refc_octree_new(_p(rec), _u64(rec.size))

    def free(self, h):
        self.lib.refc_octree_free(_vp(h))

    def newton(self, h, table, length, nsteps, src6, init2=None, accuracy=1e-8, jacobiantype="infinitesimal",
               plane="sachs", stop_bundle="redshift", opening=1e-4, use_prev=0, vobs=(0.0, 0.0, 0.0), hubble=0.7,
               mpc=3.08568e22, order=2, nthreads=0):
        """Catalogues::iterateNewtonMethod per source: n x 16 (result 2 x 2,
        jacobian 2 x 2, redshifts 6, interpRef, 0)."""
        tab = np.ascontiguousarray(table, dtype=np.float64)
        src6 = np.ascontiguousarray(src6, dtype=np.float64).reshape(-1, 6)
        n = src6.shape[0]
        i2 = None if init2 is None else np.ascontiguousarray(init2, dtype=np.float64).reshape(-1, 2)
        vo = np.ascontiguousarray(vobs, dtype=np.float64)
        out = np.zeros((n, 16))
        self.lib.refc_newton(_vp(h), int(order), _p(tab), _u64(tab.shape[1]), ctypes.c_double(length),
                             ctypes.c_double(hubble), ctypes.c_double(mpc), ctypes.c_uint(nsteps),
                             ctypes.c_double(accuracy), int(JACTYPES[jacobiantype]), int(PLANES[plane]),
                             stop_bundle.encode(), ctypes.c_double(opening), int(use_prev), _p(vo), _p(src6),
                             _p(i2), _u64(n), _p(out), int(nthreads))
        return out

    def newton_once(self, h, table, length, nsteps, src6, ang2, iteration, jac4=None, accuracy=1e-8,
                    jacobiantype="infinitesimal", plane="sachs", stop_bundle="redshift", opening=1e-4, use_prev=0,
                    vobs=(0.0, 0.0, 0.0), hubble=0.7, mpc=3.08568e22):
        """One Catalogues::newtonMethod2d call per source: n x 16 as newton()."""
        tab = np.ascontiguousarray(table, dtype=np.float64)
        src6 = np.ascontiguousarray(src6, dtype=np.float64).reshape(-1, 6)
        a2 = np.ascontiguousarray(ang2, dtype=np.float64).reshape(-1, 2)
        j4 = None if jac4 is None else np.ascontiguousarray(jac4, dtype=np.float64).reshape(-1, 4)
        vo = np.ascontiguousarray(vobs, dtype=np.float64)
        out = np.zeros((src6.shape[0], 16))
        self.lib.refc_newton_once(_vp(h), _p(tab), _u64(tab.shape[1]), ctypes.c_double(length),
                                  ctypes.c_double(hubble), ctypes.c_double(mpc), ctypes.c_uint(nsteps),
                                  ctypes.c_double(accuracy), int(JACTYPES[jacobiantype]), int(PLANES[plane]),
                                  stop_bundle.encode(), ctypes.c_double(opening), int(use_prev), _p(vo), _p(src6),
                                  _p(a2), _p(j4), ctypes.c_uint(iteration), _u64(src6.shape[0]), _p(out))
        return out

    def relcat(self, h, table, length, nsteps, targets8, filename, accuracy=1e-8, jacobiantype="infinitesimal",
               plane="sachs", stop_bundle="redshift", opening=1e-4, use_prev=0, vobs=(0.0, 0.0, 0.0), rotm1=None,
               prev18=None, hubble=0.7, mpc=3.08568e22):
        """Catalogues::relCat: writes filename.txt, _err.txt, _reject.txt."""
        tab = np.ascontiguousarray(table, dtype=np.float64)
        t8 = np.ascontiguousarray(targets8, dtype=np.float64).reshape(-1, 8)
        rot = None if rotm1 is None else np.ascontiguousarray(rotm1, dtype=np.float64).reshape(3, 3)
        pv = None if prev18 is None else np.ascontiguousarray(prev18, dtype=np.float64).reshape(-1, 18)
        vo = np.ascontiguousarray(vobs, dtype=np.float64)
        self.lib.refc_relcat(_vp(h), _p(tab), _u64(tab.shape[1]), ctypes.c_double(length), ctypes.c_double(hubble),
                             ctypes.c_double(mpc), ctypes.c_uint(nsteps), ctypes.c_double(accuracy),
                             int(JACTYPES[jacobiantype]), int(PLANES[plane]), stop_bundle.encode(),
                             ctypes.c_double(opening), int(use_prev), _p(vo), _p(rot), _p(t8), _u64(t8.shape[0]),
                             _p(pv), _u64(0 if pv is None else pv.shape[0]), filename.encode())

    def relcat_prev(self, h, table, length, nsteps, prev18, filename, jacobiantype="infinitesimal", plane="sachs",
                    stop_bundle="redshift", opening=1e-4, vobs=(0.0, 0.0, 0.0), hubble=0.7, mpc=3.08568e22):
        """Catalogues::relCat_with_previous_cat: returns the updated rows."""
        tab = np.ascontiguousarray(table, dtype=np.float64)
        pv = np.array(prev18, dtype=np.float64).reshape(-1, 18)
        vo = np.ascontiguousarray(vobs, dtype=np.float64)
        self.lib.refc_relcat_prev(_vp(h), _p(tab), _u64(tab.shape[1]), ctypes.c_double(length),
                                  ctypes.c_double(hubble), ctypes.c_double(mpc), ctypes.c_uint(nsteps),
                                  int(JACTYPES[jacobiantype]), int(PLANES[plane]), stop_bundle.encode(),
                                  ctypes.c_double(opening), _p(vo), _p(pv), _u64(pv.shape[0]), filename.encode())
        return pv

    def acquire(self, evolfile, paramfile, celldir, mpc=3.08568e22, capacity=200000):
        """Input::acquire: (h, omega_m, lboxmpch), 4 x n table."""
        hol = np.zeros(3)
        tab = np.zeros((4, capacity))
        n = ctypes.c_uint64(0)
        self.lib.refc_acquire(evolfile.encode(), paramfile.encode(), celldir.encode(), ctypes.c_double(mpc),
                              _p(hol), _p(tab), _u64(capacity), ctypes.byref(n))
        if n.value > capacity:
            raise ValueError("table larger than capacity")
        return tuple(hol), np.ascontiguousarray(tab[:, :n.value])

    def correct(self, table4, traj_t, traj_a):
        tab = np.ascontiguousarray(table4, dtype=np.float64)
        t = np.ascontiguousarray(traj_t, dtype=np.float64)
        a = np.ascontiguousarray(traj_a, dtype=np.float64)
        out = np.zeros_like(tab)
        self.lib.refc_correct(_p(tab), _u64(tab.shape[1]), _p(t), _p(a), _u64(t.size), _p(out))
        return out

    def fullsky_cones(self, ncones):
        out = np.zeros((ncones, 7))
        self.lib.refc_fullsky_cones(ctypes.c_uint32(ncones), _p(out))
        return out

    def narrow_cones(self, celldir, ncones):
        """(cones7, cones_rot7, rotm1 3 x 3, (thetay, thetaz) rad)."""
        c = np.zeros((ncones, 7))
        r = np.zeros((ncones, 7))
        m = np.zeros((3, 3))
        tyz = np.zeros(2)
        self.lib.refc_narrow_cones(celldir.encode(), ctypes.c_uint32(ncones), _p(c), _p(r), _p(m), _p(tyz))
        return c, r, m, tyz

    def prep_import(self, path, cone7, micro_radius, isfullsky=1, rotm1=None, coarseonly=0, ncoarse=0):
        """Input::importascii of one ASCII cell file with PreparationASCII's
        filter: the selected 48-byte records."""
        c7 = np.ascontiguousarray(cone7, dtype=np.float64).reshape(7)
        rot = None if rotm1 is None else np.ascontiguousarray(rotm1, dtype=np.float64).reshape(9)
        self.lib.refc_prep_import.restype = _u64
        n = self.lib.refc_prep_import(str(path).encode(), _p(c7), ctypes.c_double(micro_radius), int(isfullsky),
                                      _p(rot) if rot is not None else None, int(coarseonly), ctypes.c_uint(ncoarse),
                                      None, _u64(0))
        out = np.zeros(n, dtype=OCTREE_RECORD)
        self.lib.refc_prep_import(str(path).encode(), _p(c7), ctypes.c_double(micro_radius), int(isfullsky),
                                  _p(rot) if rot is not None else None, int(coarseonly), ctypes.c_uint(ncoarse),
                                  _p(out), _u64(n))
        return out

    def prep_correct(self, rec, filler, reserve, table4, correction=1, coarsecorrection=1, h=0.7, omegam=0.3,
                     lboxmpch=1000.0, mpc=3.08568e22, capacity=None):
        """Miscellaneous::correctOctree on the records (in their order), the
        container's spare storage (reserve records) filled with copies of
        filler: the corrected records."""
        rec = np.ascontiguousarray(rec)
        filler = np.ascontiguousarray(np.asarray(filler).reshape(1))
        tab = np.ascontiguousarray(table4, dtype=np.float64)
        cap = int(capacity or 16 * rec.size + 64)
        out = np.zeros(cap, dtype=OCTREE_RECORD)
        self.lib.refc_prep_correct.restype = _u64
        m = self.lib.refc_prep_correct(_p(rec), _u64(rec.size), _p(filler), _u64(reserve), int(correction),
                                       int(coarsecorrection), ctypes.c_double(h), ctypes.c_double(omegam),
                                       ctypes.c_double(lboxmpch), ctypes.c_double(mpc), _p(tab),
                                       _u64(tab.shape[1]), _p(out), _u64(cap))
        assert m <= cap
        return out[:m]

    def get_targets(self, src8, cones7, icone):
        s = np.ascontiguousarray(src8, dtype=np.float64).reshape(-1, 8)
        cn = np.ascontiguousarray(cones7, dtype=np.float64).reshape(-1, 7)
        out = np.zeros_like(s)
        m = self.lib.refc_get_targets(_p(s), _u64(s.shape[0]), _p(cn), ctypes.c_uint32(cn.shape[0]),
                                      ctypes.c_uint32(icone), _p(out))
        return out[:m]


def ref_available():
    return os.path.exists(REF_LIB)


REF_RAYS = os.path.join(HERE, "_ref", "rays")
CONDA_LIB = os.path.dirname(HDF5_LIB)


def rays_available():
    """The reference's own rays program (oracle/_ref/rays) and the MPI and
    HDF5 libraries it was linked against."""
    return os.path.exists(REF_RAYS) and os.path.exists(os.path.join(CONDA_LIB, "libmpi.so"))


def run_rays(params, workdir, timeout=600):
    """Runs the reference's rays program as one MPI task on a parameter file
    written from params ("key = value" lines) in workdir; returns its
    standard output.  The system C++ runtime comes first on the library
    path, the MPI and HDF5 libraries after it."""
    import subprocess
    path = os.path.join(workdir, "raytracer.txt")
    with open(path, "w") as fh:
        for k, v in params.items():
            fh.write(f"{k} = {v}\n")
    env = dict(os.environ, LD_LIBRARY_PATH="/usr/lib/x86_64-linux-gnu:" + CONDA_LIB)
    r = subprocess.run([REF_RAYS, path], cwd=workdir, env=env, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"rays exited with {r.returncode}: {r.stdout[-2000:]} {r.stderr[-2000:]}")
    return r.stdout


def refc_available():
    return os.path.exists(REFC_LIB) and os.path.exists(HDF5_LIB)


def refv_available():
    return os.path.exists(REFV_LIB)
