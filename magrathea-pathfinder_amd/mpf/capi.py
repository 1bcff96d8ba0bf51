"""ctypes binding of libmpf.so (include/mpf.h).

The HIP library is the product path: this module fails loudly when it is
missing, and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libmpf.so")

HOMOGENEOUS, NGP, CIC, TSC = -1, 0, 1, 2
PREC_REF, PREC_FP64, PREC_SPLIT = 0, 1, 2
STATIC_MAX_RAYS = 262144   # MPF_STATIC_MAX_RAYS of include/mpf.h
STOP_NONE, STOP_RADIUS, STOP_LAMBDA, STOP_REDSHIFT, STOP_T, STOP_A, STOP_PLANE, STOP_S = -1, 0, 1, 2, 3, 4, 5, 6
STOP_NAMES = {"none": -1, "r": 0, "radius": 0, "lambda": 1, "redshift": 2, "t": 3, "eta": 3, "a": 4,
              "plane": 5, "s": 6}

OUT_FIELDS = ("a", "t", "x", "y", "z", "kt", "kx", "ky", "kz", "redshift", "lambda", "s", "chi", "isw",
              "iswold", "phi", "dphidx", "dphidy", "dphidz", "dphidt", "rho", "ah", "dphidl", "laplacian",
              "dsdl2", "error", "level", "nsteps")
NOUT = len(OUT_FIELDS)
OUT = {name: i for i, name in enumerate(OUT_FIELDS)}

# every symbol declared in include/mpf.h
EXPORTS = ("mpf_create", "mpf_destroy", "mpf_last_error", "mpf_set_precision", "mpf_load_octree_records",
           "mpf_load_octree_fields64", "mpf_octree_info", "mpf_set_cosmology", "mpf_locate", "mpf_interpolate",
           "mpf_integrate", "mpf_integrate_dev", "mpf_integrate_dev_static", "mpf_last_run", "mpf_synchronize", "mpf_validate_octree",
           "mpf_dev_alloc", "mpf_dev_free", "mpf_dev_upload", "mpf_dev_download", "mpf_dev_copy",
           "mpf_integrate_surfaces", "mpf_integrate_surfaces_dev", "mpf_healpix_pix2vec", "mpf_healpix_vec2pix",
           "mpf_hmaps_launch_dev", "mpf_hmaps_fill_dev", "mpf_healpix_bin_dev", "mpf_catalogue_solve",
           "mpf_catalogue_solve_dev", "mpf_catalogue_previous", "mpf_catalogue_write",
           "mpf_catalogue_write_previous", "mpf_mlp_init", "mpf_mlp_state", "mpf_mlp_train", "mpf_mlp_gradient",
           "mpf_mlp_infer", "mpf_mlp_infer_dev", "mpf_catalogue_features", "mpf_catalogue_features_dev",
           "mpf_catalogue_seed", "mpf_catalogue_seed_dev", "mpf_mfma_probe", "mpf_mfma_probe32", "mpf_mlp_last_ms",
           "mpf_device_memory",
           "mpf_integrate_trajectory",
           "mpf_cosmology_acquire", "mpf_cosmology_correct", "mpf_cosmology_prepare", "mpf_cosmology_last_error",
           "mpf_cone_last_error", "mpf_cones_fullsky", "mpf_cones_narrow", "mpf_narrow_cone_info", "mpf_cone_inside",
           "mpf_cone_select", "mpf_cone_targets", "mpf_hmaps_cone_pixels", "mpf_launch_angles", "mpf_launch_towards",
           "mpf_launch_cone", "mpf_launch_bundle", "mpf_set_launch_config", "mpf_get_launch_config", "mpf_autotune",
           "mpf_last_kernel_ms", "mpf_propagate", "mpf_jacobian_bundle", "mpf_bundle_plane_info",
           "mpf_split_tidal_blocks",
           "mpf_ray_statistics", "mpf_ray_statistics_write", "mpf_hmaps_homogeneous", "mpf_hmaps_components_dev", "mpf_healpix_write_fits",
           "mpf_integrate_surfaces_born", "mpf_integrate_surfaces_born_dev",
           "mpf_prep_last_error", "mpf_prep_create", "mpf_prep_destroy", "mpf_prep_clear", "mpf_prep_import_ascii",
           "mpf_prep_load_records", "mpf_prep_update", "mpf_prep_interpolate", "mpf_prep_correct", "mpf_prep_size",
           "mpf_prep_records", "mpf_prep_save", "mpf_prep_keys", "mpf_prep_collide", "mpf_prepare_ascii",
           "mpf_velocity_begin", "mpf_velocity_deposit", "mpf_velocity_deposit_dev", "mpf_velocity_finish",
           "mpf_velocity_set", "mpf_velocity_get", "mpf_velocity_clear", "mpf_set_observer_velocity")

CAT_FIELDS = ("phi", "theta", "phi_obs", "theta_obs", "sep", "iter", "z0", "z1", "z2", "z3", "z4", "z5",
              "a11", "a12", "a21", "a22", "status", "nrounds")
CAT = {name: i for i, name in enumerate(CAT_FIELDS)}
NCAT = 20
CAT_OK, CAT_OUTSIDE, CAT_REJECTED = 0, 1, 2

SURF_FIELDS = ("valid", "firstid", "f", "x", "y", "z", "dist", "chi", "lambda", "t", "a", "redshift", "s", "isw",
               "rho", "phi", "rhomax", "kt", "kx", "ky", "kz", "a11", "a12", "a22")
SURF = {name: i for i, name in enumerate(SURF_FIELDS)}
NSURF = 32
CONE_N = 7          # vertex (3), base centre (3), half-angle: a cone of cone.h
MAX_SURFACES = 16

# learned first guess of the catalogue root-finder (sizes of include/mpf.h)
MLP_NIN, MLP_NOUT, MLP_NHIDDEN, MLP_KIN, MLP_KOUT = 10, 2, 128, 16, 16
MLP_NPARAM = MLP_KIN * MLP_NHIDDEN + MLP_NHIDDEN + MLP_NHIDDEN * MLP_NHIDDEN + MLP_NHIDDEN + \
    MLP_NHIDDEN * MLP_KOUT + MLP_KOUT
MLP_NNORM = 24
MLP_NSTATE = MLP_NPARAM + MLP_NNORM


class MpfError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("order", ctypes.c_int), ("precision", ctypes.c_int),
                ("nsteps", ctypes.c_uint), ("length", ctypes.c_double)]


class Surfaces(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("n", ctypes.c_uint32), ("value", ctypes.c_double * 16),
                ("jacobian", ctypes.c_int), ("ray_value", ctypes.c_void_p)]


def make_surfaces(kind, values, jacobian=True, ray_value=None):
    if isinstance(kind, str):
        kind = STOP_NAMES[kind]
    values = list(values)
    s = Surfaces()
    s.kind = int(kind)
    s.n = len(values)
    for i, v in enumerate(values):
        s.value[i] = float(v)
    s.jacobian = int(bool(jacobian))
    s.ray_value = ray_value
    return s


class NewtonCfg(ctypes.Structure):
    _fields_ = [("accuracy", ctypes.c_double), ("max_iterations", ctypes.c_uint32), ("vobs", ctypes.c_double * 3),
                ("h", ctypes.c_double), ("mpc", ctypes.c_double), ("jacobian_all_rounds", ctypes.c_uint32),
                ("jacobian_born", ctypes.c_uint32), ("jacobian_bundle", ctypes.c_uint32),
                ("bundle_kind", ctypes.c_int32), ("bundle_plane", ctypes.c_int32), ("bundle_opening", ctypes.c_double),
                ("use_previous", ctypes.c_uint32)]

BUNDLE_PLANES = {"sachs": 0, "normal": 1}


def make_newton(accuracy=1e-8, max_iterations=10, vobs=(0.0, 0.0, 0.0), h=0.7, mpc=3.08568e22,
                jacobian_all_rounds=False, jacobian_born=False, bundle=None, use_previous=0):
    """bundle: None, or (stop_bundle, plane, opening) for jacobiantype = bundle,
    e.g. ("redshift", "sachs", 1e-4); use_previous 2: the grid re-run of
    rejected sources."""
    c = NewtonCfg()
    c.use_previous = int(use_previous)
    # False/True or the schedule 0, 1, 2 of include/mpf.h
    c.jacobian_all_rounds = int(jacobian_all_rounds)
    c.jacobian_born = int(bool(jacobian_born))
    if bundle is not None:
        kind, plane, opening = bundle
        c.jacobian_bundle = 1
        c.bundle_kind = STOP_NAMES[kind] if isinstance(kind, str) else int(kind)
        c.bundle_plane = BUNDLE_PLANES[plane] if isinstance(plane, str) else int(plane)
        c.bundle_opening = float(opening)
    c.accuracy = accuracy
    c.max_iterations = max_iterations
    for i in range(3):
        c.vobs[i] = vobs[i]
    c.h = h
    c.mpc = mpc
    return c


class LaunchCfg(ctypes.Structure):
    _fields_ = [("refill", ctypes.c_uint32), ("sort_bucket", ctypes.c_uint64), ("blocks_per_cu", ctypes.c_uint32),
                ("waves_per_simd", ctypes.c_uint32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Stop(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("value", ctypes.c_double), ("kitarget", ctypes.c_double * 3)]


_lib = None


def load_library(path=LIB_PATH):
    """Load libmpf.so; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MpfError(f"libmpf.so not found at {path}: build it first (__graft_entry__.build())")
    lib = ctypes.CDLL(path)
    vp, u64, i32, dbl = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    lib.mpf_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Config)]
    lib.mpf_destroy.argtypes = [vp]
    lib.mpf_last_error.argtypes = [vp]
    lib.mpf_last_error.restype = ctypes.c_char_p
    lib.mpf_set_precision.argtypes = [vp, i32]
    lib.mpf_load_octree_records.argtypes = [vp, vp, u64]
    lib.mpf_load_octree_fields64.argtypes = [vp, vp, u64]
    lib.mpf_velocity_begin.argtypes = [vp]
    lib.mpf_velocity_deposit.argtypes = [vp, vp, vp, u64, i32]
    lib.mpf_velocity_deposit_dev.argtypes = [vp, vp, vp, u64, i32]
    lib.mpf_velocity_finish.argtypes = [vp]
    lib.mpf_velocity_set.argtypes = [vp, vp, vp, u64]
    lib.mpf_velocity_get.argtypes = [vp, vp, vp, u64]
    lib.mpf_velocity_clear.argtypes = [vp]
    lib.mpf_set_observer_velocity.argtypes = [vp, vp]
    lib.mpf_octree_info.argtypes = [vp, vp, vp, vp]
    lib.mpf_set_cosmology.argtypes = [vp, vp, vp, vp, u64]
    lib.mpf_locate.argtypes = [vp, vp, u64, vp, vp]
    lib.mpf_interpolate.argtypes = [vp, i32, vp, u64, vp]
    lib.mpf_integrate.argtypes = [vp, vp, u64, ctypes.POINTER(Stop), vp, vp]
    lib.mpf_integrate_dev.argtypes = [vp, vp, u64, ctypes.POINTER(Stop), vp, i32]
    lib.mpf_integrate_dev_static.argtypes = [vp, vp, u64, ctypes.POINTER(Stop), vp, ctypes.POINTER(u64)]
    lib.mpf_last_run.argtypes = [vp, vp, vp]
    lib.mpf_synchronize.argtypes = [vp]
    lib.mpf_validate_octree.argtypes = [vp, u64, vp, vp, ctypes.c_char_p, u64]
    lib.mpf_dev_alloc.argtypes = [vp, u64, ctypes.POINTER(vp)]
    lib.mpf_dev_free.argtypes = [vp, vp]
    lib.mpf_dev_upload.argtypes = [vp, vp, vp, u64]
    lib.mpf_dev_download.argtypes = [vp, vp, vp, u64]
    lib.mpf_dev_copy.argtypes = [vp, vp, vp, u64]
    lib.mpf_integrate_surfaces.argtypes = [vp, vp, u64, ctypes.POINTER(Surfaces), vp, vp]
    lib.mpf_integrate_surfaces_dev.argtypes = [vp, vp, u64, ctypes.POINTER(Surfaces), vp, vp, i32]
    lib.mpf_catalogue_solve.argtypes = [vp, vp, vp, u64, ctypes.POINTER(NewtonCfg), vp, vp]
    lib.mpf_catalogue_solve_dev.argtypes = [vp, vp, vp, u64, ctypes.POINTER(NewtonCfg), vp, vp]
    lib.mpf_catalogue_previous.argtypes = [vp, vp, u64, ctypes.POINTER(NewtonCfg)]
    lib.mpf_catalogue_write.argtypes = [ctypes.c_char_p, vp, u64, vp]
    lib.mpf_catalogue_write_previous.argtypes = [ctypes.c_char_p, vp, u64]
    i64 = ctypes.c_int64
    lib.mpf_healpix_pix2vec.argtypes = [i64, vp, u64, vp]
    lib.mpf_healpix_vec2pix.argtypes = [i64, vp, u64, vp]
    lib.mpf_hmaps_launch_dev.argtypes = [vp, i64, vp, u64, vp]
    lib.mpf_hmaps_fill_dev.argtypes = [vp, i64, vp, vp, u64, ctypes.c_uint32, vp]
    lib.mpf_healpix_bin_dev.argtypes = [vp, i64, vp, vp, u64, ctypes.c_uint32, ctypes.c_uint32, vp]
    f32, u32 = ctypes.c_float, ctypes.c_uint32
    lib.mpf_mlp_init.argtypes = [vp, u64]
    lib.mpf_mlp_state.argtypes = [vp, vp, vp]
    lib.mpf_mlp_train.argtypes = [vp, vp, vp, u64, u32, u32, f32, u64, vp]
    lib.mpf_mlp_gradient.argtypes = [vp, vp, vp, u64, vp, vp]
    lib.mpf_mlp_infer.argtypes = [vp, vp, u64, vp]
    lib.mpf_mlp_infer_dev.argtypes = [vp, vp, u64, vp]
    lib.mpf_catalogue_features.argtypes = [vp, vp, u64, vp]
    lib.mpf_catalogue_features_dev.argtypes = [vp, vp, u64, vp]
    lib.mpf_catalogue_seed.argtypes = [vp, vp, u64, i32, vp]
    lib.mpf_catalogue_seed_dev.argtypes = [vp, vp, u64, i32, vp]
    lib.mpf_mfma_probe.argtypes = [vp, vp, vp, vp]
    lib.mpf_mfma_probe32.argtypes = [vp, vp, vp, ctypes.c_int, vp]
    lib.mpf_mlp_last_ms.argtypes = [vp, ctypes.POINTER(dbl)]
    lib.mpf_device_memory.argtypes = [ctypes.c_int, ctypes.POINTER(u64), ctypes.POINTER(u64)]
    lib.mpf_integrate_trajectory.argtypes = [vp, vp, u64, ctypes.POINTER(Stop), u64, vp, vp]
    cp, dbl = ctypes.c_char_p, ctypes.c_double
    lib.mpf_cosmology_acquire.argtypes = [cp, cp, cp, dbl, vp, vp, vp, vp, vp, vp, vp, u64, vp]
    lib.mpf_cosmology_correct.argtypes = [vp, vp, u64, vp, vp, u64]
    lib.mpf_cosmology_prepare.argtypes = [i32, i32, cp, cp, cp, dbl, vp, vp, vp, vp, vp, vp, vp, u64, vp]
    lib.mpf_cosmology_last_error.restype = ctypes.c_char_p
    lib.mpf_hmaps_homogeneous.argtypes = [vp, u64, vp, u32, i32, vp, vp]
    lib.mpf_hmaps_components_dev.argtypes = [vp, i64, vp, vp, vp, vp, u64, u32, vp, u32, vp, vp, u64,
                                             ctypes.POINTER(u32)]
    lib.mpf_integrate_surfaces_born.argtypes = [vp, vp, u64, ctypes.POINTER(Surfaces), vp, vp]
    lib.mpf_integrate_surfaces_born_dev.argtypes = [vp, vp, u64, ctypes.POINTER(Surfaces), vp, vp, i32]
    lib.mpf_healpix_write_fits.argtypes = [cp, vp, i64, i32, ctypes.c_char, i32]
    lib.mpf_ray_statistics.argtypes = [vp, vp, vp, u64, u64, vp, vp, u64, i32, i32, u32, dbl, dbl, vp, u64, vp, vp]
    lib.mpf_ray_statistics_write.argtypes = [cp, vp, u64, u64]
    lib.mpf_jacobian_bundle.argtypes = [vp, vp, u64, vp, i32, vp, u32, vp, vp, dbl, vp]
    lib.mpf_bundle_plane_info.argtypes = [vp, ctypes.POINTER(u64)]
    lib.mpf_split_tidal_blocks.argtypes = [vp, i32, ctypes.POINTER(u64)]
    lib.mpf_propagate.argtypes = [vp, vp, u64, u32, dbl, vp, i32, dbl, u64, vp, vp, vp]
    lib.mpf_last_kernel_ms.argtypes = [vp, ctypes.POINTER(dbl)]
    lib.mpf_set_launch_config.argtypes = [vp, i32, ctypes.POINTER(LaunchCfg)]
    lib.mpf_get_launch_config.argtypes = [vp, i32, ctypes.POINTER(LaunchCfg)]
    lib.mpf_autotune.argtypes = [vp, vp, u64, ctypes.POINTER(Stop), ctypes.POINTER(Surfaces), vp, vp,
                                 ctypes.POINTER(LaunchCfg), vp, u32, ctypes.POINTER(u32)]
    # cones and launch (host)
    lib.mpf_cone_last_error.restype = ctypes.c_char_p
    lib.mpf_cones_fullsky.argtypes = [u32, dbl, vp]
    lib.mpf_cones_narrow.argtypes = [u32, dbl, dbl, dbl, dbl, dbl, vp, vp, vp, vp]
    lib.mpf_narrow_cone_info.argtypes = [cp, vp]
    lib.mpf_cone_inside.argtypes = [vp, vp, u64, u64, vp]
    lib.mpf_cone_select.argtypes = [vp, u32, u32, i32, vp, u64, u64, vp]
    lib.mpf_cone_targets.argtypes = [vp, u32, u32, vp, u64, u64, vp, vp]
    lib.mpf_hmaps_cone_pixels.argtypes = [i64, vp, u32, u32, vp, u64, vp]
    lib.mpf_launch_angles.argtypes = [vp, vp, u64, vp]
    lib.mpf_launch_towards.argtypes = [vp, vp, u64, vp]
    lib.mpf_launch_cone.argtypes = [vp, u32, u32, dbl, u32, u64, vp, vp]
    lib.mpf_launch_bundle.argtypes = [vp, u32, dbl, dbl, i32, vp]
    # octree preparation (host)
    lib.mpf_prep_last_error.restype = ctypes.c_char_p
    lib.mpf_prep_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(PrepConfig)]
    lib.mpf_prep_destroy.argtypes = [vp]
    lib.mpf_prep_clear.argtypes = [vp]
    lib.mpf_prep_import_ascii.argtypes = [vp, cp, vp, vp, ctypes.POINTER(u64)]
    lib.mpf_prep_load_records.argtypes = [vp, vp, u64]
    lib.mpf_prep_update.argtypes = [vp]
    lib.mpf_prep_interpolate.argtypes = [vp, i32, vp, u64, vp]
    lib.mpf_prep_correct.argtypes = [vp, vp, vp, vp, u64]
    lib.mpf_prep_size.argtypes = [vp, ctypes.POINTER(u64)]
    lib.mpf_prep_records.argtypes = [vp, vp, u64]
    lib.mpf_prep_save.argtypes = [vp, cp]
    lib.mpf_prep_keys.argtypes = [vp, vp, u64, vp]
    lib.mpf_prep_collide.argtypes = [vp, u64, dbl, vp, vp]
    lib.mpf_prepare_ascii.argtypes = [ctypes.POINTER(PrepConfig), cp, cp, vp, u32, vp, u32, u32, u64, vp, vp, vp, u64,
                                      ctypes.POINTER(cp)]
    _lib = lib
    return lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def validate_octree(rec):
    """Host-side structural check of cone records (no GPU): (ok, min_level, max_level, message)."""
    lib = load_library()
    rec = np.ascontiguousarray(rec)
    lo = ctypes.c_uint32(0)
    hi = ctypes.c_uint32(0)
    buf = ctypes.create_string_buffer(512)
    rc = lib.mpf_validate_octree(_p(rec), ctypes.c_uint64(rec.size), ctypes.byref(lo), ctypes.byref(hi), buf, 512)
    return rc == 0, lo.value, hi.value, buf.value.decode()


def healpix_pix2vec(nside, pix):
    lib = load_library()
    pix = np.ascontiguousarray(pix, dtype=np.int64).ravel()
    out = np.zeros((pix.size, 3))
    if lib.mpf_healpix_pix2vec(int(nside), _p(pix), ctypes.c_uint64(pix.size), _p(out)):
        raise MpfError("mpf_healpix_pix2vec: invalid arguments")
    return out


def healpix_vec2pix(nside, vec):
    lib = load_library()
    vec = np.ascontiguousarray(vec, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(vec.shape[0], dtype=np.int64)
    if lib.mpf_healpix_vec2pix(int(nside), _p(vec), ctypes.c_uint64(vec.shape[0]), _p(out)):
        raise MpfError("mpf_healpix_vec2pix: invalid arguments")
    return out


class PrepConfig(ctypes.Structure):
    """mpf_prep_cfg: the create_octree parameters the ASCII branch reads."""
    _fields_ = [("isfullsky", ctypes.c_int32), ("coarseonly", ctypes.c_int32), ("ncoarse", ctypes.c_uint32),
                ("correction", ctypes.c_int32), ("coarsecorrection", ctypes.c_int32),
                ("acorrection", ctypes.c_int32), ("order", ctypes.c_int32), ("microcoeff", ctypes.c_double),
                ("h", ctypes.c_double), ("omegam", ctypes.c_double), ("lboxmpch", ctypes.c_double),
                ("mpc", ctypes.c_double), ("allocation", ctypes.c_uint64)]


MPC_SI = 3.0856775814913673e22


def prep_config(isfullsky=1, coarseonly=0, ncoarse=0, correction=1, coarsecorrection=0, acorrection=0, order=TSC,
                microcoeff=16.0, h=0.7, omegam=0.3, lboxmpch=1000.0, mpc=MPC_SI, allocation=0):
    return PrepConfig(int(isfullsky), int(coarseonly), int(ncoarse), int(correction), int(coarsecorrection),
                      int(acorrection), int(order), float(microcoeff), float(h), float(omegam), float(lboxmpch),
                      float(mpc), int(allocation))


def _prep_check(rc, name):
    if rc != 0:
        raise MpfError(f"{name}: {load_library().mpf_prep_last_error().decode()} ({rc})")


class Prep:
    """Octree builder of the create_octree ASCII branch (mpf_prep_*)."""

    def __init__(self, **cfg):
        self.lib = load_library()
        self.cfg = prep_config(**cfg)
        self.h = ctypes.c_void_p()
        _prep_check(self.lib.mpf_prep_create(ctypes.byref(self.h), ctypes.byref(self.cfg)), "mpf_prep_create")

    def close(self):
        if self.h:
            self.lib.mpf_prep_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clear(self):
        _prep_check(self.lib.mpf_prep_clear(self.h), "mpf_prep_clear")

    def import_ascii(self, path, cone7, rotm1=None):
        n = ctypes.c_uint64(0)
        rot = None if rotm1 is None else _f64(rotm1, 9)
        c7 = _f64(cone7, 7)
        _prep_check(self.lib.mpf_prep_import_ascii(self.h, str(path).encode(), _p(c7), _p(rot), ctypes.byref(n)),
                    "mpf_prep_import_ascii")
        return n.value

    def load_records(self, rec):
        rec = np.ascontiguousarray(rec)
        _prep_check(self.lib.mpf_prep_load_records(self.h, _p(rec), rec.size), "mpf_prep_load_records")

    def update(self):
        _prep_check(self.lib.mpf_prep_update(self.h), "mpf_prep_update")

    def interpolate(self, order, xyz):
        xyz = _f64(xyz, (-1, 3))
        out = np.zeros((xyz.shape[0], 7), dtype=np.float32)
        _prep_check(self.lib.mpf_prep_interpolate(self.h, int(order), _p(xyz), xyz.shape[0], _p(out)),
                    "mpf_prep_interpolate")
        return out

    def correct(self, table):
        t, a, dadt = (_f64(table[k], -1) for k in range(3))
        _prep_check(self.lib.mpf_prep_correct(self.h, _p(t), _p(a), _p(dadt), t.size), "mpf_prep_correct")

    def records(self):
        from . import formats
        n = ctypes.c_uint64(0)
        _prep_check(self.lib.mpf_prep_size(self.h, ctypes.byref(n)), "mpf_prep_size")
        rec = np.zeros(n.value, dtype=formats.RECORD_DTYPE)
        if n.value:
            _prep_check(self.lib.mpf_prep_records(self.h, _p(rec), n.value), "mpf_prep_records")
        return rec

    def save(self, path):
        _prep_check(self.lib.mpf_prep_save(self.h, str(path).encode()), "mpf_prep_save")


def prep_keys(level, xyz):
    """Keys of cells from float positions (the ASCII importer's Index::compute<float>): n x 2 uint64 (lo, hi)."""
    level = np.ascontiguousarray(level, dtype=np.uint32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    out = np.zeros((xyz.shape[0], 2), dtype=np.uint64)
    _prep_check(load_library().mpf_prep_keys(_p(level), _p(xyz), xyz.shape[0], _p(out)), "mpf_prep_keys")
    return out


def prep_collide(key2, micro_radius, cone7):
    key2 = np.ascontiguousarray(key2, dtype=np.uint64).reshape(-1, 2)
    out = np.zeros(key2.shape[0], dtype=np.uint8)
    c7 = _f64(cone7, 7)
    _prep_check(load_library().mpf_prep_collide(_p(key2), key2.shape[0], float(micro_radius), _p(c7), _p(out)),
                "mpf_prep_collide")
    return out.astype(bool)


def prepare_ascii(celldir, cones_rot7, conefiles, table, minicone=None, rotm1=None, ntasks=1, rank=0, seed=42,
                  **cfg):
    """Create_octree::PreparationASCII: one cone file per cone of this rank."""
    c = prep_config(**cfg)
    cones = _f64(cones_rot7, (-1, CONE_N))
    t, a, dadt = (_f64(table[k], -1) for k in range(3))
    names = (ctypes.c_char_p * len(conefiles))(*[str(f).encode() for f in conefiles])
    rot = None if rotm1 is None else _f64(rotm1, 9)
    cdir, mdir = _dir(celldir), (_dir(minicone) if minicone else None)
    _prep_check(load_library().mpf_prepare_ascii(ctypes.byref(c), cdir, mdir, _p(cones), cones.shape[0], _p(rot), ntasks,
                                                 rank, seed, _p(t), _p(a), _p(dadt), t.size, names),
                "mpf_prepare_ascii")


def catalogue_write(prefix, out, id_npart):
    """relCat's three files (<prefix>.txt, _err.txt, _reject.txt) from
    mpf_catalogue_solve rows and the targets' (id, npart)."""
    out = np.ascontiguousarray(out, dtype=np.float64).reshape(-1, NCAT)
    idn = np.ascontiguousarray(id_npart, dtype=np.float64).reshape(-1, 2)
    rc = load_library().mpf_catalogue_write(str(prefix).encode(), _p(out), ctypes.c_uint64(out.shape[0]), _p(idn))
    if rc:
        raise MpfError(f"mpf_catalogue_write failed ({rc})")


def catalogue_write_previous(prefix, rows18):
    rows = np.ascontiguousarray(rows18, dtype=np.float64).reshape(-1, 18)
    rc = load_library().mpf_catalogue_write_previous(str(prefix).encode(), _p(rows), ctypes.c_uint64(rows.shape[0]))
    if rc:
        raise MpfError(f"mpf_catalogue_write_previous failed ({rc})")


def _dir(d):
    d = str(d)
    return (d if d.endswith("/") else d + "/").encode()


def _cone_check(rc, name):
    if rc != 0:
        raise RuntimeError(f"{name}: {load_library().mpf_cone_last_error().decode()} ({rc})")


def _f64(a, shape):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(shape)


def cones_fullsky(ncones, radius=2.0):
    """GenerateFullskyCones (miscellaneous.h:200-218): (ncones, 7) cones."""
    out = np.zeros((ncones, CONE_N))
    _cone_check(load_library().mpf_cones_fullsky(ncones, radius, _p(out)), "mpf_cones_fullsky")
    return out


def device_memory(device=0):
    """(free, total) bytes of device memory (hipMemGetInfo)."""
    f, t = ctypes.c_uint64(0), ctypes.c_uint64(0)
    rc = load_library().mpf_device_memory(int(device), ctypes.byref(f), ctypes.byref(t))
    if rc:
        raise MpfError(f"mpf_device_memory failed ({rc})")
    return f.value, t.value


def cones_narrow(ncones, phi_deg, theta_deg, thetay_deg, thetaz_deg, radius=2.0):
    """GenerateNarrowCones (miscellaneous.h:257-402): (cones, cones rotated
    into the octree frame, rotm1 3x3, spiral size)."""
    c, cr, rot = np.zeros((ncones, CONE_N)), np.zeros((ncones, CONE_N)), np.zeros(9)
    size = ctypes.c_uint64(0)
    _cone_check(load_library().mpf_cones_narrow(ncones, radius, phi_deg, theta_deg, thetay_deg, thetaz_deg, _p(c), _p(cr),
                                        _p(rot), ctypes.byref(size)), "mpf_cones_narrow")
    return c, cr, rot.reshape(3, 3), size.value


def narrow_cone_info(path):
    """info_narrow_cone.txt -> (phi, theta, thetay, thetaz) in degrees."""
    out = np.zeros(4)
    _cone_check(load_library().mpf_narrow_cone_info(str(path).encode(), _p(out)), "mpf_narrow_cone_info")
    return tuple(float(v) for v in out)


def cone_inside(cone7, xyz):
    xyz = _f64(xyz, (-1, 3))
    out = np.zeros(xyz.shape[0], dtype=np.uint8)
    _cone_check(load_library().mpf_cone_inside(_p(_f64(cone7, 7)), _p(xyz), xyz.shape[0], 3, _p(out)), "mpf_cone_inside")
    return out.astype(bool)


def cone_select(cones7, icone, xyz, by_value=False):
    cones7 = _f64(cones7, (-1, CONE_N))
    xyz = _f64(xyz, (-1, 3))
    out = np.zeros(xyz.shape[0], dtype=np.uint8)
    _cone_check(load_library().mpf_cone_select(_p(cones7), cones7.shape[0], icone, int(by_value), _p(xyz), xyz.shape[0], 3,
                                       _p(out)), "mpf_cone_select")
    return out.astype(bool)


def cone_targets(cones7, icone, targets):
    """getTargets (miscellaneous.h:523-600): indices of the owned targets
    (rows whose first three columns are positions)."""
    cones7 = _f64(cones7, (-1, CONE_N))
    targets = np.ascontiguousarray(targets, dtype=np.float64)
    n, stride = targets.shape
    idx = np.zeros(n, dtype=np.uint64)
    k = ctypes.c_uint64(0)
    _cone_check(load_library().mpf_cone_targets(_p(cones7), cones7.shape[0], icone, _p(targets), n, stride, _p(idx),
                                        ctypes.byref(k)), "mpf_cone_targets")
    return idx[:k.value].astype(np.int64)


def hmaps_cone_pixels(nside, cones7, icone):
    """getPixels_per_cone2 (hmaps.h:261-322): owned RING pixels, ascending."""
    cones7 = _f64(cones7, (-1, CONE_N))
    k = ctypes.c_uint64(0)
    _cone_check(load_library().mpf_hmaps_cone_pixels(nside, _p(cones7), cones7.shape[0], icone, None, 0, ctypes.byref(k)),
                "mpf_hmaps_cone_pixels")
    pix = np.zeros(max(1, k.value), dtype=np.int64)
    _cone_check(load_library().mpf_hmaps_cone_pixels(nside, _p(cones7), cones7.shape[0], icone, _p(pix), k.value,
                                             ctypes.byref(k)), "mpf_hmaps_cone_pixels")
    return pix[:k.value]


MAP_NAMES = {"lensing": 0, "lensing_born": 1, "dr": 2, "dl": 3, "dt": 4, "da": 5, "dz": 6, "ds": 7, "isw": 8,
             "dens": 9, "dens_max": 10, "steps": 11, "phi": 12, "deflection": 13}


def hmaps_homogeneous(ref_rows, z_stop, stop_ray="redshift"):
    """hmaps.cpp:297-325: (nz x (t, chi, lambda, redshift, a), surface values of stop_ray)."""
    ref_rows = np.ascontiguousarray(ref_rows, dtype=np.float64).reshape(-1, NOUT)
    z = np.ascontiguousarray(z_stop, dtype=np.float64).ravel()
    homo, iref = np.zeros((z.size, 5)), np.zeros(z.size)
    k = STOP_NAMES[stop_ray] if isinstance(stop_ray, str) else int(stop_ray)
    if load_library().mpf_hmaps_homogeneous(_p(ref_rows), ref_rows.shape[0], _p(z), z.size, k, _p(homo), _p(iref)):
        raise MpfError("mpf_hmaps_homogeneous: invalid arguments (a redshift beyond the reference ray?)")
    return homo, iref


def healpix_write_fits(path, hmap, nside, nest=False, coordsys="G", clean=True):
    hmap = np.ascontiguousarray(hmap, dtype=np.float32).ravel()
    if hmap.size != 12 * nside * nside:
        raise ValueError("map size is not 12 nside^2")
    if load_library().mpf_healpix_write_fits(str(path).encode(), _p(hmap), int(nside), int(bool(nest)),
                                             coordsys.encode(), int(bool(clean))):
        raise MpfError("mpf_healpix_write_fits failed")


STAT_NAMES = {"distance": 0, "distance2": 1, "homogeneous": 2, "inhomogeneous": 3, "invdistance": 4,
              "invdistance2": 5}


def ray_statistics(trajs, dists, ref_rows, ref_dist, interp="redshift", stat="distance", nstat=100, amin=0.0,
                   length=1.0):
    """rays.cpp statistics over propagate's central trajectories: (rows of
    (z, reference statistic, mean, std), rays used)."""
    n = len(trajs)
    cap = max([len(t) for t in trajs] + [1])
    rows = np.zeros((n, cap, NOUT))
    dist = np.zeros((n, cap))
    ne = np.zeros(n, dtype=np.uint64)
    for i, (t, d) in enumerate(zip(trajs, dists)):
        rows[i, :len(t)] = t
        dist[i, :len(t)] = d
        ne[i] = len(t)
    ref_rows = np.ascontiguousarray(ref_rows, dtype=np.float64).reshape(-1, NOUT)
    ref_dist = np.ascontiguousarray(ref_dist, dtype=np.float64).ravel()
    ik = STOP_NAMES[interp] if isinstance(interp, str) else int(interp)
    sk = STAT_NAMES[stat] if isinstance(stat, str) else int(stat)
    nout, nused = ctypes.c_uint64(0), ctypes.c_uint64(0)
    args = (_p(rows), _p(dist), _p(ne), n, cap, _p(ref_rows), _p(ref_dist), ref_rows.shape[0], ik, sk, nstat, amin,
            length)
    lib = load_library()
    if lib.mpf_ray_statistics(*args, None, 0, ctypes.byref(nout), ctypes.byref(nused)):
        raise MpfError("mpf_ray_statistics: invalid arguments")
    out = np.zeros((max(1, nout.value), 4))
    if lib.mpf_ray_statistics(*args, _p(out), nout.value, ctypes.byref(nout), ctypes.byref(nused)):
        raise MpfError("mpf_ray_statistics: invalid arguments")
    return out[:nout.value], nused.value


def ray_statistics_write(path, out4, count):
    """The statistics file of rays.cpp (Output::save) from ray_statistics' rows."""
    out4 = np.ascontiguousarray(out4, dtype=np.float64).reshape(-1, 4)
    if load_library().mpf_ray_statistics_write(str(path).encode(), _p(out4), ctypes.c_uint64(out4.shape[0]),
                                               ctypes.c_uint64(int(count))):
        raise MpfError(f"mpf_ray_statistics_write: cannot write {path}")


def launch_angles(phi_theta, observer=(0.0, 0.0, 0.0)):
    pt = _f64(phi_theta, (-1, 2))
    out = np.zeros((pt.shape[0], 9))
    _cone_check(load_library().mpf_launch_angles(_p(_f64(observer, 3)), _p(pt), pt.shape[0], _p(out)), "mpf_launch_angles")
    return out


def launch_towards(begin, end):
    b, e = _f64(begin, (-1, 3)), _f64(end, (-1, 3))
    out = np.zeros((b.shape[0], 9))
    _cone_check(load_library().mpf_launch_towards(_p(b), _p(e), b.shape[0], _p(out)), "mpf_launch_towards")
    return out


def launch_cone(cones7, icone, n, seed, micro_radius=1.0):
    """rays.cpp:233-239: (in9, bundle rotation) of n random rays in cone icone."""
    cones7 = _f64(cones7, (-1, CONE_N))
    in9, rot = np.zeros((n, 9)), np.zeros(n)
    _cone_check(load_library().mpf_launch_cone(_p(cones7), cones7.shape[0], icone, micro_radius, seed, n, _p(in9), _p(rot)),
                "mpf_launch_cone")
    return in9, rot


def launch_bundle(in9, count, angle, rotation=0.0, center=True):
    out = np.zeros((count + bool(center), 9))
    _cone_check(load_library().mpf_launch_bundle(_p(_f64(in9, 9)), count, angle, rotation, int(bool(center)), _p(out)),
                "mpf_launch_bundle")
    return out


def _cosmology_call(fn, args, name):
    """Runs an mpf_cosmology_* table function twice (size, then fill)."""
    lib = load_library()
    h, om, lb = (ctypes.c_double(0.0) for _ in range(3))
    n = ctypes.c_uint64(0)
    null = ctypes.c_void_p(None)
    rc = fn(*args, ctypes.byref(h), ctypes.byref(om), ctypes.byref(lb), null, null, null, null, ctypes.c_uint64(0),
            ctypes.byref(n))
    if rc != 0:
        raise MpfError(f"{name}: {rc} {lib.mpf_cosmology_last_error().decode()}")
    cols = np.zeros((4, n.value))
    rc = fn(*args, ctypes.byref(h), ctypes.byref(om), ctypes.byref(lb), _p(cols[0]), _p(cols[1]), _p(cols[2]),
            _p(cols[3]), ctypes.c_uint64(n.value), ctypes.byref(n))
    if rc != 0:
        raise MpfError(f"{name}: {rc} {lib.mpf_cosmology_last_error().decode()}")
    return cols, h.value, om.value, lb.value


def cosmology_acquire(evolution_file, parameter_file, cell_dir, mpc=3.08568e22):
    """Input::acquire: ((t, a, dadt, d2adt2) rows, h, omega_m, lbox_mpch)."""
    lib = load_library()
    return _cosmology_call(lib.mpf_cosmology_acquire, (str(evolution_file).encode(), str(parameter_file).encode(),
                                                       str(cell_dir).encode(), float(mpc)), "mpf_cosmology_acquire")


def cosmology_correct(t, a, traj_t, traj_a):
    """Input::correct: the corrected a column."""
    lib = load_library()
    t = np.ascontiguousarray(t, dtype=np.float64)
    a = np.array(a, dtype=np.float64)
    tt = np.ascontiguousarray(traj_t, dtype=np.float64)
    ta = np.ascontiguousarray(traj_a, dtype=np.float64)
    rc = lib.mpf_cosmology_correct(_p(t), _p(a), ctypes.c_uint64(t.size), _p(tt), _p(ta), ctypes.c_uint64(tt.size))
    if rc != 0:
        raise MpfError(f"mpf_cosmology_correct: {rc}")
    return a


def cosmology_prepare(evolution_file, parameter_file, cell_dir, mpc=3.08568e22, device=0, precision=0):
    """rays.cpp:200-215: acquire, the homogeneous reference photon, correct."""
    lib = load_library()
    return _cosmology_call(lib.mpf_cosmology_prepare, (int(device), int(precision), str(evolution_file).encode(),
                                                       str(parameter_file).encode(), str(cell_dir).encode(),
                                                       float(mpc)), "mpf_cosmology_prepare")


def make_stop(kind="none", value=0.0, kitarget=(0.0, 0.0, 0.0)):
    if isinstance(kind, str):
        kind = STOP_NAMES[kind]
    s = Stop()
    s.kind = int(kind)
    s.value = float(value)
    for i in range(3):
        s.kitarget[i] = float(kitarget[i])
    return s


class Context:
    """One HIP device, one octree, one cosmology table (mpf_ctx)."""

    def __init__(self, length, order=TSC, precision=PREC_REF, nsteps=10, device=0):
        self.lib = load_library()
        cfg = Config(device, order, precision, nsteps, float(length))
        h = ctypes.c_void_p()
        rc = self.lib.mpf_create(ctypes.byref(h), ctypes.byref(cfg))
        if rc:
            raise MpfError(f"mpf_create failed ({rc})")
        self.h = h
        self.order = order
        self.precision = precision

    def close(self):
        if getattr(self, "h", None):
            self.lib.mpf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            msg = self.lib.mpf_last_error(self.h).decode()
            raise MpfError(f"{what} failed ({rc}): {msg}")

    def set_precision(self, precision):
        self._check(self.lib.mpf_set_precision(self.h, int(precision)), "mpf_set_precision")
        self.precision = precision

    def load_records(self, rec):
        rec = np.ascontiguousarray(rec)
        self._check(self.lib.mpf_load_octree_records(self.h, _p(rec), ctypes.c_uint64(rec.size)),
                    "mpf_load_octree_records")

    def load_fields64(self, f7):
        f7 = np.ascontiguousarray(f7, dtype=np.float64)
        self._check(self.lib.mpf_load_octree_fields64(self.h, _p(f7), ctypes.c_uint64(f7.shape[0])),
                    "mpf_load_octree_fields64")

    # ---- velocity field (the reference's velocity-field build, hmaps.cpp:359-441)
    VELOCITY_SCHEMES = {"cic": 1, "tsc": 2}

    def velocity_field(self, particles, scheme="cic"):
        """Deposits particle sets into the tree and finishes the field:
        `particles` is an iterable of (pos3, vel3) float arrays (one per
        particle file), `scheme` "cic" or "tsc" (the parameter velocity_field)."""
        if scheme not in self.VELOCITY_SCHEMES:
            raise ValueError("velocity_field must be cic or tsc")
        self._check(self.lib.mpf_velocity_begin(self.h), "mpf_velocity_begin")
        for pos, vel in particles:
            pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
            vel = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 3)
            if pos.shape != vel.shape:
                raise ValueError("positions and velocities differ in count")
            self._check(self.lib.mpf_velocity_deposit(self.h, _p(pos), _p(vel), ctypes.c_uint64(pos.shape[0]),
                                                      self.VELOCITY_SCHEMES[scheme]), "mpf_velocity_deposit")
        self._check(self.lib.mpf_velocity_finish(self.h), "mpf_velocity_finish")

    def velocity_begin(self):
        self._check(self.lib.mpf_velocity_begin(self.h), "mpf_velocity_begin")

    def velocity_deposit_dev(self, pos_ptr, vel_ptr, n, scheme="cic"):
        """One particle set already in device memory (3 floats per particle each)."""
        self._check(self.lib.mpf_velocity_deposit_dev(self.h, ctypes.c_void_p(pos_ptr), ctypes.c_void_p(vel_ptr),
                                                      ctypes.c_uint64(n), self.VELOCITY_SCHEMES[scheme]),
                    "mpf_velocity_deposit_dev")

    def velocity_finish(self):
        self._check(self.lib.mpf_velocity_finish(self.h), "mpf_velocity_finish")

    def velocity_set(self, vel3, rho=None):
        vel3 = np.ascontiguousarray(vel3, dtype=np.float32).reshape(-1, 3)
        r = None if rho is None else np.ascontiguousarray(rho, dtype=np.float32)
        self._check(self.lib.mpf_velocity_set(self.h, _p(r), _p(vel3), ctypes.c_uint64(vel3.shape[0])),
                    "mpf_velocity_set")

    def velocity_get(self):
        """(rho, vel3) in record order."""
        n = self.octree_info()[0]
        rho = np.zeros(n, dtype=np.float32)
        vel = np.zeros((n, 3), dtype=np.float32)
        self._check(self.lib.mpf_velocity_get(self.h, _p(rho), _p(vel), ctypes.c_uint64(n)), "mpf_velocity_get")
        return rho, vel

    def velocity_clear(self):
        self._check(self.lib.mpf_velocity_clear(self.h), "mpf_velocity_clear")

    def set_observer_velocity(self, vobs):
        v = np.ascontiguousarray(vobs, dtype=np.float64).reshape(3)
        self._check(self.lib.mpf_set_observer_velocity(self.h, _p(v)), "mpf_set_observer_velocity")

    def octree_info(self):
        n = ctypes.c_uint64(0)
        lo = ctypes.c_uint32(0)
        hi = ctypes.c_uint32(0)
        self._check(self.lib.mpf_octree_info(self.h, ctypes.byref(n), ctypes.byref(lo), ctypes.byref(hi)),
                    "mpf_octree_info")
        return n.value, lo.value, hi.value

    def set_cosmology(self, table):
        t = np.ascontiguousarray(table[0], dtype=np.float64)
        a = np.ascontiguousarray(table[1], dtype=np.float64)
        d = np.ascontiguousarray(table[2], dtype=np.float64)
        self._check(self.lib.mpf_set_cosmology(self.h, _p(t), _p(a), _p(d), ctypes.c_uint64(t.size)),
                    "mpf_set_cosmology")

    def locate(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        n = xyz.shape[0]
        elem = np.zeros(n, dtype=np.int64)
        lvl = np.zeros(n, dtype=np.uint32)
        self._check(self.lib.mpf_locate(self.h, _p(xyz), ctypes.c_uint64(n), _p(elem), _p(lvl)), "mpf_locate")
        return elem, lvl

    def interpolate(self, order, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        n = xyz.shape[0]
        out = np.zeros((n, 7))
        self._check(self.lib.mpf_interpolate(self.h, int(order), _p(xyz), ctypes.c_uint64(n), _p(out)),
                    "mpf_interpolate")
        return out

    def integrate(self, in9, stop="none", value=0.0, kitarget=(0.0, 0.0, 0.0)):
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(-1, 9)
        n = in9.shape[0]
        out = np.zeros((n, NOUT))
        steps = ctypes.c_uint64(0)
        st = make_stop(stop, value, kitarget)
        self._check(self.lib.mpf_integrate(self.h, _p(in9), ctypes.c_uint64(n), ctypes.byref(st), _p(out),
                                           ctypes.byref(steps)), "mpf_integrate")
        return out, steps.value

    def integrate_surfaces(self, in9, kind, values, jacobian=True, ray_value=None):
        """Rays to the last surface; returns (per-ray records, per-ray per-surface records).
        ray_value: optional per-ray surface/stop value (one surface per ray)."""
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(-1, 9)
        n = in9.shape[0]
        rv = None
        if ray_value is not None:
            rv = np.ascontiguousarray(ray_value, dtype=np.float64).ravel()
            values = [0.0]
        sf = make_surfaces(kind, values, jacobian, None if rv is None else rv.ctypes.data)
        out = np.zeros((n, NOUT))
        surf = np.zeros((n, sf.n, NSURF))
        self._check(self.lib.mpf_integrate_surfaces(self.h, _p(in9), ctypes.c_uint64(n), ctypes.byref(sf), _p(out),
                                                    _p(surf)), "mpf_integrate_surfaces")
        return out, surf

    def integrate_surfaces_born(self, in9, kind, values, jacobian=True, ray_value=None):
        """Surfaces in the Born approximation (homogeneous path, Jacobian on the octree's data)."""
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(-1, 9)
        n = in9.shape[0]
        rv = None
        if ray_value is not None:
            rv = np.ascontiguousarray(ray_value, dtype=np.float64).ravel()
            values = [0.0]
        sf = make_surfaces(kind, values, jacobian, None if rv is None else rv.ctypes.data)
        out = np.zeros((n, NOUT))
        surf = np.zeros((n, sf.n, NSURF))
        self._check(self.lib.mpf_integrate_surfaces_born(self.h, _p(in9), ctypes.c_uint64(n), ctypes.byref(sf),
                                                         _p(out), _p(surf)), "mpf_integrate_surfaces_born")
        return out, surf

    def integrate_surfaces_born_dev(self, in_ptr, n, out_ptr, surf_ptr, kind, values, jacobian=True, sync=False,
                                    ray_value_ptr=None):
        sf = make_surfaces(kind, values if ray_value_ptr is None else [0.0], jacobian, ray_value_ptr)
        self._check(self.lib.mpf_integrate_surfaces_born_dev(self.h, ctypes.c_void_p(in_ptr), ctypes.c_uint64(n),
                                                             ctypes.byref(sf), ctypes.c_void_p(out_ptr),
                                                             ctypes.c_void_p(surf_ptr), int(bool(sync))),
                    "mpf_integrate_surfaces_born_dev")

    def integrate_surfaces_dev(self, in_ptr, n, out_ptr, surf_ptr, kind, values, jacobian=True, sync=False,
                               ray_value_ptr=None):
        sf = make_surfaces(kind, values if ray_value_ptr is None else [0.0], jacobian, ray_value_ptr)
        self._check(self.lib.mpf_integrate_surfaces_dev(self.h, ctypes.c_void_p(in_ptr), ctypes.c_uint64(n),
                                                        ctypes.byref(sf), ctypes.c_void_p(out_ptr),
                                                        ctypes.c_void_p(surf_ptr), int(bool(sync))),
                    "mpf_integrate_surfaces_dev")

    def catalogue_solve(self, src6, init2=None, **newton):
        """Newton root-finding per source; returns (n x NCAT records, rays integrated)."""
        src6 = np.ascontiguousarray(src6, dtype=np.float64).reshape(-1, 6)
        n = src6.shape[0]
        i2 = None if init2 is None else np.ascontiguousarray(init2, dtype=np.float64).reshape(-1, 2)
        cfg = make_newton(**newton)
        out = np.zeros((n, NCAT))
        rays = ctypes.c_uint64(0)
        self._check(self.lib.mpf_catalogue_solve(self.h, _p(src6), _p(i2), ctypes.c_uint64(n), ctypes.byref(cfg),
                                                 _p(out), ctypes.byref(rays)), "mpf_catalogue_solve")
        return out, rays.value

    def catalogue_solve_dev(self, src_ptr, init_ptr, n, out_ptr, **newton):
        cfg = make_newton(**newton)
        rays = ctypes.c_uint64(0)
        self._check(self.lib.mpf_catalogue_solve_dev(self.h, ctypes.c_void_p(src_ptr), ctypes.c_void_p(init_ptr),
                                                     ctypes.c_uint64(n), ctypes.byref(cfg), ctypes.c_void_p(out_ptr),
                                                     ctypes.byref(rays)), "mpf_catalogue_solve_dev")
        return rays.value

    def catalogue_previous(self, rows18, **newton):
        """relCat_with_previous_cat: returns a copy of the rows with the
        Jacobian columns 13-16 recomputed."""
        rows = np.array(rows18, dtype=np.float64).reshape(-1, 18)
        cfg = make_newton(**newton)
        self._check(self.lib.mpf_catalogue_previous(self.h, _p(rows), ctypes.c_uint64(rows.shape[0]),
                                                    ctypes.byref(cfg)), "mpf_catalogue_previous")
        return rows

    # learned first guess of the catalogue root-finder
    def mlp_init(self, seed=0):
        self._check(self.lib.mpf_mlp_init(self.h, ctypes.c_uint64(seed)), "mpf_mlp_init")

    def mlp_state(self, set_state=None):
        """Returns the MLP_NSTATE-float state; replaces it first when set_state is given."""
        get = np.zeros(MLP_NSTATE, dtype=np.float32)
        st = None if set_state is None else np.ascontiguousarray(set_state, dtype=np.float32).reshape(MLP_NSTATE)
        self._check(self.lib.mpf_mlp_state(self.h, _p(get), _p(st)), "mpf_mlp_state")
        return get

    def mlp_train(self, feat, target, epochs=10, batch=1024, lr=1e-3, seed=0):
        """Adam epochs over (feat, target); returns the mean loss of each epoch."""
        feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1, MLP_NIN)
        target = np.ascontiguousarray(target, dtype=np.float32).reshape(-1, MLP_NOUT)
        assert feat.shape[0] == target.shape[0]
        loss = np.zeros(max(int(epochs), 1), dtype=np.float32)
        self._check(self.lib.mpf_mlp_train(self.h, _p(feat), _p(target), ctypes.c_uint64(feat.shape[0]),
                                           int(epochs), int(batch), float(lr), ctypes.c_uint64(seed), _p(loss)),
                    "mpf_mlp_train")
        return loss[:int(epochs)]

    def mlp_gradient(self, feat, target):
        feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1, MLP_NIN)
        target = np.ascontiguousarray(target, dtype=np.float32).reshape(-1, MLP_NOUT)
        grad = np.zeros(MLP_NPARAM, dtype=np.float32)
        loss = ctypes.c_float(0.0)
        self._check(self.lib.mpf_mlp_gradient(self.h, _p(feat), _p(target), ctypes.c_uint64(feat.shape[0]),
                                              _p(grad), ctypes.byref(loss)), "mpf_mlp_gradient")
        return grad, loss.value

    def mlp_infer(self, feat):
        feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(-1, MLP_NIN)
        out = np.zeros((feat.shape[0], MLP_NOUT), dtype=np.float32)
        self._check(self.lib.mpf_mlp_infer(self.h, _p(feat), ctypes.c_uint64(feat.shape[0]), _p(out)),
                    "mpf_mlp_infer")
        return out

    def mlp_infer_dev(self, feat_ptr, n, out_ptr):
        self._check(self.lib.mpf_mlp_infer_dev(self.h, ctypes.c_void_p(feat_ptr), ctypes.c_uint64(n),
                                               ctypes.c_void_p(out_ptr)), "mpf_mlp_infer_dev")

    def catalogue_features(self, src6):
        src6 = np.ascontiguousarray(src6, dtype=np.float64).reshape(-1, 6)
        feat = np.zeros((src6.shape[0], MLP_NIN), dtype=np.float32)
        self._check(self.lib.mpf_catalogue_features(self.h, _p(src6), ctypes.c_uint64(src6.shape[0]), _p(feat)),
                    "mpf_catalogue_features")
        return feat

    def catalogue_features_dev(self, src_ptr, n, feat_ptr):
        self._check(self.lib.mpf_catalogue_features_dev(self.h, ctypes.c_void_p(src_ptr), ctypes.c_uint64(n),
                                                        ctypes.c_void_p(feat_ptr)), "mpf_catalogue_features_dev")

    def catalogue_seed(self, src6, use_network=True):
        src6 = np.ascontiguousarray(src6, dtype=np.float64).reshape(-1, 6)
        init2 = np.zeros((src6.shape[0], 2))
        self._check(self.lib.mpf_catalogue_seed(self.h, _p(src6), ctypes.c_uint64(src6.shape[0]),
                                                int(bool(use_network)), _p(init2)), "mpf_catalogue_seed")
        return init2

    def catalogue_seed_dev(self, src_ptr, n, init_ptr, use_network=True):
        self._check(self.lib.mpf_catalogue_seed_dev(self.h, ctypes.c_void_p(src_ptr), ctypes.c_uint64(n),
                                                    int(bool(use_network)), ctypes.c_void_p(init_ptr)),
                    "mpf_catalogue_seed_dev")

    def mfma_probe(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.float32).reshape(16, 16)
        b = np.ascontiguousarray(b, dtype=np.float32).reshape(16, 16)
        c = np.zeros((16, 16), dtype=np.float32)
        self._check(self.lib.mpf_mfma_probe(self.h, _p(a), _p(b), _p(c)), "mpf_mfma_probe")
        return c

    def mfma_probe32(self, a, b, layout=0):
        """16 x 32 by 32 x 16 product through one 16x16x32 BF16 MFMA with the
        given operand layout (0: the kernels'; 1: the K-halves alternative)."""
        a = np.ascontiguousarray(a, dtype=np.float32).reshape(16, 32)
        b = np.ascontiguousarray(b, dtype=np.float32).reshape(32, 16)
        c = np.zeros((16, 16), dtype=np.float32)
        self._check(self.lib.mpf_mfma_probe32(self.h, _p(a), _p(b), int(layout), _p(c)), "mpf_mfma_probe32")
        return c

    def mlp_last_ms(self):
        """Duration (ms) of the last inference kernel."""
        ms = ctypes.c_double(-1.0)
        self._check(self.lib.mpf_mlp_last_ms(self.h, ctypes.byref(ms)), "mpf_mlp_last_ms")
        return ms.value

    def hmaps_launch_dev(self, nside, pix_ptr, n, in9_ptr):
        self._check(self.lib.mpf_hmaps_launch_dev(self.h, int(nside), ctypes.c_void_p(pix_ptr), ctypes.c_uint64(n),
                                                  ctypes.c_void_p(in9_ptr)), "mpf_hmaps_launch_dev")

    def hmaps_fill_dev(self, nside, pix_ptr, surf_ptr, n, nsurf, maps_ptr):
        self._check(self.lib.mpf_hmaps_fill_dev(self.h, int(nside), ctypes.c_void_p(pix_ptr), ctypes.c_void_p(surf_ptr),
                                                ctypes.c_uint64(n), ctypes.c_uint32(nsurf), ctypes.c_void_p(maps_ptr)),
                    "mpf_hmaps_fill_dev")

    def healpix_bin_dev(self, nside, in9_ptr, surf_ptr, n, nsurf, isurf, maps_ptr):
        self._check(self.lib.mpf_healpix_bin_dev(self.h, int(nside), ctypes.c_void_p(in9_ptr),
                                                 ctypes.c_void_p(surf_ptr), ctypes.c_uint64(n), ctypes.c_uint32(nsurf),
                                                 ctypes.c_uint32(isurf), ctypes.c_void_p(maps_ptr)),
                    "mpf_healpix_bin_dev")
        return self.last_ms()

    def last_ms(self):
        ms = ctypes.c_double(0.0)
        steps = ctypes.c_uint64(0)
        self.lib.mpf_last_run(self.h, ctypes.byref(steps), ctypes.byref(ms))
        return ms.value

    def integrate_dev(self, in_ptr, n, out_ptr, stop="none", value=0.0, kitarget=(0.0, 0.0, 0.0), sync=False):
        """Device-resident integration: in_ptr/out_ptr are device addresses (ints)."""
        st = make_stop(stop, value, kitarget)
        self._check(self.lib.mpf_integrate_dev(self.h, ctypes.c_void_p(in_ptr), ctypes.c_uint64(n),
                                               ctypes.byref(st), ctypes.c_void_p(out_ptr), int(bool(sync))),
                    "mpf_integrate_dev")

    def integrate_dev_static(self, in_ptr, n, out_ptr, stop="none", value=0.0, kitarget=(0.0, 0.0, 0.0)):
        """mpf_integrate_dev_static: the device-resident integration with no
        device allocation (workspace in a fixed per-GPU array); synchronous,
        returns the total step count."""
        st = make_stop(stop, value, kitarget)
        steps = ctypes.c_uint64(0)
        self._check(self.lib.mpf_integrate_dev_static(self.h, ctypes.c_void_p(in_ptr), ctypes.c_uint64(n),
                                                      ctypes.byref(st), ctypes.c_void_p(out_ptr),
                                                      ctypes.byref(steps)), "mpf_integrate_dev_static")
        return steps.value

    def trajectory(self, in9, stop="none", value=0.0, kitarget=(0.0, 0.0, 0.0), capacity=100000):
        """Whole trajectories: (out n x NOUT, list of per-ray element arrays)."""
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(-1, 9)
        n = in9.shape[0]
        st = make_stop(stop, value, kitarget)
        out = np.zeros((n, NOUT))
        traj = np.zeros((n, capacity, NOUT))
        self._check(self.lib.mpf_integrate_trajectory(self.h, _p(in9), ctypes.c_uint64(n), ctypes.byref(st),
                                                      ctypes.c_uint64(capacity), _p(out), _p(traj)),
                    "mpf_integrate_trajectory")
        m = np.maximum(out[:, OUT["nsteps"]], -1).astype(np.int64) + 1
        return out, [traj[i, :m[i]] for i in range(n)]

    def set_launch_config(self, surfaces=False, **kw):
        """Set knobs of the persistent integrators' launch (refill, sort_bucket,
        blocks_per_cu, waves_per_simd); the others keep their value."""
        cfg = LaunchCfg()
        self._check(self.lib.mpf_get_launch_config(self.h, int(bool(surfaces)), ctypes.byref(cfg)),
                    "mpf_get_launch_config")
        for k, v in kw.items():
            setattr(cfg, k, int(v))
        self._check(self.lib.mpf_set_launch_config(self.h, int(bool(surfaces)), ctypes.byref(cfg)),
                    "mpf_set_launch_config")

    def launch_config(self, surfaces=False):
        cfg = LaunchCfg()
        self._check(self.lib.mpf_get_launch_config(self.h, int(bool(surfaces)), ctypes.byref(cfg)),
                    "mpf_get_launch_config")
        return cfg.as_dict()

    def autotune_dev(self, in_ptr, n, out_ptr, stop="none", value=0.0, surf_ptr=None, kind=None, values=None,
                     jacobian=True, ray_value_ptr=None, max_trials=64):
        """Tune the launch configuration on a device-resident batch (with
        surfaces when surf_ptr is given); returns (best config dict, trials
        array of rows refill, sort_bucket, blocks_per_cu, waves_per_simd, ms, steps)."""
        st = make_stop(stop, value)
        sf = None
        if surf_ptr is not None:
            sf = make_surfaces(kind, values if ray_value_ptr is None else [0.0], jacobian, ray_value_ptr)
        best = LaunchCfg()
        trials = np.zeros((max_trials, 6))
        nt = ctypes.c_uint32(0)
        self._check(self.lib.mpf_autotune(self.h, ctypes.c_void_p(in_ptr), ctypes.c_uint64(n), ctypes.byref(st),
                                          ctypes.byref(sf) if sf is not None else None, ctypes.c_void_p(out_ptr),
                                          ctypes.c_void_p(surf_ptr) if surf_ptr is not None else None,
                                          ctypes.byref(best), _p(trials), max_trials, ctypes.byref(nt)),
                    "mpf_autotune")
        return best.as_dict(), trials[:min(nt.value, max_trials)]

    def last_run(self):
        steps = ctypes.c_uint64(0)
        ms = ctypes.c_double(0.0)
        self._check(self.lib.mpf_last_run(self.h, ctypes.byref(steps), ctypes.byref(ms)), "mpf_last_run")
        return steps.value, ms.value

    def propagate(self, in9, count, angle, rotation=None, interpolation="redshift", amin=0.0, capacity=20000):
        """Integrator::propagate for each central photon: (list of central
        trajectories, list of their angular diameter distances); an empty
        pair for a rejected bundle."""
        in9 = np.ascontiguousarray(in9, dtype=np.float64).reshape(-1, 9)
        n = in9.shape[0]
        kind = STOP_NAMES[interpolation] if isinstance(interpolation, str) else int(interpolation)
        rot = None if rotation is None else np.ascontiguousarray(rotation, dtype=np.float64).ravel()
        traj = np.zeros((n, capacity, NOUT))
        dist = np.zeros((n, capacity))
        ne = np.zeros(n, dtype=np.uint64)
        self._check(self.lib.mpf_propagate(self.h, _p(in9), ctypes.c_uint64(n), ctypes.c_uint32(count),
                                           ctypes.c_double(angle), _p(rot), int(kind), ctypes.c_double(amin),
                                           ctypes.c_uint64(capacity), _p(traj), _p(dist), _p(ne)), "mpf_propagate")
        return [traj[i, :ne[i]] for i in range(n)], [dist[i, :ne[i]] for i in range(n)]

    def jacobian_bundle(self, phi_theta, kind, ref, ki, dist, opening, observer=(0.0, 0.0, 0.0)):
        """Lensing::dbetadtheta per direction: (n, nd, 4) = a11, a12, a21, a22.
        ref: (nd,) values shared by every direction or (n, nd); ki (n, nd, 3);
        dist (n, nd)."""
        pt = np.ascontiguousarray(phi_theta, dtype=np.float64).reshape(-1, 2)
        n = pt.shape[0]
        ref = np.asarray(ref, dtype=np.float64)
        ref = np.ascontiguousarray(np.broadcast_to(ref, (n, ref.shape[-1])) if ref.ndim == 1 else ref)
        nd = ref.shape[1]
        ki = np.ascontiguousarray(ki, dtype=np.float64).reshape(n, nd, 3)
        dist = np.ascontiguousarray(dist, dtype=np.float64).reshape(n, nd)
        k = STOP_NAMES[kind] if isinstance(kind, str) else int(kind)
        out = np.zeros((n, nd, 4))
        self._check(self.lib.mpf_jacobian_bundle(self.h, _p(pt), ctypes.c_uint64(n),
                                                 _p(np.ascontiguousarray(observer, dtype=np.float64)), int(k),
                                                 _p(ref), ctypes.c_uint32(nd), _p(ki), _p(dist),
                                                 ctypes.c_double(opening), _p(out)), "mpf_jacobian_bundle")
        return out

    def split_tidal_blocks(self, enable=-1):
        """Allow (1) or free and stop (0) the split tier's tidal blocks; -1
        only queries.  Returns the number built."""
        v = ctypes.c_uint64(0)
        self._check(self.lib.mpf_split_tidal_blocks(self.h, int(enable), ctypes.byref(v)), "mpf_split_tidal_blocks")
        return int(v.value)

    def bundle_plane_info(self):
        """Directions of the last plane bundle redone from whole trajectories."""
        v = ctypes.c_uint64(0)
        self._check(self.lib.mpf_bundle_plane_info(self.h, ctypes.byref(v)), "mpf_bundle_plane_info")
        return int(v.value)

    def hmaps_nmaps(self, components, bundle=False):
        """Maps per surface for a component list (lensing: 4, or 5 with the bundle Jacobian; deflection: 2)."""
        return sum((5 if bundle else 4) if c in ("lensing", 0) else
                   (4 if c in ("lensing_born", 1) else (2 if c in ("deflection", 13) else 1)) for c in components)

    def hmaps_components_dev(self, nside, pix_ptr, surf_ptr, n, nz, components, homo5, maps_ptr, maps_capacity,
                             jac_ptr=None, surf_born_ptr=None):
        """Every map component of FillMap into FP32 maps (maps_capacity floats
        at maps_ptr); returns maps per surface."""
        comps = np.array([MAP_NAMES[c] if isinstance(c, str) else int(c) for c in components], dtype=np.int32)
        homo5 = np.ascontiguousarray(homo5, dtype=np.float64).reshape(nz, 5)
        nm = ctypes.c_uint32(0)
        self._check(self.lib.mpf_hmaps_components_dev(self.h, int(nside), ctypes.c_void_p(pix_ptr),
                                                      ctypes.c_void_p(surf_ptr) if surf_ptr else None,
                                                      ctypes.c_void_p(surf_born_ptr) if surf_born_ptr else None,
                                                      ctypes.c_void_p(jac_ptr) if jac_ptr else None,
                                                      ctypes.c_uint64(n), nz, _p(comps), comps.size, _p(homo5),
                                                      ctypes.c_void_p(maps_ptr), ctypes.c_uint64(maps_capacity),
                                                      ctypes.byref(nm)),
                    "mpf_hmaps_components_dev")
        return nm.value

    def last_kernel_ms(self):
        """Duration of the persistent integration kernel alone in the last run."""
        ms = ctypes.c_double(0.0)
        self._check(self.lib.mpf_last_kernel_ms(self.h, ctypes.byref(ms)), "mpf_last_kernel_ms")
        return ms.value

    def dev_alloc(self, nbytes):
        """Device buffer on the context's device; returns its address (int)."""
        ptr = ctypes.c_void_p()
        self._check(self.lib.mpf_dev_alloc(self.h, ctypes.c_uint64(int(nbytes)), ctypes.byref(ptr)), "mpf_dev_alloc")
        return ptr.value

    def dev_free(self, ptr):
        self._check(self.lib.mpf_dev_free(self.h, ctypes.c_void_p(ptr)), "mpf_dev_free")

    def upload(self, ptr, array):
        a = np.ascontiguousarray(array)
        self._check(self.lib.mpf_dev_upload(self.h, ctypes.c_void_p(ptr), _p(a), ctypes.c_uint64(a.nbytes)),
                    "mpf_dev_upload")

    def download(self, ptr, shape, dtype=np.float64):
        a = np.empty(shape, dtype=dtype)
        self._check(self.lib.mpf_dev_download(self.h, _p(a), ctypes.c_void_p(ptr), ctypes.c_uint64(a.nbytes)),
                    "mpf_dev_download")
        return a

    def dev_copy(self, dst, src, nbytes):
        """Plain streaming copy kernel (counter calibration); returns its duration in ms."""
        self._check(self.lib.mpf_dev_copy(self.h, ctypes.c_void_p(dst), ctypes.c_void_p(src),
                                          ctypes.c_uint64(int(nbytes))), "mpf_dev_copy")
        return self.last_run()[1]

    def synchronize(self):
        self._check(self.lib.mpf_synchronize(self.h), "mpf_synchronize")
