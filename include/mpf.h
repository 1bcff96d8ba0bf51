/* mpf — MI355X-native photon propagation for Magrathea-Pathfinder inputs.
 *
 * C ABI of libmpf.so.  Plain pointers and sizes; no HIP or torch types.
 * Every function returns an int status: 0 = OK, negative = error; the message
 * of the last error is returned by mpf_last_error(ctx).  No function
 * terminates the process (the reference calls std::terminate on bad
 * parameters, integrator.h:807-810).  A context is single-threaded: it owns
 * one HIP stream and its device allocations, and calls on one context must
 * not overlap.  Buffers are host memory unless the name says dev_.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * the Magrathea-Pathfinder v2.0 sources).
 */
#ifndef MPF_H
#define MPF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpf_ctx mpf_ctx;

/* interpolation order: the reference ORDER macro (Makefile:35-37), and
 * Order -1 of Integrator::integrate (integrator.h:505, 573): homogeneous
 * data (a = 1, every other field 0) wherever the photon is inside the box,
 * as for the reference photon of rays.cpp:211-214 (exact tiers only) */
enum { MPF_HOMOGENEOUS = -1, MPF_NGP = 0, MPF_CIC = 1, MPF_TSC = 2 };

/* precision tiers */
enum {
    /* FP32 fields, FP64 weights, FP32 accumulation in the reference's term
     * order, FP64 ray state: bit-identical to the reference
     * (simplehyperoctree.h:1870-1910 accumulates into Gravity<float>). */
    MPF_PREC_REF = 0,
    /* FP64 fields (mpf_load_octree_fields64, else the FP32 record fields
     * promoted), FP64 accumulation, FP64 state: the plain FP64 kernel. */
    MPF_PREC_FP64 = 1,
    /* FP32 perturbation (interpolated phi, grad phi, dphi/dt and the
     * perturbation terms of the geodesic equation), FP64 background (a, t,
     * position, affine parameter, redshift and distance accumulations). */
    MPF_PREC_SPLIT = 2
};

/* stop criteria of Integrator::integrate (integrator.h:779-811) */
enum {
    MPF_STOP_NONE = -1, /* integrator.h:477-618: run until the data ends */
    MPF_STOP_RADIUS = 0, MPF_STOP_LAMBDA = 1, MPF_STOP_REDSHIFT = 2, MPF_STOP_T = 3,
    MPF_STOP_A = 4, MPF_STOP_PLANE = 5, MPF_STOP_S = 6
};

typedef struct {
    int device;       /* HIP device ordinal */
    int order;        /* MPF_NGP / MPF_CIC / MPF_TSC */
    int precision;    /* MPF_PREC_* */
    unsigned nsteps;  /* integration steps per cell (parameter file `nsteps`) */
    double length;    /* box length in metres: lboxmpch * mpc / h (rays.cpp:208) */
} mpf_config;

typedef struct {
    int kind;          /* MPF_STOP_* */
    double value;      /* interpRef */
    double kitarget[3];/* plane normal for MPF_STOP_PLANE */
} mpf_stop;

/* Per-ray output: the last element of the trajectory after the final
 * pop (integrator.h:816-821), as 28 doubles. */
enum {
    MPF_OUT_A, MPF_OUT_T, MPF_OUT_X, MPF_OUT_Y, MPF_OUT_Z, MPF_OUT_KT, MPF_OUT_KX, MPF_OUT_KY, MPF_OUT_KZ,
    MPF_OUT_REDSHIFT, MPF_OUT_LAMBDA, MPF_OUT_S, MPF_OUT_CHI, MPF_OUT_ISW, MPF_OUT_ISWOLD, MPF_OUT_PHI,
    MPF_OUT_DPHIDX, MPF_OUT_DPHIDY, MPF_OUT_DPHIDZ, MPF_OUT_DPHIDT, MPF_OUT_RHO, MPF_OUT_AH, MPF_OUT_DPHIDL,
    MPF_OUT_LAPLACIAN, MPF_OUT_DSDL2, MPF_OUT_ERROR, MPF_OUT_LEVEL, MPF_OUT_NSTEPS, MPF_NOUT
};

/* Context lifecycle.  Replaces the per-process setup of rays.cpp:127-219. */
int mpf_create(mpf_ctx** ctx, const mpf_config* cfg);
int mpf_destroy(mpf_ctx* ctx);
const char* mpf_last_error(const mpf_ctx* ctx);
int mpf_set_precision(mpf_ctx* ctx, int precision);

/* Load n cone-file records of 48 bytes, sorted by key, as read by
 * Input::load (input.h:1298-1310, datahandler.h:940).  Replaces
 * Miscellaneous::loadOctree (miscellaneous.h:423-455).  The tree must be
 * parent-closed (every stored cell below the coarsest stored level has its
 * parent stored), which the reference's cone files are; else -EINVAL. */
int mpf_load_octree_records(mpf_ctx* ctx, const void* rec48, uint64_t n);
/* Optional FP64 copy of the fields, n x 7 in record order
 * (rho, phi, dphidx, dphidy, dphidz, a, dphidt), used by MPF_PREC_FP64 and as
 * the source of the FP32 perturbation fields of MPF_PREC_SPLIT. */
int mpf_load_octree_fields64(mpf_ctx* ctx, const double* f7, uint64_t n);
int mpf_octree_info(const mpf_ctx* ctx, uint64_t* ncells, uint32_t* min_level, uint32_t* max_level);
/* Host-only structural check of cone records, no device needed: keys well
 * formed, strictly sorted (the precondition of the reference's upper_bound
 * lookups, simplehyperoctree.h:894-899) and parent-closed.  Writes the
 * coarsest and finest stored levels; on failure a message into msg. */
int mpf_validate_octree(const void* rec48, uint64_t n, uint32_t* min_level, uint32_t* max_level, char* msg,
                        uint64_t msg_size);

/* Input::acquire (input.h:1091-1233), host only: the table (t, a, da/dt,
 * d2a/dt2; t the conformal time from the observer in seconds, increasing)
 * derived from a RAMSES evolution file (rows: a, H/H0, tau H0, t_lookback H0,
 * t_proper H0), with h and Omega_m from the parameter file ("key = value"
 * lines) and the box length (Mpc/h) from the cell directory's name
 * ("...boxlen<L>...", '-' read as the decimal point).  The same arithmetic
 * as the reference, so the table is bit-identical.  n receives the number of
 * rows; the columns are written when capacity >= n (call with capacity 0 to
 * size them).  -EINVAL with a message when a file or a value is missing. */
int mpf_cosmology_acquire(const char* evolution_file, const char* parameter_file, const char* cell_dir, double mpc,
                          double* h, double* omega_m, double* lbox_mpch, double* t, double* a, double* dadt,
                          double* d2adt2, uint64_t capacity, uint64_t* n);
/* Input::correct(cosmology, trajectory) (input.h:1330-1378), host only: a is
 * replaced, inside the trajectory's time range, by the a of a photon
 * propagated in a homogeneous universe (traj_t, traj_a: m elements), and
 * smoothed outside it (4 passes) when the trajectory is shorter than the
 * table.  In place on a (n rows). */
int mpf_cosmology_correct(const double* t, double* a, uint64_t n, const double* traj_t, const double* traj_a,
                          uint64_t m);
/* The cosmology preparation of rays.cpp:200-215: mpf_cosmology_acquire, then a
 * photon from the box centre toward +x through a homogeneous level-5 octree
 * (MPF_HOMOGENEOUS, on `device`, exact tier `precision`), then
 * mpf_cosmology_correct with its trajectory.  Outputs as mpf_cosmology_acquire;
 * d2adt2 is the uncorrected column, as in the reference. */
int mpf_cosmology_prepare(int device, int precision, const char* evolution_file, const char* parameter_file,
                          const char* cell_dir, double mpc, double* h, double* omega_m, double* lbox_mpch, double* t,
                          double* a, double* dadt, double* d2adt2, uint64_t capacity, uint64_t* n);

/* Message of the last failed mpf_cosmology_* call of this thread. */
const char* mpf_cosmology_last_error(void);

/* Cosmology table (t, a, da/dt), n rows, t increasing.  The table the
 * reference builds with Input::acquire + Input::correct (input.h:1091-1233,
 * 1330-1378) and reads with Utility::interpolate (utility.h:442-449). */
int mpf_set_cosmology(mpf_ctx* ctx, const double* t, const double* a, const double* dadt, uint64_t n);

/* SimpleHyperOctree::locate (simplehyperoctree.h:894-899): index of the
 * finest stored cell containing each point, in key (record) order, or -1
 * for end(); level of that cell (0 for end()). */
int mpf_locate(mpf_ctx* ctx, const double* xyz, uint64_t n, int64_t* elem_out, uint32_t* level_out);

/* SimpleHyperOctree::ngp/cic/tsc (simplehyperoctree.h:1396, 1412, 1481):
 * interpolated fields (n x 7, FIELDS order as above) at n points with the
 * context precision.  order 3 = tsc with one 27-cell cache carried over the
 * points in sequence (simplehyperoctree.h:1554), as the integrator uses it. */
int mpf_interpolate(mpf_ctx* ctx, int order, const double* xyz, uint64_t n, double* out7);

/* Integrator::integrate (integrator.h:477-618 without a stop criterion,
 * 648-824 with one) of n rays.  in9: launch photon per ray (a, t, x, y, z,
 * dt/dl, dx/dl, dy/dl, dz/dl) as built by Integrator::launch
 * (integrator.h:264-329).  out: n x MPF_NOUT doubles.  total_steps (may be
 * NULL): sum over rays of accepted RK4 steps (trajectory length - 1). */
int mpf_integrate(mpf_ctx* ctx, const double* in9, uint64_t n, const mpf_stop* stop, double* out,
                  uint64_t* total_steps);
/* Same with device-resident buffers; returns once the work is queued on the
 * context stream unless sync != 0. */
int mpf_integrate_dev(mpf_ctx* ctx, const double* dev_in9, uint64_t n, const mpf_stop* stop, double* dev_out,
                      int sync);
/* mpf_integrate_dev without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same kernels, launch
 * configuration and results, bit for bit; synchronous.  Its workspace (queue
 * and step counters, the ray order and its sort, the split tier's launch
 * records) is a fixed device array per GPU, held until the process ends:
 * at most MPF_STATIC_MAX_RAYS rays per call (else -EINVAL).  Every call on a
 * GPU shares that array, so calls that overlap on one GPU race.  The split
 * tier needs its stencil blocks already built (any mpf_integrate_dev call on
 * the context builds them; else -EPROTO).  total_steps (may be NULL): as
 * mpf_integrate. */
#define MPF_STATIC_MAX_RAYS 262144
int mpf_integrate_dev_static(mpf_ctx* ctx, const double* dev_in9, uint64_t n, const mpf_stop* stop, double* dev_out,
                             uint64_t* total_steps);
/* Integrator::propagate (integrator.h:826-1015), the per-ray computation of
 * the rays program (rays.cpp:276-284), exact tiers, any order: around each
 * central photon i (in9, n x 9) a bundle of `count` rays on a circle of
 * half-angle `angle` turned by rotation[i] (NULL: 0) (Integrator::launch,
 * integrator.h:345-382), every ray integrated until the data ends.  A bundle
 * is rejected when one of its rays is empty, ends within EXTENT / 4 of its
 * start, fails the amin test on the data's scale factor at its end (amin 0:
 * no test), or when the end distances differ by 1/8 or more.  Accepted:
 * traj receives the central trajectory (capacity elements x MPF_NOUT doubles
 * per photon, element k with k in its MPF_OUT_NSTEPS slot) and dist its
 * angular diameter distance (Photon::distance, metres) at each element: the
 * mean distance of the other rays at the same value of `interpolation`
 * (MPF_STOP_REDSHIFT, _A (as 1/a - 1), _T, _LAMBDA, _RADIUS) times
 * length a / angle.  nelem[i]: elements of the central trajectory, 0 when
 * the bundle is rejected (the reference returns an empty trajectory).
 * -ENOSPC when a ray has more than capacity elements. */
int mpf_propagate(mpf_ctx* ctx, const double* in9, uint64_t n, uint32_t count, double angle, const double* rotation,
                  int interpolation, double amin, uint64_t capacity, double* traj, double* dist, uint64_t* nelem);
/* Statistics of the rays program (rays.cpp:300-456) over the central
 * trajectories of mpf_propagate (rows, dist, nelem, capacity as it writes
 * them; rejected bundles, nelem 0, are skipped), host only.  Each trajectory
 * gives a curve of an abscissa (interp: MPF_STOP_REDSHIFT redshift,
 * MPF_STOP_A t and MPF_STOP_T 1/a - 1 (the reference's mapping of `a` and
 * `t`), MPF_STOP_RADIUS distance from the first element, MPF_STOP_LAMBDA)
 * against a statistic (MPF_STAT_*, distances in units of `length`),
 * re-interpolated at the abscissae of every statmod-th element of the
 * homogeneous reference ray (ref_rows m x MPF_NOUT, ref_dist m; its elements
 * with a < amin dropped when amin is normal; statmod = max(1, m' / nstat)).
 * out4: per abscissa (redshift of the reference element, its statistic, the
 * mean over rays, their standard deviation); nout receives their count (out4
 * may be NULL to size it), nused the number of rays used.  One process's
 * numbers, which are the reference's rank output and, with one rank, its
 * global output. */
enum {
    MPF_STAT_DISTANCE = 0, MPF_STAT_DISTANCE2 = 1, MPF_STAT_HOMOGENEOUS = 2, MPF_STAT_INHOMOGENEOUS = 3,
    MPF_STAT_INVDISTANCE = 4, MPF_STAT_INVDISTANCE2 = 5
};
int mpf_ray_statistics(const double* rows, const double* dist, const uint64_t* nelem, uint64_t n, uint64_t capacity,
                       const double* ref_rows, const double* ref_dist, uint64_t m, int interp, int stat,
                       uint32_t nstat, double amin, double length, double* out4, uint64_t cap_out, uint64_t* nout,
                       uint64_t* nused);
/* The statistics file of rays.cpp:441-452 (Output::save, output.h:228-241)
 * from mpf_ray_statistics' out4 (nout rows): "# count", then one line per
 * abscissa with the reference element's redshift, its statistic, the mean
 * and the standard deviation at 17 significant digits; count is the number
 * of rays used (nused).  Host only; -EINVAL when the file cannot be
 * written. */
int mpf_ray_statistics_write(const char* path, const double* out4, uint64_t nout, uint64_t count);
/* Whole trajectories (the reference's Evolution<Photon>, integrator.h:605,
 * 776, after the final pop), exact tiers: traj receives up to `capacity`
 * elements per ray, element k of ray i at traj[(i*capacity + k)*MPF_NOUT]
 * with k in its MPF_OUT_NSTEPS slot; ray i has out[i*MPF_NOUT +
 * MPF_OUT_NSTEPS] + 1 elements.  -ENOSPC (and the first `capacity` elements)
 * when a ray has more. */
int mpf_integrate_trajectory(mpf_ctx* ctx, const double* in9, uint64_t n, const mpf_stop* stop, uint64_t capacity,
                             double* out, double* traj);
/* Constant-parameter surfaces: the stop criterion of hmaps (`stop_ray`,
 * surfaces `interpRefvec`, hmaps.cpp:182, hmaps.h:610-660) and of the
 * catalogue root-finder (radius = source distance, catalogues.h:431-456).
 * The ray is integrated with stop criterion `kind` at value[n-1]; at each
 * surface the quantities are interpolated linearly between the last element
 * before it and the next one, and, when `jacobian` is set, the infinitesimal
 * lensing Jacobian (lensing.h:399-480) is evaluated at the surface distance. */
#define MPF_MAX_SURFACES 16
typedef struct {
    int kind;          /* MPF_STOP_RADIUS, _LAMBDA, _REDSHIFT, _T, _A or _S */
    uint32_t n;        /* number of surfaces, 1..MPF_MAX_SURFACES */
    double value[MPF_MAX_SURFACES]; /* increasing (decreasing for _A) */
    int jacobian;      /* also compute a11, a12, a22 */
    /* optional per-ray value (n doubles; host memory for mpf_integrate_surfaces,
     * device memory for the _dev call): one surface per ray, at which the ray
     * also stops (a source distance or redshift); then n must be 1 and
     * value[] is ignored */
    const double* ray_value;
} mpf_surfaces;
/* Per ray and surface, MPF_NSURF doubles: */
enum {
    MPF_SURF_VALID, MPF_SURF_FIRSTID, MPF_SURF_F, MPF_SURF_X, MPF_SURF_Y, MPF_SURF_Z, MPF_SURF_DIST,
    MPF_SURF_CHI, MPF_SURF_LAMBDA, MPF_SURF_T, MPF_SURF_A, MPF_SURF_REDSHIFT, MPF_SURF_S, MPF_SURF_ISW,
    MPF_SURF_RHO, MPF_SURF_PHI, MPF_SURF_RHOMAX, MPF_SURF_KT, MPF_SURF_KX, MPF_SURF_KY, MPF_SURF_KZ,
    MPF_SURF_A11, MPF_SURF_A12, MPF_SURF_A22, MPF_NSURF = 32
};
/* out (n x MPF_NOUT, may be NULL): as mpf_integrate; out_surf: n x surf->n x MPF_NSURF.
 * VALID is 0 where the reference rejects the pixel or source (firstid 0 or the
 * last element, hmaps.h:716). */
int mpf_integrate_surfaces(mpf_ctx* ctx, const double* in9, uint64_t n, const mpf_surfaces* surf, double* out,
                           double* out_surf);
int mpf_integrate_surfaces_dev(mpf_ctx* ctx, const double* dev_in9, uint64_t n, const mpf_surfaces* surf,
                               double* dev_out, double* dev_out_surf, int sync);

/* Velocity field of the reference's velocity-field build (VELOCITYFIELD;
 * hmaps.cpp:359-441).  The cells then carry a velocity (gravity2.h:45-90)
 * that the exact tiers interpolate with the other fields, and the redshift
 * of every step takes the Doppler and ISW terms
 *   z = a0/a (1 + phi0/c^2 - phi/c^2 + isw + v.k/(c k_t) - v0n) - 1,
 *   v0n = vobs.k0/(c k0_t)                     (integrator.h:537-539, 594-595).
 * The split tier refuses a tree with a velocity field.
 *
 * mpf_velocity_begin: rho of every cell becomes the deposited weight, so
 *   the sums start from zero (hmaps.cpp:391).
 * mpf_velocity_deposit: Hmaps::CreateOctreeVelocityWithCIC (scheme
 *   MPF_VELOCITY_CIC, hmaps.h:439-478) or WithTSC (MPF_VELOCITY_TSC,
 *   hmaps.h:497-541) of n particles (pos3: box units in the octree frame, as
 *   Hmaps::fill_particles_vectors leaves them, hmaps.h:345-401; vel3: in units
 *   of c, its conversion unit_l 1e-2 / (c unit_t), hmaps.h:378-380),
 *   at every level from the coarsest to the finest stored; may be called once
 *   per particle file.  The weights are summed in FP64 (the reference: FP32 in
 *   thread order) and rounded once to FP32.
 * mpf_velocity_finish: v /= (w + (w == 0)), rho = w, and a cell whose w is
 *   not a normal number takes its parent's velocity, level by level
 *   (hmaps.cpp:422-441); the tree then carries the field.
 * mpf_velocity_set / _get: the field in record order (rho may be NULL: kept).
 * mpf_velocity_clear: the tree carries no field (rho keeps its values).
 * mpf_set_observer_velocity: vobs in units of c like the cell velocities
 *   (the Doppler terms divide v.k by c k_t, and |k| = c k_t); the parameters v0x, v0y, v0z of
 *   hmaps.cpp:97-99), used with a velocity field. */
#define MPF_VELOCITY_CIC 1
#define MPF_VELOCITY_TSC 2
int mpf_velocity_begin(mpf_ctx* ctx);
int mpf_velocity_deposit(mpf_ctx* ctx, const float* pos3, const float* vel3, uint64_t n, int scheme);
int mpf_velocity_deposit_dev(mpf_ctx* ctx, const float* dev_pos3, const float* dev_vel3, uint64_t n, int scheme);
int mpf_velocity_finish(mpf_ctx* ctx);
int mpf_velocity_set(mpf_ctx* ctx, const float* rho, const float* vel3, uint64_t n);
int mpf_velocity_get(mpf_ctx* ctx, float* rho, float* vel3, uint64_t n);
int mpf_velocity_clear(mpf_ctx* ctx);
int mpf_set_observer_velocity(mpf_ctx* ctx, const double* vobs3);

/* Lensing::dbetadtheta (lensing.h:140-364), the bundle Jacobian
 * (jacobiantype = bundle): for each direction i (phi_theta, n x 2) four rays
 * from observer3 towards the unit direction plus or minus tan(opening) along
 * its screen vectors e1, e2, integrated on the device to the last of the nd
 * values of the stop parameter `kind` (MPF_STOP_RADIUS, _LAMBDA, _REDSHIFT,
 * _T, _A, or _PLANE: the position along the last ki, in an exact tier, FP64
 * for a split context), each ray's position interpolated at every value
 * (along each ki for _PLANE: each ray records its crossings of the planes
 * while it is integrated, all directions in one launch; a direction whose
 * rays' projections on a normal decrease somewhere is redone from whole
 * trajectories, as upper_bound over them places it), and
 * the finite differences projected on the screen normal to ki (n x nd x 3,
 * the reference's kiTargets) over 2 dist tan(opening) (dist: n x nd).  ref:
 * n x nd values, the same for every direction (up to 16, monotonic as the
 * surfaces of mpf_integrate_surfaces), or one per direction (nd = 1).  out4:
 * n x nd x (a11, a12, a21, a22); when a ray ends before a value: every entry
 * of that direction 42, 42, 1, 0 (for `a`: that value's entry 42, 42, 1, 3
 * and the others 0), as the reference returns them. */
int mpf_jacobian_bundle(mpf_ctx* ctx, const double* phi_theta, uint64_t n, const double* observer3, int kind,
                        const double* ref, uint32_t nd, const double* ki, const double* dist, double opening,
                        double* out4);
/* Tidal blocks of the split tier (built at the first split run with the
 * infinitesimal Jacobian, when the memory allows): per qualifying cell, the
 * six differences of grad phi one cell away on either side (the samples of
 * lensing.h:421-448) as polynomials in the offset, so that a step reads one
 * block instead of looking up six neighbour cells.  enable 1: allowed (the
 * default), 0: freed and not built again (every sample looked up), -1: only
 * query.  nblocks (may be null) receives the number built.  No counterpart in
 * the reference. */
int mpf_split_tidal_blocks(mpf_ctx* ctx, int enable, uint64_t* nblocks);
/* The number of directions of the last mpf_jacobian_bundle call with
 * MPF_STOP_PLANE that were redone from whole trajectories.  No counterpart
 * in the reference (which always stores whole trajectories). */
int mpf_bundle_plane_info(const mpf_ctx* ctx, uint64_t* whole_trajectory_directions);
/* The Born approximation of the surfaces (lensing_born maps, hmaps.h:705-713;
 * infinitesimal_born catalogues, catalogues.h:670-674): the rays follow the
 * path of Integrator::integrate<-1> (homogeneous data: straight lines of the
 * background) while the fused infinitesimal Jacobian samples the octree's
 * data along them.  Runs in the context's exact tier (FP64 for a split
 * context).  Arguments as for mpf_integrate_surfaces[_dev]. */
int mpf_integrate_surfaces_born(mpf_ctx* ctx, const double* in9, uint64_t n, const mpf_surfaces* surf, double* out,
                                double* out_surf);
int mpf_integrate_surfaces_born_dev(mpf_ctx* ctx, const double* dev_in9, uint64_t n, const mpf_surfaces* surf,
                                    double* dev_out, double* dev_out_surf, int sync);
/* Catalogue root-finding: Catalogues::relCat / iterateNewtonMethod /
 * newtonMethod2d (catalogues.h:410-920) with the infinitesimal Jacobian
 * (jacobiantype = infinitesimal), for a batch of sources on the device. */
typedef struct {
    double accuracy;          /* cat_accuracy (rad), e.g. 1e-8 */
    uint32_t max_iterations;  /* 10 (catalogues.h:765) */
    double vobs[3];           /* observer velocity (SI) */
    double h;                 /* dimensionless Hubble parameter */
    double mpc;               /* one megaparsec in metres (parameter file `mpc`) */
    /* 0: the Jacobian is integrated with the rays of rounds >= 3 (where the
     * Newton update needs it, catalogues.h:509-517) and, for sources that
     * converged earlier, by one more ray at the converged direction;
     * 1: with every round's rays, no extra ray; 2: with the rays of rounds
     * >= 1, and one more ray for the sources that converge at their first
     * ray.  Same results; 1 integrates fewer rays when most sources converge
     * within three rounds, 2 spares the first round's rays the Jacobian when
     * few sources converge at their first ray (a good but not exact first
     * guess). */
    uint32_t jacobian_all_rounds;
    /* 1: jacobiantype = infinitesimal_born (catalogues.h:670-674): the
     * Jacobian columns of converged sources come from the homogeneous path
     * (Integrator::integrate<-1>) in the converged direction to the source's
     * distance, sampled on the octree's data; the Newton updates still use
     * the ray's own Jacobian.  0: jacobiantype = infinitesimal. */
    uint32_t jacobian_born;
    /* 1: jacobiantype = bundle (catalogues.h:455-491, 509-523, 666-681):
     * iterations 0-2 the identity, 3-4 the infinitesimal Jacobian, then and
     * for the converged source Lensing::dbetadtheta (mpf_jacobian_bundle)
     * from the observer at the origin, with stop_bundle bundle_kind
     * (MPF_STOP_RADIUS, _LAMBDA, _REDSHIFT, _T or _A; interpRef taken at the
     * crossing of the source radius), plane bundle_plane (0 sachs: the ray's
     * direction there; 1 normal: the crossing point) and half-opening
     * bundle_opening (rad).  A failed bundle ends the source with observed
     * angles 42, 42 and separation 22 (status MPF_CAT_OUTSIDE, the _err
     * file; iterations 0 and redshifts 0 where the reference leaves them
     * unset).  Runs the bundles' rays through the host (four per source and
     * round). */
    uint32_t jacobian_bundle;
    int32_t bundle_kind;
    int32_t bundle_plane;
    double bundle_opening;
    /* use_previous_catalogues (catalogues.h:501-623, catalogues.cpp:240-257):
     * 0 the catalogue; 2 the re-run of rejected sources from their last
     * angles (init2): a source not within the accuracy sends a grid of up to
     * 11 x 11 rays at offsets from -err to err in steps of err / 5,
     * err = 5 sep / (iteration + 1), around its angles, and takes the first
     * within the accuracy (converged: its angles, with z1 and the scale factor
     * of the central ray and, for the bundle, the bundle around the central
     * direction) or else the closest if closer than the central ray; no
     * Jacobian-based update, and no Jacobian for a rejected source.  The
     * reference allows 100 iterations there (set max_iterations).  A grid ray
     * that ends before the source radius is no candidate (the reference
     * extrapolates it). */
    uint32_t use_previous;
} mpf_newton_cfg;
/* Per source, MPF_NCAT doubles: the catalogue columns of catalogues.h:844-859
 * plus status and the number of rounds run. */
enum {
    MPF_CAT_PHI, MPF_CAT_THETA, MPF_CAT_PHI_OBS, MPF_CAT_THETA_OBS, MPF_CAT_SEP, MPF_CAT_ITER,
    MPF_CAT_Z0, MPF_CAT_Z1, MPF_CAT_Z2, MPF_CAT_Z3, MPF_CAT_Z4, MPF_CAT_Z5,
    MPF_CAT_A11, MPF_CAT_A12, MPF_CAT_A21, MPF_CAT_A22, MPF_CAT_STATUS, MPF_CAT_NROUNDS, MPF_NCAT = 20
};
/* status: converged (catalogue file), outside (the ray never reaches the
 * source radius: observed angles 42, the _err file), rejected (no convergence
 * within max_iterations: iterations 42, the _reject file) */
enum { MPF_CAT_OK = 0, MPF_CAT_OUTSIDE = 1, MPF_CAT_REJECTED = 2 };
/* src6: n x (x, y, z, vx, vy, vz), positions in box units in the octree frame
 * (already rotated for narrow cones, catalogues.h:818-824), velocities in
 * simulation units; init2 (may be NULL): n x (phi, theta) first guesses, by
 * default the unlensed direction (catalogues.h:832-836).  total_rays (may be
 * NULL): rays integrated, summed over Newton rounds. */
int mpf_catalogue_solve(mpf_ctx* ctx, const double* src6, const double* init2, uint64_t n,
                        const mpf_newton_cfg* cfg, double* out, uint64_t* total_rays);
/* Same on device buffers (sync on return). */
int mpf_catalogue_solve_dev(mpf_ctx* ctx, const double* dev_src6, const double* dev_init2, uint64_t n,
                            const mpf_newton_cfg* cfg, double* dev_out, uint64_t* total_rays);
/* Catalogues::relCat_with_previous_cat (catalogues.h:922-1044), the re-run
 * of a computed catalogue with use_previous_catalogues = 1: for each row of
 * rows18 (n x 18: the catalogue file's columns id, phi, theta, phi_obs,
 * theta_obs, separation, iterations, z0..z5, a11, a12, a21, a22, npart) a
 * ray from its observed angles (its comoving ones for jacobian_born) to the
 * scale factor 1 / (1 + z0), and the lensing Jacobian of cfg's
 * jacobiantype at the distance interpolated there replaces columns 13-16,
 * in place.  A row whose ray never reaches that scale factor gets 42, 42,
 * 42, 42 (the reference extrapolates the ray's last segment); a failed
 * bundle gives the reference's entries.  Of cfg only jacobian_born,
 * jacobian_bundle and the bundle fields are read. */
int mpf_catalogue_previous(mpf_ctx* ctx, double* rows18, uint64_t n, const mpf_newton_cfg* cfg);
/* The files of Catalogues::relCat (catalogues.h:862-912) from n rows of
 * mpf_catalogue_solve, host only: <prefix>.txt (converged sources: id, the
 * 16 catalogue columns, npart, 17 significant digits), <prefix>_err.txt (the
 * ray never reached the source, or its bundle failed: id, phi, theta,
 * phi_obs, theta_obs and the separation twice, at the stream's default 6
 * digits) and <prefix>_reject.txt (not converged: as <prefix>.txt), each
 * replaced.  id_npart: n x 2 (the targets' id and npart). */
int mpf_catalogue_write(const char* prefix, const double* out, uint64_t n, const double* id_npart);
/* The files of relCat_with_previous_cat (catalogues.h:1008-1040), host
 * only: the 18 columns of each row with 17 significant digits, in
 * <prefix>.txt when a11 and a12 are not 42, else in <prefix>_err.txt. */
int mpf_catalogue_write_previous(const char* prefix, const double* rows18, uint64_t n);

/* Learned first guess of the catalogue root-finder.  No counterpart in the
 * reference: it proposes the observed angles that seed the Newton
 * iterations (init2 of mpf_catalogue_solve; the reference starts from the
 * unlensed direction, catalogues.h:832-836), which still finish each source
 * to the accuracy.  A 10 -> 128 -> 128 -> 2 GELU network (BF16 operands, FP32
 * accumulation on the matrix cores, FP32 master weights, Adam), one per
 * context, trained on pairs from the root-finder itself.
 *
 * Features per source (MPF_MLP_NIN floats), from the octree along straight
 * lines from the observer (TSC at chi_k = (k + 1/2) D / 256, k < 256: exact
 * FP64 with one stencil cache per line, or the split tier's FP32 stencil
 * blocks when the context precision is MPF_PREC_SPLIT), invariant under
 * rotations of the sky:
 *   0 comoving distance D (box units); 1, 2 Born deflection
 *   2/c^2 sum (D - chi)/D grad_perp(phi) dchi along e_phi and e_theta
 *   (arcsec) on the line to the source; 3, 4 the same sum on the line of the
 *   same length towards the source's angles plus that deflection (theta + f2,
 *   phi + f1 / sin(theta); projected on that line's e_phi, e_theta) less
 *   f1, f2; 5, 6 the same difference over the far halves (k >= 128); 7, 8
 *   the first line's sum over its far half; 9 phi / c^2 at k = 127 on the
 *   first line (units of 1e-5).  The pairs (1, 2) .. (7, 8) and the targets
 *   are vectors in the screen basis: training turns them together by random
 *   angles and reflections (isotropy), so that the network cannot learn
 *   where in the training cones a source lies from their directions.
 * Targets (MPF_MLP_NOUT floats): the deflection left after the Born
 * estimate, (phi_obs - phi) sin(theta) - f1 and theta_obs - theta - f2
 * (arcsec).  The first guess is phi + (f1 + y0) / sin(theta),
 * theta + f2 + y1, with y the network's output.
 *
 * Parameter vector (MPF_MLP_NPARAM floats), padded to matrix-core tiles:
 * W1 [16][128] (rows 10..15 zero), b1 [128], W2 [128][128], b2 [128],
 * W3 [128][16] (columns 2..15 zero), b3 [16] (2..15 zero), y = x W + b.
 * The state (MPF_MLP_NSTATE floats) appends the normalisation: feature mean
 * [10], feature std [10], target mean [2], target std [2]; the network sees
 * (x - mean) / std and its output is scaled back by the target std and mean.
 * The normalisation is taken from the first training set unless set. */
#define MPF_MLP_NIN 10
#define MPF_MLP_NOUT 2
#define MPF_MLP_NHIDDEN 128
#define MPF_MLP_KIN 16
#define MPF_MLP_KOUT 16
#define MPF_MLP_NPARAM 20752
#define MPF_MLP_NNORM 24
#define MPF_MLP_NSTATE (MPF_MLP_NPARAM + MPF_MLP_NNORM)
/* (Re)initialise the context's network: Glorot-uniform hidden weights from
 * seed, zero output layer and biases, Adam state cleared, normalisation
 * unset. */
int mpf_mlp_init(mpf_ctx* ctx, uint64_t seed);
/* Read (get != 0) and/or replace (set != 0) the MPF_MLP_NSTATE-float state. */
int mpf_mlp_state(mpf_ctx* ctx, float* get, const float* set);
/* Adam (lr, beta 0.9 / 0.999, eps 1e-8) on the mean squared error of the
 * normalised outputs, over `epochs` passes of shuffled mini-batches (shuffle
 * from seed).  loss_per_epoch (may be null): mean loss of each epoch. */
int mpf_mlp_train(mpf_ctx* ctx, const float* feat, const float* target, uint64_t n, uint32_t epochs,
                  uint32_t batch, float lr, uint64_t seed, float* loss_per_epoch);
/* Loss and gradient (MPF_MLP_NPARAM floats) of one batch, parameters
 * unchanged: the training step without the update. */
int mpf_mlp_gradient(mpf_ctx* ctx, const float* feat, const float* target, uint64_t n, float* grad, float* loss);
/* Predicted targets (n x MPF_MLP_NOUT).  The _dev form allocates nothing:
 * one persistent kernel holds the BF16 weights in LDS and normalises the
 * features and scales the outputs itself. */
int mpf_mlp_infer(mpf_ctx* ctx, const float* feat, uint64_t n, float* out);
int mpf_mlp_infer_dev(mpf_ctx* ctx, const float* dev_feat, uint64_t n, float* dev_out);
/* Duration (ms, HIP events on the context's stream) of the last inference
 * kernel; -1 before the first. */
int mpf_mlp_last_ms(mpf_ctx* ctx, double* ms);
/* Features of sources (n x 6 doubles as in mpf_catalogue_solve). */
int mpf_catalogue_features(mpf_ctx* ctx, const double* src6, uint64_t n, float* feat);
int mpf_catalogue_features_dev(mpf_ctx* ctx, const double* dev_src6, uint64_t n, float* dev_feat);
/* First guesses init2 (n x 2: phi, theta) = unlensed direction + Born
 * estimate + (use_network != 0) the network's correction. */
int mpf_catalogue_seed(mpf_ctx* ctx, const double* src6, uint64_t n, int use_network, double* init2);
int mpf_catalogue_seed_dev(mpf_ctx* ctx, const double* dev_src6, uint64_t n, int use_network, double* dev_init2);
/* One 16 x 16 x 16 product on the matrix cores, c = bf16(a) bf16(b) with FP32
 * accumulation (row-major 16 x 16 floats): checks the operand layout the
 * network's kernels assume. */
int mpf_mfma_probe(mpf_ctx* ctx, const float* a, const float* b, float* c);
/* One 16 x 16 x 32 product on gfx950's matrix cores, c = bf16(a) bf16(b)
 * (a 16 x 32, b 32 x 16, c 16 x 16, row-major floats), with the operand
 * layout the network's kernels use (layout 0: slot j of lane l holds
 * K = 8 (l >> 4) + j in both operands), or with B's slots given the K of two
 * halves (layout 1: 4 (l >> 4) + j and 16 + 4 (l >> 4) + j) while A keeps
 * layout 0: a mismatched pairing, so that a check can show it is caught. */
int mpf_mfma_probe32(mpf_ctx* ctx, const float* a, const float* b, int layout, float* c);

/* Cones and photon launch (host, no device needed).  A cone is MPF_CONE_N
 * doubles: vertex x, y, z, base centre x, y, z, half-angle (rad), the three
 * members of the reference's Cone (cone.h).  Results are bit-identical to the
 * reference's; errors return a negative status with the message in
 * mpf_cone_last_error(). */
enum { MPF_CONE_N = 7 };
const char* mpf_cone_last_error(void);
/* Miscellaneous::GenerateFullskyCones (miscellaneous.h:200-218): cones with
 * their vertex at the origin and bases on the spiral of
 * HyperSphere::uniform (abstracthypersphere.h:365-418) on a sphere of the
 * given radius (EXTENT / 2 = 2 in the reference programs, rays.cpp:106,145). */
int mpf_cones_fullsky(uint32_t ncones, double radius, double* cones7);
/* Miscellaneous::GenerateNarrowCones (miscellaneous.h:257-402): angles in
 * degrees (phi, theta, thetay, thetaz of info_narrow_cone.txt).  cones7: the
 * cones in the simulation frame (catalogue source selection); cones_rot7: the
 * same cones with their base rotated into the octree frame (coneIfRot: ray
 * launch, map pixels); rotm1: 3 x 3 row-major rotation into the octree frame
 * (source positions, catalogues.h:818-824).  spiral_points (may be NULL):
 * the spiral size that puts exactly ncones points in the region. */
int mpf_cones_narrow(uint32_t ncones, double radius, double phi_deg, double theta_deg, double thetay_deg,
                     double thetaz_deg, double* cones7, double* cones_rot7, double* rotm1, uint64_t* spiral_points);
/* The ASCII branch of GenerateNarrowCones (miscellaneous.h:293-300): reads
 * info_narrow_cone.txt with Input::parse; angles4 = phi (read as an integer,
 * std::stoul), theta, thetay, thetaz, in degrees. */
int mpf_narrow_cone_info(const char* path, double* angles4);
/* Cone::inside (cone.h:359-373) for n points (x, y, z at every `stride`
 * doubles). */
int mpf_cone_inside(const double* cone7, const double* xyz, uint64_t n, uint64_t stride, uint8_t* out);
/* Closest-cone ownership: out[i] = 1 when point i is inside cone icone and
 * no other cone in front of it has its axis closer (the test of
 * Miscellaneous::getTargets miscellaneous.h:523-578, by_value = 1: other
 * cones are those whose members differ; of Hmaps::getPixels_per_cone2
 * hmaps.h:261-315 and Integrator::launch integrator.h:166-249, by_value = 0:
 * every other index). */
int mpf_cone_select(const double* cones7, uint32_t ncones, uint32_t icone, int by_value, const double* xyz,
                    uint64_t n, uint64_t stride, uint8_t* out);
/* Miscellaneous::getTargets (miscellaneous.h:523-600): the indices, in input
 * order, of the targets (positions at the start of every `stride` doubles,
 * 8 in the reference) that cone icone owns. */
int mpf_cone_targets(const double* cones7, uint32_t ncones, uint32_t icone, const double* targets, uint64_t n,
                     uint64_t stride, uint64_t* index_out, uint64_t* nselected);
/* Hmaps::getPixels_per_cone2 (hmaps.h:261-322): the RING pixels whose centre
 * cone icone owns, in increasing order.  npix_out receives the count; with
 * pix_out NULL only the count is returned. */
int mpf_hmaps_cone_pixels(int64_t nside, const double* cones7, uint32_t ncones, uint32_t icone, int64_t* pix_out,
                          uint64_t capacity, uint64_t* npix_out);
/* Integrator::launch from a point towards angles (integrator.h:304-329):
 * phi_theta n x 2 (rad) -> in9 n x 9. */
int mpf_launch_angles(const double* observer3, const double* phi_theta, uint64_t n, double* in9);
/* Integrator::launch from begin to end (integrator.h:264-290), n pairs. */
int mpf_launch_towards(const double* begin3, const double* end3, uint64_t n, double* in9);
/* rays.cpp:233-239 (ray_targets = random): one std::mt19937 seeded with
 * `seed` (the reference seeds seed + rank), n photons from the origin towards
 * random points of a sphere of radius micro_radius (1 / microcoeff) that
 * cone icone owns (Integrator::launch into a series of cones,
 * integrator.h:166-249, HyperSphere::random abstracthypersphere.h:323-345),
 * each followed by one uniform draw times 2 pi: rotation (may be NULL), the
 * angle by which the ray's bundle is turned. */
int mpf_launch_cone(const double* cones7, uint32_t ncones, uint32_t icone, double micro_radius, uint32_t seed,
                    uint64_t n, double* in9, double* rotation);
/* Integrator::launch around one photon (integrator.h:345-382): count photons
 * on a circle of half-angle `angle` turned by `rotation`, preceded by the
 * photon itself when center != 0; out9 holds count + (center != 0) photons. */
int mpf_launch_bundle(const double* in9, uint32_t count, double angle, double rotation, int center, double* out9);

/* Octree preparation (host, no device needed): the ASCII branch of the
 * reference's create_octree program, which turns cell files into cone files.
 * A builder holds one octree as the reference's flat container of (key,
 * Gravity<float>) cells; errors return a negative status with the message in
 * mpf_prep_last_error().  The operation order, types and searches follow the
 * reference's, so the cone files match its output byte for byte, except where
 * the reference reads outside its container (the last cell of
 * SimpleHyperOctree::update, a missing parent) or races (the coarse-density
 * mean), where the builder does what the loop states, cell by cell. */
typedef struct mpf_prep mpf_prep;
typedef struct mpf_prep_cfg {
    int32_t isfullsky;        /* 0: narrow cone, positions and forces rotated by rotm1 */
    int32_t coarseonly;       /* keep only cells of level ncoarse + log2(EXTENT) */
    uint32_t ncoarse;         /* coarse level of the simulation */
    int32_t correction;       /* complete partial octs by interpolation (else coarsen them) */
    int32_t coarsecorrection; /* non-normal coarse densities from their neighbours */
    int32_t acorrection;      /* must be 0 (the Full Universe Run a snapping is not supported) */
    int32_t order;            /* interpolation of the completed children: 0 NGP, 1 CIC, 2 TSC */
    double microcoeff;        /* central sphere of radius 1 / microcoeff */
    double h, omegam, lboxmpch, mpc;
    uint64_t allocation;      /* cells to reserve */
} mpf_prep_cfg;
const char* mpf_prep_last_error(void);
/* Create_octree::ReadParamFile's parameters (create_octree.h:120-146) */
int mpf_prep_create(mpf_prep** prep, const mpf_prep_cfg* cfg);
int mpf_prep_destroy(mpf_prep* prep);
int mpf_prep_clear(mpf_prep* prep);
/* Input::importascii (input.h:997-1069): appends the cells of one 10-column
 * file (level, x, y, z, rho, phi, dphidx, dphidy, dphidz, a; RAMSES units)
 * whose circumscribed sphere touches the central sphere or the cone
 * (Input::collide, input.h:216-226; cone7 as for the cone functions, the
 * octree-frame cone); rotm1 (narrow cones): 3 x 3 row-major. */
int mpf_prep_import_ascii(mpf_prep* prep, const char* path, const double* cone7, const double* rotm1,
                          uint64_t* added);
/* Replaces the container by n 48-byte records in the given order. */
int mpf_prep_load_records(mpf_prep* prep, const void* rec48, uint64_t n);
/* SimpleHyperOctree::update (simplehyperoctree.h:1272-1307). */
int mpf_prep_update(mpf_prep* prep);
/* SimpleHyperOctree::ngp / cic / tsc (order 0 / 1 / 2, the uncached forms,
 * simplehyperoctree.h:1396-1541) on the container as it is: out7 n x 7 in
 * the order rho, phi, dphidx, dphidy, dphidz, a, dphidt. */
int mpf_prep_interpolate(mpf_prep* prep, int order, const double* xyz, uint64_t n, float* out7);
/* Miscellaneous::correctOctree (miscellaneous.h:467-479): SI units
 * (Input::sistemize, input.h:382-394), Input::correct (input.h:1382-1520),
 * dphi/da to dphi/dt with the corrected cosmology table (t, a, da/dt). */
int mpf_prep_correct(mpf_prep* prep, const double* t, const double* a, const double* dadt, uint64_t n);
int mpf_prep_size(mpf_prep* prep, uint64_t* n);
int mpf_prep_records(mpf_prep* prep, void* rec48, uint64_t capacity);
/* Input::save (input.h:1276-1289): refuses an existing file. */
int mpf_prep_save(mpf_prep* prep, const char* path);
/* Keys of cells from float positions at their levels (the importer's
 * Index::compute<float>), key2 n x 2 (low, high 64 bits), and the
 * importer's filter for them. */
int mpf_prep_keys(const uint32_t* level, const float* xyz, uint64_t n, uint64_t* key2);
int mpf_prep_collide(const uint64_t* key2, uint64_t n, double micro_radius, const double* cone7, uint8_t* out);
/* Create_octree::PreparationASCII (create_octree.h:509-585): the .dat files
 * of celldir (and minicone, narrow cones) in the order of the reference's
 * shuffles (std::mt19937 seeded seed + rank), one cone file for each cone
 * icone with icone % ntasks == rank; cones_rot7: the octree-frame cones. */
int mpf_prepare_ascii(const mpf_prep_cfg* cfg, const char* celldir, const char* minicone, const double* cones_rot7,
                      uint32_t ncones, const double* rotm1, uint32_t ntasks, uint32_t rank, uint64_t seed,
                      const double* t, const double* a, const double* dadt, uint64_t ncos,
                      const char* const* conefiles);

/* HEALPix RING scheme (host, no device needed), restated from the published
 * algorithm (Gorski et al. 2005): the reference links HEALPix C++ 3.70 /
 * chealpix (pix2vec_ring at hmaps.h:599, nside2npix at hmaps.cpp:185). */
int mpf_healpix_pix2vec(int64_t nside, const int64_t* pix, uint64_t n, double* vec3);
int mpf_healpix_vec2pix(int64_t nside, const double* vec3, uint64_t n, int64_t* pix);
/* Launch photons at pixel centres (Hmaps::FillMap, hmaps.h:599-607), device
 * buffers: dev_in9 receives n x 9 doubles. */
int mpf_hmaps_launch_dev(mpf_ctx* ctx, int64_t nside, const int64_t* dev_pix, uint64_t n, double* dev_in9);
/* Per-pixel lensing maps of rays launched at pixel centres (hmaps.h:724-762,
 * infinitesimal Jacobian): dev_maps[(j*4 + k)*npix + pix] for surface j and
 * k = kappa, gamma1, gamma2, 1/magnification; rejected pixels are 0.  Pixels
 * not in dev_pix are left untouched. */
int mpf_hmaps_fill_dev(mpf_ctx* ctx, int64_t nside, const int64_t* dev_pix, const double* dev_surf, uint64_t n,
                       uint32_t nsurf, double* dev_maps);
/* Map components of hmaps (map_components, hmaps.cpp:187-268): codes of
 * Hmaps::FillMap's index_components.  lensing: kappa, gamma1, gamma2, 1/mu
 * (infinitesimal Jacobian) or kappa, gamma1, gamma2, w, 1/mu (bundle); dr,
 * dl, dt, da, dz, ds: the surface's chi, lambda, t, a, redshift, s over the
 * homogeneous value minus 1 (ds over chi); isw, dens, dens_max, steps, phi:
 * the interpolated value, the largest rho before the surface, the element
 * index; deflection: phi - beta1, theta - beta2 of the pixel centre and the
 * ray's position at the surface; lensing_born: kappa, gamma1, gamma2, 1/mu
 * from the Born surfaces (mpf_integrate_surfaces_born_dev, radius surfaces
 * at the homogeneous distances, hmaps.h:705-713). */
enum {
    MPF_MAP_LENSING = 0, MPF_MAP_LENSING_BORN = 1, MPF_MAP_DR = 2, MPF_MAP_DL = 3, MPF_MAP_DT = 4, MPF_MAP_DA = 5,
    MPF_MAP_DZ = 6, MPF_MAP_DS = 7, MPF_MAP_ISW = 8, MPF_MAP_DENS = 9, MPF_MAP_DENS_MAX = 10, MPF_MAP_STEPS = 11,
    MPF_MAP_PHI = 12, MPF_MAP_DEFLECTION = 13
};
/* The homogeneous quantities of hmaps.cpp:297-325 at each map redshift
 * z_stop[iz]: along the homogeneous reference ray (ref_rows, m x MPF_OUT
 * records of mpf_integrate_trajectory with MPF_HOMOGENEOUS), the element
 * before the redshift (upper_bound) and the interpolation factor give t, chi
 * and lambda; a = 1 / (1 + z); homo5: nz x (t, chi, lambda, redshift, a).
 * interp_ref (may be NULL): the surface values of stop_ray (MPF_STOP_T
 * thomo, _RADIUS rhomo, _LAMBDA lambdahomo, _A ahomo, _REDSHIFT z, _S
 * rhomo).  Host only. */
int mpf_hmaps_homogeneous(const double* ref_rows, uint64_t m, const double* z_stop, uint32_t nz, int stop_ray,
                          double* homo5, double* interp_ref);
/* Every map component of Hmaps::FillMap (hmaps.h:716-836) for rays launched
 * at pixel centres (mpf_hmaps_launch_dev) and integrated with nz surfaces
 * (dev_surf, n x nz x MPF_NSURF): dev_maps[(nmaps * iz + k) * npix + pix] in
 * FP32, the layout of the reference's map vector; pixels of rejected
 * surfaces (or of failed bundle Jacobians) are 0 in every map of that
 * surface; pixels not in dev_pix are untouched.  dev_surf_born: the Born
 * surface records (n x nz x MPF_NSURF), needed by lensing_born only, as
 * dev_surf is needed only by the other components.  dev_jac: n x nz x 4 bundle
 * Jacobians (mpf_jacobian_bundle, copied to the device), or NULL for the
 * infinitesimal one of the surface records.  homo5: host, nz x 5 of
 * mpf_hmaps_homogeneous.  nmaps receives the maps per surface; -EINVAL,
 * writing nothing, when nz x nmaps x 12 nside^2 exceeds maps_capacity
 * (floats; dev_maps may then be NULL, to size the buffer). */
int mpf_hmaps_components_dev(mpf_ctx* ctx, int64_t nside, const int64_t* dev_pix, const double* dev_surf,
                             const double* dev_surf_born, const double* dev_jac, uint64_t n, uint32_t nz,
                             const int* components, uint32_t ncomp,
                             const double* homo5, float* dev_maps, uint64_t maps_capacity, uint32_t* nmaps);
/* Write one RING (nest = 0) or NESTED map as a HEALPix FITS file: a binary
 * table of one FP32 column with the PIXTYPE, ORDERING, NSIDE, FIRSTPIX,
 * LASTPIX and COORDSYS keywords (write_healpix_map of chealpix, called at
 * hmaps.cpp:510).  clean != 0 first replaces every value that is not normal
 * (zero, subnormal, infinite, NaN) by Healpix_undef (-1.6375e30), as
 * hmaps.cpp:494-499 does before writing.  Host only; the file is replaced. */
int mpf_healpix_write_fits(const char* path, const float* map, int64_t nside, int nest, char coordsys, int clean);
/* Bin rays by the pixel of their observed direction (dev_in9 dx/dl, dy/dl,
 * dz/dl): dev_maps[k*npix + pix], k = mean kappa, mean gamma1, mean gamma2,
 * ray count, from surface `isurf` of dev_surf (n x nsurf x MPF_NSURF).
 * Deterministic: 64-bit fixed-point sums (2^-46) reduced per wavefront. */
int mpf_healpix_bin_dev(mpf_ctx* ctx, int64_t nside, const double* dev_in9, const double* dev_surf, uint64_t n,
                        uint32_t nsurf, uint32_t isurf, double* dev_maps);

/* Device buffers on the context's device, for callers that keep inputs and
 * outputs resident between calls (mpf_integrate_dev).  Copies are
 * synchronous. */
/* Free and total bytes of device memory of a HIP device (hipMemGetInfo):
 * callers size their batches with it; no counterpart in the reference. */
int mpf_device_memory(int device, uint64_t* free_bytes, uint64_t* total_bytes);
int mpf_dev_alloc(mpf_ctx* ctx, uint64_t bytes, void** dev_ptr);
int mpf_dev_free(mpf_ctx* ctx, void* dev_ptr);
int mpf_dev_upload(mpf_ctx* ctx, void* dev_dst, const void* host_src, uint64_t bytes);
int mpf_dev_download(mpf_ctx* ctx, void* host_dst, const void* dev_src, uint64_t bytes);
/* Device-to-device copy with a plain streaming kernel; its duration lands
 * in mpf_last_run.  A kernel of known traffic, used to calibrate HBM byte
 * counters (FETCH_SIZE / WRITE_SIZE) against the integration kernels.  No
 * counterpart in the reference. */
int mpf_dev_copy(mpf_ctx* ctx, void* dev_dst, const void* dev_src, uint64_t bytes);
/* Launch configuration of the persistent integrators (mpf_integrate*,
 * mpf_catalogue_solve*).  No counterpart in the reference, which splits rays
 * over threads in contiguous chunks (Utility::parallelize,
 * utility.h:127-143).  It changes no result, only the time.  kind 0: runs
 * without surfaces; kind 1: runs with surfaces (maps, catalogue rounds).  A
 * new context starts with refill 64 and sort bucket 65,536 for both kinds:
 * the lanes of a wavefront then hold rays of neighbouring directions, which
 * the kernels' block reuse needs (in input order a Jacobian batch of 4 M
 * rays ran 15 times slower). */
typedef struct {
    /* idle lanes a wavefront gathers before it takes rays from the queue
     * (1..64; 64: whole wavefronts of rays at a time) */
    uint32_t refill;
    /* 0: rays in input order; else rays ordered by decreasing path length
     * (per-ray stop or surface value), cut into buckets of this many rays,
     * each bucket ordered by direction (a bucket of n rays: direction only) */
    uint64_t sort_bucket;
    /* workgroups of 256 lanes per compute unit in the persistent grid
     * (0: 8 for the split tier, 4 for the exact tiers; at most 32) */
    uint32_t blocks_per_cu;
    /* split tier without surfaces: 1 = coefficient block kept in registers
     * (1 wavefront per SIMD), 2..4 = coefficients re-read, at least that
     * many wavefronts per SIMD; with surfaces 1 or 2; 0: 2 */
    uint32_t waves_per_simd;
} mpf_launch_cfg;
int mpf_set_launch_config(mpf_ctx* ctx, int kind, const mpf_launch_cfg* cfg);
int mpf_get_launch_config(mpf_ctx* ctx, int kind, mpf_launch_cfg* cfg);
/* Autotuner of the launch configuration: runs the given device-resident
 * batch (stop as for mpf_integrate_dev, or surfaces as for
 * mpf_integrate_surfaces_dev when surf is not NULL) under candidate
 * configurations, one knob at a time (refill, sort bucket, waves per SIMD,
 * blocks per CU, then the sort bucket again), keeps a change that saves at
 * least 1 % of the kernel time
 * (HIP events, ray ordering included), stores the result in the context for
 * that kind and returns it in best.  trials (may be NULL): up to max_trials
 * rows of 6 doubles (refill, sort_bucket, blocks_per_cu, waves_per_simd,
 * ms, accepted steps); ntrials: the number of runs.  dev_out / dev_out_surf
 * receive the results of the last run. */
int mpf_autotune(mpf_ctx* ctx, const double* dev_in9, uint64_t n, const mpf_stop* stop, const mpf_surfaces* surf,
                 double* dev_out, double* dev_out_surf, mpf_launch_cfg* best, double* trials, uint32_t max_trials,
                 uint32_t* ntrials);
/* Accepted steps and duration (HIP events on the context stream) of the
 * last integration kernel. */
int mpf_last_run(mpf_ctx* ctx, uint64_t* total_steps, double* kernel_ms);
/* Duration (HIP events on the context stream) of the persistent integration
 * kernel alone in the last integration: without the ray ordering, the launch
 * records and the final values that mpf_last_run's time includes. */
int mpf_last_kernel_ms(mpf_ctx* ctx, double* ms);
int mpf_synchronize(mpf_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
