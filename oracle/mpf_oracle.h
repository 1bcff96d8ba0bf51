/* mpf oracle: CPU restatement of the reference photon-propagation path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline.  The product (libmpf.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Pinning (what
 * each function is checked against) is listed per function in mpf_oracle.c.
 */
#ifndef MPF_ORACLE_H
#define MPF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpfo_octree mpfo_octree;

/* Per-ray output record, doubles (same layout as libmpf's mpf_ray_out). */
enum {
    MPFO_A, MPFO_T, MPFO_X, MPFO_Y, MPFO_Z, MPFO_KT, MPFO_KX, MPFO_KY, MPFO_KZ,
    MPFO_REDSHIFT, MPFO_LAMBDA, MPFO_S, MPFO_CHI, MPFO_ISW, MPFO_ISWOLD, MPFO_PHI,
    MPFO_DPHIDX, MPFO_DPHIDY, MPFO_DPHIDZ, MPFO_DPHIDT, MPFO_RHO, MPFO_AH, MPFO_DPHIDL,
    MPFO_LAPLACIAN, MPFO_DSDL2, MPFO_ERROR, MPFO_LEVEL, MPFO_NSTEPS, MPFO_NREC
};

/* precision: 0 = reference (FP32 fields, FP32 accumulation),
 *            1 = FP64 (fields64 if loaded, else FP32 fields promoted; FP64 accumulation) */
mpfo_octree* mpfo_octree_new(const void* rec48, uint64_t n);
int mpfo_octree_set_fields64(mpfo_octree* o, const double* f7, uint64_t n);
void mpfo_octree_free(mpfo_octree* o);
/* velocity-field build (gravity2.h cells): rho replaced, velocity per record
 * (units of c), observer velocity for the Doppler term of the redshift */
int mpfo_octree_set_velocity(mpfo_octree* o, const float* rho, const float* vel3, uint64_t n, const double* vobs3);
/* particle deposit (scheme 1 CIC, 2 TSC), normalisation and parent fill of
 * hmaps.h:439-541 and hmaps.cpp:389-441, on record-order rho / vel3 */
int mpfo_velocity_deposit(const mpfo_octree* o, int scheme, const float* pos, const float* vel, uint64_t np,
                          int accumulate, int finish, float* rho, float* vel3);
void mpfo_interp10(const mpfo_octree* o, int order, int precision, const double* xyz, uint64_t m, double* out10);

void mpfo_compute_key(unsigned level, const double* xyz, uint64_t m, uint64_t* lohi);
void mpfo_key_info(const uint64_t* lohi, uint64_t m, uint32_t* level, double* center3, double* ext, uint64_t* parent);
int mpfo_key_containing(const uint64_t* a, const uint64_t* b);
void mpfo_locate(const mpfo_octree* o, const double* xyz, uint64_t m, int64_t* elem, uint32_t* level);
/* order 0 ngp, 1 cic, 2 uncached tsc, 3 cached tsc over the points in sequence */
void mpfo_interp(const mpfo_octree* o, int order, int precision, const double* xyz, uint64_t m, double* out7);
uint64_t mpfo_integrate(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                        double length, unsigned nsteps, int stop_kind, double stop_value, const double* kitarget,
                        const double* in9, uint64_t n, double* out, int nthreads);
uint64_t mpfo_trajectory(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                         double length, unsigned nsteps, int stop_kind, double stop_value, const double* kitarget,
                         const double* in9, double* rows, uint64_t maxlen);
uint64_t mpfo_jacobian_infinitesimal(const mpfo_octree* o, int order, int precision, const double* cosmo_tab,
                                     uint64_t ncos, double length, unsigned nsteps, int stop_kind, double stop_value,
                                     const double* kitarget, const double* in9, const double* dist, uint64_t nd,
                                     double* out4);
/* Surfaces of Hmaps::FillMap (hmaps.h:615-660, 724-840) on the stored
 * trajectory, and the infinitesimal Jacobian at each surface distance
 * (lensing.h:399-480).  kind: 0 radius, 1 lambda, 2 redshift, 3 t, 4 a, 6 s;
 * the ray stops at vals[nsurf-1].  out: nsurf x 32 doubles (layout of
 * libmpf's MPF_SURF_*).  Returns the trajectory length. */
uint64_t mpfo_surfaces(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                       double length, unsigned nsteps, int kind, const double* vals, uint32_t nsurf, int jacobian,
                       const double* in9, double* out);
/* Catalogues::iterateNewtonMethod + newtonMethod2d (catalogues.h:410-800),
 * jacobiantype = infinitesimal, use_previous_catalogues = 0, observer at the
 * origin.  src6: x y z vx vy vz per source; init2 (may be NULL) first guesses;
 * cfg: accuracy, max_iterations, vobs[3], h, mpc.  out: 20 doubles per source
 * (libmpf's MPF_CAT_* layout).  Returns the rays integrated. */
uint64_t mpfo_catalogue(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                        double length, unsigned nsteps, const double* src6, const double* init2, uint64_t n,
                        double accuracy, unsigned max_iterations, const double* vobs, double h, double mpc,
                        double* out, int nthreads);
/* The same with jacobiantype = bundle (catalogues.h:455-491, 509-523,
 * 666-681): iterations >= 5 and the converged Jacobian from
 * Lensing::dbetadtheta with stop_bundle bundle_kind (0 radius, 1 lambda,
 * 2 redshift, 3 t, 4 a), plane 0 sachs / 1 normal and the bundle opening;
 * bundle_kind -1 is jacobiantype = infinitesimal. */
uint64_t mpfo_catalogue_bundle(const mpfo_octree* o, int order, int precision, const double* cosmo_tab,
                               uint64_t ncos, double length, unsigned nsteps, const double* src6, const double* init2,
                               uint64_t n, double accuracy, unsigned max_iterations, const double* vobs, double h,
                               double mpc, int bundle_kind, int plane, double opening, double* out, int nthreads);
/* Every mode of the root-finder: jactype 0 infinitesimal, 1
 * infinitesimal_born (the converged Jacobian on the homogeneous path of the
 * last photon launched, catalogues.h:670-674), 2 bundle (bundle_kind 0
 * radius, 1 lambda, 2 redshift, 3 t, 4 a, 5 plane); use_prev 0, or 2: the
 * grid re-run of rejected sources (catalogues.h:527-623; max_iterations 100
 * in the reference).  out[18]: grid rays that ended before the source
 * radius (extrapolated as the reference does; one-element trajectories are
 * no candidate). */
uint64_t mpfo_catalogue_ex(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                           double length, unsigned nsteps, const double* src6, const double* init2, uint64_t n,
                           double accuracy, unsigned max_iterations, const double* vobs, double h, double mpc,
                           int jactype, int bundle_kind, int plane, double opening, int use_prev, double* out,
                           int nthreads);
/* Catalogues::relCat_with_previous_cat (catalogues.h:922-1044): new Jacobian
 * columns (n x 4) of n previous-catalogue rows (18 columns); flag per row:
 * 1 one-element trajectory (42s), 2 the ray ended before the scale factor
 * (extrapolated), 0 otherwise. */
void mpfo_catalogue_prev(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                         double length, unsigned nsteps, const double* prev18, uint64_t n, int jactype,
                         int bundle_kind, int plane, double opening, double* jac4, int* flag);
/* Integrator::launch around one photon (integrator.h:345-382): out9 holds
 * the photon and its count bundle photons. */
void mpfo_launch_bundle(const double* in9, unsigned count, double angle, double rotation, double* out9);
/* Integrator::propagate (integrator.h:826-1015): central trajectory (rows,
 * MPFO_NREC doubles per element) and its angular diameter distance (dist) at
 * every element, up to maxlen elements; returns the element count, 0 when
 * the bundle is rejected.  kind: interpolation parameter (2 redshift, 4 a,
 * 3 t, 1 lambda, 0 r). */
uint64_t mpfo_propagate(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                        double length, unsigned nsteps, const double* in9, unsigned count, double angle,
                        double rotation, int kind, double amin, double* rows, double* dist, uint64_t maxlen);
/* Lensing::dbetadtheta (lensing.h:140-364), the bundle Jacobian of one
 * direction at nd values of the stop parameter `kind`. */
void mpfo_dbetadtheta(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                      double length, unsigned nsteps, int kind, double opening, const double* observer, double phi,
                      double theta, const double* ref, const double* ki, const double* dist, uint64_t nd,
                      double* out4);
void mpfo_interpolate(const double* x, const double* y, uint64_t n, const double* x0, uint64_t m, double* out);

#ifdef __cplusplus
}
#endif
#endif
