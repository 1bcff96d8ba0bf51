// Kernel launch interface shared by the C ABI and the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/mpf.h"
#include "mpf_device.h"
#include "mpf_surfaces.h"

namespace mpf {

struct IntegrateArgs {
    const double* in9 = nullptr;     // n x 9 launch photons
    double* out = nullptr;           // n x MPF_NOUT
    uint64_t n = 0;
    unsigned long long* queue = nullptr;  // work-queue counter (zeroed before launch)
    // queue position -> ray index (null: input order), and the number of idle
    // lanes a wavefront gathers before it takes rays from the queue (1..64;
    // a wavefront with no active lane always takes rays)
    const uint32_t* queue_order = nullptr;
    unsigned refill = 1;
    unsigned long long* steps = nullptr;  // accepted-step counter
    const double* ct = nullptr;      // cosmology table columns
    const double* ca = nullptr;
    const double* cdadt = nullptr;
    const double* cslope = nullptr;  // d(da/dt)/dt of each table segment (split tier)
    int ncos = 0;
    int order = 2;
    unsigned nsteps = 10;
    double length = 1.0;
    int stop_kind = -1;
    double stop_value = 0.0;
    double kit[3] = {0.0, 0.0, 0.0};
    double vobs[3] = {0.0, 0.0, 0.0};  // observer velocity / c, trees with a velocity field
    double* scratch = nullptr;       // per-ray launch records (split tier)
    // constant-parameter surfaces (hmaps.h:615-660) and the infinitesimal Jacobian
    int surf_kind = 0;               // 0 r, 1 lambda, 2 redshift, 3 t, 4 a, 6 s
    int nsurf = 0;
    double surf_val[16] = {0};
    double* surf_out = nullptr;      // n x nsurf x SF_NREC
    const double* ray_value = nullptr; // per-ray stop/surface value (one surface), or null
    int jac = 0;
    // the distance the Jacobian is taken at: 0 the source distance of a
    // per-ray radius (catalogues.h:516), else the interpolated position's
    // (hmaps.h:667); 1 the interpolated chi (catalogues.h:955)
    int jac_at_chi = 0;
    // whole trajectories (exact tiers): element k of ray i at
    // traj[(i * traj_cap + k) * MPF_NOUT], k < traj_cap; traj_over counts
    // rays with more elements than that
    double* traj = nullptr;
    uint64_t traj_cap = 0;
    unsigned long long* traj_over = nullptr;
    // planes crossed by bundle rays (Lensing::dbetadtheta with stop_bundle =
    // plane, exact tiers): the rays ray_group * plane_div .. + plane_div - 1
    // share nplane planes, values plane_val[group * nplane + j] and normals
    // plane_kit[3 * (group * nplane + j)]; each ray stops on its group's last
    // plane (stop_kind 5) and records in plane_out[(ray * nplane + j) *
    // PL_NREC] where its trajectory crosses plane j
    int nplane = 0;
    unsigned plane_div = 1;
    const double* plane_val = nullptr;
    const double* plane_kit = nullptr;
    double* plane_out = nullptr;
};

// per ray and plane record of IntegrateArgs::plane_out (doubles).  PL_STATE:
// 1 the crossing is in PL_X..PL_Z, 0 the ray never reaches the plane (the
// reference's failure entries), 2 the projections along the normal were not
// non-decreasing along the trajectory, or the trajectory ended on the plane
// after a popped element: the crossing needs the whole trajectory
enum { PL_STATE = 0, PL_FIRSTID, PL_F, PL_X, PL_Y, PL_Z, PL_NREC };

hipError_t launch_integrate_exact(int precision, const IntegrateArgs& A, const DevTree& t, int grid, hipStream_t st);
// the cell-major copies DevTree::q32 / q64 from the structure-of-arrays fields (q64 when t.f64)
hipError_t launch_field_quads(const DevTree& t, float* q32, double* q64, hipStream_t st);
// velocity field (mpf_velocity.hip): one deposit pass at level lvl into the
// n x 4 FP64 sums acc; then the normalisation and the parent fill, with
// level_first[l - min_level] the first level-sorted cell of level l
// (max_level - min_level + 2 entries) and parent an n-entry work array
// particles in Morton order of their position: sorted copies of pos and vel
// in the work area (velocity_order_work_bytes(n) bytes, n < 2^31)
size_t velocity_order_work_bytes(uint64_t n);
hipError_t launch_velocity_order(const float* pos, const float* vel, uint64_t n, void* work, const float** pos_s,
                                 const float** vel_s, hipStream_t st);
hipError_t launch_velocity_deposit(const DevTree& t, int scheme, const float* pos, const float* vel, uint64_t np,
                                   unsigned lvl, double* acc, hipStream_t st);
hipError_t launch_velocity_finish(const DevTree& t, const double* acc, float* rho32, double* rho64, float* vel,
                                  uint32_t* parent, const uint64_t* level_first, unsigned min_level,
                                  unsigned max_level, hipStream_t st);
hipError_t launch_locate(const DevTree& t, const double* xyz, uint64_t n, int64_t* elem, uint32_t* level,
                         hipStream_t st);
hipError_t launch_interp(int precision, const DevTree& t, int order, const double* xyz, uint64_t n, double* out7,
                         hipStream_t st);
hipError_t launch_surface_final(const IntegrateArgs& A, hipStream_t st);
hipError_t launch_bin(int64_t nside, const double* in9, const double* surf, int nsurf, int isurf, uint64_t n,
                      unsigned long long* acc, double* maps, hipStream_t st);
hipError_t launch_pixel_fill(int64_t nside, const int64_t* pix, const double* surf, int nsurf, uint64_t n,
                             double* maps, hipStream_t st);
hipError_t launch_component_fill(int64_t nside, const int64_t* pix, const double* surf, const double* surf_born,
                                 const double* jac, const double* homo, uint64_t n, int nz, const int* comps,
                                 int ncomp, int nmaps, float* maps, hipStream_t st);
hipError_t launch_pixel_launch(int64_t nside, const int64_t* pix, uint64_t n, double* in9, hipStream_t st);
hipError_t launch_vec2pix(int64_t nside, const double* vec, uint64_t n, int64_t* pix, hipStream_t st);
hipError_t launch_pix2vec(int64_t nside, const int64_t* pix, uint64_t n, double* vec, hipStream_t st);
struct CatRound {
    const double* src = nullptr;
    const double* dist = nullptr;
    double* ang = nullptr;
    double* out = nullptr;
    const double* surf = nullptr;
    const uint32_t* act = nullptr;
    uint64_t nact = 0;
    uint32_t* next = nullptr;
    unsigned int* nnext = nullptr;
    uint32_t* need_jac = nullptr;
    unsigned int* nneed = nullptr;
    int round = 0, jac = 0, max_iter = 10;
    double acc = 1e-8, length = 1.0, h = 0.7, mpc = 3.08568e22, vobs[3] = {0, 0, 0}, a0 = 1.0, phi0 = 0.0;
    // jacobiantype = bundle: stop_bundle (MPF_STOP_*), plane (0 sachs, 1 normal), and the
    // per-source crossing context (n x CAT_BCTX: finalPoint, kiTarget, interpRef, the bundle's
    // direction) the host reads
    int bundle = 0, bkind = 0, bplane = 0;
    double* bctx = nullptr;
    // jacobiantype = infinitesimal_born: the converged Jacobian comes from a Born pass
    int born = 0;
    // use_previous_catalogues = 2: non-converged sources go to the grid list (gsrc, ngrid) with
    // their grid half-size gerr and the central ray's scale factor gas, per source
    int rerun = 0;
    uint32_t* gsrc = nullptr;
    unsigned int* ngrid = nullptr;
    double* gerr = nullptr;
    double* gas = nullptr;
};
constexpr int CAT_BCTX = 10;
// the grid of the re-run mode: rays per grid source (cnt), their launch
// photons and radii from the slot offsets (off, an exclusive sum of cnt)
hipError_t launch_cat_grid_count(const CatRound& R, uint64_t ngrid, uint32_t* cnt, hipStream_t st);
hipError_t launch_cat_grid_launch(const CatRound& R, uint64_t ngrid, const uint32_t* off, const uint32_t* cnt,
                                  double* in9, double* dist_slot, hipStream_t st);
// the choice among each source's grid rays (surf: their radius surfaces), survivors to R.next
hipError_t launch_cat_grid_select(const CatRound& R, uint64_t ngrid, const uint32_t* off, const uint32_t* cnt,
                                  const double* gsurf, hipStream_t st);
hipError_t launch_cat_init(const double* src, const double* init, uint64_t n, double* dist, double* ang, double* out,
                           hipStream_t st);
hipError_t launch_cat_launch(const double* ang, const uint32_t* act, uint64_t nact, double* in9, double* dist_act,
                             const double* dist, hipStream_t st);
hipError_t launch_newton(const CatRound& R, hipStream_t st);
// rounds >= 5 with the bundle Jacobian: J (4 per pending source), survivors to next2
hipError_t launch_newton_bundle(const CatRound& R, const uint32_t* pend, uint64_t npend, const double* J,
                                uint32_t* next2, unsigned int* nnext2, hipStream_t st);
hipError_t launch_cat_jac_bundle(const uint32_t* ids, uint64_t m, const double* J, double* out, hipStream_t st);
hipError_t launch_cat_jac(const uint32_t* act, uint64_t nact, const double* surf, double* out, hipStream_t st);
hipError_t launch_cat_jac_born(uint64_t n, const double* surf, double* out, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v, hipStream_t st);

}  // namespace mpf
