// Catalogue root-finding: Catalogues::iterateNewtonMethod / newtonMethod2d
// (catalogues.h:410-800) for a batch of sources, infinitesimal Jacobian.
//
// Each Newton round integrates the still-active sources to their comoving
// distance (a radius surface per ray), then newton_kernel, per source:
//   * no crossing of the source radius -> failure (42, 42; catalogues.h:435-440)
//   * separation |x_src - x_ray(D)| / D (catalogues.h:495-496) above the
//     accuracy -> Newton update in the (e1, e2) plane at (phi, theta)
//     (catalogues.h:627-660) with the identity for iterations 0-2 and the
//     inverse infinitesimal Jacobian afterwards (catalogues.h:509-517);
//     more than max_iterations rounds -> rejected (iterations = 42)
//   * otherwise converged: observed angles, Jacobian and the six redshifts
//     of catalogues.h:682-719.
// The redshift z1 of every crossing is kept (the reference's redshifts[1]
// is written at each call, catalogues.h:451, so a rejected source carries the
// last ray's).  With use_previous_catalogues = 2 (the re-run of rejected
// sources, catalogues.h:527-623) a source that has not converged sends a grid
// of rays around its angles; the grid kernels below integrate them in one
// batch and pick the reference's choice.
#include "mpf_kernels.h"

namespace mpf {

namespace {

enum { CA = 0, CT, CX, CY, CZ, CKT, CKX, CKY, CKZ };

// launch photons from the current angles (Integrator::launch, integrator.h:304-329)
__global__ void cat_launch_kernel(const double* __restrict__ ang, const uint32_t* __restrict__ act, uint64_t nact,
                                  double* __restrict__ in9, double* __restrict__ dist_act,
                                  const double* __restrict__ dist) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nact) return;
    const uint32_t i = act[k];
    const double phi = ang[2 * i], theta = ang[2 * i + 1];
    double* r = in9 + 9 * k;
    r[CA] = 1.0; r[CT] = 0.0; r[CX] = 0.0; r[CY] = 0.0; r[CZ] = 0.0; r[CKT] = 1.0;
    r[CKX] = cos(phi) * sin(theta);
    r[CKY] = sin(phi) * sin(theta);
    r[CKZ] = cos(theta);
    dist_act[k] = dist[i];
}

// comoving distance and angles of each source (catalogues.h:834-836)
__global__ void cat_init_kernel(const double* __restrict__ src, const double* __restrict__ init, uint64_t n,
                                double* __restrict__ dist, double* __restrict__ ang, double* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* s = src + 6 * i;
    const double D = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    const double theta = acos(s[2] / D);
    const double phi = atan2(s[1], s[0]);
    dist[i] = D;
    ang[2 * i] = init ? init[2 * i] : phi;
    ang[2 * i + 1] = init ? init[2 * i + 1] : theta;
    double* o = out + (uint64_t)MPF_NCAT * i;
    for (int k = 0; k < MPF_NCAT; ++k) o[k] = 0.0;
    o[MPF_CAT_PHI] = phi;
    o[MPF_CAT_THETA] = theta;
}

struct NewtonArgs {
    const double* src;       // n x 6
    const double* dist;      // n
    double* ang;             // n x 2 current angles
    double* out;             // n x MPF_NCAT
    const double* surf;      // nact x SF_NREC
    const uint32_t* act;     // active source ids
    uint64_t nact;
    uint32_t* next;          // next round's active ids
    unsigned int* nnext;
    uint32_t* need_jac;      // converged without a Jacobian: final pass
    unsigned int* nneed;
    int round, jac, max_iter;
    double acc, length, h, mpc, vobs[3], a0, phi0;
    // jacobiantype = bundle: stop_bundle kind, plane (0 sachs, 1 normal) and
    // per source the crossing's finalPoint, kiTarget and interpRef
    int bundle, bkind, bplane;
    double* bctx;            // n x BCTX
    int born, rerun;
    uint32_t* gsrc;          // re-run mode: sources that send a grid
    unsigned int* ngrid;
    double* gerr;            // per source: the grid's half-size
    double* gas;             // per source: the central ray's scale factor at the crossing
};

enum { BCTX_FP = 0, BCTX_KI = 3, BCTX_REF = 6, BCTX_ANG = 7, BCTX_PAD = 9, BCTX = 10 };
static_assert(BCTX == CAT_BCTX, "bundle context size");

// kiTarget and interpRef of the crossing (catalogues.h:458-491, 586-616):
// s the record the plane and t / lambda come from, z1 and as the redshift
// and scale factor the reference takes there (the central ray's, also when a
// grid ray was chosen), (phi, theta) the direction the bundle is launched around
__device__ void bundle_context(const NewtonArgs& N, uint32_t i, const double* s, double D, double z1, double as,
                               double phi, double theta) {
    double* b = N.bctx + (uint64_t)BCTX * i;
    b[BCTX_FP] = s[SF_X]; b[BCTX_FP + 1] = s[SF_Y]; b[BCTX_FP + 2] = s[SF_Z];
    if (N.bplane == 0) { b[BCTX_KI] = s[SF_KX]; b[BCTX_KI + 1] = s[SF_KY]; b[BCTX_KI + 2] = s[SF_KZ]; }
    else { b[BCTX_KI] = s[SF_X]; b[BCTX_KI + 1] = s[SF_Y]; b[BCTX_KI + 2] = s[SF_Z]; }
    double ref;
    switch (N.bkind) {
        case MPF_STOP_REDSHIFT: ref = z1; break;
        case MPF_STOP_A: ref = as; break;
        case MPF_STOP_T: ref = s[SF_T]; break;
        case MPF_STOP_LAMBDA: ref = s[SF_LAMBDA]; break;
        case MPF_STOP_PLANE:
            ref = b[BCTX_KI] * b[BCTX_FP] + b[BCTX_KI + 1] * b[BCTX_FP + 1] + b[BCTX_KI + 2] * b[BCTX_FP + 2];
            break;
        default: ref = D; break;
    }
    b[BCTX_REF] = ref;
    b[BCTX_ANG] = phi;
    b[BCTX_ANG + 1] = theta;
}

// the six redshifts of a converged source (catalogues.h:682-719): as and z1
// at the central ray's crossing, the potential, ISW and wave vector of the
// record s of the ray kept (the central one, or the grid ray chosen), whose
// launch direction is (phi, theta)
__device__ void write_redshifts(const NewtonArgs& N, double* o, const double* v, double as, double z1,
                                const double* s, double phi, double theta) {
    const double c = C_LIGHT, c2 = C2;
    const double potential = s[SF_PHI], isw = s[SF_ISW];
    const double unit_t = as * as * N.mpc / (N.h * 100.0 * 1000.0);
    const double unit_l = as * N.length * 100;
    const double vx = v[0] * unit_l * 1e-2 / (unit_t * c);
    const double vy = v[1] * unit_l * 1e-2 / (unit_t * c);
    const double vz = v[2] * unit_l * 1e-2 / (unit_t * c);
    // launch element: k normalised with the potential at the observer (integrator.h:673-676)
    const double k0[3] = {cos(phi) * sin(theta), sin(phi) * sin(theta), cos(theta)};
    const double norm = sqrt((c2 * (1.0 + 2.0 * (N.phi0 / c2)) * 1.0 * 1.0) /
                             ((1.0 - 2.0 * (N.phi0 / c2)) * (k0[0] * k0[0] + k0[1] * k0[1] + k0[2] * k0[2])));
    const double kx0 = k0[0] * norm, ky0 = k0[1] * norm, kz0 = k0[2] * norm, kt0 = 1.0;
    const double* vo = N.vobs;
    const double gref = -N.a0 * c * kt0 * (1. + N.phi0 / c2 + (vo[0] * kx0 + vo[1] * ky0 + vo[2] * kz0) / (c * kt0) +
                                          0.5 * (vo[0] * vo[0] + vo[1] * vo[1] + vo[2] * vo[2]));
    const double dtdl = s[SF_KT], dxdl = s[SF_KX], dydl = s[SF_KY], dzdl = s[SF_KZ];
    const double gK = -as * c * dtdl * (1. + potential / c2 + (vx * dxdl + vy * dydl + vz * dzdl) / (c * dtdl) +
                                        0.5 * (vx * vx + vy * vy + vz * vz));
    o[MPF_CAT_Z0] = -1 + 1. / as;
    o[MPF_CAT_Z1] = z1;
    o[MPF_CAT_Z2] = z1 + (vx * dxdl + vy * dydl + vz * dzdl) / (as * c * dtdl) -
                    (vo[0] * kx0 + vo[1] * ky0 + vo[2] * kz0) / (as * c * kt0);
    o[MPF_CAT_Z3] = o[MPF_CAT_Z2] + 0.5 * (vx * vx + vy * vy + vz * vz - vo[0] * vo[0] - vo[1] * vo[1] - vo[2] * vo[2]) / as;
    o[MPF_CAT_Z4] = o[MPF_CAT_Z3] + isw / as;
    o[MPF_CAT_Z5] = gK / gref - 1;
}

// a converged source: observed angles, iterations, then its Jacobian (a
// final pass, or this round's) and redshifts
__device__ void converge(const NewtonArgs& N, uint32_t i, double* o, const double* s, double D, double z1,
                         double as, double phi, double theta, double bphi, double btheta) {
    o[MPF_CAT_PHI_OBS] = phi;
    o[MPF_CAT_THETA_OBS] = theta;
    o[MPF_CAT_ITER] = N.round + 1;
    o[MPF_CAT_STATUS] = MPF_CAT_OK;
    N.ang[2 * i] = phi;
    N.ang[2 * i + 1] = theta;
    if (N.bundle) {  // the converged Jacobian is the bundle's (catalogues.h:666-667): final pass
        bundle_context(N, i, s, D, z1, as, bphi, btheta);
        N.need_jac[atomicAdd(N.nneed, 1u)] = i;
    } else if (!N.jac && !N.born) {
        // the ray's own Jacobian, from one more pass (the Born pass replaces it)
        N.need_jac[atomicAdd(N.nneed, 1u)] = i;
    }
    write_redshifts(N, o, N.src + 6 * i + 3, as, z1, s, phi, theta);
}

// the Newton update in the (e1, e2) plane at (phi, theta) (catalogues.h:627-660)
// with the inverse Jacobian j; false when the source is rejected
__device__ bool newton_update(const NewtonArgs& N, uint32_t i, double fx, double fy, double fz, double j00,
                              double j01, double j10, double j11) {
    const double* tt = N.src + 6 * i;
    double* o = N.out + (uint64_t)MPF_NCAT * i;
    const double D = N.dist[i];
    const double phi = N.ang[2 * i], theta = N.ang[2 * i + 1];
    const double cp = cos(phi), sp = sin(phi), ct = cos(theta), st = sin(theta);
    const double e1[3] = {-sp, cp, 0.0};
    const double e2[3] = {-cp * ct, -sp * ct, st};
    const double tg[3] = {D * cp * st, D * sp * st, D * ct};
    const double sv[3] = {fx - tt[0], fy - tt[1], fz - tt[2]};
    const double pv[3] = {tg[0] - tt[0], tg[1] - tt[1], tg[2] - tt[2]};
    const double sepx = sv[0] * e1[0] + sv[1] * e1[1];
    const double sepy = sv[0] * e2[0] + sv[1] * e2[1] + sv[2] * e2[2];
    const double posx = pv[0] * e1[0] + pv[1] * e1[1];
    const double posy = pv[0] * e2[0] + pv[1] * e2[1] + pv[2] * e2[2];
    const double sepxnew = j00 * sepx + j01 * sepy;
    const double sepynew = j10 * sepx + j11 * sepy;
    const double posxnew = posx - sepxnew;
    const double posynew = posy - sepynew;
    const double xn = tt[0] + (posxnew * e1[0] + posynew * e2[0]);
    const double yn = tt[1] + (posxnew * e1[1] + posynew * e2[1]);
    const double zn = tt[2] + posynew * e2[2];
    const double nphi = atan2(yn, xn);
    const double ntheta = acos(zn / sqrt(xn * xn + yn * yn + zn * zn));
    N.ang[2 * i] = nphi;
    N.ang[2 * i + 1] = ntheta;
    if (N.round + 1 > N.max_iter) {  // rejected (catalogues.h:775-778)
        o[MPF_CAT_PHI_OBS] = nphi; o[MPF_CAT_THETA_OBS] = ntheta;
        o[MPF_CAT_ITER] = 42;
        o[MPF_CAT_STATUS] = MPF_CAT_REJECTED;
        return false;
    }
    return true;
}

__global__ void newton_kernel(NewtonArgs N) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N.nact) return;
    const uint32_t i = N.act[k];
    const double* s = N.surf + (uint64_t)SF_NREC * k;
    const double* tt = N.src + 6 * i;
    double* o = N.out + (uint64_t)MPF_NCAT * i;
    const double D = N.dist[i];
    const double phi = N.ang[2 * i], theta = N.ang[2 * i + 1];
    if (s[SF_CROSSED] == 0.0) {  // the ray never reached the source radius
        o[MPF_CAT_PHI_OBS] = 42; o[MPF_CAT_THETA_OBS] = 42; o[MPF_CAT_SEP] = 0;
        // firstid of catalogues.h:433-439: the last element minus one
        o[MPF_CAT_ITER] = s[SF_FIRSTID] - (s[SF_FIRSTID] > 0.0 ? 1.0 : 0.0);
        o[MPF_CAT_STATUS] = MPF_CAT_OUTSIDE;
        o[MPF_CAT_NROUNDS] = N.round + 1;
        return;
    }
    o[MPF_CAT_Z1] = s[SF_REDSHIFT];   // redshifts[1] of every call (catalogues.h:451)
    if (N.jac) {  // the Jacobian of this round's ray at the source distance
        o[MPF_CAT_A11] = s[SF_A11]; o[MPF_CAT_A12] = s[SF_A12];
        o[MPF_CAT_A21] = s[SF_A12]; o[MPF_CAT_A22] = s[SF_A22];
    }
    const double fx = s[SF_X], fy = s[SF_Y], fz = s[SF_Z];
    const double sep = sqrt(pow(tt[0] - fx, 2) + pow(tt[1] - fy, 2) + pow(tt[2] - fz, 2)) / D;
    o[MPF_CAT_SEP] = sep;
    o[MPF_CAT_NROUNDS] = N.round + 1;
    if (sep > N.acc) {
        if (N.rerun) {  // a grid of rays around (phi, theta) (catalogues.h:527-531)
            N.gerr[i] = 5 * sep / (N.round + 1);
            N.gas[i] = s[SF_A];
            if (N.bundle) bundle_context(N, i, s, D, s[SF_REDSHIFT], s[SF_A], phi, theta);
            N.gsrc[atomicAdd(N.ngrid, 1u)] = i;
            return;
        }
        if (N.bundle && N.round >= 5) {
            // the bundle Jacobian of this crossing, computed on the host, then newton_bundle_kernel
            bundle_context(N, i, s, D, s[SF_REDSHIFT], s[SF_A], phi, theta);
            N.next[atomicAdd(N.nnext, 1u)] = i;
            return;
        }
        double j00 = 1, j01 = 0, j10 = 0, j11 = 1;
        if (N.round >= 3) {  // inverse of the infinitesimal Jacobian (utility.h:687-698), a21 = a12
            const double a11 = s[SF_A11], a12 = s[SF_A12], a21 = s[SF_A12], a22 = s[SF_A22];
            const double invdet = 1. / (a11 * a22 - a12 * a21);
            j00 = a22 * invdet; j01 = -a12 * invdet; j10 = -a21 * invdet; j11 = a11 * invdet;
        }
        if (newton_update(N, i, fx, fy, fz, j00, j01, j10, j11)) N.next[atomicAdd(N.nnext, 1u)] = i;
        return;
    }
    // converged (catalogues.h:665-719)
    converge(N, i, o, s, D, s[SF_REDSHIFT], s[SF_A], phi, theta, phi, theta);
}

// ---- the grid of the re-run mode (catalogues.h:527-623)

// for (double i = -err; i <= err; i += 0.2 * err): the number of values and
// the k-th value, accumulated as the reference's loop does.  The count is
// held to 11: a grid whose step vanishes against err (err below the smallest
// normal numbers) would not end.
constexpr int GRID_MAX = 11;
__device__ int grid_count(double err) {
    int k = 0;
    for (double v = -err; v <= err && k < GRID_MAX; v += 0.2 * err) ++k;
    return k;
}
__device__ double grid_value(double err, int k) {
    double v = -err;
    for (int q = 0; q < k; ++q) v += 0.2 * err;
    return v;
}

__global__ void grid_count_kernel(NewtonArgs N, uint64_t ngrid, uint32_t* __restrict__ cnt) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngrid) return;
    const int m = grid_count(N.gerr[N.gsrc[g]]);
    cnt[g] = (uint32_t)(m * m);
}

// launch photons of every grid ray, (i, j) in the reference's loop order:
// slot off[g] + a * m + b holds i = value a, j = value b
__global__ void grid_launch_kernel(NewtonArgs N, uint64_t ngrid, const uint32_t* __restrict__ off,
                                   const uint32_t* __restrict__ cnt, double* __restrict__ in9,
                                   double* __restrict__ dist_slot) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngrid) return;
    const uint32_t i = N.gsrc[g];
    const double err = N.gerr[i];
    const int m = grid_count(err);
    const double phi = N.ang[2 * i], theta = N.ang[2 * i + 1], D = N.dist[i];
    for (int a = 0; a < m; ++a) {
        const double di = grid_value(err, a);
        for (int b = 0; b < m; ++b) {
            const double dj = grid_value(err, b);
            const uint64_t q = (uint64_t)off[g] + (uint64_t)(a * m + b);
            const double ph = phi + di, th = theta + dj;
            double* r = in9 + 9 * q;
            r[CA] = 1.0; r[CT] = 0.0; r[CX] = 0.0; r[CY] = 0.0; r[CZ] = 0.0; r[CKT] = 1.0;
            r[CKX] = cos(ph) * sin(th);
            r[CKY] = sin(ph) * sin(th);
            r[CKZ] = cos(th);
            dist_slot[q] = D;
        }
    }
    (void)cnt;
}

// The reference scans the grid in loop order: the first ray within the
// accuracy ends the search (catalogues.h:556-603); before it, a ray closer
// than the best so far (the central ray first) becomes the best.  A grid ray
// that ends before the source radius is no candidate (the reference
// extrapolates its last segment, or reads past a one-element trajectory).
__global__ void grid_select_kernel(NewtonArgs N, uint64_t ngrid, const uint32_t* __restrict__ off,
                                   const uint32_t* __restrict__ cnt, const double* __restrict__ gsurf) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngrid) return;
    const uint32_t i = N.gsrc[g];
    const double* tt = N.src + 6 * i;
    double* o = N.out + (uint64_t)MPF_NCAT * i;
    const double D = N.dist[i];
    const double err = N.gerr[i];
    const int m = grid_count(err);
    const double phi = N.ang[2 * i], theta = N.ang[2 * i + 1];
    double best = o[MPF_CAT_SEP];
    int bk = -1;
    bool good = false;
    for (int k = 0; k < (int)cnt[g]; ++k) {
        const double* s = gsurf + (uint64_t)SF_NREC * ((uint64_t)off[g] + (uint64_t)k);
        if (s[SF_CROSSED] == 0.0) continue;
        const double sp = sqrt(pow(tt[0] - s[SF_X], 2) + pow(tt[1] - s[SF_Y], 2) + pow(tt[2] - s[SF_Z], 2)) / D;
        if (sp < N.acc) {
            best = sp;
            bk = k;
            good = true;
            break;
        }
        if (sp < best) {
            best = sp;
            bk = k;
        }
    }
    double nphi = phi, ntheta = theta;
    if (bk >= 0) {
        nphi = phi + grid_value(err, bk / m);
        ntheta = theta + grid_value(err, bk % m);
    }
    o[MPF_CAT_SEP] = best;
    if (good) {
        const double* s = gsurf + (uint64_t)SF_NREC * ((uint64_t)off[g] + (uint64_t)bk);
        // z1 and the scale factor stay the central ray's; the bundle is
        // launched around the central direction (catalogues.h:586-616, 666)
        converge(N, i, o, s, D, o[MPF_CAT_Z1], N.gas[i], nphi, ntheta, phi, theta);
        return;
    }
    N.ang[2 * i] = nphi;
    N.ang[2 * i + 1] = ntheta;
    if (N.round + 1 > N.max_iter) {  // rejected (catalogues.h:775-778)
        o[MPF_CAT_PHI_OBS] = nphi; o[MPF_CAT_THETA_OBS] = ntheta;
        o[MPF_CAT_ITER] = 42;
        o[MPF_CAT_STATUS] = MPF_CAT_REJECTED;
        return;
    }
    N.next[atomicAdd(N.nnext, 1u)] = i;
}

// rounds >= 5 with jacobiantype = bundle: the update with the inverse of the
// bundle Jacobian J (4 per pending source, from mpf_jacobian_bundle), or the
// failure 42, 42, 22 (catalogues.h:518-523)
__global__ void newton_bundle_kernel(NewtonArgs N, const uint32_t* __restrict__ pend, uint64_t npend,
                                     const double* __restrict__ J, uint32_t* __restrict__ next2,
                                     unsigned int* __restrict__ nnext2) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npend) return;
    const uint32_t i = pend[k];
    const double* j = J + 4 * k;
    double* o = N.out + (uint64_t)MPF_NCAT * i;
    o[MPF_CAT_A11] = j[0]; o[MPF_CAT_A12] = j[1]; o[MPF_CAT_A21] = j[2]; o[MPF_CAT_A22] = j[3];
    if (j[0] == 42 && j[1] == 42) {
        o[MPF_CAT_PHI_OBS] = 42; o[MPF_CAT_THETA_OBS] = 42; o[MPF_CAT_SEP] = 22;
        o[MPF_CAT_ITER] = 0;  // left unset by the reference
        o[MPF_CAT_STATUS] = MPF_CAT_OUTSIDE;
        return;
    }
    const double a11 = j[0], a12 = j[1], a21 = j[2], a22 = j[3];
    const double invdet = 1. / (a11 * a22 - a12 * a21);
    const double* b = N.bctx + (uint64_t)BCTX * i;
    if (newton_update(N, i, b[BCTX_FP], b[BCTX_FP + 1], b[BCTX_FP + 2], a22 * invdet, -a12 * invdet, -a21 * invdet,
                      a11 * invdet))
        next2[atomicAdd(nnext2, 1u)] = i;
}

// the converged sources' bundle Jacobian (catalogues.h:666-681)
__global__ void cat_jac_bundle_kernel(const uint32_t* __restrict__ ids, uint64_t m, const double* __restrict__ J,
                                      double* __restrict__ out) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const double* j = J + 4 * k;
    double* o = out + (uint64_t)MPF_NCAT * ids[k];
    o[MPF_CAT_A11] = j[0]; o[MPF_CAT_A12] = j[1]; o[MPF_CAT_A21] = j[2]; o[MPF_CAT_A22] = j[3];
    if (j[0] == 42 && j[1] == 42) {
        o[MPF_CAT_PHI_OBS] = 42; o[MPF_CAT_THETA_OBS] = 42; o[MPF_CAT_SEP] = 22;
        o[MPF_CAT_ITER] = 0;  // left unset by the reference
        o[MPF_CAT_STATUS] = MPF_CAT_OUTSIDE;
        // only redshifts[1] was written by the call that failed
        for (int q = MPF_CAT_Z0; q <= MPF_CAT_Z5; ++q)
            if (q != MPF_CAT_Z1) o[q] = 0.0;
    }
}

// Jacobian of the converged ray, from a pass with the Jacobian on
__global__ void cat_jac_kernel(const uint32_t* __restrict__ act, uint64_t nact, const double* __restrict__ surf,
                               double* __restrict__ out) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nact) return;
    const double* s = surf + (uint64_t)SF_NREC * k;
    double* o = out + (uint64_t)MPF_NCAT * act[k];
    o[MPF_CAT_A11] = s[SF_A11]; o[MPF_CAT_A12] = s[SF_A12];
    o[MPF_CAT_A21] = s[SF_A12]; o[MPF_CAT_A22] = s[SF_A22];
}

// the Born Jacobian of converged sources (jacobiantype = infinitesimal_born,
// catalogues.h:670-674): k runs over every source
__global__ void cat_jac_born_kernel(uint64_t n, const double* __restrict__ surf, double* __restrict__ out) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double* o = out + (uint64_t)MPF_NCAT * k;
    if (o[MPF_CAT_STATUS] != (double)MPF_CAT_OK) return;
    const double* s = surf + (uint64_t)SF_NREC * k;
    o[MPF_CAT_A11] = s[SF_A11]; o[MPF_CAT_A12] = s[SF_A12];
    o[MPF_CAT_A21] = s[SF_A12]; o[MPF_CAT_A22] = s[SF_A22];
}

inline unsigned blocks_for(uint64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

hipError_t launch_cat_jac_born(uint64_t n, const double* surf, double* out, hipStream_t st) {
    if (n) hipLaunchKernelGGL(cat_jac_born_kernel, dim3(blocks_for(n)), dim3(256), 0, st, n, surf, out);
    return hipGetLastError();
}

hipError_t launch_cat_init(const double* src, const double* init, uint64_t n, double* dist, double* ang, double* out,
                           hipStream_t st) {
    if (n) hipLaunchKernelGGL(cat_init_kernel, dim3(blocks_for(n)), dim3(256), 0, st, src, init, n, dist, ang, out);
    return hipGetLastError();
}

hipError_t launch_cat_launch(const double* ang, const uint32_t* act, uint64_t nact, double* in9, double* dist_act,
                             const double* dist, hipStream_t st) {
    if (nact) hipLaunchKernelGGL(cat_launch_kernel, dim3(blocks_for(nact)), dim3(256), 0, st, ang, act, nact, in9,
                                 dist_act, dist);
    return hipGetLastError();
}

static NewtonArgs newton_args(const CatRound& R) {
    NewtonArgs N;
    N.src = R.src; N.dist = R.dist; N.ang = R.ang; N.out = R.out; N.surf = R.surf; N.act = R.act; N.nact = R.nact;
    N.next = R.next; N.nnext = R.nnext; N.need_jac = R.need_jac; N.nneed = R.nneed;
    N.round = R.round; N.jac = R.jac; N.max_iter = R.max_iter; N.acc = R.acc; N.length = R.length; N.h = R.h;
    N.mpc = R.mpc; N.a0 = R.a0; N.phi0 = R.phi0;
    for (int k = 0; k < 3; ++k) N.vobs[k] = R.vobs[k];
    N.bundle = R.bundle; N.bkind = R.bkind; N.bplane = R.bplane; N.bctx = R.bctx;
    N.born = R.born; N.rerun = R.rerun; N.gsrc = R.gsrc; N.ngrid = R.ngrid; N.gerr = R.gerr; N.gas = R.gas;
    return N;
}

hipError_t launch_cat_grid_count(const CatRound& R, uint64_t ngrid, uint32_t* cnt, hipStream_t st) {
    const NewtonArgs N = newton_args(R);
    if (ngrid) hipLaunchKernelGGL(grid_count_kernel, dim3(blocks_for(ngrid)), dim3(256), 0, st, N, ngrid, cnt);
    return hipGetLastError();
}

hipError_t launch_cat_grid_launch(const CatRound& R, uint64_t ngrid, const uint32_t* off, const uint32_t* cnt,
                                  double* in9, double* dist_slot, hipStream_t st) {
    const NewtonArgs N = newton_args(R);
    if (ngrid)
        hipLaunchKernelGGL(grid_launch_kernel, dim3(blocks_for(ngrid)), dim3(256), 0, st, N, ngrid, off, cnt, in9,
                           dist_slot);
    return hipGetLastError();
}

hipError_t launch_cat_grid_select(const CatRound& R, uint64_t ngrid, const uint32_t* off, const uint32_t* cnt,
                                  const double* gsurf, hipStream_t st) {
    const NewtonArgs N = newton_args(R);
    if (ngrid)
        hipLaunchKernelGGL(grid_select_kernel, dim3(blocks_for(ngrid)), dim3(256), 0, st, N, ngrid, off, cnt, gsurf);
    return hipGetLastError();
}

hipError_t launch_newton(const CatRound& R, hipStream_t st) {
    const NewtonArgs N = newton_args(R);
    if (R.nact) hipLaunchKernelGGL(newton_kernel, dim3(blocks_for(R.nact)), dim3(256), 0, st, N);
    return hipGetLastError();
}

hipError_t launch_newton_bundle(const CatRound& R, const uint32_t* pend, uint64_t npend, const double* J,
                                uint32_t* next2, unsigned int* nnext2, hipStream_t st) {
    const NewtonArgs N = newton_args(R);
    if (npend)
        hipLaunchKernelGGL(newton_bundle_kernel, dim3(blocks_for(npend)), dim3(256), 0, st, N, pend, npend, J, next2,
                           nnext2);
    return hipGetLastError();
}

hipError_t launch_cat_jac_bundle(const uint32_t* ids, uint64_t m, const double* J, double* out, hipStream_t st) {
    if (m) hipLaunchKernelGGL(cat_jac_bundle_kernel, dim3(blocks_for(m)), dim3(256), 0, st, ids, m, J, out);
    return hipGetLastError();
}

hipError_t launch_cat_jac(const uint32_t* act, uint64_t nact, const double* surf, double* out, hipStream_t st) {
    if (nact) hipLaunchKernelGGL(cat_jac_kernel, dim3(blocks_for(nact)), dim3(256), 0, st, act, nact, surf, out);
    return hipGetLastError();
}

}  // namespace mpf
