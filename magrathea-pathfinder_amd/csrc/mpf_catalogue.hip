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
};

enum { BCTX_FP = 0, BCTX_KI = 3, BCTX_REF = 6, BCTX = 8 };

// kiTarget and interpRef of the crossing (catalogues.h:458-491)
__device__ void bundle_context(const NewtonArgs& N, uint32_t i, const double* s, double D) {
    double* b = N.bctx + (uint64_t)BCTX * i;
    b[BCTX_FP] = s[SF_X]; b[BCTX_FP + 1] = s[SF_Y]; b[BCTX_FP + 2] = s[SF_Z];
    if (N.bplane == 0) { b[BCTX_KI] = s[SF_KX]; b[BCTX_KI + 1] = s[SF_KY]; b[BCTX_KI + 2] = s[SF_KZ]; }
    else { b[BCTX_KI] = s[SF_X]; b[BCTX_KI + 1] = s[SF_Y]; b[BCTX_KI + 2] = s[SF_Z]; }
    double ref;
    switch (N.bkind) {
        case MPF_STOP_REDSHIFT: ref = s[SF_REDSHIFT]; break;
        case MPF_STOP_A: ref = s[SF_A]; break;
        case MPF_STOP_T: ref = s[SF_T]; break;
        case MPF_STOP_LAMBDA: ref = s[SF_LAMBDA]; break;
        default: ref = D; break;
    }
    b[BCTX_REF] = ref;
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
    const double c = C_LIGHT, c2 = C2;
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
    if (N.jac) {  // the Jacobian of this round's ray at the source distance
        o[MPF_CAT_A11] = s[SF_A11]; o[MPF_CAT_A12] = s[SF_A12];
        o[MPF_CAT_A21] = s[SF_A12]; o[MPF_CAT_A22] = s[SF_A22];
    }
    const double fx = s[SF_X], fy = s[SF_Y], fz = s[SF_Z];
    const double sep = sqrt(pow(tt[0] - fx, 2) + pow(tt[1] - fy, 2) + pow(tt[2] - fz, 2)) / D;
    o[MPF_CAT_SEP] = sep;
    o[MPF_CAT_NROUNDS] = N.round + 1;
    if (sep > N.acc) {
        if (N.bundle && N.round >= 5) {
            // the bundle Jacobian of this crossing, computed on the host, then newton_bundle_kernel
            bundle_context(N, i, s, D);
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
    o[MPF_CAT_PHI_OBS] = phi;
    o[MPF_CAT_THETA_OBS] = theta;
    o[MPF_CAT_ITER] = N.round + 1;
    o[MPF_CAT_STATUS] = MPF_CAT_OK;
    if (N.bundle) {  // the converged Jacobian is the bundle's (catalogues.h:666-667): final pass
        bundle_context(N, i, s, D);
        N.need_jac[atomicAdd(N.nneed, 1u)] = i;
    } else if (!N.jac) N.need_jac[atomicAdd(N.nneed, 1u)] = i;
    const double as = s[SF_A];
    const double z1 = s[SF_REDSHIFT];
    const double potential = s[SF_PHI], isw = s[SF_ISW];
    const double* v = tt + 3;
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
        for (int q = MPF_CAT_Z0; q <= MPF_CAT_Z5; ++q) o[q] = 0.0;
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
    return N;
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
