// Constant-parameter surfaces and the infinitesimal lensing Jacobian,
// accumulated while a ray is integrated.
//
// Hmaps::FillMap (hmaps.h:615-660) and Catalogues::newtonMethod2d
// (catalogues.h:431-456) locate, on the stored trajectory, the last element
// before a surface (first element whose parameter passes the surface value,
// minus one) and interpolate linearly between it and the next element with
// f = (v[i+1] - ref) / (v[i+1] - v[i]).  Here the crossing is detected when the
// next element is appended and the interpolated values are written at once.
//
// Lensing::dbetadtheta_infinitesimal (lensing.h:399-480) sums, over elements
// i >= 1 up to the last one with chi <= D,
//     a_c(D) = sum (D - chi_i) chi_i / D * T_c,i * (chi_i - chi_i-1)
// with T the angular tidal components (d1d1, d1d2, d2d2) from 6 TSC samples
// of grad phi at +-1/2 dl around the element.  Writing
// (D - chi) chi / D = chi - chi^2 / D, two running sums per component,
// S1 = sum chi T dr and S2 = sum chi^2 T dr, give a_c(D) = S1 - S2 / D for
// any D known only when the surface is reached.
#pragma once
#include "mpf_device.h"

namespace mpf {

constexpr int MAX_SURF = 16;

// per ray and surface output record (doubles)
enum {
    SF_VALID = 0, SF_FIRSTID, SF_F, SF_X, SF_Y, SF_Z, SF_DIST, SF_CHI, SF_LAMBDA, SF_T, SF_A, SF_REDSHIFT, SF_S,
    SF_ISW, SF_RHO, SF_PHI, SF_RHOMAX, SF_KT, SF_KX, SF_KY, SF_KZ, SF_A11, SF_A12, SF_A22,
    SF_S1 = 24,  // 24..26: S1 snapshot (11, 12, 22); 27..29: S2 snapshot
    SF_REF = 30,      // the surface value
    SF_CROSSED = 31,  // the surface was passed (valid or not)
    SF_NREC = 32
};

// the quantities of one trajectory element the surfaces need
struct SurfElem {
    double x, y, z, chi, lambda, t, a, redshift, s, isw, rho, phi, kt, kx, ky, kz;
};

// Accumulators of one ray in registers, with the running maximum of rho
struct SurfAcc {
    double S1[3], S2[3];
    double rmax_prev;    // max rho over elements 0 .. i-2
    double rmax_cur;     // max rho over elements 0 .. i-1
    uint32_t crossed;    // bit j: surface j passed
    __device__ double s1(int c) const { return S1[c]; }
    __device__ double s2(int c) const { return S2[c]; }
    __device__ void add(int c, double d1, double d2) { S1[c] += d1; S2[c] += d2; }
    __device__ double rho_max() const { return rmax_prev; }
    __device__ void push_rho(double rho) {
        rmax_prev = rmax_cur;
        rmax_cur = (rho > rmax_cur) ? rho : rmax_cur;
    }
};

__device__ __forceinline__ void surf_init(SurfAcc& s, double rho0) {
#pragma unroll
    for (int c = 0; c < 3; ++c) { s.S1[c] = 0.0; s.S2[c] = 0.0; }
    s.rmax_prev = -1.0e300;
    s.rmax_cur = rho0;
    s.crossed = 0;
}

// The same with the sums and the rho maxima in LDS (S[k * stride], k = S1
// 0..2, S2 0..2, the maximum of rho over elements 0 .. i-2, over 0 .. i-1)
struct SurfAccLds {
    double* S;
    unsigned stride;
    uint32_t crossed;
    __device__ double s1(int c) const { return S[c * stride]; }
    __device__ double s2(int c) const { return S[(3 + c) * stride]; }
    __device__ void add(int c, double d1, double d2) { S[c * stride] += d1; S[(3 + c) * stride] += d2; }
    __device__ double rho_max() const { return S[6 * stride]; }
    __device__ void push_rho(double rho) {
        const double cur = S[7 * stride];
        S[6 * stride] = cur;
        S[7 * stride] = (rho > cur) ? rho : cur;
    }
};
constexpr int SURF_LDS_DOUBLES = 8;

__device__ __forceinline__ void surf_init(SurfAccLds& s, double rho0) {
#pragma unroll
    for (int k = 0; k < 6; ++k) s.S[k * s.stride] = 0.0;
    s.S[6 * s.stride] = -1.0e300;
    s.S[7 * s.stride] = rho0;
    s.crossed = 0;
}

// The surface parameter of an element, chosen among scalars: chosen among
// the fields of an element struct with a runtime kind, the compiler reads it
// at a variable offset and keeps the whole struct in scratch memory.
__device__ __forceinline__ double surf_param(int kind, double chi, double lambda, double redshift, double t, double a,
                                             double s) {
    double v = s;
    v = (kind == 4) ? a : v;
    v = (kind == 3) ? t : v;
    v = (kind == 2) ? redshift : v;
    v = (kind == 1) ? lambda : v;
    v = (kind == 0) ? chi : v;
    return v;
}

// Called after element i (>= 1) is appended, with the previous element.
// `tidal` (d1d1, d1d2, d2d2) of element i is added after the crossings are
// recorded, so a snapshot holds the sums up to element i-1 = firstid.
// vals: the surface values, or, when per_ray, the single value rv; vp, vn:
// the surface parameter (surf_param) of the previous and the new element.
template <class Acc>
__device__ __forceinline__ void surf_append(Acc& s, int kind, int nsurf, const double* vals, bool per_ray,
                                            double rv, double* rec, uint32_t i, const SurfElem& pv,
                                            const SurfElem& nw, double vp, double vn, const double* tidal,
                                            bool jac) {
    for (int j = 0; j < nsurf; ++j) {
        if (s.crossed & (1u << j)) continue;
        const double ref = per_ray ? rv : vals[j];
        const bool pass = (kind == 4) ? (vn < ref) : (vn > ref);
        if (!pass) continue;
        s.crossed |= (1u << j);
        double* o = rec + (uint64_t)SF_NREC * j;
        const double f = (vn - ref) / (vn - vp);
        o[SF_VALID] = (i >= 2) ? 1.0 : 0.0;   // firstid == 0 is rejected (hmaps.h:716)
        o[SF_REF] = ref;
        o[SF_CROSSED] = 1.0;
        o[SF_FIRSTID] = (double)(i - 1);
        o[SF_F] = f;
        o[SF_X] = pv.x * f + nw.x * (1 - f);
        o[SF_Y] = pv.y * f + nw.y * (1 - f);
        o[SF_Z] = pv.z * f + nw.z * (1 - f);
        o[SF_DIST] = sqrt(o[SF_X] * o[SF_X] + o[SF_Y] * o[SF_Y] + o[SF_Z] * o[SF_Z]);
        o[SF_CHI] = pv.chi * f + nw.chi * (1 - f);
        o[SF_LAMBDA] = pv.lambda * f + nw.lambda * (1 - f);
        o[SF_T] = pv.t * f + nw.t * (1 - f);
        o[SF_A] = pv.a * f + nw.a * (1 - f);
        o[SF_REDSHIFT] = pv.redshift * f + nw.redshift * (1 - f);
        o[SF_S] = pv.s * f + nw.s * (1 - f);
        o[SF_ISW] = pv.isw * f + nw.isw * (1 - f);
        o[SF_RHO] = pv.rho * f + nw.rho * (1 - f);
        o[SF_PHI] = pv.phi * f + nw.phi * (1 - f);
        o[SF_RHOMAX] = s.rho_max();
        o[SF_KT] = pv.kt * f + nw.kt * (1 - f);
        o[SF_KX] = pv.kx * f + nw.kx * (1 - f);
        o[SF_KY] = pv.ky * f + nw.ky * (1 - f);
        o[SF_KZ] = pv.kz * f + nw.kz * (1 - f);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            o[SF_S1 + c] = s.s1(c);
            o[SF_S1 + 3 + c] = s.s2(c);
        }
    }
    if (jac) {
        const double dr = nw.chi - pv.chi;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double td = tidal[c] * dr;
            s.add(c, nw.chi * td, nw.chi * nw.chi * td);
        }
    }
    s.push_rho(nw.rho);
}

// Angular tidal components (lensing.h:421-448) from the differences of the
// 6 samples of grad phi at +-dl/2 along each axis: D = (d_x gx, d_y gx,
// d_z gx, d_y gy, d_z gy, d_z gz), each the + sample minus the - sample.
// The angles' sines and cosines come from the position itself.
__device__ __forceinline__ void tidal_from_differences(double x, double y, double z, double r, double dl,
                                                       const double* D, double* T) {
    const double rho = sqrt(x * x + y * y);
    const double cp = (rho > 0.0) ? x / rho : 1.0, sp = (rho > 0.0) ? y / rho : 0.0;
    const double ct = z / r, st = rho / r;
    const double s2p = 2.0 * sp * cp, s2t = 2.0 * st * ct;
    const double dxxp = D[0] / dl;
    const double dxyp = D[1] / dl;
    const double dxzp = D[2] / dl;
    const double dyyp = D[3] / dl;
    const double dyzp = D[4] / dl;
    const double dzzp = D[5] / dl;
    T[0] = sp * sp * dxxp + cp * cp * dyyp - s2p * dxyp;
    T[2] = cp * cp * ct * ct * dxxp + sp * sp * ct * ct * dyyp + st * st * dzzp + s2p * ct * ct * dxyp -
           sp * s2t * dyzp - cp * s2t * dxzp;
    T[1] = ct * cp * sp * (dyyp - dxxp) + (cp * cp - sp * sp) * ct * dxyp + sp * st * dxzp - cp * st * dyzp;
}

// The same from the samples themselves: g[k][d], sample k in (x-, x+, y-, y+,
// z-, z+), component d of grad phi.
__device__ __forceinline__ void tidal_from_samples(double x, double y, double z, double r, double dl,
                                                   const double (*g)[3], double* T) {
    const double D[6] = {g[1][0] - g[0][0], g[3][0] - g[2][0], g[5][0] - g[4][0],
                         g[3][1] - g[2][1], g[5][1] - g[4][1], g[5][2] - g[4][2]};
    tidal_from_differences(x, y, z, r, dl, D, T);
}

// dl = 2^-(level - 1 - log2(EXTENT)) (lensing.h:429)
__device__ __forceinline__ double jacobian_step(unsigned level) { return ldexp(1.0, 3 - (int)level); }

}  // namespace mpf
