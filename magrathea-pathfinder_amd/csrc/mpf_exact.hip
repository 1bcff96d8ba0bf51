// Exact-tier kernels: locate, NGP/CIC/TSC interpolation and the RK4 null
// geodesic integrator, reproducing the reference's arithmetic operation by
// operation (this file is compiled with -ffp-contract=off).
//
//   P = 0 (MPF_PREC_REF):  FP32 fields, FP32 accumulation -> bit-identical
//                          to the reference (Gravity<float> accumulator).
//   P = 1 (MPF_PREC_FP64): FP64 fields and accumulation, same operations.
//
// One photon per lane; lanes refill from a device work queue (one atomic per
// wave) when their ray ends, so uneven path lengths do not idle a wave.
#include "mpf_device.h"
#include "mpf_exact_dev.h"
#include "mpf_kernels.h"

namespace mpf {

namespace {

using namespace exact;

template <int P, int NF = 7>
__device__ __forceinline__ void interp(int order, const DevTree& t, TscCache& cache, double x, double y, double z,
                                       double* d, int* loc) {
    if (order == 0) ngp<P, NF>(t, x, y, z, d, loc);
    else if (order == 1) cic<P, NF>(t, x, y, z, d, loc);
    else tsc<P, true, NF>(t, cache, x, y, z, d, loc);
}

// Utility::interpolate (utility.h:442-449) with a per-lane search hint
__device__ __forceinline__ double cosmo_interp(double x0, const double* __restrict__ x, const double* __restrict__ y,
                                               int n, int& hint) {
    if (n <= 0) return 0.0;
    if (n == 1) return y[0];
    int i = hint;
    while (i < n && !(x0 < x[i])) ++i;
    while (i > 0 && x0 < x[i - 1]) --i;
    hint = i;
    i += (i <= 0) - (i >= n);
    return y[i - 1] + ((y[i] - y[i - 1]) * (x0 - x[i - 1]) / (x[i] - x[i - 1]));
}

enum { CA = 0, CT, CX, CY, CZ, CKT, CKX, CKY, CKZ };

// Integrator::dphotondl (integrator.h:409-445) given the interpolated fields
__device__ __forceinline__ void rhs(const double* in, const double* data, double dadt, double scale, double* out) {
    const double two = 2.0;
    const double dphidl = data[F_DT] * in[CKT] + (data[F_DX] * in[CKX] + data[F_DY] * in[CKY] + data[F_DZ] * in[CKZ]);
    out[CA] = in[CKT] * dadt;
    out[CT] = in[CKT];
    out[CX] = in[CKX] / scale;
    out[CY] = in[CKY] / scale;
    out[CZ] = in[CKZ] / scale;
    out[CKT] = -(two * dadt / in[CA] * in[CKT] * in[CKT]) - (two / C2 * in[CKT]) * (dphidl - data[F_DT] * in[CKT]);
    out[CKX] = -(two * dadt / in[CA] * in[CKT] * in[CKX]) + (two / C2 * dphidl * in[CKX]) - (two * data[F_DX] * in[CKT] * in[CKT]);
    out[CKY] = -(two * dadt / in[CA] * in[CKT] * in[CKY]) + (two / C2 * dphidl * in[CKY]) - (two * data[F_DY] * in[CKT] * in[CKT]);
    out[CKZ] = -(two * dadt / in[CA] * in[CKT] * in[CKZ]) + (two / C2 * dphidl * in[CKZ]) - (two * data[F_DZ] * in[CKT] * in[CKT]);
}

// one trajectory element (photon.h:44-47 core + the extras the output keeps)
struct Rec {
    double v[MPF_NOUT];
};

__device__ __forceinline__ double level_extent(int lvl) {
    // extent<Type, Position, Extent>() of the located cell: E / (1 * 2^l)
    const unsigned l = lvl < 0 ? 0u : (unsigned)lvl;
    return EXTENT / (1.0 * (double)(1ull << l));
}

// NF = 10: the tree carries a velocity field and the redshift takes the
// Doppler and ISW terms of the velocity-field build
template <int NF>
struct RayState {
    SurfAcc surf;
    double stopv;     // this ray's stop value (and its surface when per-ray)
    Rec back, prev;
    double data[NF];  // fields at back's position
    double dl, a0, phi0c2;
    double v0n;       // observer velocity along the launch direction / (c dt/dl)
    int hint;
    uint32_t appended;
    bool has_prev;
    uint32_t pl_crossed;  // bit j: plane j passed (IntegrateArgs::nplane)
    bool pl_nonmono;      // a projection on some plane normal decreased
};

// projection of an element's position on a plane normal, in the reference's
// order (lensing.h:281-293)
__device__ __forceinline__ double along_normal(const double* r, const double* k) {
    return r[CX] * k[0] + r[CY] * k[1] + r[CZ] * k[2];
}

// After element appended (>= 1) of a bundle ray: the first element whose
// projection exceeds a plane's value is where std::upper_bound over the whole
// trajectory lands when the projections do not decrease (checked here, and
// the crossing is left to the host when they do); the crossing interpolates
// between it and the element before (lensing.h:281-304).
template <int NF>
__device__ void plane_append(const IntegrateArgs& A, RayState<NF>& s, uint64_t ray) {
    const uint64_t g = (ray / A.plane_div) * (uint64_t)A.nplane;
    const double* pv = s.prev.v;
    const double* nw = s.back.v;
    for (int j = 0; j < A.nplane; ++j) {
        const double* k = A.plane_kit + 3 * (g + j);
        const double v = A.plane_val[g + j];
        const double an = along_normal(nw, k), ap = along_normal(pv, k);
        if (an < ap) s.pl_nonmono = true;
        if ((s.pl_crossed >> j) & 1u) continue;
        if (!(v < an)) continue;
        s.pl_crossed |= 1u << j;
        double* o = A.plane_out + ((uint64_t)ray * A.nplane + j) * PL_NREC;
        const double f = (an - v) / (an - ap);
        o[PL_FIRSTID] = (double)(s.appended - 1);
        o[PL_F] = f;
        o[PL_X] = pv[CX] * f + nw[CX] * (1 - f);
        o[PL_Y] = pv[CY] * f + nw[CY] * (1 - f);
        o[PL_Z] = pv[CZ] * f + nw[CZ] * (1 - f);
    }
}

// The state of each plane once the trajectory is complete (after the pop of
// integrator.h:816-821): the reference's failure test on the last element
// (lensing.h:281-290), the crossing found while integrating, or the plane
// met exactly by the last element (upper_bound finds nothing: the last
// segment, f = 0, as the host's whole-trajectory path takes it).
template <int NF>
__device__ void plane_finish(const IntegrateArgs& A, const RayState<NF>& s, uint64_t ray) {
    const uint64_t g = (ray / A.plane_div) * (uint64_t)A.nplane;
    const bool pop = signbit(s.back.v[MPF_OUT_REDSHIFT]) || signbit(s.back.v[CA]);
    const int64_t len = (int64_t)s.appended + 1 - (pop ? 1 : 0);
    const double* last = pop ? s.prev.v : s.back.v;
    for (int j = 0; j < A.nplane; ++j) {
        const double* k = A.plane_kit + 3 * (g + j);
        const double v = A.plane_val[g + j];
        double* o = A.plane_out + ((uint64_t)ray * A.nplane + j) * PL_NREC;
        double st;
        if (len < 2) {
            st = 0.0;
        } else {
            const double al = along_normal(last, k);
            if (al < v) {
                st = 0.0;
            } else if (((s.pl_crossed >> j) & 1u) && o[PL_FIRSTID] + 1.0 <= (double)(len - 1)) {
                st = 1.0;
            } else if (!pop) {
                const double* pv = s.prev.v;
                const double* nw = s.back.v;
                const double f = (al - v) / (al - along_normal(pv, k));
                o[PL_FIRSTID] = (double)(len - 2);
                o[PL_F] = f;
                o[PL_X] = pv[CX] * f + nw[CX] * (1 - f);
                o[PL_Y] = pv[CY] * f + nw[CY] * (1 - f);
                o[PL_Z] = pv[CZ] * f + nw[CZ] * (1 - f);
                st = 1.0;
            } else {
                st = 2.0;
            }
            if (st == 1.0 && s.pl_nonmono) st = 2.0;
        }
        o[PL_STATE] = st;
    }
}

// element k of a ray's trajectory, the element index in the NSTEPS slot
__device__ __forceinline__ void traj_store(const IntegrateArgs& A, uint64_t ray, uint32_t k, const double* rec) {
    if (k >= A.traj_cap) {
        if (k == A.traj_cap) atomicAdd(A.traj_over, 1ull);
        return;
    }
    double* o = A.traj + ((uint64_t)ray * A.traj_cap + k) * MPF_NOUT;
#pragma unroll
    for (int i = 0; i < MPF_NOUT; ++i) o[i] = rec[i];
    o[MPF_OUT_NSTEPS] = (double)k;
}

template <int P, int NF>
__device__ void ray_launch(const IntegrateArgs& A, const DevTree& t, TscCache& cache, const double* in9,
                           RayState<NF>& s) {
    const double zero = 0.0, one = 1.0, two = 2.0;
    const double c = C_LIGHT, c2 = C2;
    const double scale = A.length;
    double* b = s.back.v;
#pragma unroll
    for (int k = 0; k < MPF_NOUT; ++k) b[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) b[k] = in9[k];
    cache_init(cache);
    s.hint = 0;
    int loc = -1;
    if (A.order >= 0) interp<P, NF>(A.order, t, cache, b[CX], b[CY], b[CZ], s.data, &loc);
    else { zerof<NF>(s.data); s.data[F_A] = one; }
    double* d = s.data;
    const double norm = sqrt((c2 * (one + two * (d[F_PHI] / c2)) * b[CKT] * b[CKT]) /
                             ((one - two * (d[F_PHI] / c2)) * (b[CKX] * b[CKX] + b[CKY] * b[CKY] + b[CKZ] * b[CKZ])));
    b[CKX] *= norm;
    b[CKY] *= norm;
    b[CKZ] *= norm;
    b[CA] = cosmo_interp(b[CT], A.ct, A.ca, A.ncos, s.hint);
    unsigned fl = 0;
    const uint32_t gl = locate_point(t, b[CX], b[CY], b[CZ], &fl);
    const int lvl0 = (gl == NIL) ? -1 : (int)fl;
    b[MPF_OUT_LEVEL] = (double)(lvl0 < 0 ? 0 : lvl0);
    b[MPF_OUT_AH] = zero;
    b[MPF_OUT_RHO] = d[F_RHO];
    b[MPF_OUT_PHI] = d[F_PHI];
    b[MPF_OUT_DPHIDX] = d[F_DX];
    b[MPF_OUT_DPHIDY] = d[F_DY];
    b[MPF_OUT_DPHIDZ] = d[F_DZ];
    b[MPF_OUT_DPHIDT] = d[F_DT];
    b[MPF_OUT_DPHIDL] = zero;
    b[MPF_OUT_LAPLACIAN] = zero;
    b[MPF_OUT_REDSHIFT] = zero;
    const double phi = b[MPF_OUT_PHI];
    b[MPF_OUT_DSDL2] = (b[CA] * b[CA]) * (-(c2 * (one + two * (phi / c2)) * b[CKT] * b[CKT]) +
                                          ((one - two * (phi / c2)) * (b[CKX] * b[CKX] + b[CKY] * b[CKY] + b[CKZ] * b[CKZ])));
    b[MPF_OUT_ERROR] = one - ((one - two * (phi / c2)) * (b[CKX] * b[CKX] + b[CKY] * b[CKY] + b[CKZ] * b[CKZ])) /
                                 (c2 * (one + two * (phi / c2)) * b[CKT] * b[CKT]);
    const double ratio = b[CA] * b[CA] * (scale / c) / A.nsteps;
    s.dl = level_extent(lvl0) * ratio;
    s.a0 = b[CA];
    s.phi0c2 = phi / c2;
    // integrator.h:537-539 (velocity-field build)
    s.v0n = (NF == 10) ? (A.vobs[0] * b[CKX] + A.vobs[1] * b[CKY] + A.vobs[2] * b[CKZ]) / (c * b[CKT]) : 0.0;
    s.appended = 0;
    s.has_prev = false;
    if (A.nsurf) surf_init(s.surf, d[F_RHO]);
    s.stopv = A.stop_value;
    s.pl_crossed = 0;
    s.pl_nonmono = false;
}

__device__ __forceinline__ SurfElem surf_elem(const double* r) {
    SurfElem e;
    e.x = r[CX]; e.y = r[CY]; e.z = r[CZ]; e.chi = r[MPF_OUT_CHI]; e.lambda = r[MPF_OUT_LAMBDA]; e.t = r[CT];
    e.a = r[CA]; e.redshift = r[MPF_OUT_REDSHIFT]; e.s = r[MPF_OUT_S]; e.isw = r[MPF_OUT_ISW];
    e.rho = r[MPF_OUT_RHO]; e.phi = r[MPF_OUT_PHI]; e.kt = r[CKT]; e.kx = r[CKX]; e.ky = r[CKY]; e.kz = r[CKZ];
    return e;
}

// tidal components at an element from 6 TSC samples of grad phi (lensing.h:421-448)
template <int P>
__device__ void element_tidal(const DevTree& t, const double* r, double* T) {
    const double x = r[CX], y = r[CY], z = r[CZ];
    const double dl = jacobian_step((unsigned)r[MPF_OUT_LEVEL]);
    const double pts[6][3] = {{x - 0.5 * dl, y, z}, {x + 0.5 * dl, y, z}, {x, y - 0.5 * dl, z},
                              {x, y + 0.5 * dl, z}, {x, y, z - 0.5 * dl}, {x, y, z + 0.5 * dl}};
    double g[6][3];
    TscCache cache;
    cache_init(cache);
    for (int k = 0; k < 6; ++k) {
        double d[7];
        int loc;
        tsc<P, false>(t, cache, pts[k][0], pts[k][1], pts[k][2], d, &loc);
        g[k][0] = d[F_DX]; g[k][1] = d[F_DY]; g[k][2] = d[F_DZ];
    }
    if (P == 0) {
        // Gravity<float> members: the differences are FP32 subtractions
        for (int k = 0; k < 6; k += 2)
            for (int c = 0; c < 3; ++c) {
                g[k + 1][c] = (double)((float)g[k + 1][c] - (float)g[k][c]);
                g[k][c] = 0.0;
            }
    }
    tidal_from_samples(x, y, z, r[MPF_OUT_CHI], dl, g, T);
}

// One iteration of the integration loop (integrator.h:683-812).  Returns
// true while the ray continues.
template <int P, int NF>
__device__ bool ray_step(const IntegrateArgs& A, const DevTree& t, TscCache& cache, RayState<NF>& s, uint64_t ray) {
    const double zero = 0.0, one = 1.0, two = 2.0, six = 6.0;
    const double c = C_LIGHT, c2 = C2;
    const double min = 0.0 - (EXTENT / two), max = 0.0 + (EXTENT / two);
    const double scale = A.length;
    const double* bk = s.back.v;
    const double dl = s.dl;
    double kin[9], ksum[9], kout[9], stage[7];
    int loc = -1;
    // k1 at back: the fields there are the ones interpolated after the
    // previous step (the reference recomputes them through its cache, with
    // the same result)
    if (A.order >= 0) {
#pragma unroll
        for (int k = 0; k < 7; ++k) stage[k] = s.data[k];
    } else zero7(stage);
    double dadt = cosmo_interp(bk[CT], A.ct, A.cdadt, A.ncos, s.hint);
    rhs(bk, stage, dadt, scale, kout);
#pragma unroll
    for (int i = 0; i < 9; ++i) { ksum[i] = kout[i]; kin[i] = bk[i] + dl / two * kout[i]; }
    // k2
    if (A.order >= 0) interp<P>(A.order, t, cache, kin[CX], kin[CY], kin[CZ], stage, &loc); else zero7(stage);
    dadt = cosmo_interp(kin[CT], A.ct, A.cdadt, A.ncos, s.hint);
    rhs(kin, stage, dadt, scale, kout);
#pragma unroll
    for (int i = 0; i < 9; ++i) { ksum[i] = ksum[i] + two * kout[i]; kin[i] = bk[i] + dl / two * kout[i]; }
    // k3
    if (A.order >= 0) interp<P>(A.order, t, cache, kin[CX], kin[CY], kin[CZ], stage, &loc); else zero7(stage);
    dadt = cosmo_interp(kin[CT], A.ct, A.cdadt, A.ncos, s.hint);
    rhs(kin, stage, dadt, scale, kout);
#pragma unroll
    for (int i = 0; i < 9; ++i) { ksum[i] = ksum[i] + two * kout[i]; kin[i] = bk[i] + dl * kout[i]; }
    // k4
    if (A.order >= 0) interp<P>(A.order, t, cache, kin[CX], kin[CY], kin[CZ], stage, &loc); else zero7(stage);
    dadt = cosmo_interp(kin[CT], A.ct, A.cdadt, A.ncos, s.hint);
    rhs(kin, stage, dadt, scale, kout);
    Rec ph;
    double* p = ph.v;
#pragma unroll
    for (int i = 0; i < 9; ++i) p[i] = bk[i] + (dl / six) * (ksum[i] + kout[i]);
    // fields at the new position
    double* data = s.data;
    if (A.order >= 0) interp<P, NF>(A.order, t, cache, p[CX], p[CY], p[CZ], data, &loc);
    else {
        zerof<NF>(data);
        data[F_A] = one;
        unsigned fl = 0;
        const uint32_t g = locate_point(t, p[CX], p[CY], p[CZ], &fl);
        loc = (g == NIL) ? -1 : (int)fl;
    }
    if (A.order == 0 || A.order == 1 || A.order == 2) {
        // ngp/cic/tsc located the finest cell at the same point: that is
        // locate()'s answer (the TSC call at the box centre goes through CIC,
        // which locates the same point)
    }
    if (!(!(p[CA] < zero) && ((p[CX] > min) && (p[CX] < max) && (p[CY] > min) && (p[CY] < max) && (p[CZ] > min) &&
                              (p[CZ] < max))))
        zerof<NF>(data);
    p[MPF_OUT_LEVEL] = (double)(loc < 0 ? 0 : loc);
    p[MPF_OUT_AH] = data[F_A];
    p[MPF_OUT_RHO] = data[F_RHO];
    p[MPF_OUT_PHI] = data[F_PHI];
    p[MPF_OUT_DPHIDX] = data[F_DX];
    p[MPF_OUT_DPHIDY] = data[F_DY];
    p[MPF_OUT_DPHIDZ] = data[F_DZ];
    p[MPF_OUT_DPHIDT] = -data[F_DT];
    p[MPF_OUT_DPHIDL] = data[F_DT] * p[CKT] + (p[MPF_OUT_DPHIDX] * p[CKX] + p[MPF_OUT_DPHIDY] * p[CKY] + p[MPF_OUT_DPHIDZ] * p[CKZ]);
    p[MPF_OUT_LAPLACIAN] = (data[F_PHI] - bk[MPF_OUT_PHI]) / dl;
    const double phi = p[MPF_OUT_PHI];
    p[MPF_OUT_DSDL2] = (p[CA] * p[CA]) * ((-c2 * (one + two * (phi / c2)) * p[CKT] * p[CKT]) +
                                          ((one - two * (phi / c2)) * (p[CKX] * p[CKX] + p[CKY] * p[CKY] + p[CKZ] * p[CKZ])));
    p[MPF_OUT_ERROR] = one - ((one - two * (phi / c2)) * (p[CKX] * p[CKX] + p[CKY] * p[CKY] + p[CKZ] * p[CKZ])) /
                                 (c2 * (one + two * (phi / c2)) * p[CKT] * p[CKT]);
    p[MPF_OUT_ISW] = bk[MPF_OUT_ISW] + (bk[MPF_OUT_DPHIDT] + p[MPF_OUT_DPHIDT]) * (p[CT] - bk[CT]) / c2;
    p[MPF_OUT_ISWOLD] = bk[MPF_OUT_ISWOLD] -
                        (2 * (p[MPF_OUT_PHI] - bk[MPF_OUT_PHI]) / c2 -
                         ((bk[MPF_OUT_DPHIDX] + p[MPF_OUT_DPHIDX]) * (p[CX] - bk[CX]) * scale +
                          (bk[MPF_OUT_DPHIDY] + p[MPF_OUT_DPHIDY]) * (p[CY] - bk[CY]) * scale +
                          (bk[MPF_OUT_DPHIDZ] + p[MPF_OUT_DPHIDZ]) * (p[CZ] - bk[CZ]) * scale) /
                             c2);
    p[MPF_OUT_CHI] = sqrt(p[CX] * p[CX] + p[CY] * p[CY] + p[CZ] * p[CZ]);
    p[MPF_OUT_LAMBDA] = bk[MPF_OUT_LAMBDA] + dl;
    p[MPF_OUT_S] = bk[MPF_OUT_S] + sqrt((p[CX] - bk[CX]) * (p[CX] - bk[CX]) + (p[CY] - bk[CY]) * (p[CY] - bk[CY]) +
                                        (p[CZ] - bk[CZ]) * (p[CZ] - bk[CZ]));
    if (NF == 10)  // Doppler and ISW terms (integrator.h:594-595, 765-766)
        p[MPF_OUT_REDSHIFT] =
            s.a0 / p[CA] *
                (one + s.phi0c2 - p[MPF_OUT_PHI] / c2 + p[MPF_OUT_ISW] +
                 (data[F_VX] * p[CKX] + data[F_VY] * p[CKY] + data[F_VZ] * p[CKZ]) / (c * p[CKT]) - s.v0n) -
            one;
    else p[MPF_OUT_REDSHIFT] = s.a0 / p[CA] * (one + s.phi0c2 - p[MPF_OUT_PHI] / c2) - one;
    const bool cont = !emptyf<NF>(data);
    if (cont) {
        const double ratio = p[CA] * p[CA] * (scale / c) / A.nsteps;
        s.dl = level_extent(loc) * ratio;
        p[MPF_OUT_NSTEPS] = (double)(s.appended + 1);
        s.prev = s.back;
        s.back = ph;
        s.has_prev = true;
        s.appended++;
        if (A.traj) traj_store(A, ray, s.appended, s.back.v);
        if (A.nsurf) {
            double T[3] = {0.0, 0.0, 0.0};
            if (A.jac) element_tidal<P>(t, s.back.v, T);
            const double* pr = s.prev.v;
            const double* bk2 = s.back.v;
            const double vp = surf_param(A.surf_kind, pr[MPF_OUT_CHI], pr[MPF_OUT_LAMBDA], pr[MPF_OUT_REDSHIFT], pr[CT],
                                         pr[CA], pr[MPF_OUT_S]);
            const double vn = surf_param(A.surf_kind, bk2[MPF_OUT_CHI], bk2[MPF_OUT_LAMBDA], bk2[MPF_OUT_REDSHIFT],
                                         bk2[CT], bk2[CA], bk2[MPF_OUT_S]);
            surf_append(s.surf, A.surf_kind, A.nsurf, A.surf_val, A.ray_value != nullptr, s.stopv,
                        A.surf_out + (uint64_t)A.nsurf * SF_NREC * ray, s.appended, surf_elem(s.prev.v),
                        surf_elem(s.back.v), vp, vn, T, A.jac != 0);
        }
        if (A.nplane) plane_append(A, s, ray);
    }
    if (A.stop_kind >= 0) {
        const double* b = s.back.v;
        bool brk;
        switch (A.stop_kind) {
            case 0: brk = b[MPF_OUT_CHI] > s.stopv; break;
            case 1: brk = b[MPF_OUT_LAMBDA] > s.stopv; break;
            case 2: brk = b[MPF_OUT_REDSHIFT] > s.stopv; break;
            case 3: brk = b[CT] > s.stopv; break;
            case 4: brk = b[CA] < s.stopv; break;
            case 5:
                if (A.nplane) {  // the group's last plane (lensing.h:201)
                    const double* k = A.plane_kit + 3 * ((ray / A.plane_div) * (uint64_t)A.nplane + A.nplane - 1);
                    brk = (p[CX] * k[0] + p[CY] * k[1] + p[CZ] * k[2]) > s.stopv;
                } else {
                    brk = (p[CX] * A.kit[0] + p[CY] * A.kit[1] + p[CZ] * A.kit[2]) > s.stopv;
                }
                break;
            case 6: brk = b[MPF_OUT_S] > s.stopv; break;
            default: brk = true; break;
        }
        if (brk) return false;
    }
    return cont;
}

template <int NF>
__device__ __forceinline__ void surf_finish(const IntegrateArgs& A, const RayState<NF>& s, uint64_t ray,
                                            uint32_t size_minus_one) {
    // surfaces never passed point past the end (firstid = size - 1), and a
    // surface whose next element was popped ends on the last element: both
    // are rejected (hmaps.h:716)
    double* rec = A.surf_out + (uint64_t)A.nsurf * SF_NREC * ray;
    for (int j = 0; j < A.nsurf; ++j) {
        double* o = rec + (uint64_t)SF_NREC * j;
        if (!(s.surf.crossed & (1u << j))) {
            for (int k = 0; k < SF_NREC; ++k) o[k] = 0.0;
            o[SF_FIRSTID] = (double)size_minus_one;
        } else if (o[SF_FIRSTID] >= (double)size_minus_one) {
            o[SF_VALID] = 0.0;
            o[SF_CROSSED] = 0.0;
        }
    }
}

template <int NF>
__device__ __forceinline__ void ray_finish(const IntegrateArgs& A, RayState<NF>& s, uint64_t ray) {
    // erase the last element if its redshift or a is negative (integrator.h:816-821)
    const Rec* r = &s.back;
    uint32_t steps = s.appended;
    if (A.nplane) plane_finish(A, s, ray);
    if (A.nsurf) {
        const bool pop = signbit(s.back.v[MPF_OUT_REDSHIFT]) || signbit(s.back.v[CA]);
        surf_finish(A, s, ray, (pop && s.appended) ? s.appended - 1 : s.appended);
    }
    if (signbit(s.back.v[MPF_OUT_REDSHIFT]) || signbit(s.back.v[CA])) {
        if (s.has_prev) {
            r = &s.prev;
            steps -= 1;
        } else {
            // the trajectory becomes empty
            double* o = A.out + (uint64_t)MPF_NOUT * ray;
            for (int k = 0; k < MPF_NOUT; ++k) o[k] = 0.0;
            o[MPF_OUT_NSTEPS] = -1.0;
            return;
        }
    }
    double* o = A.out + (uint64_t)MPF_NOUT * ray;
#pragma unroll
    for (int k = 0; k < MPF_NOUT; ++k) o[k] = r->v[k];
    o[MPF_OUT_NSTEPS] = (double)steps;
    atomicAdd(A.steps, (unsigned long long)steps);
}

template <int P, int NF>
__global__ void __launch_bounds__(256) integrate_exact_kernel(IntegrateArgs A, DevTree t) {
    RayState<NF> s;
    TscCache cache;
    cache_init(cache);
    bool active = false, drained = false;
    uint64_t ray = 0;
    const unsigned lane = __lane_id();
    while (true) {
        // refill idle lanes from the queue, one atomic per wave
        const bool need = !active && !drained;
        const unsigned long long m = __ballot(need);
        if (m && ((unsigned)__popcll(m) >= A.refill || !__any(active))) {
            const int leader = __ffsll((long long)m) - 1;
            unsigned long long base = 0;
            if ((int)lane == leader) base = atomicAdd(A.queue, (unsigned long long)__popcll(m));
            base = __shfl(base, leader);
            if (need) {
                const uint64_t my = base + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
                if (my < A.n) {
                    ray = A.queue_order ? (uint64_t)A.queue_order[my] : my;
                    ray_launch<P, NF>(A, t, cache, A.in9 + 9 * ray, s);
                    if (A.traj) traj_store(A, ray, 0, s.back.v);
                    if (A.ray_value) s.stopv = A.ray_value[ray];
                    if (A.nplane) s.stopv = A.plane_val[(ray / A.plane_div) * (uint64_t)A.nplane + A.nplane - 1];
                    active = !emptyf<NF>(s.data);
                    if (!active) ray_finish(A, s, ray);
                } else {
                    drained = true;
                }
            }
        }
        if (!__any(active)) {
            if (__all(drained)) break;
            continue;
        }
        if (active) {
            if (!ray_step<P, NF>(A, t, cache, s, ray)) {
                active = false;
                ray_finish(A, s, ray);
            }
        }
    }
}

// ---- the lean integrator: plain integration (no surfaces, no stored
// trajectory), the path of mpf_integrate[_dev].  Only the recurrence of the
// trajectory stays in registers; the record of the last element is derived
// from it when the ray ends.  The final pop (integrator.h:816-821) needs the
// element before the last one only when the last one has a negative
// redshift or scale factor: that element is written to the ray's output slot
// at the step that appends such an element, and the last one overwrites it
// at the end when it is compatible.  The four interpolations of a step run
// in one loop (one copy of the TSC code), and the 27 cell indices of the TSC
// cache live in LDS (two sets per lane: a refill that fails keeps the old
// cache, as the reference's does).  Every expression is ray_step's.

constexpr int LEAN_BLOCK = 256;

// TscCache with its 27 cell indices in LDS: set `cur` of this lane is
// g[(cur * 27 + i) * LEAN_BLOCK]
struct LeanCache {
    unsigned lvl;
    uint32_t cx[3], cy[3], cz[3];
    bool dflt;
    bool leaf;      // the central cell (13) stores no children
    unsigned cur;
    uint32_t* g;
};

__device__ __forceinline__ void lean_cache_init(LeanCache& c) {
    c.dflt = true;
    c.leaf = false;
    c.lvl = 0;
    c.cur = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) { c.cx[i] = 0; c.cy[i] = 0; c.cz[i] = 0; }
}

template <typename T>
__device__ __forceinline__ T pick3(const T* a, int i) { return (i == 0) ? a[0] : ((i == 1) ? a[1] : a[2]); }

// the fields a stage of the RK4 step reads (Integrator::dphotondl,
// integrator.h:409-445): grad phi and dphi/dt; each field is summed on its
// own (mac(), simplehyperoctree.h:1870-1910), so leaving the others out
// changes none of these
__device__ __forceinline__ bool stage_field(int k) { return k == F_DX || k == F_DY || k == F_DZ || k == F_DT; }

// The lean TSC reads the cell-major copies (integrate_impl builds them
// before every lean TSC launch): FP64 ones when the tier is FP64 and the tree
// has FP64 fields, else FP32 ones.
template <int P, int NF>
__device__ __forceinline__ void loadf_sel(const DevTree& t, uint32_t g, double* d, bool full) {
    if (P == 1 && t.q64) {
        const double2* q = t.q64 + 4 * (uint64_t)g;
        const double2 a = q[0], b = q[1];
        d[F_DX] = a.x; d[F_DY] = a.y; d[F_DZ] = b.x; d[F_DT] = b.y;
        if (full) {
            const double2 e = q[2], f = q[3];
            d[F_RHO] = e.x; d[F_PHI] = e.y; d[F_A] = f.x;
        }
    } else {
        const float4* q = t.q32 + 2 * (uint64_t)g;
        const float4 a = q[0];
        d[F_DX] = (double)a.x; d[F_DY] = (double)a.y; d[F_DZ] = (double)a.z; d[F_DT] = (double)a.w;
        if (full) {
            const float4 b = q[1];
            d[F_RHO] = (double)b.x; d[F_PHI] = (double)b.y; d[F_A] = (double)b.z;
        }
    }
    if (NF == 10 && full) {
#pragma unroll
        for (int k = 0; k < 3; ++k) d[7 + k] = t.vel ? (double)t.vel[(uint64_t)k * t.n + g] : 0.0;
    }
}

template <int P, int NF>
__device__ __forceinline__ void macf_sel(double* acc, const double* f, double w, bool full) {
#pragma unroll
    for (int k = 0; k < NF; ++k) {
        if (full || stage_field(k)) {
            if (P == 0) acc[k] = (double)(float)(acc[k] + f[k] * w);
            else acc[k] = acc[k] + f[k] * w;
        }
    }
}

// exact::tsc<P, true, NF> with the cache above (simplehyperoctree.h:1554-1644)
// full = false: only the fields stage_field names are summed (the others
// keep what they held)
template <int P, int NF>
__device__ void tsc_lean(const DevTree& t, LeanCache& cache, double x, double y, double z, double* result, int* loc,
                         bool full) {
    const double two = 2.0;
    const double point[3] = {x, y, z};
    if (point[0] == 0.0 && point[1] == 0.0 && point[2] == 0.0) {
        cic<P, NF>(t, 0., 0., 0., result, loc);
        return;
    }
    uint32_t ek[3] = {axis_coord(x, 32), axis_coord(y, 32), axis_coord(z, 32)};
    unsigned fl = 0;
    uint32_t g0;
    // a point inside the cached central cell, a leaf: that cell is the finest
    // stored one containing it (what the descent from the base returns)
    if (!cache.dflt && cache.leaf && (ek[0] >> (32 - cache.lvl)) == cache.cx[1] &&
        (ek[1] >> (32 - cache.lvl)) == cache.cy[1] && (ek[2] >> (32 - cache.lvl)) == cache.cz[1]) {
        g0 = cache.g[((uint64_t)cache.cur * 27 + 13) * LEAN_BLOCK];
        fl = cache.lvl;
    } else {
        g0 = find_cell(t, 32, ek[0], ek[1], ek[2], &fl);
    }
    *loc = (g0 == NIL) ? -1 : (int)fl;
    bool invalid = (g0 == NIL);
    unsigned idl = invalid ? 32u : fl;
    uint32_t idk[3] = {ek[0] >> (32 - idl), ek[1] >> (32 - idl), ek[2] >> (32 - idl)};
    unsigned ilvl = idl, lvl = idl;
    double half = EXTENT / 1.0 * (1.0 / (1.0 * (double)(1ull << idl))) / two;
    double cell[NF];
    zerof<NF>(result);
    do {
        ilvl = lvl;
        const double twohalves = two * half;
        bool hit;
        if (cache.dflt) hit = !invalid && idl == 0;  // 27 default elements hold the root key
        else hit = !invalid && cache.lvl == idl && cache.cx[1] == idk[0] && cache.cy[1] == idk[1] &&
                   cache.cz[1] == idk[2];
        if (hit) {
            if (cache.dflt) {  // default elements carry Data(): the sum stays zero
                zerof<NF>(result);
                return;
            }
            double wx[3], wy[3], wz[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const int a = (i == 1) ? 0 : 1;
                wx[i] = tsc_axis_w(a, fabs((cell_center(idl, cache.cx[i]) - point[0]) / (twohalves)));
                wy[i] = tsc_axis_w(a, fabs((cell_center(idl, cache.cy[i]) - point[1]) / (twohalves)));
                wz[i] = tsc_axis_w(a, fabs((cell_center(idl, cache.cz[i]) - point[2]) / (twohalves)));
            }
            const uint32_t* g = cache.g + (uint64_t)cache.cur * 27 * LEAN_BLOCK;
            // cells in the reference's order (x, then y, then z), the weight
            // (wx wy) wz as it associates; one (x, y) pair per iteration, its 3
            // z cells unrolled (the whole 27 unrolled hoists loads and spills)
#pragma unroll 1
            for (int ixy = 0; ixy < 9; ++ixy) {
                const double wxy = pick3(wx, ixy / 3) * pick3(wy, ixy % 3);
#pragma unroll
                for (int iz = 0; iz < 3; ++iz) {
                    const int ic = 3 * ixy + iz;
                    loadf_sel<P, NF>(t, g[ic * LEAN_BLOCK], cell, full);
                    macf_sel<P, NF>(result, cell, wxy * wz[iz], full);
                }
            }
            return;
        }
        if (!invalid) {
            uint32_t nx[3], ny[3], nz[3];
            double wx[3], wy[3], wz[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const int a = (i == 1) ? 0 : 1;
                nx[i] = axis_coord(point[0] + (i - 1) * twohalves, ilvl);
                ny[i] = axis_coord(point[1] + (i - 1) * twohalves, ilvl);
                nz[i] = axis_coord(point[2] + (i - 1) * twohalves, ilvl);
                wx[i] = tsc_axis_w(a, fabs((cell_center(ilvl, nx[i]) - point[0]) / (twohalves)));
                wy[i] = tsc_axis_w(a, fabs((cell_center(ilvl, ny[i]) - point[1]) / (twohalves)));
                wz[i] = tsc_axis_w(a, fabs((cell_center(ilvl, nz[i]) - point[2]) / (twohalves)));
            }
            // the other set receives the new cells; it becomes the cache only when all 27 are found
            uint32_t* gn = cache.g + (uint64_t)(cache.cur ^ 1u) * 27 * LEAN_BLOCK;
            int ic = 0;
#pragma unroll 1  // one cell at a time: unrolled, the loads hoisted across cells spill registers
            for (; ic < 27; ++ic) {
                const int ix = ic / 9, iy = (ic / 3) % 3, iz = ic % 3;
                unsigned vl = 0;
                const uint32_t gv = find_cell(t, ilvl, pick3(nx, ix), pick3(ny, iy), pick3(nz, iz), &vl);
                if (gv == NIL) {
                    zerof<NF>(result);
                    break;
                }
                lvl = lvl < vl ? lvl : vl;
                if (lvl < ilvl) {
                    zerof<NF>(result);
                    break;
                }
                const double w = pick3(wx, ix) * pick3(wy, iy) * pick3(wz, iz);
                loadf_sel<P, NF>(t, gv, cell, full);
                macf_sel<P, NF>(result, cell, w, full);
                gn[ic * LEAN_BLOCK] = gv;
            }
            if (ic == 27) {
                cache.dflt = false;
                cache.lvl = ilvl;
                cache.cur ^= 1u;
                cache.leaf = t.child_mask[gn[13 * LEAN_BLOCK]] == 0;
#pragma unroll
                for (int i = 0; i < 3; ++i) { cache.cx[i] = nx[i]; cache.cy[i] = ny[i]; cache.cz[i] = nz[i]; }
            }
        }
        if (lvl < ilvl) {
            const unsigned el = lvl + 1;
            idl = lvl;
            idk[0] = axis_coord(x, el) >> 1;
            idk[1] = axis_coord(y, el) >> 1;
            idk[2] = axis_coord(z, el) >> 1;
            invalid = false;
            zerof<NF>(result);
            half = EXTENT / 1.0 * (1.0 / (1.0 * (double)(1ull << idl))) / two;
        }
    } while (lvl < ilvl);
}

template <int P, int NF>
__device__ __forceinline__ void interp_lean(int order, const DevTree& t, LeanCache& cache, double x, double y,
                                            double z, double* d, int* loc, bool full = true) {
    if (order == 0) ngp<P, NF>(t, x, y, z, d, loc);
    else if (order == 1) cic<P, NF>(t, x, y, z, d, loc);
    else tsc_lean<P, NF>(t, cache, x, y, z, d, loc, full);
}

template <int NF>
struct LeanState {
    double c[9];   // a, t, x, y, z, kt, kx, ky, kz of the last element
    double d[NF];  // the fields interpolated there
    double lambda, s, isw, iswold, laplacian, redshift;
    double dl, a0, phi0c2, v0n, stopv;
    int level, hint;
    uint32_t appended;
};

// the MPF_NOUT record of the last element (ray_launch and ray_step write the
// same values)
template <int NF>
__device__ void lean_record(const LeanState<NF>& st, double* o, uint32_t nsteps) {
    const double zero = 0.0, one = 1.0, two = 2.0;
    const double c2 = C2;
    const double* c = st.c;
    const double* d = st.d;
#pragma unroll
    for (int k = 0; k < 9; ++k) o[k] = c[k];
    const double phi = d[F_PHI];
    o[MPF_OUT_LEVEL] = (double)st.level;
    o[MPF_OUT_RHO] = d[F_RHO];
    o[MPF_OUT_PHI] = phi;
    o[MPF_OUT_DPHIDX] = d[F_DX];
    o[MPF_OUT_DPHIDY] = d[F_DY];
    o[MPF_OUT_DPHIDZ] = d[F_DZ];
    o[MPF_OUT_ISW] = st.isw;
    o[MPF_OUT_ISWOLD] = st.iswold;
    o[MPF_OUT_LAMBDA] = st.lambda;
    o[MPF_OUT_S] = st.s;
    o[MPF_OUT_REDSHIFT] = st.redshift;
    o[MPF_OUT_LAPLACIAN] = st.laplacian;
    o[MPF_OUT_ERROR] = one - ((one - two * (phi / c2)) * (c[CKX] * c[CKX] + c[CKY] * c[CKY] + c[CKZ] * c[CKZ])) /
                                 (c2 * (one + two * (phi / c2)) * c[CKT] * c[CKT]);
    if (st.appended == 0) {  // the launch element (ray_launch)
        o[MPF_OUT_AH] = zero;
        o[MPF_OUT_DPHIDT] = d[F_DT];
        o[MPF_OUT_DPHIDL] = zero;
        o[MPF_OUT_CHI] = zero;
        o[MPF_OUT_DSDL2] = (c[CA] * c[CA]) * (-(c2 * (one + two * (phi / c2)) * c[CKT] * c[CKT]) +
                                              ((one - two * (phi / c2)) * (c[CKX] * c[CKX] + c[CKY] * c[CKY] + c[CKZ] * c[CKZ])));
    } else {                 // an appended element (ray_step)
        o[MPF_OUT_AH] = d[F_A];
        o[MPF_OUT_DPHIDT] = -d[F_DT];
        o[MPF_OUT_DPHIDL] = d[F_DT] * c[CKT] + (d[F_DX] * c[CKX] + d[F_DY] * c[CKY] + d[F_DZ] * c[CKZ]);
        o[MPF_OUT_CHI] = sqrt(c[CX] * c[CX] + c[CY] * c[CY] + c[CZ] * c[CZ]);
        o[MPF_OUT_DSDL2] = (c[CA] * c[CA]) * ((-c2 * (one + two * (phi / c2)) * c[CKT] * c[CKT]) +
                                              ((one - two * (phi / c2)) * (c[CKX] * c[CKX] + c[CKY] * c[CKY] + c[CKZ] * c[CKZ])));
    }
    o[MPF_OUT_NSTEPS] = (double)nsteps;
}

template <int P, int NF>
__device__ void lean_launch(const IntegrateArgs& A, const DevTree& t, LeanCache& cache, const double* in9,
                            LeanState<NF>& st) {
    const double zero = 0.0, one = 1.0, two = 2.0;
    const double c = C_LIGHT, c2 = C2;
    double* b = st.c;
#pragma unroll
    for (int k = 0; k < 9; ++k) b[k] = in9[k];
    lean_cache_init(cache);
    st.hint = 0;
    int loc = -1;
    if (A.order >= 0) interp_lean<P, NF>(A.order, t, cache, b[CX], b[CY], b[CZ], st.d, &loc);
    else { zerof<NF>(st.d); st.d[F_A] = one; }
    const double* d = st.d;
    const double norm = sqrt((c2 * (one + two * (d[F_PHI] / c2)) * b[CKT] * b[CKT]) /
                             ((one - two * (d[F_PHI] / c2)) * (b[CKX] * b[CKX] + b[CKY] * b[CKY] + b[CKZ] * b[CKZ])));
    b[CKX] *= norm;
    b[CKY] *= norm;
    b[CKZ] *= norm;
    b[CA] = cosmo_interp(b[CT], A.ct, A.ca, A.ncos, st.hint);
    unsigned fl = 0;
    const uint32_t gl = locate_point(t, b[CX], b[CY], b[CZ], &fl);
    const int lvl0 = (gl == NIL) ? -1 : (int)fl;
    st.level = lvl0 < 0 ? 0 : lvl0;
    st.lambda = zero; st.s = zero; st.isw = zero; st.iswold = zero; st.laplacian = zero; st.redshift = zero;
    const double ratio = b[CA] * b[CA] * (A.length / c) / A.nsteps;
    st.dl = level_extent(lvl0) * ratio;
    st.a0 = b[CA];
    st.phi0c2 = d[F_PHI] / c2;
    st.v0n = (NF == 10) ? (A.vobs[0] * b[CKX] + A.vobs[1] * b[CKY] + A.vobs[2] * b[CKZ]) / (c * b[CKT]) : 0.0;
    st.appended = 0;
    st.stopv = A.stop_value;
}

// one iteration of the integration loop (ray_step without surfaces or a
// stored trajectory); true while the ray continues
template <int P, int NF>
__device__ bool lean_step(const IntegrateArgs& A, const DevTree& t, LeanCache& cache, LeanState<NF>& st,
                          uint64_t ray) {
    const double zero = 0.0, one = 1.0, two = 2.0, six = 6.0;
    const double c = C_LIGHT, c2 = C2;
    const double min = 0.0 - (EXTENT / two), max = 0.0 + (EXTENT / two);
    const double scale = A.length;
    const double* bk = st.c;
    const double dl = st.dl;
    double kin[9], ksum[9], kout[9], data[NF];
    int loc = -1;
    // k1 at the last element: its fields are the ones interpolated when it
    // was appended (the reference recomputes them through its cache, with the
    // same result)
    if (A.order >= 0) {
#pragma unroll
        for (int k = 0; k < NF; ++k) data[k] = st.d[k];
    } else zerof<NF>(data);
    double dadt = cosmo_interp(bk[CT], A.ct, A.cdadt, A.ncos, st.hint);
    rhs(bk, data, dadt, scale, kout);
#pragma unroll
    for (int i = 0; i < 9; ++i) { ksum[i] = kout[i]; kin[i] = bk[i] + dl / two * kout[i]; }
    // stages 1-3 interpolate at the k2, k4 and k4 points and advance the RK4
    // sums; stage 4 interpolates at the new point (kin then holds it)
    for (int stg = 1; stg <= 4; ++stg) {
        if (A.order >= 0) interp_lean<P, NF>(A.order, t, cache, kin[CX], kin[CY], kin[CZ], data, &loc, stg == 4);
        else {
            zerof<NF>(data);
            if (stg == 4) {
                data[F_A] = one;
                unsigned fl = 0;
                const uint32_t g = locate_point(t, kin[CX], kin[CY], kin[CZ], &fl);
                loc = (g == NIL) ? -1 : (int)fl;
            }
        }
        if (stg == 4) break;
        dadt = cosmo_interp(kin[CT], A.ct, A.cdadt, A.ncos, st.hint);
        rhs(kin, data, dadt, scale, kout);
        if (stg == 1) {
#pragma unroll
            for (int i = 0; i < 9; ++i) { ksum[i] = ksum[i] + two * kout[i]; kin[i] = bk[i] + dl / two * kout[i]; }
        } else if (stg == 2) {
#pragma unroll
            for (int i = 0; i < 9; ++i) { ksum[i] = ksum[i] + two * kout[i]; kin[i] = bk[i] + dl * kout[i]; }
        } else {
#pragma unroll
            for (int i = 0; i < 9; ++i) kin[i] = bk[i] + (dl / six) * (ksum[i] + kout[i]);
        }
    }
    const double* p = kin;
    if (!(!(p[CA] < zero) && ((p[CX] > min) && (p[CX] < max) && (p[CY] > min) && (p[CY] < max) && (p[CZ] > min) &&
                              (p[CZ] < max))))
        zerof<NF>(data);
    if (emptyf<NF>(data)) return false;   // not appended: the ray ends on the last element
    // the new element (ray_step's expressions)
    const double* od = st.d;
    const double bk_dphidt = (st.appended == 0) ? od[F_DT] : -od[F_DT];
    const double p_dphidt = -data[F_DT];
    const double laplacian = (data[F_PHI] - od[F_PHI]) / dl;
    const double isw = st.isw + (bk_dphidt + p_dphidt) * (p[CT] - bk[CT]) / c2;
    const double iswold = st.iswold - (2 * (data[F_PHI] - od[F_PHI]) / c2 -
                                       ((od[F_DX] + data[F_DX]) * (p[CX] - bk[CX]) * scale +
                                        (od[F_DY] + data[F_DY]) * (p[CY] - bk[CY]) * scale +
                                        (od[F_DZ] + data[F_DZ]) * (p[CZ] - bk[CZ]) * scale) /
                                           c2);
    const double chi = sqrt(p[CX] * p[CX] + p[CY] * p[CY] + p[CZ] * p[CZ]);
    const double lambda = st.lambda + dl;
    const double sv = st.s + sqrt((p[CX] - bk[CX]) * (p[CX] - bk[CX]) + (p[CY] - bk[CY]) * (p[CY] - bk[CY]) +
                                  (p[CZ] - bk[CZ]) * (p[CZ] - bk[CZ]));
    double redshift;
    if (NF == 10)
        redshift = st.a0 / p[CA] *
                       (one + st.phi0c2 - data[F_PHI] / c2 + isw +
                        (data[F_VX] * p[CKX] + data[F_VY] * p[CKY] + data[F_VZ] * p[CKZ]) / (c * p[CKT]) - st.v0n) -
                   one;
    else redshift = st.a0 / p[CA] * (one + st.phi0c2 - data[F_PHI] / c2) - one;
    if (signbit(redshift) || signbit(p[CA])) {
        // the final pop would erase this element: keep the current last one
        lean_record<NF>(st, A.out + (uint64_t)MPF_NOUT * ray, st.appended);
    }
    const double ratio = p[CA] * p[CA] * (scale / c) / A.nsteps;
    st.dl = level_extent(loc) * ratio;
    const double pkx = p[CX], pky = p[CY], pkz = p[CZ], pt = p[CT], pa = p[CA];
#pragma unroll
    for (int i = 0; i < 9; ++i) st.c[i] = p[i];
#pragma unroll
    for (int k = 0; k < NF; ++k) st.d[k] = data[k];
    st.level = loc < 0 ? 0 : loc;
    st.laplacian = laplacian;
    st.isw = isw;
    st.iswold = iswold;
    st.lambda = lambda;
    st.s = sv;
    st.redshift = redshift;
    st.appended++;
    if (A.stop_kind >= 0) {
        bool brk;
        switch (A.stop_kind) {
            case 0: brk = chi > st.stopv; break;
            case 1: brk = lambda > st.stopv; break;
            case 2: brk = redshift > st.stopv; break;
            case 3: brk = pt > st.stopv; break;
            case 4: brk = pa < st.stopv; break;
            case 5: brk = (pkx * A.kit[0] + pky * A.kit[1] + pkz * A.kit[2]) > st.stopv; break;
            case 6: brk = sv > st.stopv; break;
            default: brk = true; break;
        }
        if (brk) return false;
    }
    return true;
}

template <int NF>
__device__ void lean_finish(const IntegrateArgs& A, const LeanState<NF>& st, uint64_t ray) {
    double* o = A.out + (uint64_t)MPF_NOUT * ray;
    if (signbit(st.redshift) || signbit(st.c[CA])) {
        if (st.appended == 0) {  // the trajectory becomes empty
            for (int k = 0; k < MPF_NOUT; ++k) o[k] = 0.0;
            o[MPF_OUT_NSTEPS] = -1.0;
            return;
        }
        // the element before the last was written when the last was appended
        atomicAdd(A.steps, (unsigned long long)(st.appended - 1));
        return;
    }
    lean_record<NF>(st, o, st.appended);
    atomicAdd(A.steps, (unsigned long long)st.appended);
}

template <int P, int NF>
__global__ void __launch_bounds__(LEAN_BLOCK, 2) integrate_lean_kernel(IntegrateArgs A, DevTree t) {
    __shared__ uint32_t cache_g[2 * 27 * LEAN_BLOCK];
    LeanState<NF> st;
    LeanCache cache;
    cache.g = cache_g + threadIdx.x;
    lean_cache_init(cache);
    bool active = false, drained = false;
    uint64_t ray = 0;
    const unsigned lane = __lane_id();
    while (true) {
        const bool need = !active && !drained;
        const unsigned long long m = __ballot(need);
        if (m && ((unsigned)__popcll(m) >= A.refill || !__any(active))) {
            const int leader = __ffsll((long long)m) - 1;
            unsigned long long base = 0;
            if ((int)lane == leader) base = atomicAdd(A.queue, (unsigned long long)__popcll(m));
            base = __shfl(base, leader);
            if (need) {
                const uint64_t my = base + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
                if (my < A.n) {
                    ray = A.queue_order ? (uint64_t)A.queue_order[my] : my;
                    lean_launch<P, NF>(A, t, cache, A.in9 + 9 * ray, st);
                    if (A.ray_value) st.stopv = A.ray_value[ray];
                    active = !emptyf<NF>(st.d);
                    if (!active) lean_finish<NF>(A, st, ray);
                } else {
                    drained = true;
                }
            }
        }
        if (!__any(active)) {
            if (__all(drained)) break;
            continue;
        }
        if (active) {
            if (!lean_step<P, NF>(A, t, cache, st, ray)) {
                active = false;
                lean_finish<NF>(A, st, ray);
            }
        }
    }
}

template <int P>
__global__ void locate_kernel(DevTree t, const double* xyz, uint64_t n, int64_t* elem, uint32_t* level) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned fl = 0;
    const uint32_t g = locate_point(t, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], &fl);
    elem[i] = (g == NIL) ? -1 : (int64_t)t.elem[g];
    level[i] = (g == NIL) ? 0u : fl;
}

template <int P>
__global__ void interp_kernel(DevTree t, int order, const double* xyz, uint64_t n, double* out7) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    TscCache cache;
    cache_init(cache);
    double d[7];
    int loc;
    const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    if (order == 0) ngp<P>(t, x, y, z, d, &loc);
    else if (order == 1) cic<P>(t, x, y, z, d, &loc);
    else tsc<P, false>(t, cache, x, y, z, d, &loc);
#pragma unroll
    for (int k = 0; k < 7; ++k) out7[7 * i + k] = d[k];
}

// cached TSC over the points in sequence: one lane walks all points
template <int P>
__global__ void interp_seq_kernel(DevTree t, const double* xyz, uint64_t n, double* out7) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    TscCache cache;
    cache_init(cache);
    double d[7];
    int loc;
    for (uint64_t i = 0; i < n; ++i) {
        tsc<P, true>(t, cache, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], d, &loc);
        for (int k = 0; k < 7; ++k) out7[7 * i + k] = d[k];
    }
}

__global__ void field_quads_kernel(DevTree t, float* __restrict__ q32, double* __restrict__ q64) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= t.n) return;
    const float* f = t.f32;
    float* o = q32 + 8 * g;
    o[0] = f[F_DX * t.n + g]; o[1] = f[F_DY * t.n + g]; o[2] = f[F_DZ * t.n + g]; o[3] = f[F_DT * t.n + g];
    o[4] = f[F_RHO * t.n + g]; o[5] = f[F_PHI * t.n + g]; o[6] = f[F_A * t.n + g]; o[7] = 0.0f;
    if (q64) {
        const double* d = t.f64;
        double* p = q64 + 8 * g;
        p[0] = d[F_DX * t.n + g]; p[1] = d[F_DY * t.n + g]; p[2] = d[F_DZ * t.n + g]; p[3] = d[F_DT * t.n + g];
        p[4] = d[F_RHO * t.n + g]; p[5] = d[F_PHI * t.n + g]; p[6] = d[F_A * t.n + g]; p[7] = 0.0;
    }
}

// a_c = S1 - S2 / D at each valid surface, then A = 1 - 2 a L / c^2 (lensing.h:470-477)
__global__ void surface_final_kernel(IntegrateArgs A) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n * (uint64_t)A.nsurf) return;
    double* o = A.surf_out + (uint64_t)SF_NREC * i;
    if (!A.jac || o[SF_CROSSED] == 0.0) {
        o[SF_A11] = 1.0; o[SF_A12] = 0.0; o[SF_A22] = 1.0;
        return;
    }
    // the catalogue's Jacobian is taken at the source distance
    // (catalogues.h:516), the maps' at the interpolated position's (hmaps.h:667),
    // the re-run catalogue's at the interpolated chi (catalogues.h:955)
    const double D = A.jac_at_chi ? o[SF_CHI] : (A.ray_value && A.surf_kind == 0) ? o[SF_REF] : o[SF_DIST];
    double a[3];
    for (int c = 0; c < 3; ++c) a[c] = o[SF_S1 + c] - o[SF_S1 + 3 + c] / D;
    o[SF_A11] = 1 - 2 * a[0] / C2 * A.length;
    o[SF_A12] = -2 * a[1] / C2 * A.length;
    o[SF_A22] = 1 - 2 * a[2] / C2 * A.length;
}

}  // namespace

hipError_t launch_field_quads(const DevTree& t, float* q32, double* q64, hipStream_t st) {
    if (t.n == 0) return hipSuccess;
    hipLaunchKernelGGL(field_quads_kernel, dim3((unsigned)((t.n + 255) / 256)), dim3(256), 0, st, t, q32, q64);
    return hipGetLastError();
}

hipError_t launch_surface_final(const IntegrateArgs& A, hipStream_t st) {
    const uint64_t m = A.n * (uint64_t)A.nsurf;
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(surface_final_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, A);
    return hipGetLastError();
}

hipError_t launch_integrate_exact(int precision, const IntegrateArgs& A, const DevTree& t, int grid, hipStream_t st) {
    if (A.nsurf == 0 && A.traj == nullptr && A.nplane == 0) {  // plain integration: the lean kernel
        if (t.vel) {
            if (precision == 0) hipLaunchKernelGGL((integrate_lean_kernel<0, 10>), dim3(grid), dim3(LEAN_BLOCK), 0, st, A, t);
            else hipLaunchKernelGGL((integrate_lean_kernel<1, 10>), dim3(grid), dim3(LEAN_BLOCK), 0, st, A, t);
        } else if (precision == 0) {
            hipLaunchKernelGGL((integrate_lean_kernel<0, 7>), dim3(grid), dim3(LEAN_BLOCK), 0, st, A, t);
        } else {
            hipLaunchKernelGGL((integrate_lean_kernel<1, 7>), dim3(grid), dim3(LEAN_BLOCK), 0, st, A, t);
        }
        return hipGetLastError();
    }
    // a tree with a velocity field: 10 interpolated fields and the Doppler redshift
    if (t.vel) {
        if (precision == 0) hipLaunchKernelGGL((integrate_exact_kernel<0, 10>), dim3(grid), dim3(256), 0, st, A, t);
        else hipLaunchKernelGGL((integrate_exact_kernel<1, 10>), dim3(grid), dim3(256), 0, st, A, t);
    } else if (precision == 0) hipLaunchKernelGGL((integrate_exact_kernel<0, 7>), dim3(grid), dim3(256), 0, st, A, t);
    else hipLaunchKernelGGL((integrate_exact_kernel<1, 7>), dim3(grid), dim3(256), 0, st, A, t);
    return hipGetLastError();
}

hipError_t launch_locate(const DevTree& t, const double* xyz, uint64_t n, int64_t* elem, uint32_t* level,
                         hipStream_t st) {
    if (n == 0) return hipSuccess;
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(locate_kernel<0>, dim3(grid), dim3(256), 0, st, t, xyz, n, elem, level);
    return hipGetLastError();
}

hipError_t launch_interp(int precision, const DevTree& t, int order, const double* xyz, uint64_t n, double* out7,
                         hipStream_t st) {
    if (n == 0) return hipSuccess;
    const unsigned grid = (unsigned)((n + 255) / 256);
    if (order == 3) {
        if (precision == 0) hipLaunchKernelGGL(interp_seq_kernel<0>, dim3(1), dim3(64), 0, st, t, xyz, n, out7);
        else hipLaunchKernelGGL(interp_seq_kernel<1>, dim3(1), dim3(64), 0, st, t, xyz, n, out7);
    } else if (precision == 0) {
        hipLaunchKernelGGL(interp_kernel<0>, dim3(grid), dim3(256), 0, st, t, order, xyz, n, out7);
    } else {
        hipLaunchKernelGGL(interp_kernel<1>, dim3(grid), dim3(256), 0, st, t, order, xyz, n, out7);
    }
    return hipGetLastError();
}

}  // namespace mpf
