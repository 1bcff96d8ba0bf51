// HEALPix maps from per-ray lensing Jacobians.
//
//  * Pixel-centre mode (Hmaps::FillMap, hmaps.h:599-840): one ray per RING
//    pixel centre; the per-pixel values of each surface are written to
//    map[surface][component][pixel] (no two rays share a pixel).
//  * Binning mode (arbitrary ray directions): each ray's convergence and shear
//    are added to the pixel of its observed direction.  Lanes of a wavefront
//    that hit the same pixel are summed first, and the leader issues one
//    atomic per map.  The sums are 64-bit fixed point (2^-46 resolution, range
//    +-2^17 per pixel), so the integer atomics are associative: maps are
//    bit-identical from run to run whatever the order of the rays.
#include "mpf_healpix.h"
#include "mpf_kernels.h"

namespace mpf {

namespace {

constexpr double FIX_SCALE = 70368744177664.0;  // 2^46

// infinitesimal-Jacobian lensing quantities (hmaps.h:759-762)
__device__ __forceinline__ void lensing_from_jacobian(const double* s, double* q) {
    const double a11 = s[SF_A11], a12 = s[SF_A12], a22 = s[SF_A22];
    q[0] = 1 - (a11 + a22) * 0.5;       // kappa
    q[1] = -0.5 * (a11 - a22);          // gamma1
    q[2] = -a12;                        // gamma2
    q[3] = a11 * a22 - a12 * a12;       // inverse magnification
}

__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

__global__ void __launch_bounds__(256) bin_kernel(int64_t nside, const double* __restrict__ in9,
                                                  const double* __restrict__ surf, int nsurf, int isurf, uint64_t n,
                                                  unsigned long long* __restrict__ acc, uint64_t npix) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool mine = false;
    int64_t pix = -1;
    long long v[4] = {0, 0, 0, 0};
    if (i < n) {
        const double* s = surf + ((uint64_t)nsurf * i + isurf) * SF_NREC;
        if (s[SF_VALID] != 0.0) {
            const double* r = in9 + 9 * i;
            pix = healpix::vec2pix_ring(nside, r[6], r[7], r[8]);
            double q[4];
            lensing_from_jacobian(s, q);
            v[0] = __double2ll_rn(q[0] * FIX_SCALE);
            v[1] = __double2ll_rn(q[1] * FIX_SCALE);
            v[2] = __double2ll_rn(q[2] * FIX_SCALE);
            v[3] = 1;
            mine = true;
        }
    }
    // wavefront reduction per distinct pixel, one atomic per pixel and map
    bool pending = mine;
    while (__any(pending)) {
        const unsigned long long m = __ballot(pending);
        const int leader = __ffsll((long long)m) - 1;
        const int64_t p0 = __shfl(pix, leader);
        const bool take = pending && pix == p0;
        long long c[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) c[k] = wave_sum_i64(take ? v[k] : 0);
        if ((int)__lane_id() == leader) {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                atomicAdd(acc + (uint64_t)k * npix + (uint64_t)p0, (unsigned long long)c[k]);
        }
        pending = pending && !take;
    }
}

__global__ void convert_kernel(const unsigned long long* __restrict__ acc, uint64_t npix, double* __restrict__ maps) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const long long cnt = (long long)acc[3 * npix + p];
    for (int k = 0; k < 3; ++k) {
        const long long s = (long long)acc[(uint64_t)k * npix + p];
        maps[(uint64_t)k * npix + p] = cnt ? ((double)s / FIX_SCALE) / (double)cnt : 0.0;
    }
    maps[3 * npix + p] = (double)cnt;
}

__global__ void fill_zero_u64(unsigned long long* p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// pixel-centre mode: map[(j * 4 + k) * npix + pix] for surface j, component k
__global__ void pixel_fill_kernel(int64_t nside, const int64_t* __restrict__ pix, const double* __restrict__ surf,
                                  int nsurf, uint64_t n, double* __restrict__ maps, uint64_t npix) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = pix[i];
    if (p < 0 || (uint64_t)p >= npix) return;
    for (int j = 0; j < nsurf; ++j) {
        const double* s = surf + ((uint64_t)nsurf * i + j) * SF_NREC;
        double q[4] = {0.0, 0.0, 0.0, 0.0};
        if (s[SF_VALID] != 0.0) lensing_from_jacobian(s, q);   // rejected pixels are 0 (hmaps.h:716-720)
        for (int k = 0; k < 4; ++k) maps[((uint64_t)j * 4 + k) * npix + p] = q[k];
    }
}

// Every map component of Hmaps::FillMap (hmaps.h:716-836) for the pixels of
// the rays, from their surface records: maps[(nmaps * iz + icomp) * npix + p]
// in FP32 (the reference's Healpix_Map<float>).  comps: component codes
// (MPF_MAP_*); homo: per surface (t, chi, lambda, redshift, a) of the
// homogeneous reference ray; jac: bundle Jacobians (n x nz x 4) or null for
// the infinitesimal one carried by the surface records.
struct CompArgs {
    int64_t nside = 0;
    const int64_t* pix = nullptr;
    const double* surf = nullptr;
    const double* surf_born = nullptr;   // Born surface records (lensing_born), or null
    const double* jac = nullptr;
    const double* homo = nullptr;
    bool real = true;                    // a component needs the ray's own surface records
    uint64_t n = 0, npix = 0;
    int nz = 0, ncomp = 0, nmaps = 0;
    int comps[16] = {0};
    float* maps = nullptr;
};

__global__ void __launch_bounds__(256) component_fill_kernel(CompArgs A) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const int64_t p = A.pix[i];
    if (p < 0 || (uint64_t)p >= A.npix) return;
    double v[3];
    healpix::pix2vec_ring(A.nside, p, v);
    const double phi = atan2(v[1], v[0]);
    const double theta = acos(v[2]);
    for (int iz = 0; iz < A.nz; ++iz) {
        const double* s = A.surf + ((uint64_t)A.nz * i + iz) * SF_NREC;
        const double* sb = A.surf_born ? A.surf_born + ((uint64_t)A.nz * i + iz) * SF_NREC : nullptr;
        float* m0 = A.maps + (uint64_t)A.nmaps * iz * A.npix + (uint64_t)p;
        bool zero = (A.real && s[SF_VALID] == 0.0) || (sb && sb[SF_VALID] == 0.0);   // hmaps.h:716-720
        if (!zero && A.jac) zero = A.jac[4 * ((uint64_t)A.nz * i + iz)] == 42.0;
        if (zero) {
            for (int k = 0; k < A.nmaps; ++k) m0[(uint64_t)k * A.npix] = 0.0f;
            continue;
        }
        const double* h = A.homo + 5 * iz;   // t, chi, lambda, redshift, a
        int icomp = 0;
        for (int j = 0; j < A.ncomp; ++j) {
            float* m = m0 + (uint64_t)icomp * A.npix;
            switch (A.comps[j]) {
                case MPF_MAP_LENSING:
                    if (A.jac) {
                        const double* a = A.jac + 4 * ((uint64_t)A.nz * i + iz);
                        const double a11 = a[0], a12 = a[1], a21 = a[2], a22 = a[3];
                        const double w = -atan2(a12 - a21, a11 + a22);
                        const double cw = cos(w), sw = sin(w);
                        m[0] = (float)(1 - (a11 + a22) * 0.5 / cw);
                        m[A.npix] = (float)(-0.5 * (cw * (a11 - a22) + sw * (a12 + a21)));
                        m[2 * A.npix] = (float)(-0.5 * (cw * (a12 + a21) + sw * (a22 - a11)));
                        m[3 * A.npix] = (float)w;
                        m[4 * A.npix] = (float)(a11 * a22 - a12 * a21);
                        icomp += 4;
                    } else {
                        double q[4];
                        lensing_from_jacobian(s, q);
                        for (int k = 0; k < 4; ++k) m[(uint64_t)k * A.npix] = (float)q[k];
                        icomp += 3;
                    }
                    break;
                case MPF_MAP_LENSING_BORN: {
                    double q[4];
                    lensing_from_jacobian(sb, q);
                    for (int k = 0; k < 4; ++k) m[(uint64_t)k * A.npix] = (float)q[k];
                    icomp += 3;
                    break;
                }
                case MPF_MAP_DR: m[0] = (float)(s[SF_CHI] / h[1] - 1); break;
                case MPF_MAP_DL: m[0] = (float)(s[SF_LAMBDA] / h[2] - 1); break;
                case MPF_MAP_DT: m[0] = (float)(s[SF_T] / h[0] - 1); break;
                case MPF_MAP_DA: m[0] = (float)(s[SF_A] / h[4] - 1); break;
                case MPF_MAP_DZ: m[0] = (float)(s[SF_REDSHIFT] / h[3] - 1); break;
                case MPF_MAP_DS: m[0] = (float)(s[SF_S] / h[1] - 1); break;
                case MPF_MAP_ISW: m[0] = (float)s[SF_ISW]; break;
                case MPF_MAP_DENS: m[0] = (float)s[SF_RHO]; break;
                case MPF_MAP_DENS_MAX: m[0] = (float)s[SF_RHOMAX]; break;
                case MPF_MAP_STEPS: m[0] = (float)s[SF_FIRSTID]; break;
                case MPF_MAP_PHI: m[0] = (float)s[SF_PHI]; break;
                case MPF_MAP_DEFLECTION: {
                    const double x = s[SF_X], y = s[SF_Y], z = s[SF_Z];
                    const double r = sqrt(x * x + y * y + z * z);
                    m[0] = (float)(phi - atan2(y, x));
                    m[A.npix] = (float)(theta - acos(z / r));
                    icomp++;
                    break;
                }
                default: break;
            }
            icomp++;
        }
    }
}

// launch photons at pixel centres (hmaps.h:599, 607; integrator.h:264-290)
__global__ void pixel_launch_kernel(int64_t nside, const int64_t* __restrict__ pix, uint64_t n,
                                    double* __restrict__ in9) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[3];
    healpix::pix2vec_ring(nside, pix[i], v);
    const double norm = sqrt((v[0] - 0.0) * (v[0] - 0.0) + (v[1] - 0.0) * (v[1] - 0.0) + (v[2] - 0.0) * (v[2] - 0.0));
    double* r = in9 + 9 * i;
    r[0] = 1.0; r[1] = 0.0; r[2] = 0.0; r[3] = 0.0; r[4] = 0.0; r[5] = 1.0;
    r[6] = (v[0] - 0.0) / norm;
    r[7] = (v[1] - 0.0) / norm;
    r[8] = (v[2] - 0.0) / norm;
}

__global__ void vec2pix_kernel(int64_t nside, const double* __restrict__ vec, uint64_t n, int64_t* __restrict__ pix) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pix[i] = healpix::vec2pix_ring(nside, vec[3 * i], vec[3 * i + 1], vec[3 * i + 2]);
}

__global__ void pix2vec_kernel(int64_t nside, const int64_t* __restrict__ pix, uint64_t n, double* __restrict__ vec) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    healpix::pix2vec_ring(nside, pix[i], vec + 3 * i);
}

inline unsigned blocks_for(uint64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

hipError_t launch_bin(int64_t nside, const double* in9, const double* surf, int nsurf, int isurf, uint64_t n,
                      unsigned long long* acc, double* maps, hipStream_t st) {
    const uint64_t npix = 12ull * (uint64_t)nside * (uint64_t)nside;
    hipLaunchKernelGGL(fill_zero_u64, dim3(blocks_for(4 * npix)), dim3(256), 0, st, acc, 4 * npix);
    if (n) hipLaunchKernelGGL(bin_kernel, dim3(blocks_for(n)), dim3(256), 0, st, nside, in9, surf, nsurf, isurf, n,
                              acc, npix);
    hipLaunchKernelGGL(convert_kernel, dim3(blocks_for(npix)), dim3(256), 0, st, acc, npix, maps);
    return hipGetLastError();
}

hipError_t launch_pixel_fill(int64_t nside, const int64_t* pix, const double* surf, int nsurf, uint64_t n,
                             double* maps, hipStream_t st) {
    const uint64_t npix = 12ull * (uint64_t)nside * (uint64_t)nside;
    if (n) hipLaunchKernelGGL(pixel_fill_kernel, dim3(blocks_for(n)), dim3(256), 0, st, nside, pix, surf, nsurf, n,
                              maps, npix);
    return hipGetLastError();
}

hipError_t launch_component_fill(int64_t nside, const int64_t* pix, const double* surf, const double* surf_born,
                                 const double* jac, const double* homo, uint64_t n, int nz, const int* comps,
                                 int ncomp, int nmaps, float* maps, hipStream_t st) {
    CompArgs A;
    A.nside = nside;
    A.pix = pix;
    A.surf = surf;
    A.surf_born = surf_born;
    A.real = false;
    for (int j = 0; j < ncomp; ++j) A.real |= comps[j] != MPF_MAP_LENSING_BORN;
    A.jac = jac;
    A.homo = homo;
    A.n = n;
    A.npix = 12ull * (uint64_t)nside * (uint64_t)nside;
    A.nz = nz;
    A.ncomp = ncomp;
    A.nmaps = nmaps;
    for (int j = 0; j < ncomp && j < 16; ++j) A.comps[j] = comps[j];
    A.maps = maps;
    if (n) hipLaunchKernelGGL(component_fill_kernel, dim3(blocks_for(n)), dim3(256), 0, st, A);
    return hipGetLastError();
}

hipError_t launch_pixel_launch(int64_t nside, const int64_t* pix, uint64_t n, double* in9, hipStream_t st) {
    if (n) hipLaunchKernelGGL(pixel_launch_kernel, dim3(blocks_for(n)), dim3(256), 0, st, nside, pix, n, in9);
    return hipGetLastError();
}

hipError_t launch_vec2pix(int64_t nside, const double* vec, uint64_t n, int64_t* pix, hipStream_t st) {
    if (n) hipLaunchKernelGGL(vec2pix_kernel, dim3(blocks_for(n)), dim3(256), 0, st, nside, vec, n, pix);
    return hipGetLastError();
}

hipError_t launch_pix2vec(int64_t nside, const int64_t* pix, uint64_t n, double* vec, hipStream_t st) {
    if (n) hipLaunchKernelGGL(pix2vec_kernel, dim3(blocks_for(n)), dim3(256), 0, st, nside, pix, n, vec);
    return hipGetLastError();
}

}  // namespace mpf
