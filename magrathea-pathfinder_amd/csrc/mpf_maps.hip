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
