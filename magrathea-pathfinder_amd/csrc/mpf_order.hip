// Ray order of the persistent integrators: the work queue hands out rays in
// the order this file computes (queue position -> ray index).  Rays are
// sorted by path length (their per-ray stop or surface value, longest
// first), cut into buckets of `bucket` consecutive rays, and each bucket is
// sorted by direction (a Morton code of the octahedral map of the wave
// vector).  A wavefront that takes neighbouring queue positions then holds
// rays of similar length whose stencils overlap.  The order changes no
// result: every ray is integrated on its own and written at its own index.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>

#include "mpf_order.h"

namespace mpf {

namespace {

__device__ inline uint32_t spread16(uint32_t v) {
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// octahedral map of a direction to [-1, 1]^2, 16 bits per coordinate,
// interleaved
__device__ inline uint32_t direction_code(double x, double y, double z) {
    const double s = fabs(x) + fabs(y) + fabs(z);
    if (!(s > 0.0)) return 0u;
    double u = x / s, v = y / s;
    if (z < 0.0) {
        const double uu = (1.0 - fabs(v)) * (u < 0.0 ? -1.0 : 1.0);
        const double vv = (1.0 - fabs(u)) * (v < 0.0 ? -1.0 : 1.0);
        u = uu;
        v = vv;
    }
    const uint32_t qu = (uint32_t)fmin(65535.0, fmax(0.0, (u + 1.0) * 32767.5));
    const uint32_t qv = (uint32_t)fmin(65535.0, fmax(0.0, (v + 1.0) * 32767.5));
    return spread16(qu) | (spread16(qv) << 1);
}

// a double as an unsigned key of the same order (NaN last)
__device__ inline uint64_t ordered_bits(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// key of the path-length sort: ascending key = longest path first (larger
// stop values are farther, except for the scale factor, where smaller is)
__global__ void length_keys_kernel(const double* __restrict__ value, int smaller_is_longer, uint64_t n,
                                   uint64_t* __restrict__ key, uint32_t* __restrict__ idx) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = ordered_bits(value[i]);
    key[i] = smaller_is_longer ? k : ~k;
    idx[i] = (uint32_t)i;
}

__global__ void iota_kernel(uint64_t n, uint32_t* __restrict__ idx) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) idx[i] = (uint32_t)i;
}

// key of the direction sort: bucket of the queue position, then direction
__global__ void direction_keys_kernel(const double* __restrict__ in9, const uint32_t* __restrict__ idx, uint64_t n,
                                      uint64_t bucket, uint64_t* __restrict__ key) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double* r = in9 + 9 * (uint64_t)idx[j];
    key[j] = ((j / bucket) << 32) | (uint64_t)direction_code(r[6], r[7], r[8]);
}

// key of the source sort: the direction of the position
__global__ void source_keys_kernel(const double* __restrict__ src6, uint64_t n, uint64_t* __restrict__ key,
                                   uint32_t* __restrict__ idx) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* s = src6 + 6 * i;
    key[i] = (uint64_t)direction_code(s[0], s[1], s[2]);
    idx[i] = (uint32_t)i;
}

unsigned grid_for(uint64_t n) { return (unsigned)((n + 255) / 256); }

int bits_for(uint64_t v) {
    int b = 0;
    while (b < 64 && (v >> b)) ++b;
    return b;
}

}  // namespace

size_t ray_order_work_bytes(uint64_t n) {
    size_t a = 0, b = 0;
    uint64_t* k = nullptr;
    uint32_t* v = nullptr;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, a, k, k, v, v, (int)n, 0, 64);
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, b, k, k, v, v, (int)n, 0, 64);
    const size_t temp = (a > b ? a : b);
    // two key arrays, one extra index array, the sort's temporary storage
    return 2 * 8 * n + 4 * n + ((temp + 255) / 256) * 256 + 256;
}

hipError_t launch_ray_order(const double* in9, const double* ray_value, int smaller_is_longer, uint64_t n,
                            uint64_t bucket, uint32_t* order, void* work, size_t work_bytes, hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (bucket == 0) bucket = n;
    unsigned char* w = static_cast<unsigned char*>(work);
    uint64_t* key_a = reinterpret_cast<uint64_t*>(w);
    uint64_t* key_b = key_a + n;
    uint32_t* idx = reinterpret_cast<uint32_t*>(key_b + n);
    unsigned char* temp = reinterpret_cast<unsigned char*>(idx + n);
    temp += (256 - (reinterpret_cast<uintptr_t>(temp) & 255)) & 255;
    size_t temp_bytes = work_bytes - (size_t)(temp - w);
    hipError_t e;
    if (ray_value && bucket < n) {
        // longest paths first
        hipLaunchKernelGGL(length_keys_kernel, dim3(grid_for(n)), dim3(256), 0, st, ray_value, smaller_is_longer, n,
                           key_a, order);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, key_a, key_b, order, idx, (int)n, 0, 64, st);
        if (e != hipSuccess) return e;
    } else {
        hipLaunchKernelGGL(iota_kernel, dim3(grid_for(n)), dim3(256), 0, st, n, idx);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(direction_keys_kernel, dim3(grid_for(n)), dim3(256), 0, st, in9, idx, n, bucket, key_a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int end_bit = 32 + bits_for((n - 1) / bucket);
    return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, key_a, key_b, idx, order, (int)n, 0, end_bit, st);
}

hipError_t launch_source_order(const double* src6, uint64_t n, uint32_t* order, void* work, size_t work_bytes,
                               hipStream_t st) {
    if (n == 0) return hipSuccess;
    unsigned char* w = static_cast<unsigned char*>(work);
    uint64_t* key_a = reinterpret_cast<uint64_t*>(w);
    uint64_t* key_b = key_a + n;
    uint32_t* idx = reinterpret_cast<uint32_t*>(key_b + n);
    unsigned char* temp = reinterpret_cast<unsigned char*>(idx + n);
    temp += (256 - (reinterpret_cast<uintptr_t>(temp) & 255)) & 255;
    size_t temp_bytes = work_bytes - (size_t)(temp - w);
    hipLaunchKernelGGL(source_keys_kernel, dim3(grid_for(n)), dim3(256), 0, st, src6, n, key_a, idx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, key_a, key_b, idx, order, (int)n, 0, 32, st);
}

}  // namespace mpf
