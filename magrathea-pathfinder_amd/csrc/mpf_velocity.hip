// Velocity field of the reference's velocity-field build (VELOCITYFIELD):
// particles deposited into the cells of the cone octree with CIC or TSC
// weights at every stored level (Hmaps::CreateOctreeVelocityWithCIC / WithTSC,
// hmaps.h:439-478, 497-541), the sums divided by the deposited weight, and
// cells without weight given their parent's velocity level by level
// (hmaps.cpp:389-441).  The weight replaces the cell's rho, as in the
// reference, where rho is cleared before the deposit (hmaps.cpp:391).
//
// Layout: the sums are n x 4 FP64 (weight, vx, vy, vz of a cell in one 32-byte
// segment) indexed by the level-sorted cell index; the finished field is 3 x n FP32
// (the reference's Gravity<float> cell).  The particles of a call are first
// put in Morton order of their position (a 63-bit key at level 21, radix
// sorted), so that the lanes of a wavefront descend the same branches of the
// tree and add into the same cache lines; then one thread per particle and
// level finds the 8 (CIC) or 27 (TSC) cells by the tree descent and adds
// into them with FP64 atomics.  The reference sums in FP32 in an unspecified
// thread order; here the sums are FP64 and rounded once to FP32.
#include <hipcub/hipcub.hpp>

#include "mpf_device.h"
#include "mpf_kernels.h"

namespace mpf {

namespace {

__device__ __forceinline__ double tsc_weight(int ai, double d) {
    return ai * 0.5 * (1.5 - d) * (1.5 - d) + (1 - ai) * (0.75 - d * d);
}

// SCHEME 1: CIC, the cells holding the 8 points at +-half a cell from the
// particle (hmaps.h:459-474); SCHEME 2: TSC, the 27 cells holding the points
// at 0 and +-one cell (hmaps.h:516-539).  The particles are in Morton order,
// so neighbouring lanes often add into the same cell: for each stencil point
// the lanes first sum over runs of equal cells (a segmented sum across the
// wavefront) and the first lane of each run adds the run's sums with FP64
// atomics.  Every lane takes part in the cross-lane steps (no early exit).
template <int SCHEME>
__global__ void __launch_bounds__(256) deposit_kernel(DevTree t, const float* __restrict__ pos,
                                                      const float* __restrict__ vel, uint64_t np, unsigned lvl,
                                                      double* __restrict__ acc) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < np;
    const uint64_t ii = live ? i : 0;
    const unsigned lane = __lane_id();
    const double p[3] = {(double)pos[3 * ii], (double)pos[3 * ii + 1], (double)pos[3 * ii + 2]};
    const float v[3] = {vel[3 * ii], vel[3 * ii + 1], vel[3 * ii + 2]};
    const double cells = (double)(1ull << lvl);
    const double half = 0.5 * EXTENT / cells;
    const double invextension = 1 / (EXTENT / cells);
    const double twohalves = 2 * half;
    constexpr int STEP = (SCHEME == 1) ? 2 : 1;
    const double off = (SCHEME == 1) ? half : twohalves;
    for (int iz = -1; iz <= 1; iz += STEP)
        for (int iy = -1; iy <= 1; iy += STEP)
            for (int ix = -1; ix <= 1; ix += STEP) {
                const uint32_t kx = axis_coord(p[0] + off * ix, lvl);
                const uint32_t ky = axis_coord(p[1] + off * iy, lvl);
                const uint32_t kz = axis_coord(p[2] + off * iz, lvl);
                unsigned fl = 0;
                uint32_t g = live ? find_cell(t, lvl, kx, ky, kz, &fl) : NIL;
                if (g != NIL && fl != lvl) g = NIL;  // no cell of this key stored
                double s[4] = {0.0, 0.0, 0.0, 0.0};
                if (g != NIL) {
                    double w;
                    if (SCHEME == 1) {
                        w = (1 - fabs(p[0] - cell_center(lvl, kx)) * invextension) *
                            (1 - fabs(p[1] - cell_center(lvl, ky)) * invextension) *
                            (1 - fabs(p[2] - cell_center(lvl, kz)) * invextension);
                    } else {
                        w = tsc_weight(ix != 0, fabs((cell_center(lvl, kx) - p[0]) / (twohalves))) *
                            tsc_weight(iy != 0, fabs((cell_center(lvl, ky) - p[1]) / (twohalves))) *
                            tsc_weight(iz != 0, fabs((cell_center(lvl, kz) - p[2]) / (twohalves)));
                    }
                    s[0] = w;
                    s[1] = (double)v[0] * w;
                    s[2] = (double)v[1] * w;
                    s[3] = (double)v[2] * w;
                }
                // runs of lanes with the same cell; lane k of a run ends up
                // with the sums of lanes k .. end-1
                const uint32_t gp = __shfl_up(g, 1);
                const bool head = (lane == 0) || (gp != g);
                const unsigned long long heads = __ballot(head);
                const unsigned long long later = (lane >= 63) ? 0ull : (heads & (~0ull << (lane + 1)));
                const unsigned end = later ? (unsigned)__builtin_ctzll(later) : 64u;
#pragma unroll
                for (unsigned d = 1; d < 64; d <<= 1) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const double o = __shfl_down(s[k], d);
                        if (lane + d < end) s[k] += o;
                    }
                }
                if (head && g != NIL) {
                    double* a = acc + 4 * (uint64_t)g;
                    atomicAdd(a, s[0]);
                    atomicAdd(a + 1, s[1]);
                    atomicAdd(a + 2, s[2]);
                    atomicAdd(a + 3, s[3]);
                }
            }
}

// hmaps.cpp:423-430: v /= (rho + (rho == 0)), rho = the FP32 weight
__global__ void __launch_bounds__(256) velocity_normalise_kernel(uint64_t n, const double* __restrict__ acc,
                                                                 float* __restrict__ rho32, double* __restrict__ rho64,
                                                                 float* __restrict__ vel) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const float w = (float)acc[4 * g];
    double mass = (double)w;
    mass = mass + (mass == 0);
#pragma unroll
    for (int k = 0; k < 3; ++k) vel[k * n + g] = (float)((double)(float)acc[4 * g + 1 + k] / mass);
    rho32[g] = w;
    if (rho64) rho64[g] = (double)w;
}

// parent of every stored child, from the child masks
__global__ void __launch_bounds__(256) parent_kernel(DevTree t, uint32_t* __restrict__ parent) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= t.n) return;
    const unsigned m = t.child_mask[g];
    if (!m) return;
    const uint32_t c0 = t.child_first[g];
    unsigned r = 0;
    for (unsigned s = 0; s < 8; ++s)
        if (m & (1u << s)) parent[c0 + r++] = (uint32_t)g;
}

// hmaps.cpp:432-441 for the cells [first, last) of one level: a cell whose
// weight is not a normal number takes its parent's velocity
__global__ void __launch_bounds__(256) parent_fill_kernel(uint64_t n, uint64_t first, uint64_t last,
                                                          const float* __restrict__ rho32,
                                                          const uint32_t* __restrict__ parent,
                                                          float* __restrict__ vel) {
    const uint64_t g = first + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= last) return;
    if (__builtin_isnormal(rho32[g])) return;
    const uint32_t q = parent[g];
    if (q >= n) return;  // no stored parent (the tree is parent-closed: not reached)
#pragma unroll
    for (int k = 0; k < 3; ++k) vel[k * n + g] = vel[k * n + q];
}

inline unsigned blocks_of(uint64_t n) { return (unsigned)((n + 255) / 256); }

// the low 21 bits of v to every third bit
__device__ inline uint64_t spread21(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x001f00000000ffffull;
    x = (x | (x << 16)) & 0x001f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__global__ void __launch_bounds__(256) particle_keys_kernel(const float* __restrict__ pos, uint64_t n,
                                                            uint64_t* __restrict__ key, uint32_t* __restrict__ idx) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t kx = axis_coord((double)pos[3 * i], 21), ky = axis_coord((double)pos[3 * i + 1], 21),
                   kz = axis_coord((double)pos[3 * i + 2], 21);
    key[i] = spread21(kx) | (spread21(ky) << 1) | (spread21(kz) << 2);
    idx[i] = (uint32_t)i;
}

__global__ void __launch_bounds__(256) gather_particles_kernel(const float* __restrict__ pos,
                                                               const float* __restrict__ vel,
                                                               const uint32_t* __restrict__ idx, uint64_t n,
                                                               float* __restrict__ pos_s, float* __restrict__ vel_s) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t j = idx[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        pos_s[3 * i + k] = pos[3 * j + k];
        vel_s[3 * i + k] = vel[3 * j + k];
    }
}

struct OrderWork {
    uint64_t *key_a, *key_b;
    uint32_t *idx_a, *idx_b;
    float *pos_s, *vel_s;
    unsigned char* temp;
    size_t temp_bytes;
};

size_t sort_temp_bytes(uint64_t n) {
    size_t t = 0;
    uint64_t* k = nullptr;
    uint32_t* v = nullptr;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, t, k, k, v, v, (int)n, 0, 63);
    return t;
}

OrderWork carve(void* work, uint64_t n) {
    OrderWork w;
    unsigned char* p = static_cast<unsigned char*>(work);
    auto take = [&](size_t bytes) {
        unsigned char* q = p;
        p += (bytes + 255) / 256 * 256;
        return q;
    };
    w.key_a = reinterpret_cast<uint64_t*>(take(8 * n));
    w.key_b = reinterpret_cast<uint64_t*>(take(8 * n));
    w.idx_a = reinterpret_cast<uint32_t*>(take(4 * n));
    w.idx_b = reinterpret_cast<uint32_t*>(take(4 * n));
    w.pos_s = reinterpret_cast<float*>(take(12 * n));
    w.vel_s = reinterpret_cast<float*>(take(12 * n));
    w.temp_bytes = sort_temp_bytes(n);
    w.temp = take(w.temp_bytes);
    return w;
}

}  // namespace

size_t velocity_order_work_bytes(uint64_t n) {
    const size_t r = 256;
    auto up = [&](size_t b) { return (b + r - 1) / r * r; };
    return 2 * up(8 * n) + 2 * up(4 * n) + 2 * up(12 * n) + up(sort_temp_bytes(n));
}

hipError_t launch_velocity_order(const float* pos, const float* vel, uint64_t n, void* work, const float** pos_s,
                                 const float** vel_s, hipStream_t st) {
    OrderWork w = carve(work, n);
    *pos_s = w.pos_s;
    *vel_s = w.vel_s;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(particle_keys_kernel, dim3(blocks_of(n)), dim3(256), 0, st, pos, n, w.key_a, w.idx_a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipcub::DeviceRadixSort::SortPairs(w.temp, w.temp_bytes, w.key_a, w.key_b, w.idx_a, w.idx_b, (int)n, 0, 63,
                                           st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(gather_particles_kernel, dim3(blocks_of(n)), dim3(256), 0, st, pos, vel, w.idx_b, n, w.pos_s,
                       w.vel_s);
    return hipGetLastError();
}

hipError_t launch_velocity_deposit(const DevTree& t, int scheme, const float* pos, const float* vel, uint64_t np,
                                   unsigned lvl, double* acc, hipStream_t st) {
    if (np == 0) return hipSuccess;
    if (scheme == 1)
        hipLaunchKernelGGL(deposit_kernel<1>, dim3(blocks_of(np)), dim3(256), 0, st, t, pos, vel, np, lvl, acc);
    else hipLaunchKernelGGL(deposit_kernel<2>, dim3(blocks_of(np)), dim3(256), 0, st, t, pos, vel, np, lvl, acc);
    return hipGetLastError();
}

hipError_t launch_velocity_finish(const DevTree& t, const double* acc, float* rho32, double* rho64, float* vel,
                                  uint32_t* parent, const uint64_t* level_first, unsigned min_level,
                                  unsigned max_level, hipStream_t st) {
    const uint64_t n = t.n;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(velocity_normalise_kernel, dim3(blocks_of(n)), dim3(256), 0, st, n, acc, rho32, rho64, vel);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(parent_kernel, dim3(blocks_of(n)), dim3(256), 0, st, t, parent);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    // level by level: a level reads its parents' finished velocities
    for (unsigned l = min_level + 1; l <= max_level; ++l) {
        const uint64_t first = level_first[l - min_level], last = level_first[l - min_level + 1];
        if (last <= first) continue;
        hipLaunchKernelGGL(parent_fill_kernel, dim3(blocks_of(last - first)), dim3(256), 0, st, n, first, last, rho32,
                           parent, vel);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace mpf
