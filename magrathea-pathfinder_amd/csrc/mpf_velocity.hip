// Velocity field of the reference's velocity-field build (VELOCITYFIELD):
// particles deposited into the cells of the cone octree with CIC or TSC
// weights at every stored level (Hmaps::CreateOctreeVelocityWithCIC / WithTSC,
// hmaps.h:439-478, 497-541), the sums divided by the deposited weight, and
// cells without weight given their parent's velocity level by level
// (hmaps.cpp:389-441).  The weight replaces the cell's rho, as in the
// reference, where rho is cleared before the deposit (hmaps.cpp:391).
//
// Layout: the sums are a 4 x n FP64 structure of arrays (weight, vx, vy, vz)
// indexed by the level-sorted cell index; the finished field is 3 x n FP32
// (the reference's Gravity<float> cell).  One thread per particle and level;
// the 8 (CIC) or 27 (TSC) cells of a particle are found by the tree descent
// and summed with FP64 atomics.  The reference sums in FP32 in an unspecified
// thread order; here the sums are FP64 and rounded once to FP32.
#include "mpf_device.h"
#include "mpf_kernels.h"

namespace mpf {

namespace {

__device__ __forceinline__ double tsc_weight(int ai, double d) {
    return ai * 0.5 * (1.5 - d) * (1.5 - d) + (1 - ai) * (0.75 - d * d);
}

// SCHEME 1: CIC, the cells holding the 8 points at +-half a cell from the
// particle (hmaps.h:459-474); SCHEME 2: TSC, the 27 cells holding the points
// at 0 and +-one cell (hmaps.h:516-539)
template <int SCHEME>
__global__ void __launch_bounds__(256) deposit_kernel(DevTree t, const float* __restrict__ pos,
                                                      const float* __restrict__ vel, uint64_t np, unsigned lvl,
                                                      double* __restrict__ acc) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double p[3] = {(double)pos[3 * i], (double)pos[3 * i + 1], (double)pos[3 * i + 2]};
    const float v[3] = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
    const double cells = (double)(1ull << lvl);
    const double half = 0.5 * EXTENT / cells;
    const double invextension = 1 / (EXTENT / cells);
    const double twohalves = 2 * half;
    const uint64_t n = t.n;
    constexpr int STEP = (SCHEME == 1) ? 2 : 1;
    const double off = (SCHEME == 1) ? half : twohalves;
    for (int iz = -1; iz <= 1; iz += STEP)
        for (int iy = -1; iy <= 1; iy += STEP)
            for (int ix = -1; ix <= 1; ix += STEP) {
                const uint32_t kx = axis_coord(p[0] + off * ix, lvl);
                const uint32_t ky = axis_coord(p[1] + off * iy, lvl);
                const uint32_t kz = axis_coord(p[2] + off * iz, lvl);
                unsigned fl = 0;
                const uint32_t g = find_cell(t, lvl, kx, ky, kz, &fl);
                if (g == NIL || fl != lvl) continue;  // no cell of this key stored
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
                atomicAdd(acc + g, w);
                atomicAdd(acc + n + g, (double)v[0] * w);
                atomicAdd(acc + 2 * n + g, (double)v[1] * w);
                atomicAdd(acc + 3 * n + g, (double)v[2] * w);
            }
}

// hmaps.cpp:423-430: v /= (rho + (rho == 0)), rho = the FP32 weight
__global__ void __launch_bounds__(256) velocity_normalise_kernel(uint64_t n, const double* __restrict__ acc,
                                                                 float* __restrict__ rho32, double* __restrict__ rho64,
                                                                 float* __restrict__ vel) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const float w = (float)acc[g];
    double mass = (double)w;
    mass = mass + (mass == 0);
#pragma unroll
    for (int k = 0; k < 3; ++k) vel[k * n + g] = (float)((double)(float)acc[(k + 1) * n + g] / mass);
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

}  // namespace

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
