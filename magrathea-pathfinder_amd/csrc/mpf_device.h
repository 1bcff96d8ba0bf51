// Device view of the level-sorted octree and the exact lookup/interpolation
// primitives shared by the kernels.
//
// Everything in this header reproduces the reference's floating-point
// expressions operation by operation; translation units that include it are
// compiled with -ffp-contract=off so no multiply-add is fused.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace mpf {

constexpr uint32_t NIL = 0xFFFFFFFFu;
constexpr double EXTENT = 4.0;
constexpr double C_LIGHT = 299792458.0;              // constants.h:441-444
constexpr double C2 = 299792458.0 * 299792458.0;     // constants.h:481-484

struct DevTree {
    uint32_t B = 0, maxL = 0;                        // coarsest / finest stored level
    uint32_t bx0 = 0, by0 = 0, bz0 = 0, bnx = 0, bny = 0, bnz = 0;
    const uint32_t* base = nullptr;                  // dense level-B table -> cell
    const uint32_t* child_first = nullptr;
    const uint8_t* child_mask = nullptr;
    const uint32_t* elem = nullptr;                  // cell -> record index
    const float* f32 = nullptr;                      // 7 x n, structure of arrays
    const double* f64 = nullptr;                     // 7 x n or null
    const float* vel = nullptr;                      // 3 x n cell velocity / c, or null
    // cell-major copies of the fields for the lean integrator's TSC, or null:
    // q32[2 g] = (dphidx, dphidy, dphidz, dphidt), q32[2 g + 1] = (rho, phi, a, 0);
    // q64[4 g .. 4 g + 3] = (dphidx, dphidy), (dphidz, dphidt), (rho, phi), (a, 0)
    const float4* q32 = nullptr;
    const double2* q64 = nullptr;
    uint64_t n = 0;
};

// Field slots (gravity.h:45-87 tuple order)
enum { F_RHO = 0, F_PHI = 1, F_DX = 2, F_DY = 3, F_DZ = 4, F_A = 5, F_DT = 6 };

// Per-axis integer coordinate of compute(L, ...) (simplehyperoctreeindex.h:1167-1170):
// Type((1/E) * 2^L * (x - (-E/2))), a signed truncation, then the low L bits.
__device__ __forceinline__ uint32_t axis_coord(double x, unsigned L) {
    const double scale = (1.0 / EXTENT) * (double)(1ull << L);
    const double v = scale * (x - (-EXTENT / 2.0));
    const long long t = (long long)v;
    return (uint32_t)((unsigned long long)t & ((1ull << L) - 1ull));
}

// center() of a level-L cell (simplehyperoctreeindex.h:890-893)
__device__ __forceinline__ double cell_center(unsigned L, uint32_t k) {
    return (-EXTENT / 2.0) + (EXTENT * ((double)(2ull * k + 1ull) / (double)(1ull << (L + 1))));
}

// The descent of find_cell below from cell g, stored at level lvl, whose
// region contains the level-L key (kx, ky, kz): the finest stored cell that
// contains the key, or NIL when the element upper_bound would find does not
// contain it.  *flvl receives the level of the cell found.
__device__ __forceinline__ uint32_t descend_cell(const DevTree& t, uint32_t g, unsigned lvl, unsigned L, uint32_t kx,
                                                 uint32_t ky, uint32_t kz, unsigned* flvl) {
    while (lvl < L) {
        const unsigned m = t.child_mask[g];
        const unsigned b = L - lvl - 1;
        const unsigned s = (((kz >> b) & 1u) << 2) | (((ky >> b) & 1u) << 1) | ((kx >> b) & 1u);
        if (m & (1u << s)) {
            g = t.child_first[g] + (uint32_t)__popc(m & ((1u << s) - 1u));
            ++lvl;
        } else {
            // g is the finest stored ancestor; upper_bound lands on it only if
            // no earlier sibling subtree is stored
            if (m & ((1u << s) - 1u)) return NIL;
            break;
        }
    }
    *flvl = lvl;
    return g;
}

// cell of the dense base table (level B) whose region holds the level-L
// key, or NIL
__device__ __forceinline__ uint32_t base_cell(const DevTree& t, unsigned L, uint32_t kx, uint32_t ky, uint32_t kz) {
    if (L < t.B) return NIL;
    const unsigned sh = L - t.B;
    const uint32_t bx = (uint32_t)((uint64_t)kx >> sh) - t.bx0;
    const uint32_t by = (uint32_t)((uint64_t)ky >> sh) - t.by0;
    const uint32_t bz = (uint32_t)((uint64_t)kz >> sh) - t.bz0;
    if (bx >= t.bnx || by >= t.bny || bz >= t.bnz) return NIL;
    return t.base[((uint64_t)bz * t.bny + by) * t.bnx + bx];
}

// Exact replacement of upper_bound(key(L, k)) - 1 followed by containing():
// returns the level-sorted cell index, or NIL when the element found does not
// contain the key.  *flvl receives the level of the cell found.
__device__ __forceinline__ uint32_t find_cell(const DevTree& t, unsigned L, uint32_t kx, uint32_t ky, uint32_t kz,
                                              unsigned* flvl) {
    const uint32_t g = base_cell(t, L, kx, ky, kz);
    if (g == NIL) return NIL;
    return descend_cell(t, g, t.B, L, kx, ky, kz, flvl);
}

__device__ __forceinline__ uint32_t locate_point(const DevTree& t, double x, double y, double z, unsigned* flvl) {
    return find_cell(t, 32, axis_coord(x, 32), axis_coord(y, 32), axis_coord(z, 32), flvl);
}

}  // namespace mpf
