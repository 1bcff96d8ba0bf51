// Split-precision tier (MPF_PREC_SPLIT): FP32 perturbation, FP64 background.
//
// Load-time preparation (split_build):
//   1. For every cell, the cell whose 27-cell TSC stencil serves points whose
//      finest cell it is: the reference evaluates the stencil at the level of
//      the finest cell and, when the first bad neighbour in its ix/iy/iz loop
//      order is coarser, redoes it at that neighbour's level; a missing
//      neighbour yields Data() (simplehyperoctree.h:1596-1640).  That chain
//      depends on the cell only, so it is resolved once per cell here.
//   2. For every stencil-centre ("target") cell, a block with its FP64 centre,
//      1/size and the TSC interpolant of (dphidt, dphidx, dphidy, dphidz, phi)
//      as tri-quadratic polynomials in the offset (u, v, w) in [-1/2, 1/2]^3
//      from the centre, 27 FP32 coefficients per field.  With
//      W-1(u) = (1/2 - u)^2 / 2, W0(u) = 3/4 - u^2, W+1(u) = (1/2 + u)^2 / 2
//      (the weights of simplehyperoctree.h:1587-1589 for d = |centre - x|/size),
//      sum_ijk f_ijk Wi(u) Wj(v) Wk(w) = sum_pqr C_pqr u^p v^q w^r.
//
// Kernel: one photon per lane, FP64 state (a, t, position, wave vector) and
// FP64 accumulations (affine parameter, distance, redshift, ISW); the
// interpolated fields, dphi/dlambda and the perturbation terms of the
// geodesic equation in FP32.  A lane keeps its finest cell and the current
// coefficient block in registers and only searches the tree when a point
// leaves that cell.  The fields after a step are reused as the first stage
// of the next step (the reference interpolates the same point twice).
#include "mpf_split.h"

#include "mpf_exact_dev.h"
#include "mpf_features.h"
#include "mpf_surfaces.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace mpf {

namespace {

constexpr int NCOEF = 27;
constexpr int NFIELD = 5;  // dphidt, dphidx, dphidy, dphidz, phi
constexpr int FSTRIDE = 28;  // floats per field (27 coefficients + pad): 7 x 16-byte loads
constexpr int BLOCK_DOUBLES = 4;
constexpr int BLOCK_BYTES = BLOCK_DOUBLES * 8 + NFIELD * FSTRIDE * 4;  // 592

// tslot word of a cell: NIL, or the block slot with PARTIAL_BIT set when the
// cell stores some of its children (then it is the finest cell only of the
// points outside them, and a lane may not reuse it for the next point)
constexpr uint32_t PARTIAL_BIT = 0x80000000u;
constexpr uint32_t SLOT_MASK = 0x7FFFFFFFu;

// the stencil blocks through the constant address space: loads at an
// address uniform across the wave become scalar loads (the blocks are not
// written while the integrators run)
typedef const __attribute__((address_space(4))) float ConstF;
typedef const __attribute__((address_space(4))) double ConstD;
// blocks per evaluation the integrator reads with scalar loads
#ifndef MPF_SCALAR_BLOCKS
#define MPF_SCALAR_BLOCKS 1
#endif
constexpr int SCALAR_BLOCKS = MPF_SCALAR_BLOCKS;
// ... for the Jacobian samples: 0, their lanes spread over more blocks than
// the ray's own cell, and the scalar walk then costs more than it saves (the
// 4 M-ray Jacobian batch: 402 ms with vector loads only, 437 ms with one
// scalar block, 469 ms with two; bit-identical results)
#ifndef MPF_JAC_SCALAR_BLOCKS
#define MPF_JAC_SCALAR_BLOCKS 0
#endif
constexpr int JAC_SCALAR_BLOCKS = MPF_JAC_SCALAR_BLOCKS;
// ... and for the ray's own fields in the kernel with surfaces: 0 (with it
// the surfaces kernel spills 76 B/lane and the same batch takes 548 ms)
#ifndef MPF_SURF_SCALAR_BLOCKS
#define MPF_SURF_SCALAR_BLOCKS 0
#endif
constexpr int SURF_SCALAR_BLOCKS = MPF_SURF_SCALAR_BLOCKS;
// ... and instead, in that kernel, a cache of blocks per wave in LDS
// (WaveCache): entries per wave, 0 for none
#ifndef MPF_WAVE_CACHE
#define MPF_WAVE_CACHE 8
#endif
constexpr int WAVE_CACHE = MPF_WAVE_CACHE;
constexpr unsigned WAVE_CACHE_MOD = WAVE_CACHE > 0 ? (unsigned)WAVE_CACHE : 1u;
static_assert(WAVE_CACHE >= 0 && WAVE_CACHE <= 32, "the entries in use are a 32-bit mask");
// device memory left free by the optional face-neighbour table: room for
// 16 M rays with a surface (about 11 GB) and nside-4096 maps (12.9 GB)
constexpr uint64_t NBR_RESERVE = 24ull << 30;
// ... and by the rho blocks: the launch records and ray order of 16 M rays
constexpr uint64_t RHO_RESERVE = 4ull << 30;
// ... and by the tidal blocks
constexpr uint64_t TIDAL_RESERVE = 4ull << 30;
// tidal blocks: the differences (+ sample minus - sample, one cell away) of
// (dphidx along x, dphidx along y, dphidx along z, dphidy along y, dphidy
// along z, dphidz along z), FSTRIDE coefficients each
constexpr int TIDAL_COMPONENTS = 6;
constexpr int TIDAL_FLOATS = TIDAL_COMPONENTS * FSTRIDE;   // 168: 672 bytes

struct SplitView {
    const uint32_t* tslot;        // cell -> block slot (| PARTIAL_BIT) or NIL
    const unsigned char* blocks;  // slot -> BLOCK_BYTES
    const uint2* nbr;             // cell x 6 faces -> neighbour entry
    const float* rho;             // slot -> FSTRIDE coefficients of rho, or null
    const uint32_t* tidx;         // slot -> tidal block index or NIL, or null
    const float* tidal;           // tidal block -> TIDAL_FLOATS coefficients
};

// Face-neighbour entries (neighbour_kernel): for cell g at level l and face
// d = 2 * axis + (0 for -1, 1 for +1), the level-l cell h across that face:
//   {h, tslot word}  h is stored at level l without children (word < PARTIAL_BIT)
//   {h, NBR_DESCEND} h is stored at level l with children: the lookup
//                    continues the descent from h
//   {NIL, NIL}       anything else (a coarser neighbour, outside the tree or
//                    the domain, no stencil): the lookup starts from the root
constexpr uint32_t NBR_DESCEND = PARTIAL_BIT;

// -------------------------------------------------------------- build

// PARTIAL_BIT of the cells that store some (not all) of their children
__global__ void partial_kernel(DevTree t, uint32_t* __restrict__ tslot) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= t.n) return;
    const unsigned m = t.child_mask[g];
    if (m != 0u && tslot[g] != NIL) tslot[g] |= PARTIAL_BIT;
}

__global__ void neighbour_kernel(DevTree t, const uint8_t* __restrict__ lev, const uint32_t* __restrict__ kx,
                                 const uint32_t* __restrict__ ky, const uint32_t* __restrict__ kz,
                                 const uint32_t* __restrict__ tslot, uint2* __restrict__ nbr) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= t.n) return;
    const unsigned l = lev[g];
    const long long lim = (long long)((1ull << l) - 1ull);
    for (unsigned d = 0; d < 6; ++d) {
        long long c[3] = {(long long)kx[g], (long long)ky[g], (long long)kz[g]};
        c[d >> 1] += (d & 1u) ? 1 : -1;
        uint2 e = make_uint2(NIL, NIL);
        if (c[d >> 1] >= 0 && c[d >> 1] <= lim) {
            unsigned fl = 0;
            const uint32_t h = find_cell(t, l, (uint32_t)c[0], (uint32_t)c[1], (uint32_t)c[2], &fl);
            if (h != NIL && fl == l) {
                if (t.child_mask[h] != 0u) e = make_uint2(h, NBR_DESCEND);
                else if (tslot[h] != NIL) e = make_uint2(h, tslot[h]);
            }
        }
        nbr[6 * g + d] = e;
    }
}

__global__ void target_kernel(DevTree t, const uint8_t* __restrict__ lev, const uint32_t* __restrict__ kx,
                              const uint32_t* __restrict__ ky, const uint32_t* __restrict__ kz, uint32_t* target) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= t.n) return;
    unsigned l = lev[g];
    uint32_t cx = kx[g], cy = ky[g], cz = kz[g];
    uint32_t cur = (uint32_t)g;
    for (int guard = 0; guard < 40; ++guard) {
        const uint32_t lim = (l >= 32) ? 0xFFFFFFFFu : ((1u << l) - 1u);
        unsigned bad_level = l;
        bool missing = false;
        for (int ic = 0; ic < 27 && !missing && bad_level == l; ++ic) {
            const int ox = ic / 9 - 1, oy = (ic / 3) % 3 - 1, oz = ic % 3 - 1;
            const long long nx = (long long)cx + ox, ny = (long long)cy + oy, nz = (long long)cz + oz;
            // stencils that wrap round the octree domain are treated as missing
            if (nx < 0 || ny < 0 || nz < 0 || nx > lim || ny > lim || nz > lim) { missing = true; break; }
            unsigned fl = 0;
            const uint32_t gv = find_cell(t, l, (uint32_t)nx, (uint32_t)ny, (uint32_t)nz, &fl);
            if (gv == NIL) missing = true;
            else if (fl < l) bad_level = fl;
        }
        if (missing) { target[g] = NIL; return; }
        if (bad_level == l) { target[g] = cur; return; }
        // drop to the ancestor at the coarser level (stored: the tree is parent-closed)
        const unsigned sh = l - bad_level;
        cx >>= sh; cy >>= sh; cz >>= sh;
        l = bad_level;
        unsigned fl = 0;
        cur = find_cell(t, l, cx, cy, cz, &fl);
        if (cur == NIL || fl != l) { target[g] = NIL; return; }
    }
    target[g] = NIL;
}

__device__ __forceinline__ double field_value(const DevTree& t, uint32_t g, int f) {
    // f: 0 dphidt, 1 dphidx, 2 dphidy, 3 dphidz, 4 phi
    const int slot = (f == 0) ? F_DT : (f == 4) ? F_PHI : (F_DX + f - 1);
    return t.f64 ? t.f64[(uint64_t)slot * t.n + g] : (double)t.f32[(uint64_t)slot * t.n + g];
}

// The 27 coefficients C_pqr of sum_ijk f_ijk W_i(u) W_j(v) W_k(w), the TSC
// interpolant of the 27 cell values v (index 9 i + 3 j + k, offsets -1..1)
// as a tri-quadratic polynomial in the offset from the centre (FP32, the
// 28th entry a zero pad)
__device__ __forceinline__ void tsc_poly(const double* v, float* co) {
    // M[p][i]: coefficient of u^p in W_i, i = -1, 0, +1
    const double M[3][3] = {{0.125, 0.75, 0.125}, {-0.5, 0.0, 0.5}, {0.5, -1.0, 0.5}};
    // transform along z, then y, then x
    double a1[27], a2[27];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int r = 0; r < 3; ++r) {
                double acc = 0.0;
                for (int k = 0; k < 3; ++k) acc += M[r][k] * v[i * 9 + j * 3 + k];
                a1[i * 9 + j * 3 + r] = acc;
            }
    for (int i = 0; i < 3; ++i)
        for (int q = 0; q < 3; ++q)
            for (int r = 0; r < 3; ++r) {
                double acc = 0.0;
                for (int j = 0; j < 3; ++j) acc += M[q][j] * a1[i * 9 + j * 3 + r];
                a2[i * 9 + q * 3 + r] = acc;
            }
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q)
            for (int r = 0; r < 3; ++r) {
                double acc = 0.0;
                for (int i = 0; i < 3; ++i) acc += M[p][i] * a2[i * 9 + q * 3 + r];
                co[p * 9 + q * 3 + r] = (float)acc;
            }
    co[27] = 0.0f;
}

__global__ void __launch_bounds__(128) block_kernel(DevTree t, const uint8_t* __restrict__ lev, const uint32_t* __restrict__ kx,
                             const uint32_t* __restrict__ ky, const uint32_t* __restrict__ kz,
                             const uint32_t* __restrict__ slot_cell, uint64_t nslots, unsigned char* blocks) {
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const uint32_t g = slot_cell[s];
    const unsigned l = lev[g];
    const uint32_t cx = kx[g], cy = ky[g], cz = kz[g];
    uint32_t nb[27];
    for (int ic = 0; ic < 27; ++ic) {
        const int ox = ic / 9 - 1, oy = (ic / 3) % 3 - 1, oz = ic % 3 - 1;
        unsigned fl = 0;
        nb[ic] = find_cell(t, l, cx + ox, cy + oy, cz + oz, &fl);
    }
    unsigned char* blk = blocks + s * (uint64_t)BLOCK_BYTES;
    double* hd = reinterpret_cast<double*>(blk);
    hd[0] = cell_center(l, cx);
    hd[1] = cell_center(l, cy);
    hd[2] = cell_center(l, cz);
    hd[3] = (double)(1ull << l) / EXTENT;  // 1 / cell size
    float* co = reinterpret_cast<float*>(blk + BLOCK_DOUBLES * 8);
    for (int f = 0; f < NFIELD; ++f) {
        double v[27];
        for (int ic = 0; ic < 27; ++ic) v[ic] = field_value(t, nb[ic], f);
        tsc_poly(v, co + f * FSTRIDE);
    }
}

// the rho coefficients of each slot (blocks for the surfaces' densities)
__global__ void __launch_bounds__(128) rho_block_kernel(DevTree t, const uint8_t* __restrict__ lev,
                                                        const uint32_t* __restrict__ kx, const uint32_t* __restrict__ ky,
                                                        const uint32_t* __restrict__ kz,
                                                        const uint32_t* __restrict__ slot_cell, uint64_t nslots,
                                                        float* __restrict__ rho) {
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const uint32_t g = slot_cell[s];
    const unsigned l = lev[g];
    double v[27];
    for (int ic = 0; ic < 27; ++ic) {
        const int ox = ic / 9 - 1, oy = (ic / 3) % 3 - 1, oz = ic % 3 - 1;
        unsigned fl = 0;
        const uint32_t h = find_cell(t, l, kx[g] + ox, ky[g] + oy, kz[g] + oz, &fl);
        v[ic] = t.f64 ? t.f64[(uint64_t)F_RHO * t.n + h] : (double)t.f32[(uint64_t)F_RHO * t.n + h];
    }
    tsc_poly(v, rho + s * (uint64_t)FSTRIDE);
}

// A slot qualifies for a tidal block when its cell is a leaf that is its
// own stencil centre and its six same-level face neighbours are leaves and
// their own stencil centres: then the Jacobian samples one cell away from a
// point of the cell lie in those neighbours at the point's own offset, and
// each sample is the neighbour's TSC polynomial at that offset.
__device__ __forceinline__ bool own_centre(const DevTree& t, const uint32_t* __restrict__ tslot,
                                           const uint32_t* __restrict__ slot_cell, uint32_t g) {
    if (g == NIL || t.child_mask[g] != 0u) return false;
    const uint32_t w = tslot[g];
    return w != NIL && (w & PARTIAL_BIT) == 0u && slot_cell[w] == g;
}

__global__ void tidal_flag_kernel(DevTree t, const uint8_t* __restrict__ lev, const uint32_t* __restrict__ kx,
                                  const uint32_t* __restrict__ ky, const uint32_t* __restrict__ kz,
                                  const uint32_t* __restrict__ tslot, const uint32_t* __restrict__ slot_cell,
                                  uint64_t nslots, uint32_t* __restrict__ flag) {
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const uint32_t g = slot_cell[s];
    bool ok = own_centre(t, tslot, slot_cell, g);
    const unsigned l = lev[g];
    const long long lim = (long long)((1ull << l) - 1ull);
    for (unsigned d = 0; d < 6 && ok; ++d) {
        long long c[3] = {(long long)kx[g], (long long)ky[g], (long long)kz[g]};
        c[d >> 1] += (d & 1u) ? 1 : -1;
        if (c[d >> 1] < 0 || c[d >> 1] > lim) { ok = false; break; }
        unsigned fl = 0;
        const uint32_t h = find_cell(t, l, (uint32_t)c[0], (uint32_t)c[1], (uint32_t)c[2], &fl);
        ok = h != NIL && fl == l && own_centre(t, tslot, slot_cell, h);
    }
    flag[s] = ok ? 1u : 0u;
}

__global__ void __launch_bounds__(128) tidal_block_kernel(DevTree t, const uint8_t* __restrict__ lev,
                                                          const uint32_t* __restrict__ kx,
                                                          const uint32_t* __restrict__ ky,
                                                          const uint32_t* __restrict__ kz,
                                                          const uint32_t* __restrict__ slot_cell,
                                                          const uint32_t* __restrict__ tidx, uint64_t nslots,
                                                          float* __restrict__ tidal) {
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots || tidx[s] == NIL) return;
    const uint32_t g = slot_cell[s];
    const unsigned l = lev[g];
    // (field of grad phi: 1 dphidx, 2 dphidy, 3 dphidz; axis of the shift)
    const int comp[TIDAL_COMPONENTS][2] = {{1, 0}, {1, 1}, {1, 2}, {2, 1}, {2, 2}, {3, 2}};
    float* out = tidal + (uint64_t)tidx[s] * TIDAL_FLOATS;
    for (int k = 0; k < TIDAL_COMPONENTS; ++k) {
        const int f = comp[k][0], axis = comp[k][1];
        double v[27];
        for (int ic = 0; ic < 27; ++ic) {
            const int o[3] = {ic / 9 - 1, (ic / 3) % 3 - 1, ic % 3 - 1};
            uint32_t cp[3], cm[3];
            cp[0] = kx[g] + o[0]; cp[1] = ky[g] + o[1]; cp[2] = kz[g] + o[2];
            for (int a = 0; a < 3; ++a) cm[a] = cp[a];
            cp[axis] += 1u;
            cm[axis] -= 1u;
            unsigned fl = 0;
            const uint32_t hp = find_cell(t, l, cp[0], cp[1], cp[2], &fl);
            const uint32_t hm = find_cell(t, l, cm[0], cm[1], cm[2], &fl);
            v[ic] = field_value(t, hp, f) - field_value(t, hm, f);
        }
        tsc_poly(v, out + k * FSTRIDE);
    }
}

// ------------------------------------------------------------- kernel

// sum_pqr C_pqr u^p v^q w^r by Horner: 26 FMA
__device__ __forceinline__ float horner27(const float* c, float u, float v, float w) {
    float g[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        float h[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const float* cc = c + p * 9 + q * 3;
            h[q] = __builtin_fmaf(__builtin_fmaf(cc[2], w, cc[1]), w, cc[0]);
        }
        g[p] = __builtin_fmaf(__builtin_fmaf(h[2], v, h[1]), v, h[0]);
    }
    return __builtin_fmaf(__builtin_fmaf(g[2], u, g[1]), u, g[0]);
}

// sum_pqr C_pqr u^p v^q w^r with the coefficients in scalar registers and
// m[3q + r] = v^q w^r: each product takes one coefficient operand (a VALU
// instruction reads one scalar register), the constant terms are added last.
// 29 instructions against Horner's 26 FMA plus the 9 copies a two-scalar
// first FMA of each inner triple needs; the 8 products m are shared by the
// fields.
__device__ __forceinline__ float mono27(const float* c, const float* m, float u) {
    float s[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const float* cp = c + 9 * p;
        float a = cp[1] * m[1];
#pragma unroll
        for (int i = 2; i < 9; ++i) a = __builtin_fmaf(cp[i], m[i], a);
        s[p] = a + cp[0];
    }
    return __builtin_fmaf(__builtin_fmaf(s[2], u, s[1]), u, s[0]);
}

__device__ __forceinline__ void monomials9(float v, float w, float* m) {
    const float w2 = w * w, v2 = v * v;
    m[0] = 1.0f; m[1] = w; m[2] = w2;
    m[3] = v; m[4] = v * w; m[5] = v * w2;
    m[6] = v2; m[7] = v2 * w; m[8] = v2 * w2;
}

// RC = true keeps the current coefficient block in registers; RC = false
// keeps only its address and reads the coefficients at every evaluation
// (from L1/L2: a lane reads the same block for ~10 consecutive steps).
template <bool RC>
struct Interp {
    // finest cell of the last point: level and coordinates
    unsigned lvl;
    uint32_t kx, ky, kz;
    bool valid;          // the last point had data
    bool reuse;          // ... and its cell stores no children: every point inside has the same finest cell
    uint32_t cell;       // ... that cell
    uint32_t slot;
    int ent;             // WaveCache entry last found for the slot (-1: none)
    double cx, cy, cz, inv;
    float c[RC ? NFIELD * NCOEF : 1];
};

// A cache in LDS of the coefficient blocks of one wave's lanes: WAVE_CACHE
// entries of a block's coefficients (ENTRY_F4 16-byte chunks) and their
// slots.  A lane keeps the entry of its block (Interp::ent); when that
// entry's tag no longer names the lane's slot, the wave finds the slot among
// the tags, or loads the block into an entry that no lane of the current
// evaluation uses (the active lanes load it together, one 16-byte chunk
// each).  A block then crosses the vector memory path once per wave instead
// of once per lane; the lanes read their coefficients from LDS.  With more
// distinct blocks in one evaluation than entries the wave evaluates in
// rounds: the lanes left without an entry wait for the next round, in which
// the entries of the lanes done are free.  Only the wave's own lanes touch
// its entries, and a wave's LDS operations complete in program order.
constexpr int ENTRY_F4 = NFIELD * FSTRIDE / 4;  // 35
typedef float F4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) F4v LdsF4;
typedef const __attribute__((address_space(1))) F4v GlobalF4;
typedef __attribute__((address_space(3))) uint32_t LdsU32;
struct WaveCache {
    LdsF4* data;    // WAVE_CACHE x ENTRY_F4 chunks
    LdsU32* tags;   // WAVE_CACHE slots (NIL: empty), then the next entry to replace
};

// The entry that holds slot's coefficients for the lanes that need one and
// whose entry `ent` does not hold it (need && !hit), or -1 when every entry
// is read by other lanes of this round; every active lane of the wave calls
// it.  Out of line: it runs when some lane of the wave moved to another block.
__device__ __noinline__ int wave_cache_miss(const unsigned char* blocks, LdsF4* data, LdsU32* tags, int ent,
                                            uint32_t slot, bool hit, bool need) {
    unsigned long long miss = __ballot(need && !hit);
    // entries that lanes of this evaluation read
    unsigned used = 0u;
#pragma unroll
    for (int j = 0; j < WAVE_CACHE; ++j)
        if (__ballot(hit && ent == j)) used |= 1u << j;
    const unsigned long long act = __ballot(true);
    const unsigned n = (unsigned)__popcll(act);
    const unsigned r = (unsigned)__popcll(act & ((1ull << __lane_id()) - 1ull));
    while (miss) {
        const uint32_t s0 = __builtin_amdgcn_readlane(slot, __builtin_ctzll(miss));
        int e = -1;
#pragma unroll
        for (int j = 0; j < WAVE_CACHE; ++j)
            if (tags[j] == s0) e = j;
        e = __builtin_amdgcn_readfirstlane(e);
        if (e < 0) {
            const unsigned start = __builtin_amdgcn_readfirstlane(tags[WAVE_CACHE]);
            for (int k = 0; k < WAVE_CACHE; ++k) {
                const unsigned j = (start + (unsigned)k) % WAVE_CACHE_MOD;
                if (!((used >> j) & 1u)) {
                    e = (int)j;
                    break;
                }
            }
            if (e < 0) break;  // every entry is read in this evaluation
            GlobalF4* src = (GlobalF4*)(blocks + (uint64_t)s0 * BLOCK_BYTES + BLOCK_DOUBLES * 8);
            LdsF4* dst = data + e * ENTRY_F4;
            for (unsigned c = r; c < (unsigned)ENTRY_F4; c += n) dst[c] = src[c];
            if (r == 0) {
                tags[e] = s0;
                tags[WAVE_CACHE] = (uint32_t)(e + 1) % WAVE_CACHE_MOD;
            }
        }
        used |= 1u << e;
        const bool mine = need && !hit && slot == s0;
        if (mine) ent = e;
        miss &= ~__ballot(mine);
    }
    if ((miss >> __lane_id()) & 1ull) ent = -1;
    return ent;
}

// For the lanes that need one, sets ent to the entry holding slot's
// coefficients, or -1 (none free this round); every active lane of the wave
// calls it.
__device__ __forceinline__ void wave_cache_find(const SplitView& sv, const WaveCache& W, int& ent, uint32_t slot,
                                                bool need) {
    bool hit = need && ent >= 0 && W.tags[ent] == slot;
    if (__ballot(need && !hit)) {
        // a lane that moved to another block looks for it among the tags
        // first (some lane of 64 moves at most evaluations, and its new block
        // is mostly cached already); only a block that is not calls the wave
        // out of line to load it
        if (need && !hit) {
#pragma unroll
            for (int j = 0; j < WAVE_CACHE; ++j) {
                if (W.tags[j] == slot) {
                    ent = j;
                    hit = true;
                }
            }
        }
        if (__ballot(need && !hit)) ent = wave_cache_miss(sv.blocks, W.data, W.tags, ent, slot, hit, need);
    }
}

template <bool RC>
__device__ __forceinline__ void load_block(const SplitView& sv, uint32_t slot, Interp<RC>& I) {
    const unsigned char* blk = sv.blocks + (uint64_t)slot * BLOCK_BYTES;
    const double4 h = *reinterpret_cast<const double4*>(blk);
    I.cx = h.x; I.cy = h.y; I.cz = h.z; I.inv = h.w;
    if (RC) {
        const float4* cf = reinterpret_cast<const float4*>(blk + BLOCK_DOUBLES * 8);
#pragma unroll
        for (int f = 0; f < NFIELD; ++f) {
#pragma unroll
            for (int i = 0; i < FSTRIDE / 4; ++i) {
                const float4 q = cf[f * (FSTRIDE / 4) + i];
                const int b = f * NCOEF + 4 * i;
                if (4 * i + 0 < NCOEF) I.c[b + 0] = q.x;
                if (4 * i + 1 < NCOEF) I.c[b + 1] = q.y;
                if (4 * i + 2 < NCOEF) I.c[b + 2] = q.z;
                if (4 * i + 3 < NCOEF) I.c[b + 3] = q.w;
            }
        }
    }
    I.slot = slot;
}

// axis_coord(x, 32): inside the domain (0 <= v < 2^32) the truncation is one
// conversion; outside it the full signed truncation and wrap of axis_coord
__device__ __forceinline__ uint32_t axis_key32(double x) {
    const double v = ((1.0 / EXTENT) * 4294967296.0) * (x - (-EXTENT / 2.0));
    if (v >= 0.0 && v < 4294967296.0) return (uint32_t)v;
    return axis_coord(x, 32);
}

// Tslot word of the finest cell of the level-32 key (kx32, ky32, kz32)
// (NIL: no data), with the cell in *g and its level in *fl.  (ax, ay, az) is
// the key's level-l prefix minus the last cell's coordinates plus one, when
// the lane has a last cell (valid); the lookup then starts from that cell or
// its face neighbour when the key lies in their regions.
__device__ __forceinline__ uint32_t lookup_near(const DevTree& t, const SplitView& sv, bool valid, uint32_t cell,
                                                unsigned lvl, uint32_t ax, uint32_t ay, uint32_t az, uint32_t kx32,
                                                uint32_t ky32, uint32_t kz32, uint32_t* g, unsigned* fl) {
    uint32_t from = NIL;
    unsigned from_lvl = t.B;
    *g = NIL;
    if (valid && ax <= 2u && ay <= 2u && az <= 2u) {
        const unsigned moved = (ax != 1u) + (ay != 1u) + (az != 1u);
        if (moved == 0u) {
            from = cell;
            from_lvl = lvl;
        } else if (moved == 1u && sv.nbr) {
            const unsigned d = (ax != 1u) ? (ax >> 1) : (ay != 1u) ? 2u + (ay >> 1) : 4u + (az >> 1);
            const uint2 e = sv.nbr[(uint64_t)cell * 6u + d];
            if (e.y < PARTIAL_BIT) {
                *g = e.x;
                *fl = lvl;
                return e.y;
            }
            if (e.y == NBR_DESCEND) {
                from = e.x;
                from_lvl = lvl;
            }
        }
    }
    if (from == NIL) from = base_cell(t, 32, kx32, ky32, kz32);
    if (from != NIL) *g = descend_cell(t, from, from_lvl, 32, kx32, ky32, kz32, fl);
    return (*g == NIL) ? NIL : sv.tslot[*g];
}

// Fields at (x, y, z): out[0..3] = dphidt, dphidx, dphidy, dphidz; out[4] = phi
// when NF == 5.  Returns false when the reference would return Data().
//
// The finest cell is the one find_cell returns for the level-32 key of the
// point.  A lane keeps its last cell (level l, coordinates k): a key whose
// level-l prefix is k lies in that cell's region, and one whose prefix is
// k +- 1 along one axis in the region of the neighbour across that face.
// The lookup then starts from the last cell (it stores no children: it is
// the answer; else the descent continues from it) or from the neighbour
// entry, and ends in the cell the root descent finds.
// SB: distinct blocks of the wave read with scalar loads (0: vector loads only)
// F0: the first field evaluated (out[0 .. F0 - 1] are left as they are)
// WC: the vector loads go through the wave's block cache W (SB = 0)
template <int NF, bool RC, int SB = 0, int F0 = 0, bool WC = false>
__device__ __forceinline__ bool split_eval(const DevTree& t, const SplitView& sv, Interp<RC>& I, double x, double y,
                                          double z, float* out, const WaveCache* W = nullptr) {
    const uint32_t kx32 = axis_key32(x), ky32 = axis_key32(y), kz32 = axis_key32(z);
    uint32_t ax = 0u, ay = 0u, az = 0u;
    if (I.valid) {
        const unsigned sh = 32u - I.lvl;
        ax = (uint32_t)((uint64_t)kx32 >> sh) - I.kx + 1u;
        ay = (uint32_t)((uint64_t)ky32 >> sh) - I.ky + 1u;
        az = (uint32_t)((uint64_t)kz32 >> sh) - I.kz + 1u;
    }
    const bool same = I.valid && I.reuse && ax == 1u && ay == 1u && az == 1u;
    if (!same) {
        unsigned fl = 0;
        uint32_t g;
        const uint32_t word = lookup_near(t, sv, I.valid, I.cell, I.lvl, ax, ay, az, kx32, ky32, kz32, &g, &fl);
        if (word == NIL) {
            I.valid = false;
            return false;
        }
        I.cell = g;
        const uint32_t slot = word & SLOT_MASK;
        I.lvl = fl;
        I.kx = (uint32_t)((uint64_t)kx32 >> (32 - fl));
        I.ky = (uint32_t)((uint64_t)ky32 >> (32 - fl));
        I.kz = (uint32_t)((uint64_t)kz32 >> (32 - fl));
        I.valid = true;
        // a cell with some children stored (a partial oct, e.g. at a cone's
        // edge) is the finest cell only of the points outside those children
        I.reuse = (word & PARTIAL_BIT) == 0;
        if (slot != I.slot) load_block<RC>(sv, slot, I);
    }
    const float u = (float)((x - I.cx) * I.inv);
    const float v = (float)((y - I.cy) * I.inv);
    const float w = (float)((z - I.cz) * I.inv);
    // every path sums in the order of mono27, so that a lane's fields do not
    // depend on which path its wave takes
    if (RC) {
        float m[9];
        monomials9(v, w, m);
#pragma unroll
        for (int f = F0; f < NF; ++f) out[f] = mono27(I.c + f * NCOEF, m, u);
    } else {
        // The wave walks the distinct blocks of its lanes (SB at most) with
        // scalar loads of the coefficients: the vector memory path returns
        // 1 KiB per wave and 16-byte load, and its address and data units
        // bound the kernel when every lane reads its block.  The lanes of a
        // ray patch mostly share one or two blocks.
        unsigned long long pend = SB ? __ballot(true) : 0ull;
        for (int it = 0; it < SB && pend; ++it) {
            const uint32_t s0 = __builtin_amdgcn_readlane(I.slot, __builtin_ctzll(pend));
            ConstF* cs = (ConstF*)(sv.blocks + (uint64_t)s0 * BLOCK_BYTES + BLOCK_DOUBLES * 8);
            const bool mine = I.slot == s0;
            float m[9];
            monomials9(v, w, m);
#pragma unroll
            for (int f = F0; f < NF; ++f) {
                float c[NCOEF];
#pragma unroll
                for (int i = 0; i < NCOEF; ++i) c[i] = cs[f * FSTRIDE + i];
                const float r = mono27(c, m, u);
                out[f] = mine ? r : out[f];
            }
            pend &= ~__ballot(mine);
        }
        if (!SB || ((pend >> __lane_id()) & 1ull)) {
            // the coefficients of field f + 1 are in flight while field f is
            // evaluated: two round trips to L2 per evaluation instead of NF
            const float4* cf = reinterpret_cast<const float4*>(sv.blocks + (uint64_t)I.slot * BLOCK_BYTES + BLOCK_DOUBLES * 8);
            float4 q[2][FSTRIDE / 4];
            float m[9];
            if (WC) {
                // from the wave's block cache (LDS, short latency: one field
                // at a time), in rounds while some lane has no entry
                monomials9(v, w, m);
                bool done = false;
                do {
                    wave_cache_find(sv, *W, I.ent, I.slot, !done);
                    if (!done && I.ent >= 0) {
                        const LdsF4* cl = W->data + I.ent * ENTRY_F4;
#pragma unroll
                        for (int f = F0; f < NF; ++f) {
                            float c[FSTRIDE];
#pragma unroll
                            for (int i = 0; i < FSTRIDE / 4; ++i) {
                                const F4v e = cl[f * (FSTRIDE / 4) + i];
                                c[4 * i] = e.x; c[4 * i + 1] = e.y; c[4 * i + 2] = e.z; c[4 * i + 3] = e.w;
                            }
                            out[f] = mono27(c, m, u);
                        }
                        done = true;
                    }
                } while (__ballot(!done));
                return true;
            }
#pragma unroll
            for (int i = 0; i < FSTRIDE / 4; ++i) q[F0 & 1][i] = cf[F0 * (FSTRIDE / 4) + i];
            monomials9(v, w, m);
#pragma unroll
            for (int f = F0; f < NF; ++f) {
                if (f + 1 < NF) {
#pragma unroll
                    for (int i = 0; i < FSTRIDE / 4; ++i) q[(f + 1) & 1][i] = cf[(f + 1) * (FSTRIDE / 4) + i];
                }
                float c[FSTRIDE];
#pragma unroll
                for (int i = 0; i < FSTRIDE / 4; ++i) {
                    const float4 e = q[f & 1][i];
                    c[4 * i] = e.x; c[4 * i + 1] = e.y; c[4 * i + 2] = e.z; c[4 * i + 3] = e.w;
                }
                out[f] = mono27(c, m, u);
            }
        }
    }
    return true;
}

// Jacobian samples: grad phi components 1..NG (dphidx, dphidy, dphidz) at
// (x, y, z), with the sample's finest cell cached in LDS (one entry per lane
// and sample direction: cell coordinates + block slot, the cell, and its
// level | JC_REUSE | JC_VALID).  The samples sit one cell away from the ray
// along each axis and stay in the same neighbour cell for about as many steps
// as the ray stays in its own; a hit costs three key compares, a miss starts
// from the cached cell's face neighbour (lookup_near).
constexpr uint32_t JC_REUSE = 0x100u, JC_VALID = 0x200u;

template <int NG, int SB>
__device__ __forceinline__ bool jac_eval(const DevTree& t, const SplitView& sv, uint4& cell, uint32_t& cid,
                                         uint32_t& lw, double x, double y, double z, float* out) {
    const uint32_t kx32 = axis_key32(x), ky32 = axis_key32(y), kz32 = axis_key32(z);
    const unsigned l0 = lw & 0x3Fu;
    const unsigned sh = 32u - l0;
    uint4 c = cell;
    const uint32_t ax = (uint32_t)((uint64_t)kx32 >> sh) - c.x + 1u;
    const uint32_t ay = (uint32_t)((uint64_t)ky32 >> sh) - c.y + 1u;
    const uint32_t az = (uint32_t)((uint64_t)kz32 >> sh) - c.z + 1u;
    const bool hit = (lw & JC_REUSE) && ax == 1u && ay == 1u && az == 1u;
    if (!hit) {
        unsigned fl = 0;
        uint32_t g;
        const uint32_t word = lookup_near(t, sv, (lw & JC_VALID) != 0u, cid, l0, ax, ay, az, kx32, ky32, kz32, &g,
                                          &fl);
        if (word == NIL) {
            lw = 0u;
            return false;
        }
        c = make_uint4((uint32_t)((uint64_t)kx32 >> (32 - fl)), (uint32_t)((uint64_t)ky32 >> (32 - fl)),
                       (uint32_t)((uint64_t)kz32 >> (32 - fl)), word & SLOT_MASK);
        cell = c;
        cid = g;
        lw = fl | JC_VALID | ((word & PARTIAL_BIT) ? 0u : JC_REUSE);
    }
    // The wave evaluates the distinct blocks of its lanes (SB at most) from
    // scalar loads, as split_eval does; the rest read their block with vector
    // loads.  Both paths compute the same offsets and horner27 sums.
    unsigned long long pend = SB ? __ballot(true) : 0ull;
    for (int it = 0; it < SB && pend; ++it) {
        const uint32_t s0 = __builtin_amdgcn_readlane(c.w, __builtin_ctzll(pend));
        const unsigned char* b0 = sv.blocks + (uint64_t)s0 * BLOCK_BYTES;
        ConstD* hd = (ConstD*)b0;
        const double hx = hd[0], hy = hd[1], hz = hd[2], hw = hd[3];
        ConstF* cs = (ConstF*)(b0 + BLOCK_DOUBLES * 8);
        const bool mine = c.w == s0;
        const float u = (float)((x - hx) * hw);
        const float v = (float)((y - hy) * hw);
        const float w = (float)((z - hz) * hw);
#pragma unroll
        for (int f = 1; f <= NG; ++f) {
            float cc[FSTRIDE];
#pragma unroll
            for (int i = 0; i < NCOEF; ++i) cc[i] = cs[f * FSTRIDE + i];
            const float r = horner27(cc, u, v, w);
            out[f - 1] = mine ? r : out[f - 1];
        }
        pend &= ~__ballot(mine);
    }
    if (!SB || ((pend >> __lane_id()) & 1ull)) {
        const unsigned char* blk = sv.blocks + (uint64_t)c.w * BLOCK_BYTES;
        const double4 h = *reinterpret_cast<const double4*>(blk);
        const float u = (float)((x - h.x) * h.w);
        const float v = (float)((y - h.y) * h.w);
        const float w = (float)((z - h.z) * h.w);
        const float4* cf = reinterpret_cast<const float4*>(blk + BLOCK_DOUBLES * 8);
#pragma unroll
        for (int f = 1; f <= NG; ++f) {
            float cc[FSTRIDE];
#pragma unroll
            for (int i = 0; i < FSTRIDE / 4; ++i) {
                const float4 q = cf[f * (FSTRIDE / 4) + i];
                cc[4 * i] = q.x; cc[4 * i + 1] = q.y; cc[4 * i + 2] = q.z; cc[4 * i + 3] = q.w;
            }
            out[f - 1] = horner27(cc, u, v, w);
        }
    }
    return true;
}

__device__ __forceinline__ double cosmo_lin(double x0, const double* __restrict__ x, const double* __restrict__ y,
                                            int n, int& hint) {
    int i = hint;
    while (i < n && !(x0 < x[i])) ++i;
    while (i > 0 && x0 < x[i - 1]) --i;
    hint = i;
    i += (i <= 0) - (i >= n);
    return y[i - 1] + ((y[i] - y[i - 1]) * (x0 - x[i - 1]) / (x[i] - x[i - 1]));
}

// da/dt of the table at x0 with the segment slopes computed once per table
// (no division per call; the exact tiers keep Utility::interpolate's).  The
// segment of the last call is kept in registers: the hint walk ends at
// upper_bound(x0) whatever the hint, so a point inside that segment takes it
// without reading the table.
struct CosmoSeg {
    double lo, hi, y, slope;
    int hint;
};

__device__ __forceinline__ void cosmo_seg_init(CosmoSeg& S) {
    S.lo = 1.0;
    S.hi = 0.0;  // empty: the first call walks
    S.hint = 0;
}

__device__ __forceinline__ double cosmo_slope(double x0, const double* __restrict__ x, const double* __restrict__ y,
                                              const double* __restrict__ slope, int n, CosmoSeg& S) {
    if (!(x0 < S.lo) && x0 < S.hi) return fma(S.slope, x0 - S.lo, S.y);
    int i = S.hint;
    while (i < n && !(x0 < x[i])) ++i;
    while (i > 0 && x0 < x[i - 1]) --i;
    S.hint = i;
    i += (i <= 0) - (i >= n);
    S.lo = x[i - 1];
    S.hi = x[i];
    S.y = y[i - 1];
    S.slope = slope[i];
    return fma(S.slope, x0 - S.lo, S.y);
}

// 1 / a: the hardware reciprocal refined by one Newton step (within an ulp
// or two of the quotient; the FP64 background of the split tier does not need
// the correctly rounded division)
__device__ __forceinline__ double recip(double a) {
    double r = __builtin_amdgcn_rcp(a);
    const double e = fma(-a, r, 1.0);
    r = fma(r, e, r);
    return r;
}

constexpr double INV_C2 = 1.0 / C2;

enum { CA = 0, CT, CX, CY, CZ, CKT, CKX, CKY, CKZ };

// geodesic right-hand side: FP64 background, FP32 perturbation
__device__ __forceinline__ void rhs_split(const double* in, const float* F, double dadt, double inv_scale,
                                          double* out) {
    const float ktf = (float)in[CKT], kxf = (float)in[CKX], kyf = (float)in[CKY], kzf = (float)in[CKZ];
    const float c2inv2 = (float)(2.0 / C2);
    const float grad = F[1] * kxf + F[2] * kyf + F[3] * kzf;  // grad phi . k
    const float dphidl = F[0] * ktf + grad;
    const float kt2 = ktf * ktf;
    const float pert_t = -(c2inv2 * ktf) * grad;                 // -(2/c^2) kt (dphidl - dphidt kt)
    const float s = c2inv2 * dphidl;
    const double bg = -((2.0 * dadt) * recip(in[CA])) * in[CKT];
    out[CA] = in[CKT] * dadt;
    out[CT] = in[CKT];
    out[CX] = in[CKX] * inv_scale;
    out[CY] = in[CKY] * inv_scale;
    out[CZ] = in[CKZ] * inv_scale;
    out[CKT] = bg * in[CKT] + (double)pert_t;
    out[CKX] = bg * in[CKX] + (double)(s * kxf - 2.0f * F[1] * kt2);
    out[CKY] = bg * in[CKY] + (double)(s * kyf - 2.0f * F[2] * kt2);
    out[CKZ] = bg * in[CKZ] + (double)(s * kzf - 2.0f * F[3] * kt2);
}

struct SplitRay {
    double core[9];
    double lambda, s, isw, iswold, redshift, chi;
    double a0, phi0c2, dl;
    float phi, dx, dy, dz, dt;   // fields at the current point (dt: +dphidt at launch, -dphidt after)
    float F[4];                  // interpolated dphidt, dphidx, dphidy, dphidz at the current point
    uint32_t appended;
    CosmoSeg cs;
    double stopv;     // this ray's stop value (kept with surfaces only; re-read otherwise)
    float rho;        // rho at the current point (surfaces with rho blocks)
};

__device__ __forceinline__ double level_extent(unsigned l) { return ldexp((double)EXTENT, -(int)l); }

// Launch records (integrator.h:669-705) with exact FP64 fields at the start
// point: normalised core (9), level of the located cell (+ LAUNCH_EMPTY for
// an empty launch element), phi, dphidx, dphidy, dphidz, dphidt, rho.
constexpr int LAUNCH_DOUBLES = 16;
constexpr double LAUNCH_EMPTY = 1024.0;

// Records are written in queue order (record q is the ray the queue hands
// out at position q), so that lanes that take rays together read adjacent
// records.
__global__ void __launch_bounds__(256) launch_kernel(IntegrateArgs A, DevTree t) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const uint64_t ray = A.queue_order ? (uint64_t)A.queue_order[i] : i;
    const double c2 = C2;
    double core[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) core[k] = A.in9[9 * ray + k];
    exact::TscCache cache;
    exact::cache_init(cache);
    double d[7];
    int loc = -1;
    exact::tsc<1, false>(t, cache, core[CX], core[CY], core[CZ], d, &loc);
    const double phi = d[F_PHI];
    const double norm = sqrt((c2 * (1.0 + 2.0 * (phi / c2)) * core[CKT] * core[CKT]) /
                             ((1.0 - 2.0 * (phi / c2)) *
                              (core[CKX] * core[CKX] + core[CKY] * core[CKY] + core[CKZ] * core[CKZ])));
    core[CKX] *= norm;
    core[CKY] *= norm;
    core[CKZ] *= norm;
    int hint = 0;
    core[CA] = cosmo_lin(core[CT], A.ct, A.ca, A.ncos, hint);
    double* L = A.scratch + (uint64_t)LAUNCH_DOUBLES * i;
#pragma unroll
    for (int k = 0; k < 9; ++k) L[k] = core[k];
    // the located cell's level, + LAUNCH_EMPTY when the reference's launch element is empty
    L[9] = (double)(loc < 0 ? 0 : loc) + (exact::empty7(d) ? LAUNCH_EMPTY : 0.0);
    L[10] = phi;
    L[11] = d[F_DX]; L[12] = d[F_DY]; L[13] = d[F_DZ]; L[14] = d[F_DT];
    L[15] = d[F_RHO];
}

// rho and a (the `ah` output) of each final point, exact FP64 TSC
__global__ void __launch_bounds__(256) final_kernel(IntegrateArgs A, DevTree t) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    double* o = A.out + (uint64_t)MPF_NOUT * i;
    if (o[MPF_OUT_NSTEPS] < 0.0) return;
    exact::TscCache cache;
    exact::cache_init(cache);
    double d[7];
    int loc = -1;
    exact::tsc<1, false>(t, cache, o[MPF_OUT_X], o[MPF_OUT_Y], o[MPF_OUT_Z], d, &loc);
    o[MPF_OUT_RHO] = d[F_RHO];
    o[MPF_OUT_AH] = o[MPF_OUT_NSTEPS] > 0.0 ? d[F_A] : 0.0;
}

template <bool RC, int W, bool SURF>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, 8)))
integrate_split_kernel(IntegrateArgs A, DevTree t, SplitView sv) {
#pragma clang fp contract(fast)
    const double c = C_LIGHT, c2 = C2;
    const double inv_scale = 1.0 / A.length;
    const double ratio_k = (A.length / c) / (double)A.nsteps;
    Interp<RC> I;
    I.valid = false;
    I.slot = NIL;
    SplitRay R;
    // per-lane Jacobian sums and rho maxima in LDS (SurfAccLds)
    __shared__ double surf_sums[SURF ? SURF_LDS_DOUBLES : 1][SURF ? 256 : 1];
    SurfAccLds SA;
    SA.S = &surf_sums[0][SURF ? threadIdx.x : 0];
    SA.stride = SURF ? 256u : 0u;
    SA.crossed = 0u;
    // surface values in LDS: a runtime index into the by-value argument
    // would make the compiler copy it to scratch
    __shared__ double surf_vals[MAX_SURF];
    // cells of the 6 Jacobian samples of each lane (jac_eval)
    __shared__ uint4 jcell[SURF ? 6 : 1][SURF ? 256 : 1];
    __shared__ uint32_t jlw[SURF ? 6 : 1][SURF ? 256 : 1];
    __shared__ uint32_t jcid[SURF ? 6 : 1][SURF ? 256 : 1];
    // the block cache of each of the 4 waves (WaveCache)
    constexpr bool WCE = SURF && !RC && WAVE_CACHE > 0;
    __shared__ F4v wc_data[WCE ? 4 * WAVE_CACHE * ENTRY_F4 : 1];
    __shared__ uint32_t wc_tags[WCE ? 4 * (WAVE_CACHE + 1) : 1];
    WaveCache WCa;
    I.ent = -1;
    if (WCE) {
        const unsigned wv = threadIdx.x / 64u;
        WCa.data = (LdsF4*)(wc_data + wv * WAVE_CACHE * ENTRY_F4);
        WCa.tags = (LdsU32*)(wc_tags + wv * (WAVE_CACHE + 1));
        if (__lane_id() <= (unsigned)WAVE_CACHE) WCa.tags[__lane_id()] = __lane_id() < (unsigned)WAVE_CACHE ? NIL : 0u;
    }
    if (SURF) {
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < MAX_SURF; ++k) surf_vals[k] = A.surf_val[k];
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) jlw[k][threadIdx.x] = 0u;
        __syncthreads();
    }
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
                    // launch record prepared by launch_kernel (integrator.h:669-705)
                    const double* L = A.scratch + (uint64_t)LAUNCH_DOUBLES * my;
#pragma unroll
                    for (int k = 0; k < 9; ++k) R.core[k] = L[k];
                    const bool empty = L[9] >= LAUNCH_EMPTY;
                    const double phi = L[10];
                    cosmo_seg_init(R.cs);
                    R.lambda = R.s = R.isw = R.iswold = R.redshift = R.chi = 0.0;
                    R.a0 = R.core[CA];
                    R.phi0c2 = phi / c2;
                    R.phi = (float)phi;
                    R.dx = (float)L[11]; R.dy = (float)L[12]; R.dz = (float)L[13];
                    R.dt = (float)L[14];
                    R.F[0] = R.dt; R.F[1] = R.dx; R.F[2] = R.dy; R.F[3] = R.dz;
                    R.appended = 0;
                    if (SURF) R.stopv = A.ray_value ? A.ray_value[ray] : A.stop_value;
                    if (SURF) {
                        surf_init(SA, L[15]);
                        R.rho = (float)L[15];
                    }
                    const unsigned l0 = (unsigned)(empty ? L[9] - LAUNCH_EMPTY : L[9]);
                    R.dl = level_extent(l0) * (R.core[CA] * R.core[CA] * ratio_k);
                    I.valid = false;
                    active = !empty;
                    if (!active) {
                        double* o = A.out + (uint64_t)MPF_NOUT * ray;
                        for (int k = 0; k < MPF_NOUT; ++k) o[k] = 0.0;
#pragma unroll
                        for (int k = 0; k < 9; ++k) o[k] = R.core[k];
                        o[MPF_OUT_LEVEL] = (double)l0;
                    }
                } else {
                    drained = true;
                }
            }
        }
        if (!__any(active)) {
            if (__all(drained)) break;
            continue;
        }
        if (!active) continue;
        // ---------------------------------------------------------- step
        const double dl = R.dl;
        double ksum[9], kin[9], kout[9];
        double dadt = cosmo_slope(R.core[CT], A.ct, A.cdadt, A.cslope, A.ncos, R.cs);
        rhs_split(R.core, R.F, dadt, inv_scale, kout);
#pragma unroll
        for (int i = 0; i < 9; ++i) { ksum[i] = kout[i]; kin[i] = R.core[i] + (0.5 * dl) * kout[i]; }
        float F[5];
        constexpr int EB = SURF ? SURF_SCALAR_BLOCKS : SCALAR_BLOCKS;
        bool ok = split_eval<4, RC, EB, 0, WCE>(t, sv, I, kin[CX], kin[CY], kin[CZ], F, &WCa);
        if (!ok) { F[0] = F[1] = F[2] = F[3] = 0.f; }
        dadt = cosmo_slope(kin[CT], A.ct, A.cdadt, A.cslope, A.ncos, R.cs);
        rhs_split(kin, F, dadt, inv_scale, kout);
#pragma unroll
        for (int i = 0; i < 9; ++i) { ksum[i] += 2.0 * kout[i]; kin[i] = R.core[i] + (0.5 * dl) * kout[i]; }
        ok = split_eval<4, RC, EB, 0, WCE>(t, sv, I, kin[CX], kin[CY], kin[CZ], F, &WCa);
        if (!ok) { F[0] = F[1] = F[2] = F[3] = 0.f; }
        dadt = cosmo_slope(kin[CT], A.ct, A.cdadt, A.cslope, A.ncos, R.cs);
        rhs_split(kin, F, dadt, inv_scale, kout);
#pragma unroll
        for (int i = 0; i < 9; ++i) { ksum[i] += 2.0 * kout[i]; kin[i] = R.core[i] + dl * kout[i]; }
        ok = split_eval<4, RC, EB, 0, WCE>(t, sv, I, kin[CX], kin[CY], kin[CZ], F, &WCa);
        if (!ok) { F[0] = F[1] = F[2] = F[3] = 0.f; }
        dadt = cosmo_slope(kin[CT], A.ct, A.cdadt, A.cslope, A.ncos, R.cs);
        rhs_split(kin, F, dadt, inv_scale, kout);
        double p[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) p[i] = R.core[i] + (dl * (1.0 / 6.0)) * (ksum[i] + kout[i]);
        // fields at the new point
        const bool inside = !(p[CA] < 0.0) && p[CX] > -2.0 && p[CX] < 2.0 && p[CY] > -2.0 && p[CY] < 2.0 &&
                            p[CZ] > -2.0 && p[CZ] < 2.0;
        bool cont = inside && split_eval<5, RC, EB, 0, WCE>(t, sv, I, p[CX], p[CY], p[CZ], F, &WCa);
        if (!cont) { F[0] = F[1] = F[2] = F[3] = F[4] = 0.f; }
        // extras (integrator.h:756-776) with FP64 accumulation
        const double nphi = (double)F[4];
        const double ndt = -(double)F[0];
        const double pdx = (double)F[1], pdy = (double)F[2], pdz = (double)F[3];
        const double isw = R.isw + ((double)R.dt + ndt) * (p[CT] - R.core[CT]) * INV_C2;
        const double ddx = p[CX] - R.core[CX], ddy = p[CY] - R.core[CY], ddz = p[CZ] - R.core[CZ];
        const double iswold = R.iswold - (2.0 * (nphi - (double)R.phi) * INV_C2 -
                                          ((((double)R.dx + pdx) * ddx + ((double)R.dy + pdy) * ddy +
                                            ((double)R.dz + pdz) * ddz) * A.length) * INV_C2);
        const double chi = sqrt(p[CX] * p[CX] + p[CY] * p[CY] + p[CZ] * p[CZ]);
        const double lambda = R.lambda + dl;
        const double s = R.s + sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
        const double redshift = R.a0 * recip(p[CA]) * (1.0 + R.phi0c2 - nphi * INV_C2) - 1.0;
        bool provisional = false;
        if (cont) {
            // a photon that the final pop could remove: keep its predecessor as the provisional output
            if (signbit(redshift) || signbit(p[CA])) {
                provisional = true;
                double* o = A.out + (uint64_t)MPF_NOUT * ray;
#pragma unroll
                for (int k = 0; k < 9; ++k) o[k] = R.core[k];
                o[MPF_OUT_REDSHIFT] = R.redshift; o[MPF_OUT_LAMBDA] = R.lambda; o[MPF_OUT_S] = R.s;
                o[MPF_OUT_CHI] = R.chi; o[MPF_OUT_ISW] = R.isw; o[MPF_OUT_ISWOLD] = R.iswold;
                o[MPF_OUT_PHI] = R.phi; o[MPF_OUT_DPHIDX] = R.dx; o[MPF_OUT_DPHIDY] = R.dy; o[MPF_OUT_DPHIDZ] = R.dz;
                o[MPF_OUT_DPHIDT] = R.dt; o[MPF_OUT_NSTEPS] = (double)R.appended;
                o[MPF_OUT_LEVEL] = (double)(I.valid ? I.lvl : 0u);
            }
            (void)provisional;
            if (SURF) {
                // rho at the new point: the rho block of its cell's slot at the
                // offsets split_eval used (FP32), when the rho blocks are built
                float rho_n = 0.f;
                if (sv.rho) {
                    const float u = (float)((p[CX] - I.cx) * I.inv);
                    const float v = (float)((p[CY] - I.cy) * I.inv);
                    const float w = (float)((p[CZ] - I.cz) * I.inv);
                    const float4* rq = reinterpret_cast<const float4*>(sv.rho + (uint64_t)I.slot * FSTRIDE);
                    float rc[FSTRIDE];
#pragma unroll
                    for (int i = 0; i < FSTRIDE / 4; ++i) {
                        const float4 q = rq[i];
                        rc[4 * i] = q.x; rc[4 * i + 1] = q.y; rc[4 * i + 2] = q.z; rc[4 * i + 3] = q.w;
                    }
                    rho_n = horner27(rc, u, v, w);
                }
                SurfElem pv, nw;
                pv.x = R.core[CX]; pv.y = R.core[CY]; pv.z = R.core[CZ]; pv.chi = R.chi; pv.lambda = R.lambda;
                pv.t = R.core[CT]; pv.a = R.core[CA]; pv.redshift = R.redshift; pv.s = R.s; pv.isw = R.isw;
                pv.rho = (double)R.rho; pv.phi = R.phi; pv.kt = R.core[CKT]; pv.kx = R.core[CKX]; pv.ky = R.core[CKY];
                pv.kz = R.core[CKZ];
                nw.x = p[CX]; nw.y = p[CY]; nw.z = p[CZ]; nw.chi = chi; nw.lambda = lambda; nw.t = p[CT];
                nw.a = p[CA]; nw.redshift = redshift; nw.s = s; nw.isw = isw; nw.rho = (double)rho_n; nw.phi = nphi;
                nw.kt = p[CKT]; nw.kx = p[CKX]; nw.ky = p[CKY]; nw.kz = p[CKZ];
                double T[3] = {0.0, 0.0, 0.0};
                if (A.jac) {
                    // 6 samples of grad phi at +-dl/2 (lensing.h:429-436), FP32
                    // interpolation; per axis pair only the components the tidal
                    // tensor uses, differenced at once (+ sample minus - sample)
                    const double jdl = jacobian_step(I.lvl);
                    const double hj = 0.5 * jdl;
                    const unsigned tl = threadIdx.x;
                    double D[6];
                    float gm[3], gp[3];
                    // the cell's tidal block, when it has one: the six
                    // differences at the point's own offset, one block read
                    const uint32_t ti = sv.tidx ? sv.tidx[I.slot] : NIL;
                    if (ti != NIL) {
                        const float u = (float)((p[CX] - I.cx) * I.inv);
                        const float v = (float)((p[CY] - I.cy) * I.inv);
                        const float w = (float)((p[CZ] - I.cz) * I.inv);
                        float mv[9];
                        monomials9(v, w, mv);
                        const float4* tq = reinterpret_cast<const float4*>(sv.tidal + (uint64_t)ti * TIDAL_FLOATS);
#pragma unroll
                        for (int k = 0; k < TIDAL_COMPONENTS; ++k) {
                            float cf[FSTRIDE];
#pragma unroll
                            for (int i = 0; i < FSTRIDE / 4; ++i) {
                                const float4 q = tq[k * (FSTRIDE / 4) + i];
                                cf[4 * i] = q.x; cf[4 * i + 1] = q.y; cf[4 * i + 2] = q.z; cf[4 * i + 3] = q.w;
                            }
                            D[k] = (double)mono27(cf, mv, u);
                        }
                    } else {
                        // x pair: d/dx of dphidx
                        if (!jac_eval<1, JAC_SCALAR_BLOCKS>(t, sv, jcell[0][tl], jcid[0][tl], jlw[0][tl], p[CX] - hj, p[CY], p[CZ], gm)) gm[0] = 0.f;
                        if (!jac_eval<1, JAC_SCALAR_BLOCKS>(t, sv, jcell[1][tl], jcid[1][tl], jlw[1][tl], p[CX] + hj, p[CY], p[CZ], gp)) gp[0] = 0.f;
                        D[0] = (double)gp[0] - (double)gm[0];
                        // y pair: d/dy of dphidx, dphidy
                        if (!jac_eval<2, JAC_SCALAR_BLOCKS>(t, sv, jcell[2][tl], jcid[2][tl], jlw[2][tl], p[CX], p[CY] - hj, p[CZ], gm)) gm[0] = gm[1] = 0.f;
                        if (!jac_eval<2, JAC_SCALAR_BLOCKS>(t, sv, jcell[3][tl], jcid[3][tl], jlw[3][tl], p[CX], p[CY] + hj, p[CZ], gp)) gp[0] = gp[1] = 0.f;
                        D[1] = (double)gp[0] - (double)gm[0];
                        D[3] = (double)gp[1] - (double)gm[1];
                        // z pair: d/dz of dphidx, dphidy, dphidz
                        if (!jac_eval<3, JAC_SCALAR_BLOCKS>(t, sv, jcell[4][tl], jcid[4][tl], jlw[4][tl], p[CX], p[CY], p[CZ] - hj, gm))
                            gm[0] = gm[1] = gm[2] = 0.f;
                        if (!jac_eval<3, JAC_SCALAR_BLOCKS>(t, sv, jcell[5][tl], jcid[5][tl], jlw[5][tl], p[CX], p[CY], p[CZ] + hj, gp))
                            gp[0] = gp[1] = gp[2] = 0.f;
                        D[2] = (double)gp[0] - (double)gm[0];
                        D[4] = (double)gp[1] - (double)gm[1];
                        D[5] = (double)gp[2] - (double)gm[2];
                    }
                    tidal_from_differences(p[CX], p[CY], p[CZ], chi, jdl, D, T);
                }
                const double vp = surf_param(A.surf_kind, R.chi, R.lambda, R.redshift, R.core[CT], R.core[CA], R.s);
                const double vn = surf_param(A.surf_kind, chi, lambda, redshift, p[CT], p[CA], s);
                surf_append(SA, A.surf_kind, A.nsurf, surf_vals, A.ray_value != nullptr, R.stopv,
                            A.surf_out + (uint64_t)A.nsurf * SF_NREC * ray, R.appended + 1, pv, nw, vp, vn, T,
                            A.jac != 0);
                R.rho = rho_n;
            }
#pragma unroll
            for (int i = 0; i < 9; ++i) R.core[i] = p[i];
            R.lambda = lambda; R.s = s; R.isw = isw; R.iswold = iswold; R.redshift = redshift; R.chi = chi;
            R.phi = F[4]; R.dx = F[1]; R.dy = F[2]; R.dz = F[3]; R.dt = -F[0];
            R.F[0] = F[0]; R.F[1] = F[1]; R.F[2] = F[2]; R.F[3] = F[3];
            R.dl = level_extent(I.lvl) * (p[CA] * p[CA] * ratio_k);
            R.appended++;
        }
        bool brk = false;
        const double stopv = SURF ? R.stopv : (A.ray_value ? A.ray_value[ray] : A.stop_value);
        switch (A.stop_kind) {
            case 0: brk = R.chi > stopv; break;
            case 1: brk = R.lambda > stopv; break;
            case 2: brk = R.redshift > stopv; break;
            case 3: brk = R.core[CT] > stopv; break;
            case 4: brk = R.core[CA] < stopv; break;
            case 5: brk = (p[CX] * A.kit[0] + p[CY] * A.kit[1] + p[CZ] * A.kit[2]) > stopv; break;
            case 6: brk = R.s > stopv; break;
            default: brk = false; break;
        }
        if (!cont || brk) {
            active = false;
            double* o = A.out + (uint64_t)MPF_NOUT * ray;
            uint32_t steps = R.appended;
            if (signbit(R.redshift) || signbit(R.core[CA])) {
                // pop: the provisional output already holds the previous element
                steps = steps ? steps - 1 : 0;
                if (R.appended == 0) o[MPF_OUT_NSTEPS] = -1.0;
            } else {
                // final element with the remaining outputs of integrator.h:756-776
                // (rho and a of the final point are added by final_kernel)
#pragma unroll
                for (int k = 0; k < 9; ++k) o[k] = R.core[k];
                const double phi = R.phi;
                const double kk = R.core[CKX] * R.core[CKX] + R.core[CKY] * R.core[CKY] + R.core[CKZ] * R.core[CKZ];
                o[MPF_OUT_REDSHIFT] = R.redshift; o[MPF_OUT_LAMBDA] = R.lambda; o[MPF_OUT_S] = R.s;
                o[MPF_OUT_CHI] = R.chi; o[MPF_OUT_ISW] = R.isw; o[MPF_OUT_ISWOLD] = R.iswold;
                o[MPF_OUT_PHI] = phi; o[MPF_OUT_DPHIDX] = R.dx; o[MPF_OUT_DPHIDY] = R.dy; o[MPF_OUT_DPHIDZ] = R.dz;
                o[MPF_OUT_DPHIDT] = R.dt; 
                o[MPF_OUT_DPHIDL] = (double)R.F[0] * R.core[CKT] + ((double)R.dx * R.core[CKX] +
                                                                    (double)R.dy * R.core[CKY] +
                                                                    (double)R.dz * R.core[CKZ]);
                o[MPF_OUT_LAPLACIAN] = 0.0;
                o[MPF_OUT_DSDL2] = (R.core[CA] * R.core[CA]) *
                                   ((-c2 * (1.0 + 2.0 * (phi / c2)) * R.core[CKT] * R.core[CKT]) +
                                    ((1.0 - 2.0 * (phi / c2)) * kk));
                o[MPF_OUT_ERROR] = 1.0 - ((1.0 - 2.0 * (phi / c2)) * kk) /
                                             (c2 * (1.0 + 2.0 * (phi / c2)) * R.core[CKT] * R.core[CKT]);
                o[MPF_OUT_LEVEL] = (double)I.lvl;
                o[MPF_OUT_NSTEPS] = (double)steps;
            }
            if (SURF) {
                // surfaces never reached, or whose next element was popped, are rejected (hmaps.h:716);
                // rho and rhomax come from the rho blocks (NaN without them)
                const bool pop = signbit(R.redshift) || signbit(R.core[CA]);
                const uint32_t last = (pop && R.appended) ? R.appended - 1 : R.appended;
                double* rec = A.surf_out + (uint64_t)A.nsurf * SF_NREC * ray;
                for (int j = 0; j < A.nsurf; ++j) {
                    double* so = rec + (uint64_t)SF_NREC * j;
                    if (!(SA.crossed & (1u << j))) {
                        for (int k = 0; k < SF_NREC; ++k) so[k] = 0.0;
                        so[SF_FIRSTID] = (double)last;
                    } else {
                        if (so[SF_FIRSTID] >= (double)last) {
                            so[SF_VALID] = 0.0;
                            so[SF_CROSSED] = 0.0;
                        }
                        if (!sv.rho) {   // no rho blocks: the densities are not carried
                            so[SF_RHO] = __builtin_nan("");
                            so[SF_RHOMAX] = __builtin_nan("");
                        }
                    }
                }
            }
            atomicAdd(A.steps, (unsigned long long)steps);
        }
    }
}

// Features of catalogue sources from the FP32 stencil blocks: one block
// read per cell the line of sight crosses.
struct SplitSampler {
    const DevTree& t;
    const SplitView& sv;
    Interp<false> I;
    __device__ SplitSampler(const DevTree& tree, const SplitView& view) : t(tree), sv(view) { restart(); }
    __device__ void restart() {
        I.valid = false;
        I.slot = NIL;
    }
    __device__ void operator()(double x, double y, double z, double* g) {
        float o[NFIELD];   // o[0], dphi/dt, is not evaluated
        if (!split_eval<NFIELD, false, 0, 1>(t, sv, I, x, y, z, o)) {
            g[0] = g[1] = g[2] = g[3] = 0.0;
            return;
        }
        g[0] = o[4]; g[1] = o[1]; g[2] = o[2]; g[3] = o[3];
    }
};

// order (may be null): the source each thread takes (by direction)
__global__ void __launch_bounds__(256) features_split_kernel(DevTree t, SplitView sv, const double* __restrict__ src, uint64_t n,
                                      const uint32_t* __restrict__ order, double length, float* __restrict__ feat) {
    // (vector loads: the samples of a wave's lanes sit at the same fraction
    // of different source distances, so they spread over more blocks than a
    // per-wave block cache holds; with it the features took 41 ms instead of
    // 16 ms for 1 M sources)
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint64_t i = order ? (uint64_t)order[j] : j;
    SplitSampler sample(t, sv);
    source_features(src + 6 * i, length, sample, feat + (uint64_t)MPF_MLP_NIN * i);
}

}  // namespace

bool split_build(SplitTree& s, const DevTree& t, const uint8_t* level, const uint32_t* kx, const uint32_t* ky,
                 const uint32_t* kz, int order, hipStream_t st, std::string& err) {
    split_release(s);
    if (order != 2) {
        err = "the split-precision tier implements TSC (order 2)";
        return false;
    }
    const uint64_t n = t.n;
    uint8_t* d_lev = nullptr;
    uint32_t *d_kx = nullptr, *d_ky = nullptr, *d_kz = nullptr, *d_target = nullptr, *d_slot_cell = nullptr;
    auto cleanup = [&]() {
        if (d_lev) (void)hipFree(d_lev);
        if (d_kx) (void)hipFree(d_kx);
        if (d_ky) (void)hipFree(d_ky);
        if (d_kz) (void)hipFree(d_kz);
        if (d_target) (void)hipFree(d_target);
        if (d_slot_cell) (void)hipFree(d_slot_cell);
    };
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess) {
            err = std::string("split tier preparation: ") + what + ": " + hipGetErrorString(e);
            return false;
        }
        return true;
    };
    if (!chk(hipMalloc((void**)&d_lev, n), "alloc") || !chk(hipMalloc((void**)&d_kx, 4 * n), "alloc") ||
        !chk(hipMalloc((void**)&d_ky, 4 * n), "alloc") || !chk(hipMalloc((void**)&d_kz, 4 * n), "alloc") ||
        !chk(hipMalloc((void**)&d_target, 4 * n), "alloc")) {
        cleanup();
        return false;
    }
    bool ok = chk(hipMemcpyAsync(d_lev, level, n, hipMemcpyHostToDevice, st), "copy") &&
              chk(hipMemcpyAsync(d_kx, kx, 4 * n, hipMemcpyHostToDevice, st), "copy") &&
              chk(hipMemcpyAsync(d_ky, ky, 4 * n, hipMemcpyHostToDevice, st), "copy") &&
              chk(hipMemcpyAsync(d_kz, kz, 4 * n, hipMemcpyHostToDevice, st), "copy");
    if (ok) {
        hipLaunchKernelGGL(target_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, t, d_lev, d_kx, d_ky,
                           d_kz, d_target);
        ok = chk(hipGetLastError(), "target kernel");
    }
    std::vector<uint32_t> target(n);
    if (ok) ok = chk(hipMemcpyAsync(target.data(), d_target, 4 * n, hipMemcpyDeviceToHost, st), "copy") &&
                 chk(hipStreamSynchronize(st), "sync");
    if (!ok) {
        cleanup();
        return false;
    }
    // slots for the stencil-centre cells, in cell order
    std::vector<uint32_t> slot_of(n, NIL);
    for (uint64_t g = 0; g < n; ++g)
        if (target[g] != NIL) slot_of[target[g]] = 0;
    std::vector<uint32_t> slot_cell;
    slot_cell.reserve(n);
    for (uint64_t g = 0; g < n; ++g)
        if (slot_of[g] != NIL) {
            slot_of[g] = (uint32_t)slot_cell.size();
            slot_cell.push_back((uint32_t)g);
        }
    const uint64_t ns = slot_cell.size();
    if (ns >= SLOT_MASK) {
        err = "split tier: more than 2^31 - 1 stencil blocks";
        cleanup();
        return false;
    }
    std::vector<uint32_t> tslot(n);
    for (uint64_t g = 0; g < n; ++g) tslot[g] = (target[g] == NIL) ? NIL : slot_of[target[g]];
    s.nslots = ns;
    s.bytes = ((n * 4 + 255) / 256) * 256 + std::max<uint64_t>(ns, 1) * BLOCK_BYTES;
    if (!chk(hipMalloc(&s.mem, s.bytes), "alloc blocks") ||
        !chk(hipMalloc((void**)&d_slot_cell, 4 * std::max<uint64_t>(ns, 1)), "alloc")) {
        cleanup();
        split_release(s);
        return false;
    }
    s.tslot = reinterpret_cast<uint32_t*>(s.mem);
    s.blocks = reinterpret_cast<unsigned char*>(s.mem) + ((n * 4 + 255) / 256) * 256;
    ok = chk(hipMemcpyAsync(s.tslot, tslot.data(), 4 * n, hipMemcpyHostToDevice, st), "copy") &&
         chk(hipMemcpyAsync(d_slot_cell, slot_cell.data(), 4 * std::max<uint64_t>(ns, 1), hipMemcpyHostToDevice, st),
             "copy");
    if (ok && n) {
        hipLaunchKernelGGL(partial_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, t, s.tslot);
        ok = chk(hipGetLastError(), "partial-oct flags");
    }
    if (ok && n) {
        // the face-neighbour table is an accelerator (about 2 % of a launch):
        // it is built only when NBR_RESERVE bytes stay free after it for the
        // rays, surfaces and maps; otherwise every lookup of a new cell starts
        // from the root
        size_t free_b = 0, total_b = 0;
        const bool room = hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
                          (uint64_t)free_b >= 6 * sizeof(uint2) * n + NBR_RESERVE;
        if (room && hipMalloc((void**)&s.nbr, 6 * sizeof(uint2) * n) == hipSuccess) {
            s.nbr_bytes = 6 * sizeof(uint2) * n;
            hipLaunchKernelGGL(neighbour_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, t, d_lev, d_kx,
                               d_ky, d_kz, s.tslot, s.nbr);
            ok = chk(hipGetLastError(), "neighbour table");
        } else {
            (void)hipGetLastError();
            s.nbr = nullptr;
        }
    }
    if (ok && ns) {
        hipLaunchKernelGGL(block_kernel, dim3((unsigned)((ns + 127) / 128)), dim3(128), 0, st, t, d_lev, d_kx, d_ky,
                           d_kz, d_slot_cell, ns, s.blocks);
        ok = chk(hipGetLastError(), "block kernel");
    }
    if (ok) ok = chk(hipStreamSynchronize(st), "sync");
    cleanup();
    if (!ok) split_release(s);
    else s.slot_cell = std::move(slot_cell);
    return ok;
}

bool split_build_rho(SplitTree& s, const DevTree& t, const uint8_t* level, const uint32_t* kx, const uint32_t* ky,
                     const uint32_t* kz, hipStream_t st, std::string& err) {
    if (s.rho) return true;
    const uint64_t n = t.n, ns = s.nslots;
    if (ns == 0 || s.slot_cell.size() != ns) {
        err = "split tier: rho blocks need the built stencil blocks";
        return false;
    }
    // the run that needs them still allocates its launch records and ray
    // order after this: leave RHO_RESERVE bytes free
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess ||
        (uint64_t)free_b < 4ull * FSTRIDE * ns + 21ull * n + RHO_RESERVE) {
        err = "split tier: no room for the rho blocks";
        return false;
    }
    uint8_t* d_lev = nullptr;
    uint32_t *d_kx = nullptr, *d_ky = nullptr, *d_kz = nullptr, *d_sc = nullptr;
    float* rho = nullptr;
    bool ok = hipMalloc((void**)&rho, 4 * FSTRIDE * ns) == hipSuccess && hipMalloc((void**)&d_lev, n) == hipSuccess &&
              hipMalloc((void**)&d_kx, 4 * n) == hipSuccess && hipMalloc((void**)&d_ky, 4 * n) == hipSuccess &&
              hipMalloc((void**)&d_kz, 4 * n) == hipSuccess && hipMalloc((void**)&d_sc, 4 * ns) == hipSuccess;
    if (!ok) err = "split tier: rho blocks: allocation failed";
    ok = ok && hipMemcpyAsync(d_lev, level, n, hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(d_kx, kx, 4 * n, hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(d_ky, ky, 4 * n, hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(d_kz, kz, 4 * n, hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(d_sc, s.slot_cell.data(), 4 * ns, hipMemcpyHostToDevice, st) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(rho_block_kernel, dim3((unsigned)((ns + 127) / 128)), dim3(128), 0, st, t, d_lev, d_kx,
                           d_ky, d_kz, d_sc, ns, rho);
        ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
        if (!ok) err = "split tier: rho blocks: kernel failed";
    }
    for (void* q : {(void*)d_lev, (void*)d_kx, (void*)d_ky, (void*)d_kz, (void*)d_sc})
        if (q) (void)hipFree(q);
    if (!ok) {
        if (rho) (void)hipFree(rho);
        return false;
    }
    s.rho = rho;
    return true;
}

bool split_build_tidal(SplitTree& s, const DevTree& t, const uint8_t* level, const uint32_t* kx, const uint32_t* ky,
                       const uint32_t* kz, hipStream_t st, std::string& err) {
    if (s.tidal) return true;
    const uint64_t n = t.n, ns = s.nslots;
    if (ns == 0 || s.slot_cell.size() != ns) {
        err = "split tier: tidal blocks need the built stencil blocks";
        return false;
    }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || (uint64_t)free_b < 13ull * n + 8ull * ns + TIDAL_RESERVE) {
        err = "split tier: no room for the tidal blocks";
        return false;
    }
    uint8_t* d_lev = nullptr;
    uint32_t *d_kx = nullptr, *d_ky = nullptr, *d_kz = nullptr, *d_sc = nullptr, *d_flag = nullptr, *tidx = nullptr;
    float* tidal = nullptr;
    auto cleanup = [&]() {
        for (void* q : {(void*)d_lev, (void*)d_kx, (void*)d_ky, (void*)d_kz, (void*)d_sc, (void*)d_flag})
            if (q) (void)hipFree(q);
    };
    bool ok = hipMalloc((void**)&d_lev, n) == hipSuccess && hipMalloc((void**)&d_kx, 4 * n) == hipSuccess &&
              hipMalloc((void**)&d_ky, 4 * n) == hipSuccess && hipMalloc((void**)&d_kz, 4 * n) == hipSuccess &&
              hipMalloc((void**)&d_sc, 4 * ns) == hipSuccess && hipMalloc((void**)&d_flag, 4 * ns) == hipSuccess;
    if (!ok) err = "split tier: tidal blocks: allocation failed";
    ok = ok && hipMemcpyAsync(d_lev, level, n, hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(d_kx, kx, 4 * n, hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(d_ky, ky, 4 * n, hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(d_kz, kz, 4 * n, hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(d_sc, s.slot_cell.data(), 4 * ns, hipMemcpyHostToDevice, st) == hipSuccess;
    std::vector<uint32_t> flag(ns);
    if (ok) {
        hipLaunchKernelGGL(tidal_flag_kernel, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, st, t, d_lev, d_kx,
                           d_ky, d_kz, s.tslot, d_sc, ns, d_flag);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(flag.data(), d_flag, 4 * ns, hipMemcpyDeviceToHost, st) == hipSuccess &&
             hipStreamSynchronize(st) == hipSuccess;
        if (!ok) err = "split tier: tidal blocks: flag kernel failed";
    }
    // compact indices of the qualifying slots; when the memory does not hold
    // all their blocks, the first ones in slot order (coarse levels first:
    // a coarse cell is crossed by more rays, and a crossing takes nsteps
    // steps whatever the cell's size)
    uint64_t cap = 0;
    if (ok && hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
        (uint64_t)free_b > 4ull * ns + TIDAL_RESERVE)
        cap = ((uint64_t)free_b - 4ull * ns - TIDAL_RESERVE) / (4ull * TIDAL_FLOATS);
    uint64_t m = 0;
    for (uint64_t q = 0; ok && q < ns; ++q) flag[q] = (flag[q] && m < cap) ? (uint32_t)m++ : NIL;
    if (ok && m == 0) {
        err = "split tier: no room for the tidal blocks";
        ok = false;
    }
    ok = ok && hipMalloc((void**)&tidx, 4 * ns) == hipSuccess &&
         hipMalloc((void**)&tidal, 4ull * TIDAL_FLOATS * std::max<uint64_t>(m, 1)) == hipSuccess;
    if (ok) {
        ok = hipMemcpyAsync(tidx, flag.data(), 4 * ns, hipMemcpyHostToDevice, st) == hipSuccess;
        hipLaunchKernelGGL(tidal_block_kernel, dim3((unsigned)((ns + 127) / 128)), dim3(128), 0, st, t, d_lev, d_kx,
                           d_ky, d_kz, d_sc, tidx, ns, tidal);
        ok = ok && hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
        if (!ok) err = "split tier: tidal blocks: kernel failed";
    }
    cleanup();
    if (!ok) {
        if (tidx) (void)hipFree(tidx);
        if (tidal) (void)hipFree(tidal);
        return false;
    }
    s.tidx = tidx;
    s.tidal = tidal;
    s.ntidal = m;
    return true;
}

bool split_supports_surfaces() { return true; }

int split_launch_doubles() { return LAUNCH_DOUBLES; }

void split_release(SplitTree& s) {
    if (s.mem) (void)hipFree(s.mem);
    if (s.nbr) (void)hipFree(s.nbr);
    if (s.rho) (void)hipFree(s.rho);
    if (s.tidx) (void)hipFree(s.tidx);
    if (s.tidal) (void)hipFree(s.tidal);
    s = SplitTree();
}

hipError_t launch_integrate_split(const IntegrateArgs& A, const DevTree& t, const SplitTree& s, int cus,
                                  unsigned blocks_per_cu, unsigned waves_per_simd, hipStream_t st,
                                  hipEvent_t ev0, hipEvent_t ev1) {
    SplitView sv{s.tslot, s.blocks, s.nbr, s.rho, s.tidx, s.tidal};
    const uint64_t blocks = (A.n + 255) / 256;
    const int grid = (int)std::min<uint64_t>(blocks, (uint64_t)cus * blocks_per_cu);
    const unsigned pgrid = (unsigned)((A.n + 255) / 256);
    hipLaunchKernelGGL(launch_kernel, dim3(pgrid), dim3(256), 0, st, A, t);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (ev0 && (e = hipEventRecord(ev0, st)) != hipSuccess) return e;
    if (A.nsurf) {
        if (waves_per_simd == 1)
            hipLaunchKernelGGL((integrate_split_kernel<false, 1, true>), dim3(grid), dim3(256), 0, st, A, t, sv);
        else
            hipLaunchKernelGGL((integrate_split_kernel<false, 2, true>), dim3(grid), dim3(256), 0, st, A, t, sv);
    } else {
        switch (waves_per_simd) {
            case 1: hipLaunchKernelGGL((integrate_split_kernel<true, 1, false>), dim3(grid), dim3(256), 0, st, A, t, sv); break;
            case 3: hipLaunchKernelGGL((integrate_split_kernel<false, 3, false>), dim3(grid), dim3(256), 0, st, A, t, sv); break;
            case 4: hipLaunchKernelGGL((integrate_split_kernel<false, 4, false>), dim3(grid), dim3(256), 0, st, A, t, sv); break;
            default: hipLaunchKernelGGL((integrate_split_kernel<false, 2, false>), dim3(grid), dim3(256), 0, st, A, t, sv); break;
        }
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (ev1 && (e = hipEventRecord(ev1, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(final_kernel, dim3(pgrid), dim3(256), 0, st, A, t);
    return hipGetLastError();
}

}  // namespace mpf

namespace mpf {
hipError_t launch_features_split(const DevTree& t, const SplitTree& s, const double* src, uint64_t n,
                                 const uint32_t* order, double length,
                                 float* feat, hipStream_t st) {
    SplitView sv{s.tslot, s.blocks, s.nbr, s.rho, s.tidx, s.tidal};
    if (n)
        hipLaunchKernelGGL(features_split_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, t, sv, src, n,
                           order, length, feat);
    return hipGetLastError();
}
}  // namespace mpf
