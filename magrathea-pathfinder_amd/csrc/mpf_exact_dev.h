// Exact NGP / CIC / TSC interpolation (reference arithmetic, operation by
// operation).  Included by the kernels that need the reference semantics;
// those translation units are compiled with -ffp-contract=off.
//
// NF = 7 interpolates the cell fields of gravity.h; NF = 10 adds the cell
// velocity of the velocity-field build (gravity2.h:45-90), read from the
// tree's velocity arrays (zero where the tree carries none).
#pragma once
#include "mpf_device.h"

namespace mpf {
namespace exact {

enum { F_VX = 7, F_VY = 8, F_VZ = 9 };

template <int P, int NF>
__device__ __forceinline__ void loadf(const DevTree& t, uint32_t g, double* d) {
    if (P == 1 && t.f64) {
#pragma unroll
        for (int k = 0; k < 7; ++k) d[k] = t.f64[(uint64_t)k * t.n + g];
    } else {
#pragma unroll
        for (int k = 0; k < 7; ++k) d[k] = (double)t.f32[(uint64_t)k * t.n + g];
    }
    if (NF == 10) {
#pragma unroll
        for (int k = 0; k < 3; ++k) d[7 + k] = t.vel ? (double)t.vel[(uint64_t)k * t.n + g] : 0.0;
    }
}

// mac() (simplehyperoctree.h:1870-1910): accumulator += factor * coefficient
template <int P, int NF>
__device__ __forceinline__ void macf(double* acc, const double* f, double w) {
#pragma unroll
    for (int k = 0; k < NF; ++k) {
        if (P == 0) acc[k] = (double)(float)(acc[k] + f[k] * w);
        else acc[k] = acc[k] + f[k] * w;
    }
}

template <int NF>
__device__ __forceinline__ void zerof(double* d) {
#pragma unroll
    for (int k = 0; k < NF; ++k) d[k] = 0.0;
}

template <int NF>
__device__ __forceinline__ bool emptyf(const double* d) {
    bool e = true;
#pragma unroll
    for (int k = 0; k < NF; ++k) e = e && (d[k] == 0.0);
    return e;
}

__device__ __forceinline__ void zero7(double* d) { zerof<7>(d); }
__device__ __forceinline__ bool empty7(const double* d) { return emptyf<7>(d); }

// ngp() (simplehyperoctree.h:1396-1401); *loc = level of the located cell or -1
template <int P, int NF = 7>
__device__ void ngp(const DevTree& t, double x, double y, double z, double* r, int* loc) {
    unsigned fl = 0;
    const uint32_t g = locate_point(t, x, y, z, &fl);
    *loc = (g == NIL) ? -1 : (int)fl;
    if (g == NIL) zerof<NF>(r);
    else loadf<P, NF>(t, g, r);
}

// cic() (simplehyperoctree.h:1412-1469)
template <int P, int NF = 7>
__device__ void cic(const DevTree& t, double x, double y, double z, double* result, int* loc) {
    const double one = 1.0, two = 2.0;
    const double point[3] = {x, y, z};
    // elem key: cipher(point) at first, compute(lvl+1, point) after a level drop
    unsigned ekl = 32;
    uint32_t ek[3] = {axis_coord(x, 32), axis_coord(y, 32), axis_coord(z, 32)};
    unsigned fl = 0;
    uint32_t g0 = find_cell(t, 32, ek[0], ek[1], ek[2], &fl);
    *loc = (g0 == NIL) ? -1 : (int)fl;
    bool invalid = (g0 == NIL);
    unsigned idl = invalid ? 32u : fl;  // level of the invalid index is 32
    uint32_t idk[3] = {ek[0] >> (32 - idl), ek[1] >> (32 - idl), ek[2] >> (32 - idl)};
    unsigned ilvl = idl, lvl = idl;
    double half = EXTENT / 1.0 * (1.0 / (1.0 * (double)(1ull << idl))) / two;
    unsigned site = 0;
    double cell[NF];
    zerof<NF>(result);
    do {
        ilvl = lvl;
        if (!invalid) {
            double dist[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) dist[d] = fabs((cell_center(idl, idk[d]) - point[d]) / (half + half));
            // child of idx that contains the elem key
            {
                const unsigned b = ekl - idl - 1;
                site = ((ek[0] >> b) & 1u) | (((ek[1] >> b) & 1u) << 1) | (((ek[2] >> b) & 1u) << 2);
            }
            for (unsigned iv = 0; iv < 8; ++iv) {
                double weight = one;
                const unsigned diff = site ^ iv;
                double vertex[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const int bb = (iv >> d) & 1u, df = (diff >> d) & 1u;
                    vertex[d] = point[d] + (bb * half) - ((!bb) * half);
                    weight *= ((df) * (one - dist[d])) + ((!df) * dist[d]);
                }
                unsigned vl = 0;
                const uint32_t gv = find_cell(t, ilvl, axis_coord(vertex[0], ilvl), axis_coord(vertex[1], ilvl),
                                              axis_coord(vertex[2], ilvl), &vl);
                if (gv != NIL) {
                    lvl = lvl < vl ? lvl : vl;
                    loadf<P, NF>(t, gv, cell);
                    macf<P, NF>(result, cell, weight);
                } else {
                    zerof<NF>(result);
                    break;
                }
            }
        }
        if (lvl < ilvl) {
            // elem = compute(lvl+1, point); idx = parent(elem)
            ekl = lvl + 1;
            ek[0] = axis_coord(x, ekl); ek[1] = axis_coord(y, ekl); ek[2] = axis_coord(z, ekl);
            idl = lvl;
            idk[0] = ek[0] >> 1; idk[1] = ek[1] >> 1; idk[2] = ek[2] >> 1;
            invalid = false;
            zerof<NF>(result);
            half = EXTENT / 1.0 * (1.0 / (1.0 * (double)(1ull << idl))) / two;
        }
    } while (lvl < ilvl);
}

// The 27-element cache of tsc(elemsTsc, ...) (simplehyperoctree.h:1554-1644):
// the central key, the per-axis coordinates of the cached cells and their
// cell indices.  `dflt` marks the default-constructed cache (27 root keys).
struct TscCache {
    unsigned lvl;
    uint32_t cx[3], cy[3], cz[3];
    uint32_t g[27];
    bool dflt;
};

__device__ __forceinline__ void cache_init(TscCache& c) {
    c.dflt = true;
    c.lvl = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) { c.cx[i] = 0; c.cy[i] = 0; c.cz[i] = 0; }
}

__device__ __forceinline__ double tsc_axis_w(int ai, double d) {
    return ai * 0.5 * (1.5 - d) * (1.5 - d) + (1 - ai) * (0.75 - d * d);
}

// tsc() with the cache (simplehyperoctree.h:1554-1644), or without it
// (UseCache = false, simplehyperoctree.h:1481-1541).
template <int P, bool UseCache, int NF = 7>
__device__ void tsc(const DevTree& t, TscCache& cache, double x, double y, double z, double* result, int* loc) {
    const double two = 2.0;
    const double point[3] = {x, y, z};
    if (point[0] == 0.0 && point[1] == 0.0 && point[2] == 0.0) {
        cic<P, NF>(t, 0., 0., 0., result, loc);
        return;
    }
    uint32_t ek[3] = {axis_coord(x, 32), axis_coord(y, 32), axis_coord(z, 32)};
    unsigned fl = 0;
    const uint32_t g0 = find_cell(t, 32, ek[0], ek[1], ek[2], &fl);
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
        bool hit = false;
        if (UseCache) {
            if (cache.dflt) hit = !invalid && idl == 0;  // 27 default elements hold the root key
            else hit = !invalid && cache.lvl == idl && cache.cx[1] == idk[0] && cache.cy[1] == idk[1] &&
                       cache.cz[1] == idk[2];
        }
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
            for (int ic = 0; ic < 27; ++ic) {
                const int ix = ic / 9, iy = (ic / 3) % 3, iz = ic % 3;
                const double w = wx[ix] * wy[iy] * wz[iz];
                loadf<P, NF>(t, cache.g[ic], cell);
                macf<P, NF>(result, cell, w);
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
            uint32_t gl[27];
            int ic = 0;
            for (; ic < 27; ++ic) {
                const int ix = ic / 9, iy = (ic / 3) % 3, iz = ic % 3;
                unsigned vl = 0;
                const uint32_t gv = find_cell(t, ilvl, nx[ix], ny[iy], nz[iz], &vl);
                if (gv == NIL) {
                    zerof<NF>(result);
                    break;
                }
                lvl = lvl < vl ? lvl : vl;
                if (lvl < ilvl) {
                    zerof<NF>(result);
                    break;
                }
                const double w = wx[ix] * wy[iy] * wz[iz];
                loadf<P, NF>(t, gv, cell);
                macf<P, NF>(result, cell, w);
                if (UseCache) gl[ic] = gv;
            }
            if (UseCache && ic == 27) {
                cache.dflt = false;
                cache.lvl = ilvl;
#pragma unroll
                for (int i = 0; i < 3; ++i) { cache.cx[i] = nx[i]; cache.cy[i] = ny[i]; cache.cz[i] = nz[i]; }
#pragma unroll
                for (int i = 0; i < 27; ++i) cache.g[i] = gl[i];
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

}  // namespace exact
}  // namespace mpf
