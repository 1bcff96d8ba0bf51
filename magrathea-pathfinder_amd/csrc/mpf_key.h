// 128-bit octree keys of SimpleHyperOctreeIndex<__uint128_t, 3> on the host:
// 32 nibbles from the most significant end, nibble i = 8 | z<<2 | y<<1 | x
// for the bit of level i+1, zero nibbles below the cell's level
// (simplehyperoctreeindex.h).  Records store the key little-endian in their
// first 16 bytes.
#pragma once
#include <cstdint>
#include <cstring>

namespace mpf {
namespace key {

using u128 = unsigned __int128;

// marker bit of every nibble (comb(4, 3), simplehyperoctreeindex.h:1544-1547)
constexpr u128 MARK = (((u128)0x8888888888888888ULL) << 64) | (u128)0x8888888888888888ULL;
// Index::invalid(): the low 127 bits set (simplehyperoctreeindex.h:1358-1361)
constexpr u128 INVALID = (~(u128)0) >> 1;

inline u128 load(const unsigned char* r) {
    uint64_t lo, hi;
    std::memcpy(&lo, r, 8);
    std::memcpy(&hi, r + 8, 8);
    return (((u128)hi) << 64) | lo;
}

inline void store(unsigned char* r, u128 k) {
    const uint64_t lo = (uint64_t)k, hi = (uint64_t)(k >> 64);
    std::memcpy(r, &lo, 8);
    std::memcpy(r + 8, &hi, 8);
}

// level(): number of leading nibbles, from the trailing marker bit
// (simplehyperoctreeindex.h:572-575)
inline unsigned level(u128 k) {
    const u128 m = k & MARK;
    unsigned tz = 128;
    if (m) {
        const uint64_t lo = (uint64_t)m, hi = (uint64_t)(m >> 64);
        tz = lo ? (unsigned)__builtin_ctzll(lo) : 64u + (unsigned)__builtin_ctzll(hi);
    }
    return (128u - tz) / 4u + (k != 0);
}

// a well-formed key of level l: the l leading nibbles carry their marker bit
// and every bit below them is zero
inline bool wellformed(u128 k, unsigned l) {
    if (l > 32) return false;
    const u128 below = (l >= 32) ? (u128)0 : ((((u128)1) << (128 - 4 * l)) - 1);
    if (k & below) return false;
    for (unsigned i = 0; i < l; ++i)
        if (!((k >> (127 - 4 * i)) & 1)) return false;
    return true;
}

// per-axis integer coordinates at level l (coordinate(), :838-841)
inline void coords(u128 k, unsigned l, uint32_t& x, uint32_t& y, uint32_t& z) {
    x = y = z = 0;
    for (unsigned i = 0; i < l; ++i) {
        const unsigned nib = (unsigned)((k >> (124 - 4 * i)) & 0xF);
        x = (x << 1) | (nib & 1u);
        y = (y << 1) | ((nib >> 1) & 1u);
        z = (z << 1) | ((nib >> 2) & 1u);
    }
}

// make(l, x, y, z) (:1143-1146): the low l bits of each coordinate, so a
// coordinate one step outside the domain wraps round it
inline u128 make(unsigned l, uint64_t x, uint64_t y, uint64_t z) {
    u128 k = 0;
    for (unsigned i = 0; i < l && i < 32; ++i) {
        const unsigned b = l - 1 - i;
        const u128 nib = 8u | (unsigned)(((z >> b) & 1u) << 2) | (unsigned)(((y >> b) & 1u) << 1) |
                         (unsigned)((x >> b) & 1u);
        k |= nib << (124 - 4 * i);
    }
    return k;
}

// a.containing(b) (:998-1001): the coordinate bits that differ lie below the
// marker bits that differ, and a <= b
inline bool containing(u128 a, u128 b) {
    const u128 d = a ^ b;
    return ((~MARK) & d) <= (MARK & d) && a <= b;
}

// parent() (:679-682): the last nibble cleared
inline u128 parent(u128 k) {
    const unsigned l = level(k);
    if (l == 0) return k;
    const unsigned keep = 4 * (l - 1);
    return keep ? (k & ~((((u128)1) << (128 - keep)) - 1)) : (u128)0;
}

// child(site) (:690-694), unchanged at level 32
inline u128 child(u128 k, unsigned site) {
    const unsigned l = level(k);
    if (l >= 32) return k;
    return k | (((u128)((site & 7u) | 8u)) << (124 - 4 * l));
}

// center(idim) (:890-893) with EXTENT 4: -2 + 4 (2c + 1) / 2^(l+1)
inline double center(u128 k, int axis) {
    const unsigned l = level(k);
    uint32_t c[3];
    coords(k, l, c[0], c[1], c[2]);
    return (-4.0 / 2.0) + (4.0 * ((double)(2ull * c[axis] + 1ull) / (double)(((u128)1) << (l + 1))));
}

}  // namespace key
}  // namespace mpf
