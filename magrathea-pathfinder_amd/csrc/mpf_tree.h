// Host-side re-layout of a key-sorted cone-file octree into level-sorted
// arrays for the device (level-major; Morton order inside a level, which is
// the key order restricted to one level).
//
// Reference layout: one flat std::vector of (128-bit key, Gravity<float>)
// sorted by key, searched with std::upper_bound + containing()
// (simplehyperoctree.h:894-899).  Here every cell gets:
//   elem[g]        its position in the key-sorted input (the reference index),
//   child_first[g] level-sorted index of its first stored child,
//   child_mask[g]  bit s set when child s (nibble 8|s) is stored,
// plus a dense table of the coarsest stored level B over its bounding box.
// A lookup descends from the table through child masks; on a parent-closed
// tree this returns exactly the element of upper_bound(key) - 1 when that
// element contains the key, and "not containing" otherwise.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mpf {

struct HostTree {
    uint64_t n = 0;
    uint32_t min_level = 0, max_level = 0;
    std::vector<uint64_t> level_offset;  // size max_level + 2, cells of level l at [off[l], off[l+1])
    std::vector<uint32_t> elem;          // level-sorted -> record index
    std::vector<uint32_t> child_first;   // level-sorted index of the first stored child
    std::vector<uint8_t> child_mask;
    std::vector<uint32_t> kx, ky, kz;    // cell coordinates at the cell's level
    std::vector<uint8_t> level;
    // dense table of level min_level over its bounding box: level-sorted index or UINT32_MAX
    uint32_t bx0 = 0, by0 = 0, bz0 = 0, bnx = 0, bny = 0, bnz = 0;
    std::vector<uint32_t> base;
    // record-order permutation inverse: record index -> level-sorted index
    std::vector<uint32_t> where;
};

// Builds the tree from n 48-byte records.  Returns false with a message on
// unsorted, duplicate or malformed keys, on a tree that is not parent-closed,
// or when the base table would exceed max_base_entries.
bool build_tree(const void* rec48, uint64_t n, HostTree& t, std::string& err,
                uint64_t max_base_entries = (1ull << 31));

// Field order (gravity.h:45-87): rho, phi, dphidx, dphidy, dphidz, a, dphidt.
// Unpacks the record fields into level-sorted structure-of-arrays.
void unpack_fields32(const void* rec48, const HostTree& t, std::vector<float>& soa7);

}  // namespace mpf
