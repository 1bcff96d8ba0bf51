#include "mpf_tree.h"

#include "mpf_key.h"

#include <algorithm>
#include <cstring>
#include <thread>

namespace mpf {

namespace {
using key::u128;
inline u128 load_key(const unsigned char* r) { return key::load(r); }
inline unsigned key_level(u128 k) { return key::level(k); }
inline bool key_wellformed(u128 k, unsigned l) { return key::wellformed(k, l); }
inline void key_coords(u128 k, unsigned l, uint32_t& x, uint32_t& y, uint32_t& z) { key::coords(k, l, x, y, z); }
}  // namespace

bool build_tree(const void* rec48, uint64_t n, HostTree& t, std::string& err, uint64_t max_base_entries) {
    const unsigned char* rec = static_cast<const unsigned char*>(rec48);
    t = HostTree();
    if (n == 0) {
        err = "empty octree";
        return false;
    }
    if (n >= 0xFFFFFFFFull) {
        err = "octree larger than 2^32-1 cells";
        return false;
    }
    t.n = n;
    std::vector<uint8_t> lev(n);
    std::vector<uint32_t> cx(n), cy(n), cz(n);
    // keys: sorted, unique, well formed
    {
        const unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
        std::vector<std::string> errs(nt);
        std::vector<std::thread> th;
        const uint64_t chunk = (n + nt - 1) / nt;
        for (unsigned w = 0; w < nt; ++w) {
            th.emplace_back([&, w]() {
                const uint64_t b = std::min(n, chunk * w), e = std::min(n, chunk * (w + 1));
                for (uint64_t i = b; i < e; ++i) {
                    const u128 k = load_key(rec + 48 * i);
                    const unsigned l = key_level(k);
                    if (!key_wellformed(k, l)) {
                        errs[w] = "malformed key at record " + std::to_string(i);
                        return;
                    }
                    if (i > 0 && !(load_key(rec + 48 * (i - 1)) < k)) {
                        errs[w] = "records not strictly sorted by key at record " + std::to_string(i);
                        return;
                    }
                    lev[i] = (uint8_t)l;
                    key_coords(k, l, cx[i], cy[i], cz[i]);
                }
            });
        }
        for (auto& x : th) x.join();
        for (auto& e : errs)
            if (!e.empty()) {
                err = e;
                return false;
            }
    }
    uint32_t lmin = 64, lmax = 0;
    std::vector<uint64_t> count(34, 0);
    for (uint64_t i = 0; i < n; ++i) {
        lmin = std::min<uint32_t>(lmin, lev[i]);
        lmax = std::max<uint32_t>(lmax, lev[i]);
        count[lev[i]]++;
    }
    t.min_level = lmin;
    t.max_level = lmax;
    t.level_offset.assign(lmax + 2, 0);
    for (uint32_t l = 0; l <= lmax; ++l) t.level_offset[l + 1] = t.level_offset[l] + count[l];
    // stable counting sort by level: key order inside a level = Morton order
    t.elem.resize(n);
    t.where.resize(n);
    {
        std::vector<uint64_t> pos(t.level_offset.begin(), t.level_offset.end() - 1);
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t g = pos[lev[i]]++;
            t.elem[g] = (uint32_t)i;
            t.where[i] = (uint32_t)g;
        }
    }
    t.kx.resize(n);
    t.ky.resize(n);
    t.kz.resize(n);
    t.level.resize(n);
    for (uint64_t g = 0; g < n; ++g) {
        const uint32_t i = t.elem[g];
        t.kx[g] = cx[i];
        t.ky[g] = cy[i];
        t.kz[g] = cz[i];
        t.level[g] = lev[i];
    }
    // children: the children of consecutive parents are consecutive, so one
    // merge pass per level links them
    t.child_first.assign(n, 0xFFFFFFFFu);
    t.child_mask.assign(n, 0);
    for (uint32_t l = lmin + 1; l <= lmax; ++l) {
        uint64_t p = t.level_offset[l - 1];
        const uint64_t pend = t.level_offset[l];
        for (uint64_t g = t.level_offset[l]; g < t.level_offset[l + 1]; ++g) {
            const uint32_t px = t.kx[g] >> 1, py = t.ky[g] >> 1, pz = t.kz[g] >> 1;
            const u128 pk = load_key(rec + 48ull * t.elem[g]) & ~((((u128)1) << (128 - 4 * (l - 1))) - 1);
            while (p < pend && load_key(rec + 48ull * t.elem[p]) < pk) ++p;
            if (p >= pend || t.kx[p] != px || t.ky[p] != py || t.kz[p] != pz) {
                err = "octree is not parent-closed: the level " + std::to_string(l) + " cell of record " +
                      std::to_string(t.elem[g]) + " has no stored parent";
                return false;
            }
            const unsigned site = (t.kx[g] & 1u) | ((t.ky[g] & 1u) << 1) | ((t.kz[g] & 1u) << 2);
            if (t.child_first[p] == 0xFFFFFFFFu) t.child_first[p] = (uint32_t)g;
            t.child_mask[p] |= (uint8_t)(1u << site);
        }
    }
    // dense base table over the bounding box of level lmin
    {
        uint32_t x0 = 0xFFFFFFFFu, y0 = x0, z0 = x0, x1 = 0, y1 = 0, z1 = 0;
        for (uint64_t g = t.level_offset[lmin]; g < t.level_offset[lmin + 1]; ++g) {
            x0 = std::min(x0, t.kx[g]); x1 = std::max(x1, t.kx[g]);
            y0 = std::min(y0, t.ky[g]); y1 = std::max(y1, t.ky[g]);
            z0 = std::min(z0, t.kz[g]); z1 = std::max(z1, t.kz[g]);
        }
        t.bx0 = x0; t.by0 = y0; t.bz0 = z0;
        t.bnx = x1 - x0 + 1; t.bny = y1 - y0 + 1; t.bnz = z1 - z0 + 1;
        const uint64_t entries = (uint64_t)t.bnx * t.bny * t.bnz;
        if (entries > max_base_entries) {
            err = "base level " + std::to_string(lmin) + " bounding box has " + std::to_string(entries) +
                  " entries, above the table limit";
            return false;
        }
        t.base.assign(entries, 0xFFFFFFFFu);
        for (uint64_t g = t.level_offset[lmin]; g < t.level_offset[lmin + 1]; ++g) {
            const uint64_t idx = ((uint64_t)(t.kz[g] - z0) * t.bny + (t.ky[g] - y0)) * t.bnx + (t.kx[g] - x0);
            t.base[idx] = (uint32_t)g;
        }
    }
    return true;
}

void unpack_fields32(const void* rec48, const HostTree& t, std::vector<float>& soa7) {
    const unsigned char* rec = static_cast<const unsigned char*>(rec48);
    const uint64_t n = t.n;
    soa7.resize(7 * n);
    // record: dphidt@16 a@20 dphidx@24 dphidy@28 dphidz@32 phi@36 rho@40
    static const int off[7] = {40, 36, 24, 28, 32, 20, 16};
    for (uint64_t g = 0; g < n; ++g) {
        const unsigned char* r = rec + 48ull * t.elem[g];
        for (int f = 0; f < 7; ++f) std::memcpy(&soa7[f * n + g], r + off[f], 4);
    }
}

}  // namespace mpf
