// Octree preparation (host).  Every arithmetic expression keeps the
// reference's operation order and types (float where the reference reads
// or stores floats), and the file is built without multiply-add fusion.
#include "mpf_prepare.h"

#include "../../include/mpf.h"

#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <random>
#include <sstream>
#include <thread>

namespace mpf {
namespace prep {

using key::u128;

namespace {

constexpr unsigned LEVEL_SHIFT = 2;  // log2(EXTENT), input.h:1003

inline bool key_less(const Elem& a, const Elem& b) { return a.k < b.k; }

inline bool same(const Grav& a, const Grav& b) {
    return a.rho == b.rho && a.phi == b.phi && a.dx == b.dx && a.dy == b.dy && a.dz == b.dz && a.a == b.a &&
           a.dt == b.dt;
}

inline bool empty(const Grav& g) { return same(g, Grav{0, 0, 0, 0, 0, 0, 0}); }

// mac (simplehyperoctree.h:1870-1927): every member += member * weight,
// with the product and the sum in double
inline void mac(Grav& r, const Grav& f, double w) {
    r.rho = (float)((double)r.rho + (double)f.rho * w);
    r.phi = (float)((double)r.phi + (double)f.phi * w);
    r.dx = (float)((double)r.dx + (double)f.dx * w);
    r.dy = (float)((double)r.dy + (double)f.dy * w);
    r.dz = (float)((double)r.dz + (double)f.dz * w);
    r.a = (float)((double)r.a + (double)f.a * w);
    r.dt = (float)((double)r.dt + (double)f.dt * w);
}

// conversion of a non-negative position to the key's integer (negative and
// NaN values give 0, as the 128-bit conversion of the reference's compiler)
inline u128 to_u128(double v) { return (v >= 0.0) ? (u128)v : (u128)0; }
inline u128 to_u128(float v) { return (v >= 0.0f) ? (u128)v : (u128)0; }

// idx.extent() with the index's default unit extent: 1 / 2^level
inline double unit_extent(u128 k) { return 1.0 / (1.0 * (double)(((u128)1) << key::level(k))); }

}  // namespace

u128 compute_f(unsigned l, float x, float y, float z) {
    if (l == 0) return 0;
    const float s = (1.0f / 4.0f) * (float)(((u128)1) << l);
    const u128 m = (l >= 64) ? ~(u128)0 : ((((u128)1) << l) - 1);
    const u128 kx = to_u128(s * (x - (-2.0f))) & m, ky = to_u128(s * (y - (-2.0f))) & m,
               kz = to_u128(s * (z - (-2.0f))) & m;
    return key::make(l, (uint64_t)kx, (uint64_t)ky, (uint64_t)kz);
}

u128 compute_d(unsigned l, double x, double y, double z) {
    if (l == 0) return 0;
    const double s = (1.0 / 4.0) * (double)(((u128)1) << l);
    const u128 m = (l >= 64) ? ~(u128)0 : ((((u128)1) << l) - 1);
    const u128 kx = to_u128(s * (x - (-2.0))) & m, ky = to_u128(s * (y - (-2.0))) & m,
               kz = to_u128(s * (z - (-2.0))) & m;
    return key::make(l, (uint64_t)kx, (uint64_t)ky, (uint64_t)kz);
}

bool collide(u128 k, double micro_radius, const cone::Cone& c) {
    // the cell's circumscribed sphere (Input::collide, input.h:216-226)
    const double p[3] = {key::center(k, 0), key::center(k, 1), key::center(k, 2)};
    const double ext = 4.0 / (1.0 * (double)(((u128)1) << key::level(k)));
    const double r = ext * std::sqrt(3.0) / 2.0;
    // sphere against sphere: distance<3> sums from the last axis (utility.h:230-233)
    const double d2 = p[2] * p[2] + 0.0;
    const double d1 = p[1] * p[1] + d2;
    if (std::sqrt(p[0] * p[0] + d1) < r + micro_radius) return true;
    // sphere against cone (utility.h:409-419)
    double nn = 0.0;
    for (int i = 0; i < 3; ++i) nn += (c.base[i] - c.vertex[i]) * (c.base[i] - c.vertex[i]);
    const double norm = std::sqrt(nn);
    double length = 0.0, distance = 0.0;
    for (int i = 0; i < 3; ++i) length += (c.base[i] - c.vertex[i]) * (p[i] - c.vertex[i]);
    length /= norm;
    for (int i = 0; i < 3; ++i) {
        const double q = p[i] - (c.vertex[i] + (c.base[i] - c.vertex[i]) * (length / norm));
        distance += q * q;
    }
    const double ins = micro_radius;
    const double g = length - ins * std::cos(c.angle);
    return (std::sqrt(distance) <
            5 * r + length * std::tan(c.angle) + 0.4 * ins * std::exp(-0.7 * g * g / (ins * ins))) &&
           !(length < 0.0) && (length < norm + r * 2.0);
}

// ----------------------------------------------------------------- octree

uint64_t Builder::upper(u128 k) const {
    return (uint64_t)std::distance(v_.begin(), std::upper_bound(v_.begin(), v_.end(), Elem{k, Grav{}}, key_less));
}

uint64_t Builder::find(u128 k) const {
    if (v_.empty()) return 0;
    const uint64_t m = upper(k);
    const uint64_t c = m - (m > 0);
    return key::containing(v_[c].k, k) ? c : (uint64_t)v_.size();
}

bool Builder::leaf(uint64_t i) const {
    return (i + 1 < v_.size()) ? !key::containing(v_[i].k, v_[i + 1].k) : true;
}

void Builder::coarsen(uint64_t i) {
    const u128 idx = v_[i].k;
    for (uint64_t j = i + 1; j < v_.size() && key::containing(idx, v_[j].k); ++j) v_[j].k = key::INVALID;
}

// update (simplehyperoctree.h:1272-1307, MEMORY_SAVING): drop invalidated
// cells, sort by key, and for each run of equal keys keep the data of the
// smallest a, with dphidt = (phi_first - phi_last) / (a_first - a_last) over
// the run sorted by a when the a differ by more than 1e-6.  Runs are the
// cells whose key equals the next one's; the reference's scan also treats
// the last cell as a run with the element past the end of the container.
void Builder::update() {
    v_.erase(std::remove_if(v_.begin(), v_.end(), [](const Elem& e) { return e.k == key::INVALID; }), v_.end());
    std::sort(v_.begin(), v_.end(), key_less);
    const uint64_t n = v_.size();
    uint64_t i = 0;
    std::vector<Grav> run;
    while (i < n) {
        uint64_t j = i;
        while (j + 1 < n && v_[j + 1].k == v_[i].k) ++j;
        if (j > i) {
            run.clear();
            for (uint64_t q = i; q <= j; ++q) run.push_back(v_[q].d);
            std::sort(run.begin(), run.end(), [](const Grav& x, const Grav& y) { return x.a < y.a; });
            v_[i].d = run.front();
            if (std::abs(run.front().a - run.back().a) > 1e-6)
                v_[i].d.dt = (run.front().phi - run.back().phi) / (run.front().a - run.back().a);
        }
        i = j + 1;
    }
    v_.erase(std::unique(v_.begin(), v_.end(), [](const Elem& x, const Elem& y) { return x.k == y.k; }), v_.end());
}

Grav Builder::ngp(double x, double y, double z) const {
    const u128 ek = compute_d(32, x, y, z);
    const uint64_t m = upper(ek);
    if (v_.empty()) return Grav{};
    return key::containing(v_[m - (m > 0)].k, ek) ? v_[m - 1].d : Grav{};
}

// SimpleHyperOctree::cic (simplehyperoctree.h:1412-1469)
Grav Builder::cic(double x, double y, double z) const {
    if (v_.empty()) return Grav{};
    const double P[3] = {x, y, z};
    u128 ek = compute_d(32, x, y, z);
    uint64_t m = upper(ek);
    u128 idx = key::containing(v_[m - (m > 0)].k, ek) ? v_[m - 1].k : key::INVALID;
    unsigned ilvl = key::level(idx), lvl = ilvl;
    double half = 4.0 / 1.0 * unit_extent(idx) / 2.0;
    double dist[3];
    Grav result{};
    do {
        ilvl = lvl;
        if (idx != key::INVALID) {
            for (int d = 0; d < 3; ++d) dist[d] = std::abs((key::center(idx, d) - P[d]) / (half + half));
            unsigned site = 0;
            for (unsigned s = 0; s < 8; ++s)
                if (key::containing(key::child(idx, s), ek)) {
                    site = s;
                    break;
                }
            for (unsigned ivertex = 0; ivertex < 8; ++ivertex) {
                double weight = 1.0, vertex[3];
                const unsigned diff = site ^ ivertex;
                for (int d = 0; d < 3; ++d) {
                    const bool b = (ivertex >> d) & 1u, df = (diff >> d) & 1u;
                    vertex[d] = P[d] + ((double)b * half) - ((double)(!b) * half);
                    weight *= ((double)df * (1.0 - dist[d])) + ((double)(!df) * dist[d]);
                }
                ek = compute_d(ilvl, vertex[0], vertex[1], vertex[2]);
                m = upper(ek);
                if (key::containing(v_[m - (m > 0)].k, ek)) {
                    lvl = std::min(lvl, key::level(v_[m - 1].k));
                    mac(result, v_[m - 1].d, weight);
                } else {
                    result = Grav{};
                    break;
                }
            }
        }
        if (lvl < ilvl) {
            ek = compute_d(lvl + 1, x, y, z);
            idx = key::parent(ek);
            result = Grav{};
            half = 4.0 / 1.0 * unit_extent(idx) / 2.0;
        }
    } while (lvl < ilvl);
    return result;
}

// SimpleHyperOctree::tsc, uncached (simplehyperoctree.h:1481-1541)
Grav Builder::tsc(double x, double y, double z) const {
    if (v_.empty()) return Grav{};
    const double P[3] = {x, y, z};
    u128 ek = compute_d(32, x, y, z);
    uint64_t m = upper(ek);
    u128 idx = key::containing(v_[m - (m > 0)].k, ek) ? v_[m - 1].k : key::INVALID;
    unsigned ilvl = key::level(idx), lvl = ilvl;
    double half = 4.0 / 1.0 * unit_extent(idx) / 2.0;
    Grav result{};
    if (x == 0.0 && y == 0.0 && z == 0.0) return cic(0., 0., 0.);
    do {
        ilvl = lvl;
        if (idx != key::INVALID) {
            const double twohalves = 2.0 * half;
            bool stop = false;
            for (int ix = -1; ix <= 1 && !stop; ++ix) {
                const int aix = std::abs(ix);
                for (int iy = -1; iy <= 1 && !stop; ++iy) {
                    const int aiy = std::abs(iy);
                    for (int iz = -1; iz <= 1 && !stop; ++iz) {
                        const int aiz = std::abs(iz);
                        ek = compute_d(ilvl, x + ix * twohalves, y + iy * twohalves, z + iz * twohalves);
                        m = upper(ek);
                        if (key::containing(v_[m - (m > 0)].k, ek)) {
                            double dist[3];
                            for (int d = 0; d < 3; ++d) dist[d] = std::abs((key::center(ek, d) - P[d]) / twohalves);
                            const double weight =
                                (aix * 0.5 * (1.5 - dist[0]) * (1.5 - dist[0]) + (1 - aix) * (0.75 - dist[0] * dist[0])) *
                                (aiy * 0.5 * (1.5 - dist[1]) * (1.5 - dist[1]) + (1 - aiy) * (0.75 - dist[1] * dist[1])) *
                                (aiz * 0.5 * (1.5 - dist[2]) * (1.5 - dist[2]) + (1 - aiz) * (0.75 - dist[2] * dist[2]));
                            lvl = std::min(lvl, key::level(v_[m - 1].k));
                            if (lvl < ilvl) {
                                result = Grav{};
                                stop = true;
                            } else {
                                mac(result, v_[m - 1].d, weight);
                            }
                        } else {
                            result = Grav{};
                            stop = true;
                        }
                    }
                }
            }
        }
        if (lvl < ilvl) {
            ek = compute_d(lvl + 1, x, y, z);
            idx = key::parent(ek);
            result = Grav{};
            half = 4.0 / 1.0 * unit_extent(idx) / 2.0;
        }
    } while (lvl < ilvl);
    return result;
}

// ------------------------------------------------------------------ input

namespace {

// a token the stream's extraction of an unsigned or a float would read in
// full: optional sign, digits with one optional point, optional exponent
bool plain_number(const char* b, const char* e, bool integral) {
    const char* p = b;
    if (p < e && (*p == '+' || *p == '-')) ++p;
    bool digits = false, point = false;
    for (; p < e && ((*p >= '0' && *p <= '9') || (!integral && *p == '.' && !point)); ++p) {
        if (*p == '.') point = true;
        else digits = true;
    }
    if (!digits) return false;
    if (p < e && !integral && (*p == 'e' || *p == 'E')) {
        ++p;
        if (p < e && (*p == '+' || *p == '-')) ++p;
        const char* q = p;
        while (p < e && *p >= '0' && *p <= '9') ++p;
        if (p == q) return false;
    }
    return p == e && !(integral && *b == '-');
}

// one line: exactly 10 plain tokens; the values strtoul / strtof give them
// (correctly rounded, as the stream's extraction)
template <class Store>
bool parse_line(const char* b, const char* e, unsigned row, Store& store) {
    const char* tok[10][2];
    int n = 0;
    const char* p = b;
    while (p < e) {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
        if (p >= e) break;
        const char* q = p;
        while (q < e && !(*q == ' ' || *q == '\t' || *q == '\r' || *q == '\v' || *q == '\f')) ++q;
        if (n == 10) return false;
        tok[n][0] = p;
        tok[n][1] = q;
        ++n;
        p = q;
    }
    if (n != 10) return false;
    char tmp[64];
    unsigned lv = 0;
    float v[9];
    for (int k = 0; k < 10; ++k) {
        const size_t len = (size_t)(tok[k][1] - tok[k][0]);
        if (len >= sizeof(tmp) || !plain_number(tok[k][0], tok[k][1], k == 0)) return false;
        std::memcpy(tmp, tok[k][0], len);
        tmp[len] = 0;
        if (k == 0) {
            const unsigned long x = std::strtoul(tmp, nullptr, 10);
            if (x > 0xFFFFFFFFul) return false;
            lv = (unsigned)x;
        } else {
            v[k - 1] = std::strtof(tmp, nullptr);
        }
    }
    store(row, lv, v);
    return true;
}

// f(i) for i in [0, n) over up to 16 threads
template <class F>
void parallel_rows(unsigned n, F&& f) {
    const unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096 || nthreads == 1) {
        for (unsigned i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nthreads; ++w)
        pool.emplace_back([&, w]() {
            const unsigned lo = (unsigned)((uint64_t)n * w / nthreads), hi = (unsigned)((uint64_t)n * (w + 1) / nthreads);
            for (unsigned i = lo; i < hi; ++i) f(i);
        });
    for (std::thread& t : pool) t.join();
}

// every line of the buffer parsed by parse_line, in threads over line
// ranges; false when a line is not one well-formed row
template <class Store>
bool parse_rows_parallel(const std::vector<char>& buf, unsigned rows, Store& store) {
    if (rows == 0) return true;
    std::vector<size_t> starts;
    starts.reserve(rows + 1);
    starts.push_back(0);
    for (size_t i = 0; i < buf.size(); ++i)
        if (buf[i] == '\n') starts.push_back(i + 1);
    // a trailing line without a newline is not a row
    const unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<int> ok(nthreads, 1);
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nthreads; ++w)
        pool.emplace_back([&, w]() {
            const unsigned lo = (unsigned)((uint64_t)rows * w / nthreads), hi = (unsigned)((uint64_t)rows * (w + 1) / nthreads);
            for (unsigned r = lo; r < hi && ok[w]; ++r)
                if (!parse_line(buf.data() + starts[r], buf.data() + starts[r + 1] - 1, r, store)) ok[w] = 0;
        });
    for (std::thread& t : pool) t.join();
    return std::all_of(ok.begin(), ok.end(), [](int x) { return x != 0; });
}

}  // namespace

bool Builder::import_ascii(const std::string& path, const cone::Cone& conic, const double* rotm1, uint64_t& added,
                           std::string& err) {
    added = 0;
    std::vector<char> buf;
    {
        FILE* f = std::fopen(path.c_str(), "rb");
        if (!f) {
            err = "cannot open " + path;
            return false;
        }
        char chunk[1 << 16];
        size_t got;
        while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
        std::fclose(f);
    }
    // as many rows as newline characters, 10 values each (input.h:1008-1030)
    const unsigned size = (unsigned)std::count(buf.begin(), buf.end(), '\n');
    std::vector<unsigned> level(size);
    std::vector<float> center(3ull * size), force(3ull * size), a(size), phi(size), rho(size);
    auto store = [&](unsigned i, unsigned lv, const float* v) {
        level[i] = lv;
        center[3 * i] = v[0]; center[3 * i + 1] = v[1]; center[3 * i + 2] = v[2];
        rho[i] = v[3]; phi[i] = v[4];
        force[3 * i] = v[5]; force[3 * i + 1] = v[6]; force[3 * i + 2] = v[7];
        a[i] = v[8];
    };
    if (!parse_rows_parallel(buf, size, store)) {
        // not one well-formed row per line: the stream extraction itself
        std::fill(level.begin(), level.end(), 0u);
        for (std::vector<float>* v : {&center, &force, &a, &phi, &rho}) std::fill(v->begin(), v->end(), 0.0f);
        std::istringstream stream(std::string(buf.begin(), buf.end()));
        for (unsigned i = 0; i < size; ++i)
            stream >> level[i] >> center[3 * i] >> center[3 * i + 1] >> center[3 * i + 2] >> rho[i] >> phi[i] >>
                force[3 * i] >> force[3 * i + 1] >> force[3 * i + 2] >> a[i];
    }
    if (!cfg_.isfullsky) {
        if (!rotm1) {
            err = "a narrow cone needs the rotation rotm1";
            return false;
        }
        for (unsigned i = 0; i < size; ++i) {
            const float c0 = center[3 * i], c1 = center[3 * i + 1], c2 = center[3 * i + 2];
            const float f0 = force[3 * i], f1 = force[3 * i + 1], f2 = force[3 * i + 2];
            for (int r = 0; r < 3; ++r) {
                center[3 * i + r] = (float)(c0 * rotm1[3 * r] + c1 * rotm1[3 * r + 1] + c2 * rotm1[3 * r + 2]);
                force[3 * i + r] = (float)(f0 * rotm1[3 * r] + f1 * rotm1[3 * r + 1] + f2 * rotm1[3 * r + 2]);
            }
        }
    }
    // keys and the filter per row in threads, then the selected rows
    // appended in file order
    const double micro = 1.0 / cfg_.microcoeff;
    std::vector<u128> keys(size);
    std::vector<uint8_t> keep(size);
    parallel_rows(size, [&](unsigned i) {
        keys[i] = compute_f(level[i] + LEVEL_SHIFT, center[3 * i], center[3 * i + 1], center[3 * i + 2]);
        keep[i] = collide(keys[i], micro, conic) ? 1 : 0;
    });
    for (unsigned i = 0; i < size; ++i) {
        if (!keep[i]) continue;
        v_.push_back(Elem{keys[i], Grav{rho[i], phi[i], force[3 * i], force[3 * i + 1], force[3 * i + 2], a[i], 0.0f}});
        ++added;
    }
    if (cfg_.coarseonly) {
        const unsigned coarse = cfg_.ncoarse + LEVEL_SHIFT;
        v_.erase(std::remove_if(v_.begin(), v_.end(), [=](const Elem& e) { return key::level(e.k) != coarse; }),
                 v_.end());
    }
    return true;
}

namespace {

// Input::sistemize (input.h:382-394): RAMSES units to SI with the cell's a
void sistemize(Grav& g, double h, double lboxmpch, double mpc) {
    const double a = (double)g.a;
    const double t = (double)(a * a * mpc) / (double)(h * 100 * 1000);
    const double l = (double)(a * lboxmpch * mpc) / (double)(h);
    g.phi = (float)(g.phi * ((l * l) / (t * t)));
    g.dx = (float)(g.dx * ((l / (t * t)) * (-a)));
    g.dy = (float)(g.dy * ((l / (t * t)) * (-a)));
    g.dz = (float)(g.dz * ((l / (t * t)) * (-a)));
}

// Input::mean (input.h:244-277) of rho over the normal same-level neighbours
// of a coarse cell
float mean_rho(const Builder& b, u128 k) {
    const unsigned l = key::level(k);
    uint32_t cx, cy, cz;
    key::coords(k, l, cx, cy, cz);
    const std::vector<Elem>& v = b.cells();
    float result = 0.0f;
    unsigned count = 0;
    for (int ix = -1; ix <= 1; ++ix)
        for (int iy = -1; iy <= 1; ++iy)
            for (int iz = -1; iz <= 1; ++iz) {
                if (ix == 0 && iy == 0 && iz == 0) continue;
                const uint64_t n = b.find(key::make(l, (uint64_t)cx + ix, (uint64_t)cy + iy, (uint64_t)cz + iz));
                if (n < v.size() && std::isnormal(v[n].d.rho)) {
                    result += v[n].d.rho;
                    ++count;
                }
            }
    return result / (float)(count + (count == 0));
}

// Utility::interpolate2 (utility.h:461-468) with a float abscissa: the
// search compares in float, the interpolation runs in double
float interpolate2(float x0, const std::vector<double>& x, const std::vector<double>& y) {
    const long long n = (long long)std::min(x.size(), y.size());
    long long i = std::distance(x.begin(), std::upper_bound(x.begin(), x.begin() + n, x0, [](const float x1, const float x2) {
                                    return x1 > x2;
                                }));
    i += (n > 1) * ((i <= 0) - (i >= n));
    return (n > 1) ? (float)(y[i - 1] + ((y[i] - y[i - 1]) * (x0 - x[i - 1]) / (x[i] - x[i - 1])))
                   : ((n > 0) ? (float)y[0] : 0.0f);
}

}  // namespace

bool Builder::correct(const std::vector<double>& tcol, const std::vector<double>& acol,
                      const std::vector<double>& dadt, double& amin, std::string& err) {
    (void)tcol;
    if (cfg_.acorrection) {
        err = "acorrection (the scale-factor snapping of the Full Universe Run) is not supported";
        return false;
    }
    if (cfg_.order < 0 || cfg_.order > 2) {
        err = "order must be 0 (NGP), 1 (CIC) or 2 (TSC)";
        return false;
    }
    // Miscellaneous::correctOctree (miscellaneous.h:467-479)
    for (Elem& e : v_) sistemize(e.d, cfg_.h, cfg_.lboxmpch, cfg_.mpc);
    for (Elem& e : v_) e.d.dt = 0.0f;
    // Input::correct (input.h:1382-1520)
    update();
    uint64_t size = v_.size();
    unsigned ncoarse = 0, nmax = 0;
    if (!v_.empty()) {
        ncoarse = 64;
        for (const Elem& e : v_) {
            const unsigned l = key::level(e.k);
            ncoarse = std::min(ncoarse, l);
            nmax = std::max(nmax, l);
        }
    }
    std::vector<unsigned> count(size, 0);
    if (cfg_.coarsecorrection) {
        // non-normal coarse densities: mean of the normal neighbours, cells in
        // order (the reference runs this loop in parallel while it writes)
        for (uint64_t i = 0; i < size; ++i)
            if (key::level(v_[i].k) == ncoarse && !std::isnormal(v_[i].d.rho)) v_[i].d.rho = mean_rho(*this, v_[i].k);
    }
    if (cfg_.correction) {
        // non-normal dphidt, then rho, from the parent, level by level (a
        // cell whose parent is not stored keeps its value; the reference
        // reads past the container there)
        for (unsigned n = ncoarse + 1; n <= nmax; ++n)
            for (uint64_t i = 0; i < size; ++i)
                if (key::level(v_[i].k) == n && !std::isnormal(v_[i].d.dt)) {
                    const uint64_t p = find(key::parent(v_[i].k));
                    if (p < v_.size()) v_[i].d.dt = v_[p].d.dt;
                }
        for (unsigned n = ncoarse + 1; n <= nmax; ++n)
            for (uint64_t i = 0; i < size; ++i)
                if (key::level(v_[i].k) == n && !std::isnormal(v_[i].d.rho)) {
                    const uint64_t p = find(key::parent(v_[i].k));
                    if (p < v_.size()) v_[i].d.rho = v_[p].d.rho;
                }
    } else {
        // coarsen the parents of refined cells with a non-normal rho
        std::vector<uint64_t> parents;
        for (uint64_t i = 0; i < size; ++i)
            if (key::level(v_[i].k) > ncoarse && !std::isnormal(v_[i].d.rho)) parents.push_back(find(key::parent(v_[i].k)));
        std::sort(parents.begin(), parents.end());
        parents.erase(std::unique(parents.begin(), parents.end()), parents.end());
        for (uint64_t p : parents)
            if (p < v_.size() && v_[p].k != key::INVALID) coarsen(p);
        update();
        size = v_.size();
        count.assign(size, 0);
    }
    std::fill(count.begin(), count.end(), 0u);
    // cells that store some but not all of their children (a child counts
    // as missing when find lands outside the cell)
    for (uint64_t i = 0; i < size; ++i)
        if (!leaf(i))
            for (unsigned j = 0; j < 8; ++j) count[i] += (find(key::child(v_[i].k, j)) == v_.size());
    if (cfg_.correction) {
        // append all 8 children, interpolated at their centres on the
        // container as it grows (the data of the cell when that is empty)
        for (uint64_t i = 0; i < size; ++i) {
            if (!count[i]) continue;
            const u128 idx = v_[i].k;
            const Grav dat = v_[i].d;
            for (unsigned s = 0; s < 8; ++s) {
                const u128 c = key::child(idx, s);
                const double px = key::center(c, 0), py = key::center(c, 1), pz = key::center(c, 2);
                const Grav g = (cfg_.order == 0) ? ngp(px, py, pz) : (cfg_.order == 1) ? cic(px, py, pz) : tsc(px, py, pz);
                v_.push_back(Elem{c, empty(g) ? dat : g});
            }
        }
    } else {
        for (uint64_t i = 0; i < size; ++i)
            if (count[i] && v_[i].k != key::INVALID) coarsen(i);
    }
    update();
    amin = 0.0;
    // dphi/da to dphi/dt
    for (Elem& e : v_) e.d.dt *= interpolate2(e.d.a, acol, dadt);
    return true;
}

// ---------------------------------------------------------------- records

// record layout: key (16 B little-endian), then the Gravity<float> members
// as libstdc++ lays out the tuple (dphidt, a, dphidx, dphidy, dphidz, phi,
// rho), then 4 B of padding
void Builder::to_records(unsigned char* r) const {
    for (const Elem& e : v_) {
        key::store(r, e.k);
        const float f[8] = {e.d.dt, e.d.a, e.d.dx, e.d.dy, e.d.dz, e.d.phi, e.d.rho, 0.0f};
        std::memcpy(r + 16, f, 32);
        r += 48;
    }
}

void Builder::from_records(const unsigned char* r, uint64_t n) {
    v_.clear();
    v_.reserve(std::max<uint64_t>(n, cfg_.allocation));
    for (uint64_t i = 0; i < n; ++i, r += 48) {
        float f[7];
        std::memcpy(f, r + 16, 28);
        v_.push_back(Elem{key::load(r), Grav{f[6], f[5], f[2], f[3], f[4], f[1], f[0]}});
    }
}

bool Builder::save(const std::string& path, std::string& err) const {
    struct stat st;
    if (stat(path.c_str(), &st) == 0) {
        err = path + " exists";
        return false;
    }
    std::vector<unsigned char> buf(48 * v_.size());
    to_records(buf.data());
    std::ofstream out(path.c_str(), std::ios::binary);
    if (!out || !out.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)buf.size())) {
        err = "cannot write " + path;
        return false;
    }
    return true;
}

// -------------------------------------------------------------- the driver

namespace {

bool dat_files(const std::string& dir, std::vector<std::string>& out, std::string& err) {
    DIR* d = opendir(dir.c_str());
    if (!d) {
        err = "cannot open the directory " + dir;
        return false;
    }
    while (struct dirent* e = readdir(d)) {
        const std::string name = e->d_name;
        if (name.find(".dat") != std::string::npos) out.push_back(dir + name);
    }
    closedir(d);
    return true;
}

}  // namespace

bool prepare_ascii(const Config& cfg, const std::string& celldir, const std::string& minicone,
                   const std::vector<cone::Cone>& cones_rot, const double* rotm1, unsigned ntasks, unsigned rank,
                   uint64_t seed, const std::vector<double>& t, const std::vector<double>& a,
                   const std::vector<double>& dadt, const std::vector<std::string>& conefiles, double& amin,
                   std::string& err) {
    if (ntasks == 0 || rank >= ntasks) {
        err = "rank must be below ntasks";
        return false;
    }
    if (conefiles.size() < cones_rot.size()) {
        err = "one output file per cone is needed";
        return false;
    }
    std::vector<std::string> files, mini;
    if (!dat_files(celldir, files, err)) return false;
    if (!cfg.isfullsky && !dat_files(minicone, mini, err)) return false;
    std::mt19937 engine(seed > 0 ? (std::mt19937::result_type)(seed + rank) : std::random_device()());
    Builder b(cfg);
    for (size_t icone = 0; icone < cones_rot.size(); ++icone) {
        if (icone % ntasks != rank) continue;
        b.clear();
        std::shuffle(files.begin(), files.end(), engine);
        uint64_t added = 0;
        for (const std::string& f : files)
            if (!b.import_ascii(f, cones_rot[icone], rotm1, added, err)) return false;
        if (!cfg.isfullsky) {
            std::shuffle(mini.begin(), mini.end(), engine);
            for (const std::string& f : mini)
                if (!b.import_ascii(f, cones_rot[icone], rotm1, added, err)) return false;
        }
        if (!b.correct(t, a, dadt, amin, err)) return false;
        if (!b.save(conefiles[icone], err)) return false;
    }
    return true;
}

}  // namespace prep
}  // namespace mpf

// ------------------------------------------------------------------ C ABI

struct mpf_prep {
    mpf::prep::Builder b;
    explicit mpf_prep(const mpf::prep::Config& c) : b(c) {}
};

namespace {

constexpr int PREP_ERR_INVAL = -22;
constexpr int PREP_ERR_STATE = -5;
thread_local std::string g_prep_err;

int prep_fail(const std::string& msg, int code = PREP_ERR_INVAL) {
    g_prep_err = msg;
    return code;
}

bool to_config(const mpf_prep_cfg* c, mpf::prep::Config& o) {
    if (!c) return false;
    o.isfullsky = c->isfullsky;
    o.coarseonly = c->coarseonly;
    o.ncoarse = c->ncoarse;
    o.correction = c->correction;
    o.coarsecorrection = c->coarsecorrection;
    o.acorrection = c->acorrection;
    o.order = c->order;
    o.microcoeff = c->microcoeff;
    o.h = c->h;
    o.omegam = c->omegam;
    o.lboxmpch = c->lboxmpch;
    o.mpc = c->mpc;
    o.allocation = c->allocation;
    return c->order >= 0 && c->order <= 2 && c->microcoeff > 0.0 && c->h > 0.0 && c->lboxmpch > 0.0 &&
           c->mpc > 0.0 && c->allocation < (1ull << 36);
}

mpf::cone::Cone read_cone(const double* c7) {
    mpf::cone::Cone c;
    for (int i = 0; i < 3; ++i) {
        c.vertex[i] = c7[i];
        c.base[i] = c7[3 + i];
    }
    c.angle = c7[6];
    return c;
}

}  // namespace

extern "C" {

const char* mpf_prep_last_error(void) { return g_prep_err.c_str(); }

int mpf_prep_create(mpf_prep** prep, const mpf_prep_cfg* cfg) {
    mpf::prep::Config c;
    if (!prep || !to_config(cfg, c)) return prep_fail("mpf_prep_create: bad arguments");
    if (c.acorrection) return prep_fail("mpf_prep_create: acorrection (Full Universe Run a snapping) is not supported");
    try {
        *prep = new mpf_prep(c);
    } catch (...) {
        return prep_fail("mpf_prep_create: out of memory", -12);
    }
    return 0;
}

int mpf_prep_destroy(mpf_prep* prep) {
    delete prep;
    return 0;
}

int mpf_prep_clear(mpf_prep* prep) {
    if (!prep) return prep_fail("mpf_prep_clear: null builder");
    prep->b.clear();
    return 0;
}

int mpf_prep_import_ascii(mpf_prep* prep, const char* path, const double* cone7, const double* rotm1,
                          uint64_t* added) {
    if (!prep || !path || !cone7) return prep_fail("mpf_prep_import_ascii: bad arguments");
    uint64_t n = 0;
    std::string err;
    if (!prep->b.import_ascii(path, read_cone(cone7), rotm1, n, err)) return prep_fail(err, PREP_ERR_STATE);
    if (added) *added = n;
    return 0;
}

int mpf_prep_load_records(mpf_prep* prep, const void* rec48, uint64_t n) {
    if (!prep || (n && !rec48)) return prep_fail("mpf_prep_load_records: bad arguments");
    prep->b.from_records(static_cast<const unsigned char*>(rec48), n);
    return 0;
}

int mpf_prep_update(mpf_prep* prep) {
    if (!prep) return prep_fail("mpf_prep_update: null builder");
    prep->b.update();
    return 0;
}

int mpf_prep_interpolate(mpf_prep* prep, int order, const double* xyz, uint64_t n, float* out7) {
    if (!prep || order < 0 || order > 2 || (n && (!xyz || !out7))) return prep_fail("mpf_prep_interpolate: bad arguments");
    for (uint64_t i = 0; i < n; ++i) {
        const double* p = xyz + 3 * i;
        const mpf::prep::Grav g = (order == 0) ? prep->b.ngp(p[0], p[1], p[2])
                                  : (order == 1) ? prep->b.cic(p[0], p[1], p[2])
                                                 : prep->b.tsc(p[0], p[1], p[2]);
        float* o = out7 + 7 * i;
        o[0] = g.rho; o[1] = g.phi; o[2] = g.dx; o[3] = g.dy; o[4] = g.dz; o[5] = g.a; o[6] = g.dt;
    }
    return 0;
}

int mpf_prep_correct(mpf_prep* prep, const double* t, const double* a, const double* dadt, uint64_t n) {
    if (!prep || (n && (!t || !a || !dadt))) return prep_fail("mpf_prep_correct: bad arguments");
    std::vector<double> tv(t, t + n), av(a, a + n), dv(dadt, dadt + n);
    double amin = 0.0;
    std::string err;
    if (!prep->b.correct(tv, av, dv, amin, err)) return prep_fail(err, PREP_ERR_STATE);
    return 0;
}

int mpf_prep_size(mpf_prep* prep, uint64_t* n) {
    if (!prep || !n) return prep_fail("mpf_prep_size: bad arguments");
    *n = prep->b.cells().size();
    return 0;
}

int mpf_prep_records(mpf_prep* prep, void* rec48, uint64_t capacity) {
    if (!prep || !rec48) return prep_fail("mpf_prep_records: bad arguments");
    if (capacity < prep->b.cells().size()) return prep_fail("mpf_prep_records: buffer smaller than the octree");
    prep->b.to_records(static_cast<unsigned char*>(rec48));
    return 0;
}

int mpf_prep_save(mpf_prep* prep, const char* path) {
    if (!prep || !path) return prep_fail("mpf_prep_save: bad arguments");
    std::string err;
    if (!prep->b.save(path, err)) return prep_fail(err, PREP_ERR_STATE);
    return 0;
}

int mpf_prep_keys(const uint32_t* level, const float* xyz, uint64_t n, uint64_t* key2) {
    if (n && (!level || !xyz || !key2)) return prep_fail("mpf_prep_keys: bad arguments");
    for (uint64_t i = 0; i < n; ++i) {
        if (level[i] > 32) return prep_fail("mpf_prep_keys: level above 32");
        const mpf::key::u128 k = mpf::prep::compute_f(level[i], xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        key2[2 * i] = (uint64_t)k;
        key2[2 * i + 1] = (uint64_t)(k >> 64);
    }
    return 0;
}

int mpf_prep_collide(const uint64_t* key2, uint64_t n, double micro_radius, const double* cone7, uint8_t* out) {
    if (!cone7 || !(micro_radius > 0.0) || (n && (!key2 || !out))) return prep_fail("mpf_prep_collide: bad arguments");
    const mpf::cone::Cone c = read_cone(cone7);
    for (uint64_t i = 0; i < n; ++i) {
        const mpf::key::u128 k = (((mpf::key::u128)key2[2 * i + 1]) << 64) | key2[2 * i];
        out[i] = mpf::prep::collide(k, micro_radius, c) ? 1 : 0;
    }
    return 0;
}

int mpf_prepare_ascii(const mpf_prep_cfg* cfg, const char* celldir, const char* minicone, const double* cones_rot7,
                      uint32_t ncones, const double* rotm1, uint32_t ntasks, uint32_t rank, uint64_t seed,
                      const double* t, const double* a, const double* dadt, uint64_t ncos,
                      const char* const* conefiles) {
    mpf::prep::Config c;
    if (!to_config(cfg, c) || !celldir || (ncones && (!cones_rot7 || !conefiles)) || (ncos && (!t || !a || !dadt)))
        return prep_fail("mpf_prepare_ascii: bad arguments");
    if (!c.isfullsky && (!minicone || !rotm1)) return prep_fail("mpf_prepare_ascii: a narrow cone needs minicone and rotm1");
    std::vector<mpf::cone::Cone> cones;
    std::vector<std::string> files;
    for (uint32_t i = 0; i < ncones; ++i) {
        cones.push_back(read_cone(cones_rot7 + 7 * i));
        if (!conefiles[i]) return prep_fail("mpf_prepare_ascii: null output file name");
        files.push_back(conefiles[i]);
    }
    std::vector<double> tv(t, t + ncos), av(a, a + ncos), dv(dadt, dadt + ncos);
    double amin = 0.0;
    std::string err;
    if (!mpf::prep::prepare_ascii(c, celldir, minicone ? minicone : "", cones, rotm1, ntasks, rank, seed, tv, av, dv,
                                  files, amin, err))
        return prep_fail(err, PREP_ERR_STATE);
    return 0;
}

}  // extern "C"
