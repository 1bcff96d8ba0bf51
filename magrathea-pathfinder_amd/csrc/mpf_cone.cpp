// Cone geometry of the reference programs, restated on the host; see
// mpf_cone.h for what each function replaces.
#include "mpf_cone.h"
#include "mpf_healpix.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <map>
#include <random>
#include <utility>

namespace mpf {
namespace cone {

namespace {

// the constants of magrathea's AbstractShape / Constants (long double
// literals rounded once to double)
const double kPi = (double)3.1415926535897932384626433832795028841971693993751058209749445L;
const double kGolden = (double)1.6180339887498948482045868343656381177203091798057628621354486L;

// sin and cos of one variable through one sincos call: the reference build
// (g++ -O3) merges such pairs into sincos, whose last bit can differ from
// separate sin and cos calls (it keeps separate calls where the argument is
// a library call written twice, as cos(atan2(..)) and sin(atan2(..)),
// cos(acos(..)) and sin(acos(..)) in the bundle launch: the two calls may set
// errno, so they are not merged into one value)
inline void sin_cos(double x, double& s, double& c) { ::sincos(x, &s, &c); }

double axis_length2(const Cone& c) {
    double r = 0.0;
    for (int d = 0; d < 3; ++d) r += (c.base[d] - c.vertex[d]) * (c.base[d] - c.vertex[d]);
    return r;
}

bool same(const Cone& a, const Cone& b) {
    for (int d = 0; d < 3; ++d)
        if (a.vertex[d] != b.vertex[d] || a.base[d] != b.base[d]) return false;
    return a.angle == b.angle;
}

// Input::trim (input.h:158-166): cut at the comment, strip non-graphical
// characters at both ends, collapse runs of white space
std::string trim(const std::string& text, const std::string& comment) {
    std::string r = text;
    const size_t cut = r.find(comment);
    if (cut != std::string::npos) r.erase(cut);
    auto graph = [](char ch) { return std::isgraph((unsigned char)ch) != 0; };
    r.erase(r.begin(), std::find_if(r.begin(), r.end(), graph));
    r.erase(std::find_if(r.rbegin(), r.rend(), graph).base(), r.end());
    r.erase(std::unique(r.begin(), r.end(),
                        [](char x, char y) { return std::isspace((unsigned char)x) && std::isspace((unsigned char)y); }),
            r.end());
    return r;
}

// Input::parse (input.h:1237-1262) with separator "=" and comment "#": runs
// of lines with the same key keep their first value; otherwise the last
// occurrence of a key wins
std::map<std::string, std::string> parse(const std::string& path, bool& opened) {
    std::map<std::string, std::string> result;
    std::ifstream in(path);
    opened = (bool)in;
    if (!in) return result;
    std::vector<std::pair<std::string, std::string>> kv;
    std::string line;
    while (std::getline(in, line)) {
        const std::string t = trim(line, "#");
        const size_t eq = t.find('=');
        std::string k = (eq == std::string::npos) ? t : t.substr(0, eq);
        std::string v = (eq == std::string::npos) ? std::string() : t.substr(eq + 1);
        k = trim(k, "#");
        v = trim(v, "#");
        if (k.empty() || v.empty()) continue;
        if (!kv.empty() && kv.back().first == k) continue;
        kv.emplace_back(k, v);
    }
    for (auto& e : kv) result[e.first] = e.second;
    return result;
}

}  // namespace

bool inside(const Cone& c, const double p[3]) {
    double dot = 0.0, dist = 0.0, ref = 0.0;
    for (int d = 0; d < 3; ++d) {
        dot += (c.base[d] - c.vertex[d]) * (p[d] - c.vertex[d]);
        dist += (p[d] - c.vertex[d]) * (p[d] - c.vertex[d]);
        ref += (c.base[d] - c.vertex[d]) * (c.base[d] - c.vertex[d]);
    }
    return (ref > 0.0) && !(std::acos(dot / std::sqrt(dist * ref)) > c.angle) && !(dot > ref) && (dot > 0.0);
}

bool owns(const Cone* cones, uint32_t ncones, uint32_t self, const double p[3], bool by_value) {
    const Cone& c = cones[self];
    if (!inside(c, p)) return false;
    // squared distance of p from the point of the cone's axis at p's projection
    double len = 0.0;
    for (int d = 0; d < 3; ++d) len += (c.base[d] - c.vertex[d]) * (p[d] - c.vertex[d]);
    // Cone::length() squared: sqrt(S) * sqrt(S), not S
    len /= std::sqrt(axis_length2(c)) * std::sqrt(axis_length2(c));
    double reference = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double q = p[d] - (c.vertex[d] + (c.base[d] - c.vertex[d]) * len);
        reference += q * q;
    }
    for (uint32_t k = 0; k < ncones; ++k) {
        const bool other = by_value ? !same(cones[k], c) : (k != self);
        if (!other) continue;
        const Cone& o = cones[k];
        // the other cone's axis direction dotted with p from its own vertex,
        // measured from this cone's vertex as the reference writes it
        double l = 0.0;
        for (int d = 0; d < 3; ++d) l += (o.base[d] - c.vertex[d]) * (p[d] - o.vertex[d]);
        if (l < 0.0) continue;
        l /= std::sqrt(axis_length2(o)) * std::sqrt(axis_length2(o));
        double dist = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double q = p[d] - (o.vertex[d] + (o.base[d] - o.vertex[d]) * l);
            dist += q * q;
        }
        if (dist < reference) return false;
    }
    return true;
}

void spiral(uint64_t n, double radius, double* xyz) {
    const double twopi = 2.0 * kPi;
    const double turn = twopi * (2.0 - kGolden);
    const double step = 2.0 / (double)(unsigned)n;
    for (uint64_t i = 0; i < n; ++i) {
        const double y = ((double)(unsigned)i * step) - 1.0 + (step / 2.0);
        const double r = std::sqrt(1.0 - y * y);
        const double ang = (double)(unsigned)i * turn;
        xyz[3 * i + 0] = 0.0 + (radius * r * std::cos(ang));
        xyz[3 * i + 1] = 0.0 + (radius * y);
        xyz[3 * i + 2] = 0.0 + (radius * r * std::sin(ang));
    }
}

double min_distance(const double* xyz, uint64_t n, double init) {
    double best = init;
    for (uint64_t i = 0; i < n; ++i)
        for (uint64_t j = 0; j < n; ++j) {
            if (i == j) continue;
            double t = 0.0;
            for (int d = 0; d < 3; ++d) t += (xyz[3 * i + d] - xyz[3 * j + d]) * (xyz[3 * i + d] - xyz[3 * j + d]);
            t = std::sqrt(t);
            best = best < t ? best : t;
        }
    return best;
}

void fullsky(uint32_t ncones, double radius, std::vector<Cone>& cones) {
    std::vector<double> tiling(3 * (size_t)ncones);
    spiral(ncones, radius, tiling.data());
    const double diameter = 2.0 * radius;
    // uniform() returns (2 radius, 0) for fewer than two points, the pairwise
    // minimum (started from 2 radius) otherwise
    const double dmin = (ncones > 1) ? min_distance(tiling.data(), ncones, diameter) : (ncones ? diameter : 0.0);
    const double alpha = 1.8 * std::asin(dmin / diameter);
    cones.assign(ncones, Cone());
    for (uint32_t i = 0; i < ncones; ++i) {
        for (int d = 0; d < 3; ++d) {
            cones[i].vertex[d] = 0.0;
            cones[i].base[d] = tiling[3 * i + d];
        }
        cones[i].angle = alpha;
    }
}

void inv3(const double A[9], double r[9]) {
    auto a = [&](int i, int j) { return A[3 * i + j]; };
    const double invdet = 1. / (a(0, 0) * (a(1, 1) * a(2, 2) - a(2, 1) * a(1, 2)) -
                                a(0, 1) * (a(1, 0) * a(2, 2) - a(1, 2) * a(2, 0)) +
                                a(0, 2) * (a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0)));
    r[0] = (a(1, 1) * a(2, 2) - a(2, 1) * a(1, 2)) * invdet;
    r[1] = -(a(0, 1) * a(2, 2) - a(0, 2) * a(2, 1)) * invdet;
    r[2] = (a(0, 1) * a(1, 2) - a(0, 2) * a(1, 1)) * invdet;
    r[3] = -(a(1, 0) * a(2, 2) - a(1, 2) * a(2, 0)) * invdet;
    r[4] = (a(0, 0) * a(2, 2) - a(0, 2) * a(2, 0)) * invdet;
    r[5] = -(a(0, 0) * a(1, 2) - a(1, 0) * a(0, 2)) * invdet;
    r[6] = (a(1, 0) * a(2, 1) - a(2, 0) * a(1, 1)) * invdet;
    r[7] = -(a(0, 0) * a(2, 1) - a(2, 0) * a(0, 1)) * invdet;
    r[8] = (a(0, 0) * a(1, 1) - a(1, 0) * a(0, 1)) * invdet;
}

bool narrow(uint32_t ncones, double radius, double phi_deg, double theta_deg, double thetay_deg, double thetaz_deg,
            std::vector<Cone>& cones, std::vector<Cone>& cones_rot, double rotm1[9], uint64_t& spiral_size,
            std::string& err) {
    const double deg = kPi / 180;
    const double thetaz = thetaz_deg * deg, thetay = thetay_deg * deg;
    const double phr = phi_deg * deg, thr = theta_deg * deg;
    double sph, cph, sth, cth;
    sin_cos(phr, sph, cph);
    sin_cos(thr, sth, cth);
    const double rot[9] = {cth * cph, cth * sph, -sth, -sph, cph, 0.0, cph * sth, sph * sth, cth};
    inv3(rot, rotm1);
    if (ncones == 0) {
        cones.clear();
        cones_rot.clear();
        spiral_size = 0;
        return true;
    }
    // the region: |phi| < thetay and |theta - pi/2| < thetaz around +x
    auto in_region = [&](const double* p) {
        const double ph = std::atan2(p[1], p[0]);
        const double th = std::acos(p[2] / radius);
        return std::abs(ph) < thetay - 0.0 && th > kPi / 2 - thetaz + 0.0 && th < kPi / 2 + thetaz - 0.0;
    };
    const double fullsky_sr = 4 * kPi;
    const double portion = 2 * thetay * (std::sin(thetaz) - std::cos(kPi / 2. + thetaz));
    if (!(portion > 0.0) || !(fullsky_sr / portion < 4294967296.0)) {
        err = "narrow cones: thetay and thetaz must be positive";
        return false;
    }
    const unsigned fsp = (unsigned)(fullsky_sr / portion);
    std::vector<double> tiling;
    auto count = [&](size_t n) {
        tiling.resize(3 * n);
        spiral(n, radius, tiling.data());
        unsigned c = 0;
        for (size_t i = 0; i < n; ++i) c += in_region(&tiling[3 * i]);
        return c;
    };
    size_t n = (size_t)ncones * fsp;
    unsigned hits = count(n);
    // resize until exactly ncones spiral points fall in the region
    // (miscellaneous.h:351-368): proportional steps in single precision for
    // the first ten rounds, then unit steps
    float round = 0;
    unsigned rounds = 0;
    while (hits != ncones) {
        if (round < 10)
            n = n + (size_t)(long long)(int)(((float)ncones - (float)hits) * (float)fsp * (1 / (1 + round)));
        else
            n = n + ncones - hits;
        if (++rounds > 1000000 || n > ((size_t)1 << 34)) {
            err = "narrow cones: no spiral size puts exactly ncones points in the region";
            return false;
        }
        hits = count(n);
        round++;
    }
    spiral_size = n;
    std::vector<double> pick;
    pick.reserve(3 * (size_t)ncones);
    for (size_t i = 0; i < n; ++i)
        if (in_region(&tiling[3 * i])) pick.insert(pick.end(), &tiling[3 * i], &tiling[3 * i] + 3);
    const double diameter = 2.0 * radius;
    const double dmin = min_distance(pick.data(), ncones, diameter);
    double alpha = 1.8 * std::asin(dmin / diameter);
    const double anglemin = 0.01;
    alpha = (anglemin > alpha) ? anglemin : alpha;
    cones.assign(ncones, Cone());
    cones_rot.assign(ncones, Cone());
    for (uint32_t i = 0; i < ncones; ++i) {
        for (int d = 0; d < 3; ++d) {
            cones[i].vertex[d] = 0.0;
            cones[i].base[d] = pick[3 * i + d];
        }
        cones[i].angle = alpha;
        cones_rot[i] = cones[i];
        const double* b = cones[i].base;
        for (int d = 0; d < 3; ++d)
            cones_rot[i].base[d] = b[0] * rotm1[3 * d + 0] + b[1] * rotm1[3 * d + 1] + b[2] * rotm1[3 * d + 2];
    }
    return true;
}

bool narrow_info(const std::string& path, double& phi_deg, double& theta_deg, double& thetay_deg, double& thetaz_deg,
                 std::string& err) {
    bool opened = false;
    auto p = parse(path, opened);
    if (!opened) {
        err = "cannot open " + path;
        return false;
    }
    for (const char* k : {"phi", "theta", "thetay", "thetaz"})
        if (!p.count(k)) {
            err = std::string("info_narrow_cone.txt: no `") + k + "` entry";
            return false;
        }
    try {
        // phi goes through std::stoul (miscellaneous.h:299): integer degrees
        phi_deg = (double)std::stoul(p["phi"]);
        theta_deg = std::stod(p["theta"]);
        thetay_deg = std::stod(p["thetay"]);
        thetaz_deg = std::stod(p["thetaz"]);
    } catch (...) {
        err = "info_narrow_cone.txt: unreadable angle";
        return false;
    }
    return true;
}

void launch_angles(const double begin[3], double phi, double theta, double o[9]) {
    o[0] = 1.0;
    o[1] = 0.0;
    o[2] = begin[0];
    o[3] = begin[1];
    o[4] = begin[2];
    o[5] = 1.0;
    double sp, cp, st, ct;
    sin_cos(phi, sp, cp);
    sin_cos(theta, st, ct);
    o[6] = cp * st;
    o[7] = sp * st;
    o[8] = ct;
}

void launch_towards(const double b[3], const double e[3], double o[9]) {
    const double norm = std::sqrt((e[0] - b[0]) * (e[0] - b[0]) + (e[1] - b[1]) * (e[1] - b[1]) +
                                  (e[2] - b[2]) * (e[2] - b[2]));
    o[0] = 1.0;
    o[1] = 0.0;
    o[2] = b[0];
    o[3] = b[1];
    o[4] = b[2];
    o[5] = 1.0;
    o[6] = (e[0] - b[0]) / norm;
    o[7] = (e[1] - b[1]) / norm;
    o[8] = (e[2] - b[2]) / norm;
}

void launch_cone_random(const Cone* cones, uint32_t ncones, uint32_t icone, double micro_radius, uint32_t seed,
                        uint64_t n, double* out9, double* rotation) {
    std::mt19937 engine(seed);
    std::uniform_real_distribution<double> uniform(0.0, 1.0);
    const double twopi_factor = 2;  // `two` of rays.cpp times pi below
    for (uint64_t r = 0; r < n; ++r) {
        double p[3];
        do {
            // a point of the sphere's surface: three normal deviates from a
            // fresh distribution (its cached second deviate is dropped),
            // scaled to the radius
            std::normal_distribution<double> normal(0, 1);
            double g[3], s = 0.0;
            for (int d = 0; d < 3; ++d) {
                g[d] = normal(engine);
                s += g[d] * g[d];
            }
            const double coef = (micro_radius / std::sqrt(s)) * 1.0;
            for (int d = 0; d < 3; ++d) p[d] = 0.0 + coef * g[d];
        } while (!owns(cones, ncones, icone, p, false));
        double* o = out9 + 9 * r;
        o[0] = 1.0;
        o[1] = 0.0;
        o[2] = o[3] = o[4] = 0.0;
        o[5] = 1.0;
        for (int d = 0; d < 3; ++d) o[6 + d] = (p[d] - 0.0) / micro_radius;
        const double u = uniform(engine);
        if (rotation) rotation[r] = u * twopi_factor * kPi;
    }
}

void launch_bundle(const double in9[9], uint32_t count, double angle, double rotation, bool center, double* out9) {
    const double pi = 4.0 * std::atan(1.0);
    const double step = (pi + pi) / (double)(count + (!count));
    const double kx = in9[6], ky = in9[7], kz = in9[8];
    const double r = std::sqrt((kx * kx) + (ky * ky) + (kz * kz));
    double sa, ca, sintheta, costheta, sinphi, cosphi, sinpsi, cospsi;
    sin_cos(angle, sa, ca);
    costheta = std::cos(std::acos(kz / r));
    sintheta = std::sin(std::acos(kz / r));
    cosphi = std::cos(std::atan2(ky, kx));
    sinphi = std::sin(std::atan2(ky, kx));
    sin_cos(rotation, sinpsi, cospsi);
    const double rcos = r * ca;
    const double rsin = r * sa;
    const uint32_t total = count + (center ? 1u : 0u);
    for (uint32_t i = 0; i < total; ++i)
        for (int k = 0; k < 9; ++k) out9[9 * i + k] = in9[k];
    if (!(r > 0.0)) return;
    for (uint32_t i = 0; i < count; ++i) {
        double si, ci;
        sin_cos(i * step, si, ci);
        const double x = rsin * ci;
        const double y = rsin * si;
        const double z = rcos;
        double* o = out9 + 9 * (i + (center ? 1 : 0));
        o[6] = -cosphi * sinpsi * costheta * x - cosphi * cospsi * costheta * y + cosphi * sintheta * z +
               sinphi * sinpsi * y - sinphi * cospsi * x;
        o[7] = -sinphi * sinpsi * costheta * x - sinphi * cospsi * costheta * y - cosphi * sinpsi * y +
               cosphi * cospsi * x + sintheta * sinphi * z;
        o[8] = sintheta * sinpsi * x + sintheta * cospsi * y + costheta * z;
    }
}

}  // namespace cone
}  // namespace mpf

// ---------------------------------------------------------------------------
// C ABI (include/mpf.h)
#include <thread>

#include "../../include/mpf.h"

namespace {

constexpr int ERR_INVAL = -22;
thread_local std::string g_cone_err;

int cone_fail(const std::string& msg) {
    g_cone_err = msg;
    return ERR_INVAL;
}

std::vector<mpf::cone::Cone> read_cones(const double* c7, uint32_t n) {
    std::vector<mpf::cone::Cone> v(n);
    for (uint32_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            v[i].vertex[d] = c7[MPF_CONE_N * i + d];
            v[i].base[d] = c7[MPF_CONE_N * i + 3 + d];
        }
        v[i].angle = c7[MPF_CONE_N * i + 6];
    }
    return v;
}

void write_cones(const std::vector<mpf::cone::Cone>& v, double* c7) {
    for (size_t i = 0; i < v.size(); ++i) {
        for (int d = 0; d < 3; ++d) {
            c7[MPF_CONE_N * i + d] = v[i].vertex[d];
            c7[MPF_CONE_N * i + 3 + d] = v[i].base[d];
        }
        c7[MPF_CONE_N * i + 6] = v[i].angle;
    }
}

// runs f(begin, end) over [0, n) on the host cores (Utility::parallelize
// splits the same way: contiguous chunks, one per thread)
template <class F>
void parallel_chunks(uint64_t n, F f) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const uint64_t nth = std::min<uint64_t>(hw, std::max<uint64_t>(1, n / 4096));
    if (nth <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    const uint64_t chunk = (n + nth - 1) / nth;
    for (uint64_t t = 0; t < nth; ++t) {
        const uint64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b < e) th.emplace_back([=] { f(b, e); });
    }
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

const char* mpf_cone_last_error(void) { return g_cone_err.c_str(); }

int mpf_cones_fullsky(uint32_t ncones, double radius, double* cones7) {
    if (!(radius > 0.0) || (ncones && !cones7)) return cone_fail("mpf_cones_fullsky: bad arguments");
    std::vector<mpf::cone::Cone> v;
    mpf::cone::fullsky(ncones, radius, v);
    write_cones(v, cones7);
    return 0;
}

int mpf_cones_narrow(uint32_t ncones, double radius, double phi_deg, double theta_deg, double thetay_deg,
                     double thetaz_deg, double* cones7, double* cones_rot7, double* rotm1, uint64_t* spiral_points) {
    if (!(radius > 0.0) || !rotm1 || (ncones && (!cones7 || !cones_rot7)))
        return cone_fail("mpf_cones_narrow: bad arguments");
    std::vector<mpf::cone::Cone> v, vr;
    uint64_t size = 0;
    std::string err;
    if (!mpf::cone::narrow(ncones, radius, phi_deg, theta_deg, thetay_deg, thetaz_deg, v, vr, rotm1, size, err))
        return cone_fail(err);
    write_cones(v, cones7);
    write_cones(vr, cones_rot7);
    if (spiral_points) *spiral_points = size;
    return 0;
}

int mpf_narrow_cone_info(const char* path, double* angles4) {
    if (!path || !angles4) return cone_fail("mpf_narrow_cone_info: bad arguments");
    std::string err;
    if (!mpf::cone::narrow_info(path, angles4[0], angles4[1], angles4[2], angles4[3], err)) return cone_fail(err);
    return 0;
}

int mpf_cone_inside(const double* cone7, const double* xyz, uint64_t n, uint64_t stride, uint8_t* out) {
    if (!cone7 || (n && (!xyz || !out)) || stride < 3) return cone_fail("mpf_cone_inside: bad arguments");
    const auto c = read_cones(cone7, 1);
    parallel_chunks(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) out[i] = mpf::cone::inside(c[0], xyz + stride * i) ? 1 : 0;
    });
    return 0;
}

int mpf_cone_select(const double* cones7, uint32_t ncones, uint32_t icone, int by_value, const double* xyz,
                    uint64_t n, uint64_t stride, uint8_t* out) {
    if (!cones7 || icone >= ncones || (n && (!xyz || !out)) || stride < 3)
        return cone_fail("mpf_cone_select: bad arguments");
    const auto c = read_cones(cones7, ncones);
    parallel_chunks(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i)
            out[i] = mpf::cone::owns(c.data(), ncones, icone, xyz + stride * i, by_value != 0) ? 1 : 0;
    });
    return 0;
}

int mpf_cone_targets(const double* cones7, uint32_t ncones, uint32_t icone, const double* targets, uint64_t n,
                     uint64_t stride, uint64_t* index_out, uint64_t* nselected) {
    if (!nselected || (n && !index_out)) return cone_fail("mpf_cone_targets: bad arguments");
    std::vector<uint8_t> sel(n);
    const int rc = mpf_cone_select(cones7, ncones, icone, 1, targets, n, stride, sel.data());
    if (rc) return rc;
    uint64_t k = 0;
    for (uint64_t i = 0; i < n; ++i)
        if (sel[i]) index_out[k++] = i;
    *nselected = k;
    return 0;
}

int mpf_hmaps_cone_pixels(int64_t nside, const double* cones7, uint32_t ncones, uint32_t icone, int64_t* pix_out,
                          uint64_t capacity, uint64_t* npix_out) {
    if (nside < 1 || nside > (1 << 29) || !cones7 || icone >= ncones || !npix_out)
        return cone_fail("mpf_hmaps_cone_pixels: bad arguments");
    const int64_t npix = 12 * nside * nside;
    const auto c = read_cones(cones7, ncones);
    std::vector<uint8_t> sel((size_t)npix);
    parallel_chunks((uint64_t)npix, [&](uint64_t b, uint64_t e) {
        double v[3];
        for (uint64_t p = b; p < e; ++p) {
            mpf::healpix::pix2vec_ring(nside, (int64_t)p, v);
            sel[p] = mpf::cone::owns(c.data(), ncones, icone, v, false) ? 1 : 0;
        }
    });
    uint64_t k = 0;
    for (int64_t p = 0; p < npix; ++p)
        if (sel[p]) {
            if (pix_out && k < capacity) pix_out[k] = p;
            ++k;
        }
    *npix_out = k;
    if (pix_out && k > capacity) return cone_fail("mpf_hmaps_cone_pixels: capacity too small");
    return 0;
}

int mpf_launch_angles(const double* observer3, const double* phi_theta, uint64_t n, double* in9) {
    if (!observer3 || (n && (!phi_theta || !in9))) return cone_fail("mpf_launch_angles: bad arguments");
    for (uint64_t i = 0; i < n; ++i)
        mpf::cone::launch_angles(observer3, phi_theta[2 * i], phi_theta[2 * i + 1], in9 + 9 * i);
    return 0;
}

int mpf_launch_towards(const double* begin3, const double* end3, uint64_t n, double* in9) {
    if (n && (!begin3 || !end3 || !in9)) return cone_fail("mpf_launch_towards: bad arguments");
    for (uint64_t i = 0; i < n; ++i) mpf::cone::launch_towards(begin3 + 3 * i, end3 + 3 * i, in9 + 9 * i);
    return 0;
}

int mpf_launch_cone(const double* cones7, uint32_t ncones, uint32_t icone, double micro_radius, uint32_t seed,
                    uint64_t n, double* in9, double* rotation) {
    if (!cones7 || icone >= ncones || !(micro_radius > 0.0) || (n && !in9))
        return cone_fail("mpf_launch_cone: bad arguments");
    const auto c = read_cones(cones7, ncones);
    // a cone that owns no direction would make the draw loop for ever
    // (the reference's does): require a positive axis and aperture
    const mpf::cone::Cone& s = c[icone];
    double ax = 0.0;
    for (int d = 0; d < 3; ++d) ax += (s.base[d] - s.vertex[d]) * (s.base[d] - s.vertex[d]);
    if (!(ax > 0.0) || !(s.angle > 0.0)) return cone_fail("mpf_launch_cone: degenerate cone");
    mpf::cone::launch_cone_random(c.data(), ncones, icone, micro_radius, seed, n, in9, rotation);
    return 0;
}

int mpf_launch_bundle(const double* in9, uint32_t count, double angle, double rotation, int center, double* out9) {
    if (!in9 || !out9) return cone_fail("mpf_launch_bundle: bad arguments");
    mpf::cone::launch_bundle(in9, count, angle, rotation, center != 0, out9);
    return 0;
}

}  // extern "C"
