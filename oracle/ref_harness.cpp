/* Reference harness: a thin extern "C" shim over the reference's own C++
 * headers (integrator.h, lensing.h, magrathea/simplehyperoctree.h), compiled
 * from the sources where they lie by oracle/Makefile into oracle/_ref/.
 *
 * TEST INFRASTRUCTURE ONLY.  This file holds no reference source: it
 * instantiates the reference templates with the types the reference programs
 * use (rays.cpp:85-91,142; catalogues.cpp:123; hmaps.cpp:127) and marshals
 * plain arrays in and out.  It is used to pin the C restatement in
 * oracle/mpf_oracle.c, to write the golden fixtures under tests/golden/, and
 * as the CPU baseline in bench.py.  Product code never links it.
 */
#include <cstdint>
#include <cstring>
#include <vector>
#include <array>
#include <string>
#include <algorithm>
#include <thread>

#include "integrator.h"
#include "lensing.h"

using namespace magrathea;

namespace {
using real = double;
using floating = float;
using position = std::ratio<0>;
using extent = std::ratio<EXTENT>;
using indexing = __uint128_t;
using Index = SimpleHyperOctreeIndex<indexing, 3>;
using Data = Gravity<floating, 3>;
using Octree = SimpleHyperOctree<real, Index, Data, 3, position, extent>;
using Element = std::pair<Index, Data>;
using Traj = Evolution<Photon<real, 3>>;
using Cosmo = std::array<std::vector<real>, 4>;

static_assert(sizeof(Element) == 48, "record size");

struct RefOctree {
    Octree octree;
};

void data_to7(const Data& d, float* o) {
    o[0] = d.rho(); o[1] = d.phi(); o[2] = d.dphidx(); o[3] = d.dphidy();
    o[4] = d.dphidz(); o[5] = d.a(); o[6] = d.dphidt();
}

/* Per-ray record written by ref_integrate (doubles): */
enum {
    O_A, O_T, O_X, O_Y, O_Z, O_KT, O_KX, O_KY, O_KZ,
    O_REDSHIFT, O_LAMBDA, O_S, O_CHI, O_ISW, O_ISWOLD, O_PHI,
    O_DPHIDX, O_DPHIDY, O_DPHIDZ, O_DPHIDT, O_RHO, O_AH, O_DPHIDL,
    O_LAPLACIAN, O_DSDL2, O_ERROR, O_LEVEL, O_NSTEPS, O_NREC
};
static_assert(O_NREC == 28, "record");

void photon_to_rec(const Photon<real, 3>& p, double* o, double nsteps) {
    o[O_A] = p.a(); o[O_T] = p.t(); o[O_X] = p.x(); o[O_Y] = p.y(); o[O_Z] = p.z();
    o[O_KT] = p.dtdl(); o[O_KX] = p.dxdl(); o[O_KY] = p.dydl(); o[O_KZ] = p.dzdl();
    o[O_REDSHIFT] = p.redshift(); o[O_LAMBDA] = p.lambda(); o[O_S] = p.s();
    o[O_CHI] = p.chi(); o[O_ISW] = p.isw(); o[O_ISWOLD] = p.iswold(); o[O_PHI] = p.phi();
    o[O_DPHIDX] = p.dphidx(); o[O_DPHIDY] = p.dphidy(); o[O_DPHIDZ] = p.dphidz();
    o[O_DPHIDT] = p.dphidt(); o[O_RHO] = p.rho(); o[O_AH] = p.ah(); o[O_DPHIDL] = p.dphidl();
    o[O_LAPLACIAN] = p.laplacian(); o[O_DSDL2] = p.dsdl2(); o[O_ERROR] = p.error();
    o[O_LEVEL] = p.level(); o[O_NSTEPS] = nsteps;
}

const char* stop_name(int kind) {
    switch (kind) {
        case 0: return "radius";
        case 1: return "lambda";
        case 2: return "redshift";
        case 3: return "t";
        case 4: return "a";
        case 5: return "plane";
        case 6: return "s";
        default: return "";
    }
}

Cosmo make_cosmo(const double* tab, uint64_t n) {
    Cosmo c;
    for (int k = 0; k < 4; ++k) c[k].assign(tab + k * n, tab + (k + 1) * n);
    return c;
}

template <int Order>
void integrate_one(Traj& traj, int stop_kind, double stop_value, const Cosmo& cosmo,
                   const Octree& octree, double length, unsigned nsteps, const std::array<double, 3>& kit) {
    static const std::array<double, 3> vobs{{0., 0., 0.}};
    if (stop_kind < 0) {
        Integrator::integrate<Order>(traj, cosmo, octree, vobs, length, nsteps);
    } else {
        Integrator::integrate<Order>(traj, std::string(stop_name(stop_kind)), stop_value, cosmo, octree, vobs, length, nsteps, kit);
    }
}

void integrate_dispatch(int order, Traj& traj, int stop_kind, double stop_value, const Cosmo& cosmo,
                        const Octree& octree, double length, unsigned nsteps, const std::array<double, 3>& kit) {
    switch (order) {
        case -1: integrate_one<-1>(traj, stop_kind, stop_value, cosmo, octree, length, nsteps, kit); break;
        case 0: integrate_one<0>(traj, stop_kind, stop_value, cosmo, octree, length, nsteps, kit); break;
        case 1: integrate_one<1>(traj, stop_kind, stop_value, cosmo, octree, length, nsteps, kit); break;
        default: integrate_one<2>(traj, stop_kind, stop_value, cosmo, octree, length, nsteps, kit); break;
    }
}

Traj launch_traj(const double* in9) {
    Traj traj;
    Photon<real, 3> p;
    p.a() = in9[0]; p.t() = in9[1]; p.x() = in9[2]; p.y() = in9[3]; p.z() = in9[4];
    p.dtdl() = in9[5]; p.dxdl() = in9[6]; p.dydl() = in9[7]; p.dzdl() = in9[8];
    traj.append(p);
    return traj;
}
}  // namespace

extern "C" {

int ref_record_size() { return (int)sizeof(Element); }

void* ref_octree_new(const void* rec, uint64_t n) {
    RefOctree* o = new RefOctree;
    o->octree.resize(n);
    if (n) std::memcpy((void*)o->octree.data(), rec, n * sizeof(Element));
    return o;
}

void ref_octree_free(void* h) { delete static_cast<RefOctree*>(h); }

/* compute(level, x, y, z) for level in [0, 32]; level 32 is cipher(). */
void ref_compute_key(unsigned level, const double* xyz, uint64_t m, uint64_t* lohi) {
    for (uint64_t i = 0; i < m; ++i) {
        Index k = (level >= 32)
            ? Index::cipher<real, position, extent>(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2])
            : Index::compute<real, position, extent>(level, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        __uint128_t v = k();
        lohi[2 * i] = (uint64_t)v;
        lohi[2 * i + 1] = (uint64_t)(v >> 64);
    }
}

/* level, centre (3), extent in box units, parent key and containment of a second key. */
void ref_key_info(const uint64_t* lohi, uint64_t m, uint32_t* level, double* center3, double* ext, uint64_t* parent) {
    for (uint64_t i = 0; i < m; ++i) {
        Index k(((__uint128_t)lohi[2 * i + 1] << 64) | lohi[2 * i]);
        level[i] = k.level();
        for (unsigned d = 0; d < 3; ++d) center3[3 * i + d] = k.center<real, position, extent>(d);
        ext[i] = real(extent::num) / real(extent::den) * k.extent();
        __uint128_t p = k.parent()();
        parent[2 * i] = (uint64_t)p;
        parent[2 * i + 1] = (uint64_t)(p >> 64);
    }
}

int ref_key_containing(const uint64_t* a, const uint64_t* b) {
    Index ka(((__uint128_t)a[1] << 64) | a[0]);
    Index kb(((__uint128_t)b[1] << 64) | b[0]);
    return ka.containing(kb) ? 1 : 0;
}

/* locate(): element index in key order (-1 for end()) and its level. */
void ref_locate(void* h, const double* xyz, uint64_t m, int64_t* elem, uint32_t* level) {
    const Octree& oc = static_cast<RefOctree*>(h)->octree;
    for (uint64_t i = 0; i < m; ++i) {
        auto it = oc.locate(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        if (it == oc.end()) { elem[i] = -1; level[i] = 0; }
        else { elem[i] = (int64_t)(it - oc.begin()); level[i] = std::get<0>(*it).level(); }
    }
}

/* order 0 ngp, 1 cic, 2 uncached tsc, 3 cached tsc over the points in sequence */
void ref_interp(void* h, int order, const double* xyz, uint64_t m, float* out7) {
    const Octree& oc = static_cast<RefOctree*>(h)->octree;
    std::vector<Element> cache(27);
    for (uint64_t i = 0; i < m; ++i) {
        const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        Data d;
        if (order == 0) d = oc.ngp(x, y, z);
        else if (order == 1) d = oc.cic(x, y, z);
        else if (order == 2) d = oc.tsc(x, y, z);
        else d = oc.tsc(cache, x, y, z);
        data_to7(d, out7 + 7 * i);
    }
}

/* Integrate n rays.  in9: launch photon core per ray.  stop_kind -1 = no stop
 * criterion (integrator.h:477), else the criterion of integrator.h:779-811.
 * out: 28 doubles per ray (final trajectory element, see O_*).  Returns the
 * total number of appended steps. nthreads<=0: hardware concurrency. */
uint64_t ref_integrate(void* h, int order, const double* cosmo_tab, uint64_t ncos, double length,
                       unsigned nsteps, int stop_kind, double stop_value, const double* kitarget,
                       const double* in9, uint64_t n, double* out, int nthreads) {
    const Octree& oc = static_cast<RefOctree*>(h)->octree;
    const Cosmo cosmo = make_cosmo(cosmo_tab, ncos);
    std::array<double, 3> kit{{0., 0., 0.}};
    if (kitarget) kit = {{kitarget[0], kitarget[1], kitarget[2]}};
    std::vector<uint64_t> steps(n, 0);
    auto body = [&](const uint64_t i) {
        Traj traj = launch_traj(in9 + 9 * i);
        integrate_dispatch(order, traj, stop_kind, stop_value, cosmo, oc, length, nsteps, kit);
        double* o = out + O_NREC * i;
        if (traj.empty()) {
            for (int k = 0; k < O_NREC; ++k) o[k] = 0.0;
            o[O_NSTEPS] = -1.0;
        } else {
            photon_to_rec(traj.back(), o, double(traj.size() - 1));
            steps[i] = traj.size() - 1;
        }
    };
    if (nthreads == 1) {
        for (uint64_t i = 0; i < n; ++i) body(i);
    } else {
        const int nt = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
        Utility::parallelize(n, body, nt);
    }
    uint64_t total = 0;
    for (uint64_t i = 0; i < n; ++i) total += steps[i];
    return total;
}

/* One ray, full trajectory: up to maxlen elements of 28 doubles.  Returns the
 * trajectory length (may exceed maxlen, in which case only maxlen rows are written). */
uint64_t ref_trajectory(void* h, int order, const double* cosmo_tab, uint64_t ncos, double length,
                        unsigned nsteps, int stop_kind, double stop_value, const double* kitarget,
                        const double* in9, double* rows, uint64_t maxlen) {
    const Octree& oc = static_cast<RefOctree*>(h)->octree;
    const Cosmo cosmo = make_cosmo(cosmo_tab, ncos);
    std::array<double, 3> kit{{0., 0., 0.}};
    if (kitarget) kit = {{kitarget[0], kitarget[1], kitarget[2]}};
    Traj traj = launch_traj(in9);
    integrate_dispatch(order, traj, stop_kind, stop_value, cosmo, oc, length, nsteps, kit);
    for (uint64_t i = 0; i < traj.size() && i < maxlen; ++i) photon_to_rec(traj[i], rows + O_NREC * i, double(i));
    return traj.size();
}

/* Infinitesimal lensing Jacobian (lensing.h:399-480) of one ray integrated as
 * in ref_trajectory, for nd source distances (box units).  out: 4 doubles per
 * distance (a11, a12, a21, a22). Returns the trajectory length. */
uint64_t ref_jacobian_infinitesimal(void* h, int order, const double* cosmo_tab, uint64_t ncos, double length,
                                    unsigned nsteps, int stop_kind, double stop_value, const double* kitarget,
                                    const double* in9, const double* dist, uint64_t nd, double* out4) {
    const Octree& oc = static_cast<RefOctree*>(h)->octree;
    const Cosmo cosmo = make_cosmo(cosmo_tab, ncos);
    std::array<double, 3> kit{{0., 0., 0.}};
    if (kitarget) kit = {{kitarget[0], kitarget[1], kitarget[2]}};
    Traj traj = launch_traj(in9);
    integrate_dispatch(order, traj, stop_kind, stop_value, cosmo, oc, length, nsteps, kit);
    std::vector<double> d(dist, dist + nd);
    auto res = Lensing::dbetadtheta_infinitesimal<2>(d, traj, oc, length);
    for (uint64_t j = 0; j < nd; ++j) {
        out4[4 * j + 0] = res[j][0][0]; out4[4 * j + 1] = res[j][0][1];
        out4[4 * j + 2] = res[j][1][0]; out4[4 * j + 3] = res[j][1][1];
    }
    return traj.size();
}

/* Utility::interpolate (utility.h:442) on a table, for pinning. */
void ref_interpolate(const double* x, const double* y, uint64_t n, const double* x0, uint64_t m, double* out) {
    std::vector<double> vx(x, x + n), vy(y, y + n);
    for (uint64_t i = 0; i < m; ++i) out[i] = Utility::interpolate(x0[i], vx, vy);
}

/* ---- cones and launch (rays.cpp:145-146, 233-239; cone.h; integrator.h:124-382) ---- */
using Point = std::array<double, 3>;
using Sphere = HyperSphere<3, Point>;

static std::vector<Cone<Point>> make_cones(const double* c7, uint32_t n) {
    std::vector<Cone<Point>> v(n);
    for (uint32_t i = 0; i < n; ++i) {
        Point vx{{c7[7 * i + 0], c7[7 * i + 1], c7[7 * i + 2]}};
        Point bs{{c7[7 * i + 3], c7[7 * i + 4], c7[7 * i + 5]}};
        v[i].assign(vx, bs, c7[7 * i + 6]);
    }
    return v;
}

static void photon_to9(const Photon<real, 3>& p, double* o) {
    o[0] = p.a(); o[1] = p.t(); o[2] = p.x(); o[3] = p.y(); o[4] = p.z();
    o[5] = p.dtdl(); o[6] = p.dxdl(); o[7] = p.dydl(); o[8] = p.dzdl();
}

/* HyperSphere::uniform<2> on a sphere of the given radius at the origin:
 * positions and the (min, max) pair distances it returns. */
void ref_sphere_uniform(uint64_t n, double radius, double* xyz, double* minmax) {
    Sphere sphere(Point{{0., 0., 0.}}, radius);
    std::vector<Point> t(n);
    auto r = sphere.uniform<2>(std::begin(t), std::end(t));
    for (uint64_t i = 0; i < n; ++i) for (int d = 0; d < 3; ++d) xyz[3 * i + d] = t[i][d];
    minmax[0] = r.first; minmax[1] = r.second;
}

/* Cone::inside for n points. */
void ref_cone_inside(const double* cone7, const double* xyz, uint64_t n, uint8_t* out) {
    auto c = make_cones(cone7, 1);
    for (uint64_t i = 0; i < n; ++i) out[i] = c[0].inside(Point{{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}});
}

/* The launch loop of rays.cpp:233-239 with seed = parameters.seed + rank. */
void ref_launch_cone(const double* cones7, uint32_t ncones, uint32_t icone, double micro_radius, uint32_t seed,
                     uint64_t n, double* in9, double* rot) {
    auto cones = make_cones(cones7, ncones);
    Sphere micro(Point{{0., 0., 0.}}, micro_radius);
    std::mt19937 engine1(seed);
    std::uniform_real_distribution<real> distribution(0, 1);
    const real pi = Constants<real>::pi();
    const unsigned two = 2;
    for (uint64_t i = 0; i < n; ++i) {
        auto p = Integrator::launch(micro, cones[icone], cones, engine1, distribution);
        photon_to9(p, in9 + 9 * i);
        rot[i] = distribution(engine1) * two * pi;
    }
}

void ref_launch_angles(const double* obs, const double* phi_theta, uint64_t n, double* in9) {
    for (uint64_t i = 0; i < n; ++i)
        photon_to9(Integrator::launch(obs[0], obs[1], obs[2], phi_theta[2 * i], phi_theta[2 * i + 1]), in9 + 9 * i);
}

void ref_launch_towards(const double* b, const double* e, uint64_t n, double* in9) {
    for (uint64_t i = 0; i < n; ++i)
        photon_to9(Integrator::launch(b[3 * i], b[3 * i + 1], b[3 * i + 2], e[3 * i], e[3 * i + 1], e[3 * i + 2]),
                   in9 + 9 * i);
}

void ref_launch_bundle(const double* in9, uint32_t count, double angle, double rotation, int center, double* out9) {
    Photon<real, 3> p;
    p.a() = in9[0]; p.t() = in9[1]; p.x() = in9[2]; p.y() = in9[3]; p.z() = in9[4];
    p.dtdl() = in9[5]; p.dxdl() = in9[6]; p.dydl() = in9[7]; p.dzdl() = in9[8];
    if (center) {
        auto v = Integrator::launch<true>(p, count, angle, rotation);
        for (size_t i = 0; i < v.size(); ++i) photon_to9(v[i], out9 + 9 * i);
    } else {
        auto v = Integrator::launch<false>(p, count, angle, rotation);
        for (size_t i = 0; i < v.size(); ++i) photon_to9(v[i], out9 + 9 * i);
    }
}

/* Integrator::propagate<ORDER> (integrator.h:826-1015) of one central photon:
 * rows (O_NREC doubles) and distance per element of the central trajectory,
 * up to maxlen; returns the element count. */
uint64_t ref_propagate(void* h, int order, const double* cosmo_tab, uint64_t ncos, double length, unsigned nsteps,
                       const double* in9, unsigned count, double angle, double rotation, const char* interp,
                       double amin, double* rows, double* dist, uint64_t maxlen) {
    const Octree& oc = static_cast<RefOctree*>(h)->octree;
    const Cosmo cosmo = make_cosmo(cosmo_tab, ncos);
    Photon<real, 3> p;
    p.a() = in9[0]; p.t() = in9[1]; p.x() = in9[2]; p.y() = in9[3]; p.z() = in9[4];
    p.dtdl() = in9[5]; p.dxdl() = in9[6]; p.dydl() = in9[7]; p.dzdl() = in9[8];
    const std::array<double, 3> vobs{{0., 0., 0.}};
    Traj homogeneous;
    Traj r;
    switch (order) {
        case 0: r = Integrator::propagate<0>(p, count, angle, rotation, std::string(interp), cosmo, oc, vobs, length,
                                             nsteps, amin, std::string(), homogeneous); break;
        case 1: r = Integrator::propagate<1>(p, count, angle, rotation, std::string(interp), cosmo, oc, vobs, length,
                                             nsteps, amin, std::string(), homogeneous); break;
        default: r = Integrator::propagate<2>(p, count, angle, rotation, std::string(interp), cosmo, oc, vobs, length,
                                              nsteps, amin, std::string(), homogeneous); break;
    }
    for (uint64_t i = 0; i < r.size() && i < maxlen; ++i) {
        photon_to_rec(r[i], rows + O_NREC * i, double(i));
        dist[i] = r[i].distance();
    }
    return r.size();
}

/* Lensing::dbetadtheta (lensing.h:140-364) with the parameters it reads
 * (stop_bundle, openingmin, nsteps). */
struct BundleParams {
    std::string stop_bundle;
    double openingmin;
    unsigned nsteps;
};

void ref_dbetadtheta(void* h, int order, const double* cosmo_tab, uint64_t ncos, double length, unsigned nsteps,
                     const char* stop, double opening, const double* observer, double phi, double theta,
                     const double* ref, const double* ki, const double* dist, uint64_t nd, double* out4) {
    const Octree& oc = static_cast<RefOctree*>(h)->octree;
    const Cosmo cosmo = make_cosmo(cosmo_tab, ncos);
    BundleParams P{std::string(stop), opening, nsteps};
    std::vector<Point> kis(nd);
    std::vector<double> refs(ref, ref + nd), dists(dist, dist + nd);
    for (uint64_t r = 0; r < nd; ++r) kis[r] = Point{{ki[3 * r], ki[3 * r + 1], ki[3 * r + 2]}};
    const Point obs{{observer[0], observer[1], observer[2]}};
    const Point vobs{{0., 0., 0.}};
    std::vector<std::array<std::array<double, 2>, 2>> j;
    switch (order) {
        case 1: j = Lensing::dbetadtheta<1>(P, kis, refs, obs, phi, theta, dists, cosmo, oc, vobs, length); break;
        default: j = Lensing::dbetadtheta<2>(P, kis, refs, obs, phi, theta, dists, cosmo, oc, vobs, length); break;
    }
    for (uint64_t r = 0; r < nd; ++r) {
        out4[4 * r] = j[r][0][0]; out4[4 * r + 1] = j[r][0][1]; out4[4 * r + 2] = j[r][1][0]; out4[4 * r + 3] = j[r][1][1];
    }
}

/* Utility::invMatrix3d, row-major; Constants<double>::deg() and pi(). */
void ref_inv3(const double* a, double* r) {
    std::array<std::array<double, 3>, 3> m;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] = a[3 * i + j];
    auto x = Utility::invMatrix3d(m);
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r[3 * i + j] = x[i][j];
}
double ref_deg() { return magrathea::Constants<double>::deg(); }
double ref_pi() { return magrathea::Constants<double>::pi(); }

/* ---- octree preparation (create_octree) ---- */

/* Index::compute<float> of the ASCII importer (input.h:1058-1059). */
void ref_prep_keys(const uint32_t* level, const float* xyz, uint64_t n, uint64_t* key2) {
    for (uint64_t i = 0; i < n; ++i) {
        const Index k = Index::template compute<float, position, extent>(level[i], xyz[3 * i], xyz[3 * i + 1],
                                                                          xyz[3 * i + 2]);
        indexing v;
        std::memcpy(&v, &k, sizeof(v));
        key2[2 * i] = (uint64_t)v;
        key2[2 * i + 1] = (uint64_t)(v >> 64);
    }
}

/* Utility::collide of a cell's circumscribed sphere (its position and
 * extent * sqrt(3) / 2, as Input::collide builds it) with the central sphere
 * and with the cone. */
void ref_prep_collide(const uint64_t* key2, uint64_t n, double micro_radius, const double* cone7, uint8_t* out) {
    const std::vector<Cone<Point>> cone = make_cones(cone7, 1);
    const Sphere micro(Point{{0.0, 0.0, 0.0}}, micro_radius);
    for (uint64_t i = 0; i < n; ++i) {
        const indexing v = (((indexing)key2[2 * i + 1]) << 64) | key2[2 * i];
        Index k;
        std::memcpy((void*)&k, &v, sizeof(v));
        Sphere s;
        for (unsigned d = 0; d < 3; ++d) s.position(d) = k.template position<real, position, extent>(d);
        s.extent() = k.template extent<real, position, extent>() * std::sqrt(real(3)) / real(2);
        out[i] = (Utility::collide(s, micro) || Utility::collide(s, cone[0], micro.extent())) ? 1 : 0;
    }
}

/* A container of n records in the given order with one more record past its
 * end (a copy of the last one, so the update's scan past the end reads a
 * duplicate of the last cell). */
static void load_container(Octree& o, const void* rec, uint64_t n) {
    o.reserve(n + 1);
    o.resize(n + 1);
    if (n) {
        std::memcpy((void*)o.data(), rec, n * sizeof(Element));
        std::memcpy((void*)(o.data() + n), (const char*)rec + (n - 1) * sizeof(Element), sizeof(Element));
    }
    o.resize(n);
}

/* SimpleHyperOctree::update on the records; returns the new count. */
uint64_t ref_prep_update(void* rec, uint64_t n) {
    Octree o;
    load_container(o, rec, n);
    o.update();
    std::memcpy(rec, (const void*)o.data(), o.size() * sizeof(Element));
    return o.size();
}

/* coarsen(begin + i) for each listed i, then update. */
uint64_t ref_prep_coarsen(void* rec, uint64_t n, const uint64_t* idx, uint64_t m) {
    Octree o;
    load_container(o, rec, n);
    for (uint64_t j = 0; j < m; ++j)
        if (!std::get<0>(o[idx[j]]).invalidated()) o.coarsen(o.begin() + idx[j]);
    o.update();
    std::memcpy(rec, (const void*)o.data(), o.size() * sizeof(Element));
    return o.size();
}

/* find / leaf on the container as given: found[i] = position or n. */
void ref_prep_find(const void* rec, uint64_t n, const uint64_t* key2, uint64_t m, uint64_t* found, uint8_t* leaf) {
    Octree o;
    load_container(o, rec, n);
    for (uint64_t i = 0; i < m; ++i) {
        const indexing v = (((indexing)key2[2 * i + 1]) << 64) | key2[2 * i];
        Index k;
        std::memcpy((void*)&k, &v, sizeof(v));
        found[i] = (uint64_t)std::distance(o.begin(), o.find(k));
    }
    for (uint64_t i = 0; i < n; ++i) leaf[i] = o.leaf(o.begin() + i) ? 1 : 0;
}

/* Uncached ngp / cic / tsc (order 0 / 1 / 2) on the container as given. */
void ref_prep_interpolate(const void* rec, uint64_t n, int order, const double* xyz, uint64_t m, float* out7) {
    Octree o;
    load_container(o, rec, n);
    for (uint64_t i = 0; i < m; ++i) {
        const double* p = xyz + 3 * i;
        const Data d = (order == 0) ? o.ngp(p[0], p[1], p[2]) : (order == 1) ? o.cic(p[0], p[1], p[2])
                                                                              : o.tsc(p[0], p[1], p[2]);
        data_to7(d, out7 + 7 * i);
    }
}

/* Utility::interpolate2 with a float abscissa (miscellaneous.h:477). */
void ref_prep_interpolate2(const float* x0, uint64_t m, const double* x, const double* y, uint64_t n, float* out) {
    const std::vector<double> vx(x, x + n), vy(y, y + n);
    for (uint64_t i = 0; i < m; ++i) out[i] = Utility::interpolate2(x0[i], vx, vy);
}

}  // extern "C"
