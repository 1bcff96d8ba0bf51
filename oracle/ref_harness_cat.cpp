/* Reference harness for the catalogue, cosmology and cone code: a thin
 * extern "C" shim over the reference's catalogues.h, input.h and
 * miscellaneous.h, compiled from the sources where they lie by
 * oracle/Makefile into oracle/_ref/libmpfrefcat.so.
 *
 * TEST INFRASTRUCTURE ONLY.  This file holds no reference source: it
 * instantiates the reference templates with the types of catalogues.cpp:85-91
 * and 123 and marshals plain arrays in and out.  It pins the C restatement in
 * oracle/mpf_oracle.c and the host code of the product, and writes the golden
 * fixtures under tests/golden/.  Product code never links it.
 *
 * The headers include mpi.h and TReadHDF5.h: they compile against the MPICH
 * and HDF5 headers of /opt/conda, and the library links libhdf5 by full path
 * (TReadHDF5.h defines non-template functions that call it).  No MPI or HDF5
 * function is ever called here: the loader (oracle/oracle.py) maps libhdf5
 * before this library so its dependency resolves without changing the
 * search path of the C++ runtime.
 */
#include <mpi.h>

#include <cstdint>
#include <cstring>
#include <iostream>
#include <streambuf>
#include <string>
#include <vector>

#include "integrator.h"
#include "miscellaneous.h"
#include "catalogues.h"
#include "gravity.h"

using namespace magrathea;

namespace {
using real = double;
using floating = float;
using point = std::array<real, 3>;
using position = std::ratio<0>;
using extent = std::ratio<EXTENT>;
using indexing = __uint128_t;
using Index = SimpleHyperOctreeIndex<indexing, 3>;
using Data = Gravity<floating, 3>;
using Octree = SimpleHyperOctree<real, Index, Data, 3, position, extent>;
using Element = std::pair<Index, Data>;
using Traj = Evolution<Photon<real, 3>>;
using Cosmo = std::array<std::vector<real>, 4>;
using Sphere = HyperSphere<3, point>;

static_assert(sizeof(Element) == 48, "record size");

struct RefOctree {
    Octree octree;
};

Cosmo make_cosmo(const double* tab, uint64_t n) {
    Cosmo c;
    for (int k = 0; k < 4; ++k) c[k].assign(tab + k * n, tab + (k + 1) * n);
    return c;
}

/* A stream buffer that drops everything: the root-finder's re-run mode and
 * the narrow-cone generator print progress lines to std::cout. */
struct NullBuf : std::streambuf {
    int overflow(int c) override { return c; }
};

struct Quiet {
    NullBuf buf;
    std::streambuf* old;
    Quiet() : old(std::cout.rdbuf(&buf)) {}
    ~Quiet() { std::cout.rdbuf(old); }
};

const char* jactype_name(int j) {
    switch (j) {
        case 1: return "infinitesimal_born";
        case 2: return "bundle";
        default: return "infinitesimal";
    }
}

/* The fields of catalogues.h:59-96 that the root-finder, the bundle
 * Jacobian and relCat read. */
parameters_t make_params(double accuracy, unsigned nsteps, int jactype, int plane, const char* stop_bundle,
                         double opening, int use_prev, int isfullsky, double mpc) {
    parameters_t P{};
    P.cat_accuracy = accuracy;
    P.nsteps = nsteps;
    P.jacobiantype = jactype_name(jactype);
    P.plane = plane == 1 ? "normal" : "sachs";
    P.stop_bundle = stop_bundle ? stop_bundle : "redshift";
    P.openingmin = opening;
    P.use_previous_catalogues = (uint)use_prev;
    P.isfullsky = (uint)isfullsky;
    P.mpc = mpc;
    return P;
}

std::vector<Cone<point>> make_cones(const double* c7, uint32_t n) {
    std::vector<Cone<point>> v(n);
    for (uint32_t i = 0; i < n; ++i) {
        point vx{{c7[7 * i + 0], c7[7 * i + 1], c7[7 * i + 2]}};
        point bs{{c7[7 * i + 3], c7[7 * i + 4], c7[7 * i + 5]}};
        v[i].assign(vx, bs, c7[7 * i + 6]);
    }
    return v;
}

void cone_out(const Cone<point>& c, double* o) {
    for (unsigned d = 0; d < 3; ++d) {
        o[d] = c.vertex(d);
        o[3 + d] = c.base(d);
    }
    o[6] = c.angle();
}

template <int Order>
void newton_one(const Octree& oc, const Cosmo& cosmo, double length, double hubble, const parameters_t& P,
                const point& vobs, const double* s6, const double* init2, double* o) {
    const point observer{{0., 0., 0.}};
    const point target{{s6[0], s6[1], s6[2]}};
    const point velocity{{s6[3], s6[4], s6[5]}};
    std::array<std::array<double, 2>, 2> jacobian{};
    std::vector<double> redshifts(6);
    double interpRef = 0;
    double phi, theta;
    if (init2) {
        phi = init2[0];
        theta = init2[1];
    } else {  // the unlensed direction, as relCat computes it (catalogues.h:832-836)
        const double d = std::sqrt(target[0] * target[0] + target[1] * target[1] + target[2] * target[2]);
        theta = std::acos(target[2] / d);
        phi = std::atan2(target[1], target[0]);
    }
    const auto r = Catalogues::iterateNewtonMethod<Order, true, false>(
        vobs, observer, phi, theta, target, velocity, jacobian, P, cosmo, oc, length, hubble, redshifts, interpRef);
    o[0] = r[0][0]; o[1] = r[0][1]; o[2] = r[1][0]; o[3] = r[1][1];
    o[4] = jacobian[0][0]; o[5] = jacobian[0][1]; o[6] = jacobian[1][0]; o[7] = jacobian[1][1];
    for (int k = 0; k < 6; ++k) o[8 + k] = redshifts[k];
    o[14] = interpRef;
    o[15] = 0;
}
}  // namespace

extern "C" {

void* refc_octree_new(const void* rec, uint64_t n) {
    RefOctree* o = new RefOctree;
    o->octree.resize(n);
    if (n) std::memcpy((void*)o->octree.data(), rec, n * sizeof(Element));
    return o;
}

void refc_octree_free(void* h) { delete static_cast<RefOctree*>(h); }

/* Catalogues::iterateNewtonMethod (catalogues.h:757-800) per source from
 * the angles init2 (NULL: the unlensed direction), with the parameters of the root-finder: jactype 0
 * infinitesimal, 1 infinitesimal_born, 2 bundle (plane 0 sachs, 1 normal;
 * stop_bundle; openingmin), use_previous_catalogues use_prev.
 * src6: n x (x, y, z, vx, vy, vz) in the octree frame.  out: n x 16 doubles
 * (result[0][0], result[0][1], result[1][0], result[1][1], jacobian 2 x 2,
 * redshifts[0..5], interpRef, 0).  Sources run serially when nthreads == 1,
 * else through Utility::parallelize. */
void refc_newton(void* h, int order, const double* cosmo_tab, uint64_t ncos, double length, double hubble,
                 double mpc, unsigned nsteps, double accuracy, int jactype, int plane, const char* stop_bundle,
                 double opening, int use_prev, const double* vobs3, const double* src6, const double* init2,
                 uint64_t n, double* out, int nthreads) {
    Quiet quiet;
    const Octree& oc = static_cast<RefOctree*>(h)->octree;
    const Cosmo cosmo = make_cosmo(cosmo_tab, ncos);
    const parameters_t P = make_params(accuracy, nsteps, jactype, plane, stop_bundle, opening, use_prev, 1, mpc);
    const point vobs{{vobs3[0], vobs3[1], vobs3[2]}};
    auto body = [&](const uint64_t i) {
        const double* a = init2 ? init2 + 2 * i : nullptr;
        if (order == 1)
            newton_one<1>(oc, cosmo, length, hubble, P, vobs, src6 + 6 * i, a, out + 16 * i);
        else
            newton_one<2>(oc, cosmo, length, hubble, P, vobs, src6 + 6 * i, a, out + 16 * i);
    };
    if (nthreads == 1) {
        for (uint64_t i = 0; i < n; ++i) body(i);
    } else {
        const int nt = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
        Utility::parallelize(n, body, nt);
    }
}

/* One call of Catalogues::newtonMethod2d (catalogues.h:410-754) per source
 * at iteration `iteration` from the angles ang2, the Jacobian starting from
 * jac4 (n x 4, or NULL: zero), ORDER = 2.  out: n x 16 as refc_newton. */
void refc_newton_once(void* h, const double* cosmo_tab, uint64_t ncos, double length, double hubble, double mpc,
                      unsigned nsteps, double accuracy, int jactype, int plane, const char* stop_bundle,
                      double opening, int use_prev, const double* vobs3, const double* src6, const double* ang2,
                      const double* jac4, unsigned iteration, uint64_t n, double* out) {
    Quiet quiet;
    const Octree& oc = static_cast<RefOctree*>(h)->octree;
    const Cosmo cosmo = make_cosmo(cosmo_tab, ncos);
    const parameters_t P = make_params(accuracy, nsteps, jactype, plane, stop_bundle, opening, use_prev, 1, mpc);
    const point vobs{{vobs3[0], vobs3[1], vobs3[2]}};
    const point observer{{0., 0., 0.}};
    for (uint64_t i = 0; i < n; ++i) {
        const double* s6 = src6 + 6 * i;
        const point target{{s6[0], s6[1], s6[2]}};
        const point velocity{{s6[3], s6[4], s6[5]}};
        std::array<std::array<double, 2>, 2> jacobian{};
        if (jac4) jacobian = {{{{jac4[4 * i], jac4[4 * i + 1]}}, {{jac4[4 * i + 2], jac4[4 * i + 3]}}}};
        std::vector<double> redshifts(6);
        double interpRef = 0;
        const std::array<double, 2> ang{{ang2[2 * i], ang2[2 * i + 1]}};
        const auto r = Catalogues::newtonMethod2d<2, true, false>(vobs, observer, target, ang, velocity, jacobian, P,
                                                                  cosmo, oc, length, hubble, redshifts, interpRef,
                                                                  iteration);
        double* o = out + 16 * i;
        o[0] = r[0][0]; o[1] = r[0][1]; o[2] = r[1][0]; o[3] = r[1][1];
        o[4] = jacobian[0][0]; o[5] = jacobian[0][1]; o[6] = jacobian[1][0]; o[7] = jacobian[1][1];
        for (int k = 0; k < 6; ++k) o[8 + k] = redshifts[k];
        o[14] = interpRef;
        o[15] = 0;
    }
}

/* Catalogues::relCat (catalogues.h:803-920) on n targets of 8 doubles (x, y,
 * z, vx, vy, vz, id, npart), writing <filename>.txt, <filename>_err.txt and
 * <filename>_reject.txt.  rotm1 (3 x 3 row-major) rotates the targets of a
 * narrow cone (isfullsky 0); NULL: full sky.  prev18 (nprev rows of the 18
 * catalogue columns, nprev == n, or NULL): the previous catalogue of the
 * re-run mode (use_prev 2).  The root-finder runs ORDER = 2 (TSC). */
void refc_relcat(void* h, const double* cosmo_tab, uint64_t ncos, double length, double hubble, double mpc,
                 unsigned nsteps, double accuracy, int jactype, int plane, const char* stop_bundle, double opening,
                 int use_prev, const double* vobs3, const double* rotm1, const double* targets8, uint64_t n,
                 const double* prev18, uint64_t nprev, const char* filename) {
    Quiet quiet;
    const Octree& oc = static_cast<RefOctree*>(h)->octree;
    const Cosmo cosmo = make_cosmo(cosmo_tab, ncos);
    const parameters_t P =
        make_params(accuracy, nsteps, jactype, plane, stop_bundle, opening, use_prev, rotm1 ? 0 : 1, mpc);
    const point vobs{{vobs3[0], vobs3[1], vobs3[2]}};
    const point observer{{0., 0., 0.}};
    std::array<std::array<double, 3>, 3> rot{};
    if (rotm1)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) rot[i][j] = rotm1[3 * i + j];
    std::vector<std::array<double, 8>> targets(n);
    for (uint64_t i = 0; i < n; ++i)
        for (int k = 0; k < 8; ++k) targets[i][k] = targets8[8 * i + k];
    std::vector<std::array<double, 18>> prev(prev18 ? nprev : 0);
    for (uint64_t i = 0; i < prev.size(); ++i)
        for (int k = 0; k < 18; ++k) prev[i][k] = prev18[18 * i + k];
    std::string name(filename);
    Catalogues::relCat(vobs, rot, name, observer, targets, prev, P, cosmo, oc, length, hubble);
}

/* Catalogues::relCat_with_previous_cat (catalogues.h:922-1044): the
 * Jacobian columns 13-16 of each row of prev18 (n x 18, in place) recomputed
 * from its observed (or, for infinitesimal_born, comoving) angles to the
 * scale factor of its FLRW redshift; writes <filename>.txt and
 * <filename>_err.txt. */
void refc_relcat_prev(void* h, const double* cosmo_tab, uint64_t ncos, double length, double hubble, double mpc,
                      unsigned nsteps, int jactype, int plane, const char* stop_bundle, double opening,
                      const double* vobs3, double* prev18, uint64_t n, const char* filename) {
    Quiet quiet;
    const Octree& oc = static_cast<RefOctree*>(h)->octree;
    const Cosmo cosmo = make_cosmo(cosmo_tab, ncos);
    const parameters_t P = make_params(1e-8, nsteps, jactype, plane, stop_bundle, opening, 1, 1, mpc);
    const point vobs{{vobs3[0], vobs3[1], vobs3[2]}};
    const point observer{{0., 0., 0.}};
    std::vector<std::array<double, 18>> prev(n);
    for (uint64_t i = 0; i < n; ++i)
        for (int k = 0; k < 18; ++k) prev[i][k] = prev18[18 * i + k];
    std::string name(filename);
    Catalogues::relCat_with_previous_cat(vobs, name, observer, prev, P, cosmo, oc, length, hubble);
    for (uint64_t i = 0; i < n; ++i)
        for (int k = 0; k < 18; ++k) prev18[18 * i + k] = prev[i][k];
}

/* Input::acquire (input.h:1091-1233) with the parameters it reads (evolfile,
 * paramfile, celldir, mpc).  hol: h, omega_m, lboxmpch.  tab: 4 columns of
 * capacity rows (t, a, da/dt, d2a/dt2); *n the row count. */
void refc_acquire(const char* evolfile, const char* paramfile, const char* celldir, double mpc, double* hol,
                  double* tab, uint64_t capacity, uint64_t* n) {
    parameters_t P{};
    P.evolfile = evolfile;
    P.paramfile = paramfile;
    P.celldir = celldir;
    P.mpc = mpc;
    double h = 0, om = 0, lbox = 0;
    const Cosmo c = Input::acquire(P, h, om, lbox);
    hol[0] = h; hol[1] = om; hol[2] = lbox;
    *n = c[0].size();
    if (*n <= capacity)
        for (int k = 0; k < 4; ++k)
            for (uint64_t i = 0; i < *n; ++i) tab[k * capacity + i] = c[k][i];
}

/* Input::correct(cosmology, trajectory) (input.h:1330-1378) with a
 * trajectory of m elements carrying t and a.  tab4 (4 x n, column-major as
 * refc_acquire writes it with capacity n) is corrected into out4. */
void refc_correct(const double* tab4, uint64_t n, const double* traj_t, const double* traj_a, uint64_t m,
                  double* out4) {
    const Cosmo c = make_cosmo(tab4, n);
    Traj traj;
    for (uint64_t i = 0; i < m; ++i) {
        Photon<real, 3> p;
        p.t() = traj_t[i];
        p.a() = traj_a[i];
        traj.append(p);
    }
    const Cosmo r = Input::correct(c, traj);
    for (int k = 0; k < 4; ++k)
        for (uint64_t i = 0; i < n; ++i) out4[k * n + i] = r[k][i];
}

/* Miscellaneous::GenerateFullskyCones (miscellaneous.h:200-218) on the
 * sphere of catalogues.cpp:124 (centre 0, radius EXTENT / 2). */
void refc_fullsky_cones(uint32_t ncones, double* cones7) {
    Sphere sphere(point{{0., 0., 0.}}, real(EXTENT) / 2);
    std::vector<Cone<point>> cone(ncones), rot(ncones);
    Miscellaneous::GenerateFullskyCones(ncones, cone, rot, sphere);
    for (uint32_t i = 0; i < ncones; ++i) cone_out(cone[i], cones7 + 7 * i);
}

/* Miscellaneous::GenerateNarrowCones (miscellaneous.h:257-412), ASCII input
 * (typefile 2): celldir (ending in '/') holds info_narrow_cone.txt.  Writes
 * the cones, the rotated cones, rotm1 (3 x 3 row-major) and thetay, thetaz
 * (rad). */
void refc_narrow_cones(const char* celldir, uint32_t ncones, double* cones7, double* conesrot7, double* rotm1,
                       double* thetayz) {
    Quiet quiet;
    parameters_t P{};
    P.celldir = celldir;
    P.ncones = ncones;
    P.typefile = 2;
    Sphere sphere(point{{0., 0., 0.}}, real(EXTENT) / 2);
    std::vector<Cone<point>> cone(ncones), rot(ncones);
    std::array<std::array<double, 3>, 3> r{};
    double ty = 0, tz = 0;
    Miscellaneous::GenerateNarrowCones(P, cone, rot, sphere, r, ty, tz);
    for (uint32_t i = 0; i < ncones; ++i) {
        cone_out(cone[i], cones7 + 7 * i);
        cone_out(rot[i], conesrot7 + 7 * i);
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) rotm1[3 * i + j] = r[i][j];
    thetayz[0] = ty;
    thetayz[1] = tz;
}

/* Miscellaneous::getTargets (miscellaneous.h:523-600): the rows of src8
 * (n x 8) that cone icone owns, in their order, into out8; returns the
 * count. */
uint64_t refc_get_targets(const double* src8, uint64_t n, const double* cones7, uint32_t ncones, uint32_t icone,
                          double* out8) {
    std::vector<std::array<double, 8>> src(n);
    for (uint64_t i = 0; i < n; ++i)
        for (int k = 0; k < 8; ++k) src[i][k] = src8[8 * i + k];
    const std::vector<Cone<point>> cones = make_cones(cones7, ncones);
    const auto sel = Miscellaneous::getTargets(src, cones[icone], cones);
    for (uint64_t i = 0; i < sel.size(); ++i)
        for (int k = 0; k < 8; ++k) out8[8 * i + k] = sel[i][k];
    return sel.size();
}

/* ---- octree preparation (create_octree's ASCII branch) ---- */

/* The parameters Input::importascii, Input::sistemize and Input::correct read
 * (Create_octree::ReadParamFile, create_octree.h:120-146). */
struct PrepParams {
    unsigned int isfullsky = 1, coarseonly = 0, ncoarse = 0, correction = 1, coarsecorrection = 1, acorrection = 0;
    double mpc = 3.08568e22;
};

/* Input::importascii<position, extent> (input.h:997-1069) of one ASCII cell
 * file into an empty octree, with the filter of
 * Create_octree::PreparationASCII (create_octree.h:560: Input::collide with
 * the central sphere of radius micro_radius at the origin and the cone
 * cone7); rotm1: 3 x 3 row-major, used when isfullsky == 0.  The selected
 * records (at most capacity) go to out48; returns their count. */
uint64_t refc_prep_import(const char* path, const double* cone7, double micro_radius, int isfullsky,
                          const double* rotm1, int coarseonly, unsigned ncoarse, void* out48, uint64_t capacity) {
    Octree oc;
    PrepParams P;
    P.isfullsky = (unsigned)isfullsky;
    P.coarseonly = (unsigned)coarseonly;
    P.ncoarse = ncoarse;
    std::array<std::array<double, 3>, 3> R{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R[i][j] = rotm1 ? rotm1[3 * i + j] : (i == j ? 1.0 : 0.0);
    const std::vector<Cone<point>> cones = make_cones(cone7, 1);
    const Sphere micro(point{{0.0, 0.0, 0.0}}, micro_radius);
    Input::importascii<position, extent>(P, R, oc, std::string(path), [&](const Element& e) {
        return Input::collide(oc, std::get<0>(e), micro, cones[0]);
    });
    const uint64_t n = oc.size();
    if (n) std::memcpy(out48, (const void*)oc.data(), std::min<uint64_t>(n, capacity) * sizeof(Element));
    return n;
}

/* Miscellaneous::correctOctree (miscellaneous.h:467-479: Input::sistemize,
 * dphidt = 0, Input::correct(parameters, octree, amin), dphi/da to dphi/dt
 * through Utility::interpolate2) on n records in the given order, with the
 * cosmology table cosmo_tab (4 x ncos: t, a, da/dt, d2a/dt2).  The container
 * holds `reserve` records of storage, the spare ones copies of filler48
 * (SimpleHyperOctree::update reads the element past the end of its
 * container, simplehyperoctree.h:1285-1300: a copy of the last cell there
 * leaves that cell as it is).  The corrected records (at most capacity) go
 * to out48; returns their count. */
uint64_t refc_prep_correct(const void* rec48, uint64_t n, const void* filler48, uint64_t reserve, int correction,
                           int coarsecorrection, double h, double omegam, double lboxmpch, double mpc,
                           const double* cosmo_tab, uint64_t ncos, void* out48, uint64_t capacity) {
    Octree oc;
    if (reserve < n + 1) reserve = n + 1;
    oc.reserve(reserve);
    oc.resize(reserve);
    for (uint64_t i = n; i < reserve; ++i) std::memcpy((void*)(oc.data() + i), filler48, sizeof(Element));
    if (n) std::memcpy((void*)oc.data(), rec48, n * sizeof(Element));
    oc.resize(n);
    PrepParams P;
    P.correction = (unsigned)correction;
    P.coarsecorrection = (unsigned)coarsecorrection;
    P.mpc = mpc;
    const Cosmo cosmo = make_cosmo(cosmo_tab, ncos);
    double amin = 0.0;
    Miscellaneous::correctOctree(oc, cosmo, P, h, omegam, lboxmpch, amin);
    const uint64_t m = oc.size();
    if (m) std::memcpy(out48, (const void*)oc.data(), std::min<uint64_t>(m, capacity) * sizeof(Element));
    return m;
}

}  // extern "C"
