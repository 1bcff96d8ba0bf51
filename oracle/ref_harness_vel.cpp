/* Reference harness of the velocity-field build: the reference's integrator
 * (integrator.h) compiled with VELOCITYFIELD, so that its octree cells carry
 * Gravity<float> of gravity2.h (rho, phi, grad phi, a, dphidt, v) and the
 * redshift of every step takes the Doppler and ISW terms
 * (integrator.h:537-539, 594-595, 707-709, 765-766).  Built by oracle/Makefile
 * into oracle/_ref/ from the reference sources where they lie.
 *
 * TEST INFRASTRUCTURE ONLY.  This file holds no reference source: it
 * instantiates the reference templates with the types hmaps.cpp uses in that
 * build (hmaps.cpp:97-99) and marshals plain arrays in and out.  It pins the
 * velocity terms of the C restatement in oracle/mpf_oracle.c and writes the
 * fixture tests/golden/velocity.npz.  Product code never links it.
 */
#include <cstdint>
#include <cstring>
#include <vector>
#include <array>
#include <string>

#include "integrator.h"

using namespace magrathea;

namespace {
using real = double;
using floating = float;
using position = std::ratio<0>;
using extent = std::ratio<EXTENT>;
using Index = SimpleHyperOctreeIndex<__uint128_t, 3>;
using Data = Gravity<floating, 3>;
using Octree = SimpleHyperOctree<real, Index, Data, 3, position, extent>;
using Element = std::pair<Index, Data>;
using Traj = Evolution<Photon<real, 3>>;
using Cosmo = std::array<std::vector<real>, 4>;

enum { NREC = 28 };

const char* stop_name(int kind) {
    static const char* names[] = {"radius", "lambda", "redshift", "t", "a", "plane", "s"};
    return (kind >= 0 && kind <= 6) ? names[kind] : "";
}

/* the 28-double per-element record of ref_harness.cpp */
void to_rec(const Photon<real, 3>& p, double* o, double nsteps) {
    const double v[NREC] = {p.a(), p.t(), p.x(), p.y(), p.z(), p.dtdl(), p.dxdl(), p.dydl(), p.dzdl(),
                            p.redshift(), p.lambda(), p.s(), p.chi(), p.isw(), p.iswold(), p.phi(),
                            p.dphidx(), p.dphidy(), p.dphidz(), p.dphidt(), p.rho(), p.ah(), p.dphidl(),
                            p.laplacian(), p.dsdl2(), p.error(), (double)p.level(), nsteps};
    std::memcpy(o, v, sizeof(v));
}

template <int Order>
void run(Traj& traj, int stop_kind, double stop_value, const Cosmo& cosmo, const Octree& oc,
         const std::array<double, 3>& vobs, double length, unsigned nsteps, const std::array<double, 3>& kit) {
    if (stop_kind < 0) Integrator::integrate<Order>(traj, cosmo, oc, vobs, length, nsteps);
    else Integrator::integrate<Order>(traj, std::string(stop_name(stop_kind)), stop_value, cosmo, oc, vobs, length,
                                      nsteps, kit);
}
}  // namespace

extern "C" {

/* n 48-byte cone records (key, then dphidt, a, dphidx, dphidy, dphidz, phi,
 * rho, pad as floats) with rho replaced by rho_override (when given) and the
 * velocity (3 floats per record, units of c) set from vel3. */
void* refv_octree_new(const void* rec, uint64_t n, const float* rho_override, const float* vel3) {
    Octree* o = new Octree;
    o->resize(n);
    const unsigned char* r = static_cast<const unsigned char*>(rec);
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t lo, hi;
        float f[8];
        std::memcpy(&lo, r + 48 * i, 8);
        std::memcpy(&hi, r + 48 * i + 8, 8);
        std::memcpy(f, r + 48 * i + 16, 32);
        Element& e = (*o)[i];
        std::get<0>(e) = Index(((__uint128_t)hi << 64) | lo);
        Data& d = std::get<1>(e);
        d.dphidt() = f[0]; d.a() = f[1]; d.dphidx() = f[2]; d.dphidy() = f[3]; d.dphidz() = f[4];
        d.phi() = f[5]; d.rho() = rho_override ? rho_override[i] : f[6];
        d.vx() = vel3 ? vel3[3 * i] : 0.f;
        d.vy() = vel3 ? vel3[3 * i + 1] : 0.f;
        d.vz() = vel3 ? vel3[3 * i + 2] : 0.f;
    }
    return o;
}

void refv_octree_free(void* h) { delete static_cast<Octree*>(h); }

/* order 0 ngp, 1 cic, 2 tsc (uncached): 10 floats per point
 * (rho, phi, dphidx, dphidy, dphidz, a, dphidt, vx, vy, vz) */
void refv_interp(void* h, int order, const double* xyz, uint64_t m, float* out10) {
    const Octree& oc = *static_cast<Octree*>(h);
    for (uint64_t i = 0; i < m; ++i) {
        const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        Data d = (order == 0) ? oc.ngp(x, y, z) : (order == 1) ? oc.cic(x, y, z) : oc.tsc(x, y, z);
        float* o = out10 + 10 * i;
        o[0] = d.rho(); o[1] = d.phi(); o[2] = d.dphidx(); o[3] = d.dphidy(); o[4] = d.dphidz();
        o[5] = d.a(); o[6] = d.dphidt(); o[7] = d.vx(); o[8] = d.vy(); o[9] = d.vz();
    }
}

/* Integrate n rays from their launch cores (9 doubles each) with the observer
 * velocity vobs3; out: the last element of each trajectory as 28 doubles
 * (nsteps -1 for an emptied trajectory). Returns the total accepted steps. */
uint64_t refv_integrate(void* h, int order, const double* cosmo_tab, uint64_t ncos, double length, unsigned nsteps,
                        int stop_kind, double stop_value, const double* kitarget, const double* vobs3,
                        const double* in9, uint64_t n, double* out) {
    const Octree& oc = *static_cast<Octree*>(h);
    Cosmo cosmo;
    for (int k = 0; k < 4; ++k) cosmo[k].assign(cosmo_tab + k * ncos, cosmo_tab + (k + 1) * ncos);
    std::array<double, 3> kit{{0., 0., 0.}}, vobs{{vobs3[0], vobs3[1], vobs3[2]}};
    if (kitarget) kit = {{kitarget[0], kitarget[1], kitarget[2]}};
    uint64_t total = 0;
    for (uint64_t i = 0; i < n; ++i) {
        Traj traj;
        Photon<real, 3> p;
        const double* c = in9 + 9 * i;
        p.a() = c[0]; p.t() = c[1]; p.x() = c[2]; p.y() = c[3]; p.z() = c[4];
        p.dtdl() = c[5]; p.dxdl() = c[6]; p.dydl() = c[7]; p.dzdl() = c[8];
        traj.append(p);
        switch (order) {
            case -1: run<-1>(traj, stop_kind, stop_value, cosmo, oc, vobs, length, nsteps, kit); break;
            case 0: run<0>(traj, stop_kind, stop_value, cosmo, oc, vobs, length, nsteps, kit); break;
            case 1: run<1>(traj, stop_kind, stop_value, cosmo, oc, vobs, length, nsteps, kit); break;
            default: run<2>(traj, stop_kind, stop_value, cosmo, oc, vobs, length, nsteps, kit); break;
        }
        double* o = out + NREC * i;
        if (traj.empty()) {
            std::memset(o, 0, NREC * sizeof(double));
            o[NREC - 1] = -1.0;
        } else {
            to_rec(traj.back(), o, double(traj.size() - 1));
            total += traj.size() - 1;
        }
    }
    return total;
}

}  // extern "C"
