// Octree preparation of the reference's create_octree program (host): cells
// read from the 10-column ASCII files, kept when they touch the central
// sphere or the cone (Input::importascii, input.h:997-1069, with
// Input::collide, input.h:216-226), converted to SI units and corrected
// (Miscellaneous::correctOctree, miscellaneous.h:467-479; Input::correct,
// input.h:1382-1520), and written as a cone file of 48-byte records
// (Input::save, input.h:1276-1289).  The octree operations the correction
// uses (update, find, leaf, coarsen and the uncached NGP/CIC/TSC of
// simplehyperoctree.h) are restated on a flat container with the reference's
// search semantics, including an unsorted tail of appended cells.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "mpf_cone.h"
#include "mpf_key.h"

namespace mpf {
namespace prep {

// Gravity<float> (gravity.h:45-87) in tuple order
struct Grav {
    float rho, phi, dx, dy, dz, a, dt;
};

struct Elem {
    key::u128 k;
    Grav d;
};

struct Config {
    int isfullsky = 1;          // 0: positions and forces rotated by rotm1 (narrow cone)
    int coarseonly = 0;         // keep only the cells of level ncoarse + log2(EXTENT)
    unsigned ncoarse = 0;
    int correction = 1, coarsecorrection = 0, acorrection = 0;
    int order = 2;              // interpolation of missing children: 0 NGP, 1 CIC, 2 TSC
    double microcoeff = 1.0;    // central sphere of radius 1 / microcoeff
    double h = 0.7, omegam = 0.3, lboxmpch = 1.0, mpc = 3.0856775814913673e22;
    uint64_t allocation = 0;    // capacity reserved for the container
};

class Builder {
public:
    explicit Builder(const Config& c) : cfg_(c) { v_.reserve(c.allocation); }
    const Config& config() const { return cfg_; }
    std::vector<Elem>& cells() { return v_; }
    const std::vector<Elem>& cells() const { return v_; }
    void clear() { v_.clear(); }

    // Input::importascii: appends the selected cells of one file
    bool import_ascii(const std::string& path, const cone::Cone& conic, const double* rotm1, uint64_t& added,
                      std::string& err);
    // Miscellaneous::correctOctree with the corrected cosmology table (t, a, da/dt)
    bool correct(const std::vector<double>& t, const std::vector<double>& a, const std::vector<double>& dadt,
                 double& amin, std::string& err);

    // SimpleHyperOctree operations (restated)
    void update();
    uint64_t find(key::u128 k) const;               // index of the containing element, or size()
    bool leaf(uint64_t i) const;
    void coarsen(uint64_t i);
    Grav ngp(double x, double y, double z) const;
    Grav cic(double x, double y, double z) const;
    Grav tsc(double x, double y, double z) const;

    void to_records(unsigned char* rec48) const;
    void from_records(const unsigned char* rec48, uint64_t n);
    // Input::save: refuses an existing file
    bool save(const std::string& path, std::string& err) const;

private:
    uint64_t upper(key::u128 k) const;
    Config cfg_;
    std::vector<Elem> v_;
};

// Key of a cell from a position at a level, computed in float (the ASCII
// reader's Real) or double (the octree's Type): Index::compute
// (simplehyperoctreeindex.h:1167-1170)
key::u128 compute_f(unsigned level, float x, float y, float z);
key::u128 compute_d(unsigned level, double x, double y, double z);

// Input::collide: the cell's circumscribed sphere touches the central
// sphere of the given radius or the cone (Utility::collide, utility.h:392-419)
bool collide(key::u128 k, double micro_radius, const cone::Cone& conic);

// Create_octree::PreparationASCII (create_octree.h:509-585) for the cones
// this rank handles: the .dat files of celldir (and of minicone for narrow
// cones) in the order of the reference's shuffles, one cone file each.
bool prepare_ascii(const Config& cfg, const std::string& celldir, const std::string& minicone,
                   const std::vector<cone::Cone>& cones_rot, const double* rotm1, unsigned ntasks, unsigned rank,
                   uint64_t seed, const std::vector<double>& t, const std::vector<double>& a,
                   const std::vector<double>& dadt, const std::vector<std::string>& conefiles, double& amin,
                   std::string& err);

}  // namespace prep
}  // namespace mpf
