// Host restatement of the reference's cosmology table (mpf_cosmology.cpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mpf {
namespace cosmo {

struct Table {
    std::vector<double> t, a, dadt, d2adt2;
};

// Input::acquire (input.h:1091-1233); false with a message when a file or a
// parameter is missing (the reference returns an empty table and zeros)
bool acquire(const std::string& evolution_file, const std::string& parameter_file, const std::string& cell_dir,
             double mpc, double& h, double& omegam, double& lboxmpch, Table& out, std::string& err);
// Input::correct(cosmology, trajectory) (input.h:1330-1378), in place on tab.a
void correct(Table& tab, const std::vector<double>& traj_t, const std::vector<double>& traj_a);
// steps per cell of the homogeneous reference photon (rays.cpp:214)
unsigned reference_nsteps(uint64_t table_size, unsigned nreference);

}  // namespace cosmo
}  // namespace mpf
