// Integrator::propagate on the device (mpf_propagate.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/mpf.h"

namespace mpf {

struct PropagateArgs {
    const double* traj = nullptr;  // (n * nb) rays x cap elements x MPF_NOUT, bundle i = rays i*nb .. i*nb+nb-1,
                                   // its central ray first
    const double* out = nullptr;   // (n * nb) x MPF_NOUT last elements (MPF_OUT_NSTEPS = elements - 1)
    uint64_t n = 0;                // bundles
    uint64_t cap = 0;              // elements per ray in traj
    unsigned nb = 1;               // rays per bundle (count + 1)
    int kind = MPF_STOP_REDSHIFT;  // interpolation parameter
    double amin = 0.0;
    double length = 1.0;           // box length in metres
    double angle = 0.0;            // bundle half-angle (rad)
    unsigned* nacc = nullptr;      // n: rays used for the distance, 0 when the bundle is rejected
    double* dist = nullptr;        // n x cap angular diameter distances of the central rays
};

hipError_t launch_bundle_accept(const PropagateArgs& P, hipStream_t st);
hipError_t launch_bundle_distance(const PropagateArgs& P, hipStream_t st);

}  // namespace mpf
