// Split-precision tier: FP32 perturbation, FP64 background.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>

#include "mpf_kernels.h"

namespace mpf {

struct SplitTree {
    void* mem = nullptr;
    size_t bytes = 0;
    uint64_t nslots = 0;
    uint32_t* tslot = nullptr;        // cell -> stencil block slot, NIL when the stencil is missing
    unsigned char* blocks = nullptr;  // per slot: FP64 centre, 1/size, 5 x 27 FP32 coefficients
    uint2* nbr = nullptr;             // cell x 6 faces -> same-level neighbour (see neighbour_kernel)
    size_t nbr_bytes = 0;
    float* rho = nullptr;             // per slot: 27 FP32 coefficients of rho (+ pad), built for surfaces
    // tidal blocks (built for the Jacobian): slot -> index into `tidal`, or
    // NIL when the slot's cell does not qualify (split_build_tidal)
    uint32_t* tidx = nullptr;
    float* tidal = nullptr;           // per index: 6 x 28 FP32 coefficients
    uint64_t ntidal = 0;
    std::vector<uint32_t> slot_cell;  // host: slot -> its stencil-centre cell
};

// level/kx/ky/kz: host arrays of the level-sorted cells
bool split_build(SplitTree& s, const DevTree& t, const uint8_t* level, const uint32_t* kx, const uint32_t* ky,
                 const uint32_t* kz, int order, hipStream_t st, std::string& err);
// the rho blocks (FP32 TSC polynomial of rho per slot) of a built tree, for
// the density entries of surfaces; level/kx/ky/kz as for split_build
bool split_build_rho(SplitTree& s, const DevTree& t, const uint8_t* level, const uint32_t* kx, const uint32_t* ky,
                     const uint32_t* kz, hipStream_t st, std::string& err);
// the tidal blocks of a built tree, for the infinitesimal Jacobian: for a
// slot whose cell is a leaf, its own stencil centre, with the six same-level
// face neighbours leaves and their own stencil centres, the differences of
// grad phi between the samples one cell away on either side (the six
// components the tidal tensor uses) as tri-quadratic polynomials in the
// offset from the cell's centre; level/kx/ky/kz as for split_build.  As
// many as leave TIDAL_RESERVE bytes free, the first qualifying slots first.
bool split_build_tidal(SplitTree& s, const DevTree& t, const uint8_t* level, const uint32_t* kx, const uint32_t* ky,
                       const uint32_t* kz, hipStream_t st, std::string& err);
void split_release(SplitTree& s);
bool split_supports_surfaces();
// doubles of launch record per ray that the split integrator's caller
// provides in IntegrateArgs::scratch
int split_launch_doubles();
// features of catalogue sources (learned first guess) from the stencil blocks
// order (may be null): the source each thread takes (launch_source_order)
hipError_t launch_features_split(const DevTree& t, const SplitTree& s, const double* src, uint64_t n,
                                 const uint32_t* order, double length, float* feat, hipStream_t st);
// grid = cus x blocks_per_cu workgroups of 256 lanes; waves_per_simd picks
// the kernel variant: without surfaces 1 = coefficient block kept in
// registers (1 wave per SIMD), 2..4 = coefficients re-read, at least that
// many waves per SIMD; with surfaces 1 or 2 waves per SIMD
// ev0/ev1 (may be null) are recorded just before and after the persistent
// kernel, which the launch records and final values bracket
hipError_t launch_integrate_split(const IntegrateArgs& A, const DevTree& t, const SplitTree& s, int cus,
                                  unsigned blocks_per_cu, unsigned waves_per_simd, hipStream_t st,
                                  hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

}  // namespace mpf
