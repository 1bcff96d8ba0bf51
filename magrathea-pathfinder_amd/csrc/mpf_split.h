// Split-precision tier: FP32 perturbation, FP64 background.
#pragma once
#include <hip/hip_runtime.h>
#include <string>

#include "mpf_kernels.h"

namespace mpf {

struct SplitTree {
    void* mem = nullptr;
    size_t bytes = 0;
};

bool split_build(SplitTree& s, const DevTree& t, int order, hipStream_t st, std::string& err);
void split_release(SplitTree& s);
hipError_t launch_integrate_split(const IntegrateArgs& A, const DevTree& t, const SplitTree& s, int cus, hipStream_t st);

}  // namespace mpf
