#include "mpf_split.h"

namespace mpf {

bool split_build(SplitTree& s, const DevTree& t, int order, hipStream_t st, std::string& err) {
    (void)s; (void)t; (void)order; (void)st;
    err = "split precision tier is not available in this build";
    return false;
}

void split_release(SplitTree& s) {
    if (s.mem) (void)hipFree(s.mem);
    s.mem = nullptr;
    s.bytes = 0;
}

hipError_t launch_integrate_split(const IntegrateArgs&, const DevTree&, const SplitTree&, int, hipStream_t) {
    return hipErrorNotSupported;
}

}  // namespace mpf
