// Host-only check of mpf::plane_segment (magrathea-pathfinder_amd/csrc/mpf_plane.h):
// prints "id f" for each plane value given on the command line, along a
// trajectory whose projections are the remaining arguments after "--".
//   plane_segment_test v1 v2 ... -- p0 p1 p2 ...
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mpf_plane.h"

int main(int argc, char** argv) {
    std::vector<double> vals, proj;
    bool after = false;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--")) { after = true; continue; }
        (after ? proj : vals).push_back(std::strtod(argv[i], nullptr));
    }
    if (proj.size() < 2) return 2;
    auto along = [&](int64_t e) {
        if (e < 0 || e >= (int64_t)proj.size()) { std::printf("out of range %lld\n", (long long)e); std::exit(3); }
        return proj[e];
    };
    for (double v : vals) {
        int64_t id = 0;
        double f = 0.0;
        mpf::plane_segment((int64_t)proj.size(), v, along, &id, &f);
        std::printf("%lld %.17g\n", (long long)id, f);
    }
    return 0;
}
