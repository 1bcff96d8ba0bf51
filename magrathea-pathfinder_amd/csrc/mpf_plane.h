// Where a bundle ray crosses a plane (Lensing::dbetadtheta with stop_bundle
// = plane, lensing.h:205-301): the segment [id, id + 1] of the trajectory
// around the first element whose projection along the plane normal exceeds
// the plane value v (upper_bound over the len elements), and the weight f of
// element id in the interpolated point (f = 1 at element id, 0 at id + 1).
// When the last element's projection equals v, upper_bound returns len and
// the reference reads one element past the trajectory; here the last segment
// is taken (f = 0: the last element).  Callers have checked len >= 2 and
// along(len - 1) >= v.
#pragma once
#include <cstdint>

namespace mpf {

template <class Along>
inline void plane_segment(int64_t len, double v, Along along, int64_t* id, double* f) {
    int64_t first = 0, cnt = len;  // upper_bound of v
    while (cnt > 0) {
        const int64_t half = cnt / 2;
        if (v < along(first + half)) cnt = half;
        else { first += half + 1; cnt -= half + 1; }
    }
    int64_t i = first - (first > 0);
    if (i > len - 2) i = len - 2;
    const double next = along(i + 1), prev = along(i);
    *id = i;
    *f = (next - v) / (next - prev);
}

}  // namespace mpf
