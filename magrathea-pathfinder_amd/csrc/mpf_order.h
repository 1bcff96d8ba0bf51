// Ray order of the persistent integrators (mpf_order.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace mpf {

// device bytes of the work area launch_ray_order needs for n rays
size_t ray_order_work_bytes(uint64_t n);
// order[q] = index of the ray the queue hands out at position q: rays by
// decreasing path length (ray_value, or input order when ray_value is null),
// in buckets of `bucket` rays (0: one bucket), each bucket by direction.
// smaller_is_longer: the stop value is a scale factor (smaller = farther).
hipError_t launch_ray_order(const double* in9, const double* ray_value, int smaller_is_longer, uint64_t n,
                            uint64_t bucket, uint32_t* order, void* work, size_t work_bytes, hipStream_t st);
// order[q] = index of the source (n x 6: position, velocity) a kernel takes
// at position q: sources by the direction of their position from the
// observer, so that the lines of sight of a wavefront's lanes share cells.
// Work area as for launch_ray_order.
hipError_t launch_source_order(const double* src6, uint64_t n, uint32_t* order, void* work, size_t work_bytes,
                               hipStream_t st);

}  // namespace mpf
