// Integrator::propagate (integrator.h:826-1015) on the device: the checks
// that accept or reject a bundle of rays, and the angular diameter distance
// along the central ray of each accepted bundle, from the whole trajectories
// the exact integrators write.  Built with -ffp-contract=off: every operation
// rounds as the reference's does.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mpf_propagate.h"

namespace mpf {

namespace {

// parameter the bundle rays are compared at (integrator.h:916-971)
__device__ inline double prop_ref(const double* e, const double* e0, int kind) {
    switch (kind) {
        case MPF_STOP_A: return 1 / e[MPF_OUT_A] - 1;
        case MPF_STOP_T: return e[MPF_OUT_T];
        case MPF_STOP_LAMBDA: return e[MPF_OUT_LAMBDA];
        case MPF_STOP_RADIUS: {
            const double x = e[MPF_OUT_X] - e0[MPF_OUT_X];
            const double y = e[MPF_OUT_Y] - e0[MPF_OUT_Y];
            const double z = e[MPF_OUT_Z] - e0[MPF_OUT_Z];
            return sqrt(x * x + y * y + z * z);
        }
        default: return e[MPF_OUT_REDSHIFT];
    }
}

__device__ inline long long elements(const double* out_rec) {
    const double s = out_rec[MPF_OUT_NSTEPS];
    return s < 0.0 ? 0 : (long long)s + 1;
}

// accept or reject each bundle (integrator.h:894-913): every ray non-empty
// and ending beyond EXTENT / 4 from its start, the amin test on the data's
// scale factor at its end, and end distances within 1/8 of each other.
// nacc[i]: rays of bundle i used for the distance (0 = rejected).
__global__ void bundle_accept_kernel(PropagateArgs P) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const double quarter = (4.0 / (2.0 * 1.0)) / 2.0;
    const double limit = 1.0 / (2.0 * 2.0 * 2.0);
    unsigned ntraj = P.nb;
    double mx = 0.0, mn = 0.0;
    for (unsigned j = 0; j < P.nb; ++j) {
        double last = 0.0;
        if (j < ntraj) {
            const uint64_t r = i * P.nb + j;
            const long long n = elements(P.out + (uint64_t)MPF_NOUT * r);
            if (!n) {
                ntraj = 0;
            } else {
                const double* f = P.traj + r * P.cap * MPF_NOUT;
                const double* b = f + (uint64_t)(n - 1) * MPF_NOUT;
                const double cx = b[MPF_OUT_X] - f[MPF_OUT_X], cy = b[MPF_OUT_Y] - f[MPF_OUT_Y],
                             cz = b[MPF_OUT_Z] - f[MPF_OUT_Z];
                last = sqrt(cx * cx + cy * cy + cz * cz);
                ntraj *= (last > quarter);
                const double ah = b[MPF_OUT_AH];
                ntraj *= (!(isnormal(P.amin) && isnormal(ah) && (ah < 1.0))) || (!(ah > P.amin));
            }
        }
        if (j == 0 || last > mx) mx = last;
        if (j == 0 || last < mn) mn = last;
    }
    ntraj *= (fabs(mx - mn) / mx < limit);
    P.nacc[i] = ntraj;
}

// distance at element k of the central ray of bundle i: the mean over the
// other rays of their separation from it at the same parameter
// (Utility::interpolate, utility.h:442-449, with std::upper_bound's
// bisection), times length a / angle (integrator.h:974-992)
__global__ void bundle_distance_kernel(PropagateArgs P) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P.n * P.cap) return;
    const uint64_t i = t / P.cap, k = t % P.cap;
    const unsigned ntraj = P.nacc[i];
    const uint64_t rc = i * P.nb;
    const long long nc = elements(P.out + (uint64_t)MPF_NOUT * rc);
    if (ntraj == 0 || (long long)k >= nc) {
        P.dist[t] = 0.0;
        return;
    }
    const double* C = P.traj + rc * P.cap * MPF_NOUT;
    const double* e = C + k * MPF_NOUT;
    double d = 0.0;
    if (ntraj > 1) {
        const double x0 = prop_ref(e, C, P.kind);
        for (unsigned j = 1; j < ntraj; ++j) {
            const uint64_t r = rc + j;
            const double* T = P.traj + r * P.cap * MPF_NOUT;
            const long long n = elements(P.out + (uint64_t)MPF_NOUT * r);
            long long first = 0, len = n;
            while (len > 0) {
                const long long half = len >> 1;
                if (x0 < prop_ref(T + (uint64_t)(first + half) * MPF_NOUT, T, P.kind)) {
                    len = half;
                } else {
                    first = first + half + 1;
                    len = len - half - 1;
                }
            }
            long long q = first;
            q += (n > 1) * ((q <= 0) - (q >= n));
            double c[3];
            const int dims[3] = {MPF_OUT_X, MPF_OUT_Y, MPF_OUT_Z};
            for (int a = 0; a < 3; ++a) {
                double v;
                if (n > 1) {
                    const double* ea = T + (uint64_t)(q - 1) * MPF_NOUT;
                    const double* eb = T + (uint64_t)q * MPF_NOUT;
                    const double xa = prop_ref(ea, T, P.kind), xb = prop_ref(eb, T, P.kind);
                    v = ea[dims[a]] + ((eb[dims[a]] - ea[dims[a]]) * (x0 - xa) / (xb - xa));
                } else {
                    v = (n > 0) ? T[dims[a]] : 0.0;
                }
                c[a] = v - e[dims[a]];
            }
            d += sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        }
        d /= (ntraj - 1);
        d *= (P.length * e[MPF_OUT_A]) / P.angle;
    }
    P.dist[t] = d;
}

}  // namespace

hipError_t launch_bundle_accept(const PropagateArgs& P, hipStream_t st) {
    if (P.n == 0) return hipSuccess;
    hipLaunchKernelGGL(bundle_accept_kernel, dim3((unsigned)((P.n + 255) / 256)), dim3(256), 0, st, P);
    return hipGetLastError();
}

hipError_t launch_bundle_distance(const PropagateArgs& P, hipStream_t st) {
    const uint64_t total = P.n * P.cap;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(bundle_distance_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, P);
    return hipGetLastError();
}

}  // namespace mpf
