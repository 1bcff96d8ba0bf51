// Features of a catalogue source for the learned first guess of the
// root-finder (layout in include/mpf.h).  The octree is sampled on the
// straight line of sight, at chi_k = (k + 1/2) D / FEAT_SAMPLES, by a sampler
// of the context's precision tier (exact FP64 TSC with a stencil cache carried
// outward, or the split tier's FP32 stencil blocks):
//   sample(x, y, z, g) -> g = (phi, dphi/dx, dphi/dy, dphi/dz), zeros where
//   the octree has no data.
// The Born deflection 2/c^2 sum (D - chi)/D grad_perp(phi) dchi (midpoint
// rule), in arcsec along e_phi and e_theta, is the first guess itself; the
// network learns what is left.  All features are invariant under rotations of
// the sky (no absolute angles).
#pragma once
#include "mpf_device.h"

namespace mpf {

constexpr int FEAT_SAMPLES = 128;
constexpr double FEAT_ARCSEC = 206264.80624709636;   // arcsec per radian

template <class Sampler>
__device__ inline void source_features(const double* s, double length, Sampler& sample, float* f) {
    const double D = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    const double theta = acos(s[2] / D), phi = atan2(s[1], s[0]);
    const double cp = cos(phi), sp = sin(phi), ct = cos(theta), st = sin(theta);
    const double ep[3] = {-sp, cp, 0.0};          // e_phi
    const double et[3] = {cp * ct, sp * ct, -st}; // e_theta
    const double nv[3] = {s[0] / D, s[1] / D, s[2] / D};
    const double g2 = 2.0 * length / C2 * FEAT_ARCSEC;  // 2 L / c^2 per unit box length, in arcsec
    double born[2] = {0.0, 0.0}, far[2] = {0.0, 0.0}, near[2] = {0.0, 0.0}, last[2] = {0.0, 0.0}, phimid = 0.0;
    for (int k = 0; k < FEAT_SAMPLES; ++k) {
        const double chi = D * (k + 0.5) / FEAT_SAMPLES;
        double g[4];
        sample(nv[0] * chi, nv[1] * chi, nv[2] * chi, g);
        const double gp = g[1] * ep[0] + g[2] * ep[1];
        const double gt = g[1] * et[0] + g[2] * et[1] + g[3] * et[2];
        const double w = (D - chi) / D * (D / FEAT_SAMPLES) * g2;   // lensing weight x quadrature step
        born[0] += w * gp;
        born[1] += w * gt;
        if (k >= FEAT_SAMPLES / 2) { far[0] += w * gp; far[1] += w * gt; }
        if (k < FEAT_SAMPLES / 4) { near[0] += w * gp; near[1] += w * gt; }
        if (k == FEAT_SAMPLES - 1) { last[0] = gp * g2 * D; last[1] = gt * g2 * D; }
        if (k == FEAT_SAMPLES / 2 - 1) phimid = g[0] / C2 * 1.0e5;
    }
    f[0] = (float)D;
    f[1] = (float)born[0];
    f[2] = (float)born[1];
    f[3] = (float)far[0];
    f[4] = (float)far[1];
    f[5] = (float)near[0];
    f[6] = (float)near[1];
    f[7] = (float)last[0];
    f[8] = (float)last[1];
    f[9] = (float)phimid;
}

}  // namespace mpf
