// Features of a catalogue source for the learned first guess of the
// root-finder (layout in include/mpf.h).  The octree is sampled on straight
// lines from the observer, at chi_k = (k + 1/2) D / FEAT_SAMPLES, by a sampler
// of the context's precision tier (exact FP64 TSC with a stencil cache carried
// outward along each line, or the split tier's FP32 stencil blocks):
//   sample(x, y, z, g) -> g = (phi, dphi/dx, dphi/dy, dphi/dz), zeros where
//   the octree has no data; sample.restart() before a new line.
// The Born deflection 2/c^2 sum (D - chi)/D grad_perp(phi) dchi (midpoint
// rule), in arcsec along e_phi and e_theta, is the first guess itself; the
// network learns what is left.  The second line runs towards the direction
// the Born deflection points to (the source's angles plus the deflection):
// its Born sum less the first one is, to first order, the distortion matrix
// times the deflection (the coupling of the deflection with the lens), which
// the Born estimate misses; the same difference over the far half and the
// first line's far-half sum say where along the line the lensing happens.
// All features are invariant under rotations of the sky (no absolute
// angles), and the pairs (1, 2), (3, 4), (5, 6), (7, 8) and the targets turn
// together with the screen basis.
#pragma once
#include "mpf_device.h"

namespace mpf {

constexpr int FEAT_SAMPLES = 256;
constexpr double FEAT_ARCSEC = 206264.80624709636;   // arcsec per radian

// Born sums along the line from the observer towards the unit vector nv, to
// distance D, projected on e_phi / e_theta of that line: whole line (b),
// far half (far), and the middle sample's potential.
template <class Sampler>
__device__ inline void born_line(const double* nv, double D, double g2, const double* ep, const double* et,
                                 Sampler& sample, double* b, double* far, double* phimid) {
    b[0] = b[1] = 0.0;
    far[0] = far[1] = 0.0;
    *phimid = 0.0;
    sample.restart();
    for (int k = 0; k < FEAT_SAMPLES; ++k) {
        const double chi = D * (k + 0.5) / FEAT_SAMPLES;
        double g[4];
        sample(nv[0] * chi, nv[1] * chi, nv[2] * chi, g);
        const double gp = g[1] * ep[0] + g[2] * ep[1];
        const double gt = g[1] * et[0] + g[2] * et[1] + g[3] * et[2];
        const double w = (D - chi) / D * (D / FEAT_SAMPLES) * g2;   // lensing weight x quadrature step
        b[0] += w * gp;
        b[1] += w * gt;
        if (k >= FEAT_SAMPLES / 2) { far[0] += w * gp; far[1] += w * gt; }
        if (k == FEAT_SAMPLES / 2 - 1) *phimid = g[0] / C2 * 1.0e5;
    }
}

// unit vector and screen basis e_phi, e_theta of the direction (phi, theta)
__device__ inline void line_basis(double phi, double theta, double* nv, double* ep, double* et) {
    const double cp = cos(phi), sp = sin(phi), ct = cos(theta), st = sin(theta);
    nv[0] = st * cp; nv[1] = st * sp; nv[2] = ct;
    ep[0] = -sp; ep[1] = cp; ep[2] = 0.0;
    et[0] = cp * ct; et[1] = sp * ct; et[2] = -st;
}

template <class Sampler>
__device__ inline void source_features(const double* s, double length, Sampler& sample, float* f) {
    const double D = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    const double theta = acos(s[2] / D), phi = atan2(s[1], s[0]);
    double nv[3], ep[3], et[3];
    line_basis(phi, theta, nv, ep, et);
    const double g2 = 2.0 * length / C2 * FEAT_ARCSEC;  // 2 L / c^2 per unit box length, in arcsec
    double born[2], far[2], phimid;
    born_line(nv, D, g2, ep, et, sample, born, far, &phimid);
    // the line towards the Born-deflected direction
    const double theta1 = theta + born[1] / FEAT_ARCSEC;
    const double phi1 = phi + born[0] / (FEAT_ARCSEC * sin(theta));
    double nv1[3], ep1[3], et1[3];
    line_basis(phi1, theta1, nv1, ep1, et1);
    double born1[2], far1[2], phimid1;
    born_line(nv1, D, g2, ep1, et1, sample, born1, far1, &phimid1);
    f[0] = (float)D;
    f[1] = (float)born[0];
    f[2] = (float)born[1];
    f[3] = (float)(born1[0] - born[0]);
    f[4] = (float)(born1[1] - born[1]);
    f[5] = (float)(far1[0] - far[0]);
    f[6] = (float)(far1[1] - far[1]);
    f[7] = (float)far[0];
    f[8] = (float)far[1];
    f[9] = (float)phimid;
}

}  // namespace mpf
