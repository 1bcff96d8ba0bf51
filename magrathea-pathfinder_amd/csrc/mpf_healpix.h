// HEALPix RING scheme, restated from the published algorithm (Gorski et al.
// 2005, ApJ 622:759, section 4 and appendix): the third-party library the
// reference links (HEALPix C++ 3.70 / chealpix, used at hmaps.h:599 and
// hmaps.cpp:185) is not part of this build.  Host and device versions.
//
//   npix = 12 nside^2; north polar cap rings 1..nside-1 hold 4i pixels,
//   equatorial rings nside..3nside hold 4 nside pixels, the south cap mirrors
//   the north.  Pixel centres: cap ring i at z = 1 - i^2 / (3 nside^2),
//   phi = (j - 1/2) pi / (2i); equatorial ring i at z = 4/3 - 2i / (3 nside),
//   phi = (j - s/2) pi / (2 nside) with s = 1 when i - nside is even.
#pragma once
#include <cmath>
#include <cstdint>

#ifdef __HIPCC__
#define MPF_HD __host__ __device__
#else
#define MPF_HD
#endif

namespace mpf {
namespace healpix {

constexpr double PI = 3.141592653589793238462643383279502884197;
constexpr double HALFPI = 1.570796326794896619231321691639751442099;

MPF_HD inline int64_t isqrt64(int64_t v) {
    int64_t r = (int64_t)sqrt((double)v + 0.5);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

// pixel centre: z = cos(theta), phi in [0, 2 pi), and sin(theta)
MPF_HD inline void pix2zphi_ring(int64_t nside, int64_t pix, double* z, double* s, double* phi) {
    const int64_t ncap = 2 * nside * (nside - 1);
    const int64_t npix = 12 * nside * nside;
    const double fact2 = 4.0 / (double)npix;
    if (pix < ncap) {  // north polar cap
        const int64_t iring = (1 + isqrt64(1 + 2 * pix)) >> 1;
        const int64_t iphi = pix + 1 - 2 * iring * (iring - 1);
        const double tmp = (double)(iring * iring) * fact2;
        *z = 1.0 - tmp;
        *s = sqrt(tmp * (2.0 - tmp));
        *phi = ((double)iphi - 0.5) * HALFPI / (double)iring;
    } else if (pix < npix - ncap) {  // equatorial belt
        const int64_t ip = pix - ncap;
        const int64_t iring = ip / (4 * nside) + nside;
        const int64_t iphi = ip % (4 * nside) + 1;
        const double fodd = ((iring + nside) & 1) ? 1.0 : 0.5;
        *z = (double)(2 * nside - iring) * (2.0 / (3.0 * (double)nside));
        *s = sqrt((1.0 - *z) * (1.0 + *z));
        *phi = ((double)iphi - fodd) * PI / (2.0 * (double)nside);
    } else {  // south polar cap
        const int64_t ip = npix - pix;
        const int64_t iring = (1 + isqrt64(2 * ip - 1)) >> 1;
        const int64_t iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
        const double tmp = (double)(iring * iring) * fact2;
        *z = tmp - 1.0;
        *s = sqrt(tmp * (2.0 - tmp));
        *phi = ((double)iphi - 0.5) * HALFPI / (double)iring;
    }
}

MPF_HD inline void pix2vec_ring(int64_t nside, int64_t pix, double* v) {
    double z, s, phi;
    pix2zphi_ring(nside, pix, &z, &s, &phi);
    v[0] = s * cos(phi);
    v[1] = s * sin(phi);
    v[2] = z;
}

// pixel containing the direction (z = cos theta, phi)
MPF_HD inline int64_t zphi2pix_ring(int64_t nside, double z, double phi) {
    const double za = fabs(z);
    double tt = fmod(phi, 2.0 * PI);
    if (tt < 0) tt += 2.0 * PI;
    tt = tt / HALFPI;  // in [0, 4)
    if (tt >= 4.0) tt = 0.0;
    const int64_t ncap = 2 * nside * (nside - 1);
    const int64_t npix = 12 * nside * nside;
    if (za <= 2.0 / 3.0) {  // equatorial belt
        const double t1 = (double)nside * (0.5 + tt);
        const double t2 = (double)nside * z * 0.75;
        const int64_t jp = (int64_t)(t1 - t2);  // ascending edge line
        const int64_t jm = (int64_t)(t1 + t2);  // descending edge line
        const int64_t ir = nside + 1 + jp - jm; // ring from z = 2/3, in [1, 2 nside + 1]
        const int64_t kshift = 1 - (ir & 1);
        int64_t ip = (jp + jm - nside + kshift + 1) / 2;
        ip = ip % (4 * nside);
        if (ip < 0) ip += 4 * nside;
        return ncap + (ir - 1) * 4 * nside + ip;
    }
    const double tp = tt - (double)(int64_t)tt;
    const double tmp = (double)nside * sqrt(3.0 * (1.0 - za));
    const int64_t jp = (int64_t)(tp * tmp);
    const int64_t jm = (int64_t)((1.0 - tp) * tmp);
    int64_t ir = jp + jm + 1;  // ring from the closest pole
    if (ir > nside) ir = nside;  // guards the cap edge against rounding
    int64_t ip = (int64_t)(tt * (double)ir);
    ip = ip % (4 * ir);
    if (ip < 0) ip += 4 * ir;
    return (z > 0) ? 2 * ir * (ir - 1) + ip : npix - 2 * ir * (ir + 1) + ip;
}

MPF_HD inline int64_t vec2pix_ring(int64_t nside, double x, double y, double z) {
    const double r = sqrt(x * x + y * y + z * z);
    return zphi2pix_ring(nside, z / r, atan2(y, x));
}

}  // namespace healpix
}  // namespace mpf
