// Distance statistics of the rays program (rays.cpp:300-456), host: each
// accepted central trajectory of mpf_propagate gives a curve (x, y) of an
// abscissa (the interpolation parameter) and a statistic (distance,
// distance^2, homogeneous or inhomogeneous redshift, inverse distance,
// inverse distance^2); the curves are re-interpolated at the abscissae of
// every nstat-th element of the homogeneous reference ray
// (Utility::reinterpolate, utility.h:502-510) and reduced to a mean and a
// standard deviation per abscissa.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/mpf.h"

namespace {

constexpr int ERR_INVAL = -22;

// Utility::interpolate (utility.h:442-449), std::upper_bound's bisection
double interp(double x0, const std::vector<double>& x, const std::vector<double>& y) {
    const long long n = (long long)std::min(x.size(), y.size());
    long long i = std::upper_bound(x.begin(), x.begin() + n, x0) - x.begin();
    i += (n > 1) * ((i <= 0) - (i >= n));
    return (n > 1) ? (y[i - 1] + ((y[i] - y[i - 1]) * (x0 - x[i - 1]) / (x[i] - x[i - 1]))) : ((n > 0) ? y[0] : 0.0);
}

// abscissa of element e of a trajectory starting at e0 (rays.cpp:322-334;
// the reference maps `a` to t and `t` to 1/a - 1)
double abscissa(const double* e, const double* e0, int interp_kind) {
    switch (interp_kind) {
        case MPF_STOP_A: return e[MPF_OUT_T];
        case MPF_STOP_T: return 1.0 / e[MPF_OUT_A] - 1.0;
        case MPF_STOP_RADIUS:
            return std::sqrt(std::pow(e[MPF_OUT_X] - e0[MPF_OUT_X], 2.0) + std::pow(e[MPF_OUT_Y] - e0[MPF_OUT_Y], 2.0) +
                             std::pow(e[MPF_OUT_Z] - e0[MPF_OUT_Z], 2.0));
        case MPF_STOP_LAMBDA: return e[MPF_OUT_LAMBDA];
        default: return e[MPF_OUT_REDSHIFT];
    }
}

// statistic of element e (rays.cpp:336-349)
double statistic(const double* e, double d, const double* e0, int stat, double invlength) {
    switch (stat) {
        case MPF_STAT_DISTANCE2: return d * invlength * d * invlength;
        case MPF_STAT_HOMOGENEOUS: return e0[MPF_OUT_A] / e[MPF_OUT_A] - 1.0;
        case MPF_STAT_INHOMOGENEOUS: return e[MPF_OUT_REDSHIFT];
        case MPF_STAT_INVDISTANCE: return 1.0 / (d * invlength);
        case MPF_STAT_INVDISTANCE2: return 1.0 / (d * invlength * d * invlength);
        default: return d * invlength;
    }
}

}  // namespace

extern "C" int mpf_ray_statistics(const double* rows, const double* dist, const uint64_t* nelem, uint64_t n,
                                  uint64_t capacity, const double* ref_rows, const double* ref_dist, uint64_t m,
                                  int interp_kind, int stat, uint32_t nstat, double amin, double length,
                                  double* out4, uint64_t cap_out, uint64_t* nout, uint64_t* nused) {
    if ((n && (!rows || !dist || !nelem)) || (m && (!ref_rows || !ref_dist)) || !nout || !nused || !(length > 0.0))
        return ERR_INVAL;
    const double invlength = 1.0 / length;
    // reference ray without the elements below amin, every statmod-th element
    std::vector<uint64_t> keep;
    for (uint64_t j = 0; j < m; ++j)
        if (!(std::isnormal(amin) && ref_rows[(uint64_t)MPF_NOUT * j + MPF_OUT_A] < amin)) keep.push_back(j);
    const unsigned statmod = std::max(1u, (unsigned)keep.size() / std::max(nstat, 1u));
    std::vector<double> refx, refy, refz;
    const double* r0 = keep.empty() ? nullptr : ref_rows + (uint64_t)MPF_NOUT * keep[0];
    for (uint64_t q = 0; q < keep.size(); ++q) {
        if (q % statmod) continue;
        const double* e = ref_rows + (uint64_t)MPF_NOUT * keep[q];
        refx.push_back(abscissa(e, r0, interp_kind));
        refy.push_back(statistic(e, ref_dist[keep[q]], r0, stat, invlength));
        refz.push_back(e[MPF_OUT_REDSHIFT]);
    }
    const uint64_t len = refx.size();
    *nout = len;
    std::vector<double> mean(len, 0.0), sd(len, 0.0);
    std::vector<std::vector<double>> ys;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t ne = nelem[i];
        if (!ne) continue;
        if (ne > capacity) return ERR_INVAL;
        const double* T = rows + (uint64_t)MPF_NOUT * capacity * i;
        std::vector<double> x(ne), y(ne);
        for (uint64_t j = 0; j < ne; ++j) {
            x[j] = abscissa(T + (uint64_t)MPF_NOUT * j, T, interp_kind);
            y[j] = statistic(T + (uint64_t)MPF_NOUT * j, dist[capacity * i + j], T, stat, invlength);
        }
        std::vector<double> yi(len);
        for (uint64_t k = 0; k < len; ++k) yi[k] = interp(refx[k], x, y);
        ys.push_back(std::move(yi));
    }
    const uint64_t size = ys.size();
    *nused = size;
    for (uint64_t j = 0; j < size; ++j)
        for (uint64_t k = 0; k < len; ++k) mean[k] += ys[j][k];
    for (uint64_t k = 0; k < len; ++k) mean[k] /= double(size);
    for (uint64_t j = 0; j < size; ++j)
        for (uint64_t k = 0; k < len; ++k) sd[k] += (ys[j][k] - mean[k]) * (ys[j][k] - mean[k]);
    for (uint64_t k = 0; k < len; ++k) sd[k] = std::sqrt(sd[k] / double(size - 1));
    if (out4 && cap_out < len) return ERR_INVAL;
    if (out4)
        for (uint64_t k = 0; k < len; ++k) {
            out4[4 * k + 0] = refz[k];
            out4[4 * k + 1] = refy[k];
            out4[4 * k + 2] = mean[k];
            out4[4 * k + 3] = sd[k];
        }
    return 0;
}

// The statistics file of rays.cpp:441-452 (Output::save, output.h:228-241):
// "# count", then per abscissa the reference element's redshift, its
// statistic, the mean and the standard deviation, max_digits10 (17)
// significant digits in the default float format.
extern "C" int mpf_ray_statistics_write(const char* path, const double* out4, uint64_t nout, uint64_t count) {
    if (!path || (nout && !out4)) return ERR_INVAL;
    FILE* f = std::fopen(path, "w");
    if (!f) return ERR_INVAL;
    bool ok = std::fprintf(f, "# %llu\n", (unsigned long long)count) > 0;
    for (uint64_t i = 0; ok && i < nout; ++i)
        ok = std::fprintf(f, "%.17g %.17g %.17g %.17g\n", out4[4 * i], out4[4 * i + 1], out4[4 * i + 2],
                          out4[4 * i + 3]) > 0;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? 0 : ERR_INVAL;
}
