// Cosmology table of the reference (host): Input::acquire (input.h:1091-1233)
// reads a RAMSES evolution file and a parameter file and derives the
// (t, a, da/dt, d2a/dt2) table with the kernel smoother, the Fornberg
// derivative and the integrator of utility.h; Input::correct
// (input.h:1330-1378) replaces a(t) by the one of a photon propagated in a
// homogeneous universe.  Same operations in the same order, so the tables are
// bit-identical to the reference's.
#include "mpf_cosmology.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

namespace mpf {
namespace cosmo {

namespace {

using Vec = std::vector<double>;

// Utility::interpolate (utility.h:442-449): upper_bound, clamped to the ends
double interpolate(double x0, const Vec& x, const Vec& y) {
    const long long n = (long long)std::min(x.size(), y.size());
    long long i = std::upper_bound(x.begin(), x.begin() + n, x0) - x.begin();
    i += (n > 1) * ((i <= 0) - (i >= n));
    return (n > 1) ? (y[i - 1] + ((y[i] - y[i - 1]) * (x0 - x[i - 1]) / (x[i] - x[i - 1]))) : ((n > 0) ? y[0] : 0.0);
}

// Utility::reverse (utility.h:607-613): every element minus a value
Vec shift(const Vec& v, double value) {
    Vec r(v);
    for (double& e : r) e -= value;
    return r;
}

// Utility::smooth (utility.h:627-636): kernel-weighted mean over at most
// `window` neighbours on each side (all of them when window is 0)
template <class Kernel>
Vec smooth(const Vec& x, const Vec& y, Kernel kernel, unsigned window) {
    const unsigned size = (unsigned)std::min(x.size(), y.size());
    const unsigned win = window > 0 ? window : size;
    Vec r(size);
    for (unsigned i = 0; i < size; ++i) {
        const unsigned n = std::min(win, std::min(i, size - 1 - i));
        double num = 0.0, den = 0.0;
        for (unsigned j = i - n; j <= i + n; ++j) {
            const double w = kernel(x[i], x[j]);
            num += w * y[j];
            den += w;
        }
        r[i] = num / den;
    }
    return r;
}

// Utility::integrate (utility.h:647-656): cumulative rectangle rule from value
Vec integrate(const Vec& x, const Vec& y, double value) {
    const unsigned size = (unsigned)std::min(x.size(), y.size());
    Vec r(size, value);
    for (unsigned i = 1; i < size; ++i) r[i] = r[i - 1] + y[i] * (x[i] - x[i - 1]);
    return r;
}

// Utility::differentiate<0> (utility.h:527-552): 8th-order centred
// finite-difference weights (Fornberg) on the linear interpolant, with a step
// of `neighbourhood` times the local grid spacing
double differentiate(double x0, const Vec& x, const Vec& y, unsigned neighbourhood) {
    static const double w[9] = {1.0 / 280.0, -4.0 / 105.0, 1.0 / 5.0, -4.0 / 5.0, 0.0,
                                4.0 / 5.0,   -1.0 / 5.0,   4.0 / 105.0, -1.0 / 280.0};
    const long long n = (long long)std::min(x.size(), y.size());
    const long long i = std::upper_bound(x.begin(), x.begin() + n, x0) - x.begin();
    const long long j = i + ((n > 1) * ((i <= 0) - (i >= n)));
    const double h = (n > 1) ? ((x[j] - x[j - 1]) * std::max(1u, neighbourhood)) : 0.0;
    double result = 0.0;
    if (n > 1) {
        for (unsigned k = 0; k < 4; ++k)
            result += w[k] * interpolate(x0 - (4 - k) * h, x, y) + w[8 - k] * interpolate(x0 + (4 - k) * h, x, y);
        result += w[4] * interpolate(x0, x, y);
    }
    return result / h;
}

Vec derive(const Vec& x, const Vec& y, unsigned neighbourhood) {
    const unsigned size = (unsigned)std::min(x.size(), y.size());
    Vec r(size);
    for (unsigned i = 0; i < size; ++i) r[i] = differentiate(x[i], x, y, neighbourhood);
    return r;
}

std::string trim(const std::string& s) {
    auto b = std::find_if(s.begin(), s.end(), [](char c) { return std::isgraph((unsigned char)c); });
    auto e = std::find_if(s.rbegin(), s.rend(), [](char c) { return std::isgraph((unsigned char)c); }).base();
    return (b < e) ? std::string(b, e) : std::string();
}

}  // namespace

bool acquire(const std::string& evolution_file, const std::string& parameter_file, const std::string& cell_dir,
             double mpc, double& h, double& omegam, double& lboxmpch, Table& out, std::string& err) {
    const double twob2 = 2.0 * std::pow(1.0 / 1000.0, 2.0);
    const unsigned nsmooth = 500, nderive = 1000;
    out = Table();
    h = omegam = lboxmpch = 0.0;
    // box length from the "boxlen<digits>" part of the cell directory, '-'
    // read as the decimal point (input.h:1129-1134)
    {
        const std::string key = "boxlen";
        std::string s(cell_dir.begin() + std::min(cell_dir.rfind(key), cell_dir.size()), cell_dir.end());
        s.assign(s.begin() + std::min(key.size(), s.size()), s.end());
        s.assign(s.begin(), s.begin() + std::min(s.find_first_not_of("0123456789-."), s.size()));
        std::replace(s.begin(), s.end(), '-', '.');
        try {
            lboxmpch = s.empty() ? 0.0 : (double)std::stold(s);
        } catch (...) {
            err = "cell directory: unreadable box length";
            return false;
        }
    }
    // h and Omega_m from "key = value" lines (input.h:1136-1157)
    double hubble = 0.0;
    {
        std::ifstream in(parameter_file);
        std::string line;
        while (in && (h == 0.0 || omegam == 0.0) && std::getline(in, line)) {
            const size_t eq = line.find('=');
            if (eq == std::string::npos) continue;
            const std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
            try {
                if (k == "h") {
                    h = (double)std::stold(v);
                    hubble = (h * 100 * 1000) / mpc;
                } else if (k == "Omega_m") {
                    omegam = (double)std::stold(v);
                }
            } catch (...) {
                err = "parameter file: unreadable value of " + k;
                return false;
            }
        }
    }
    // evolution file: a, H/H0, tau H0, t_lookback H0, t_proper H0 per line
    std::vector<std::string> lines;
    {
        std::ifstream in(evolution_file);
        std::string line;
        while (std::getline(in, line))
            if (!line.empty()) lines.push_back(line);
    }
    unsigned size = (unsigned)lines.size();
    if (size == 0 || h == 0.0 || omegam == 0.0 || lboxmpch == 0.0) {
        err = "evolution file, h, Omega_m or the box length missing";
        h = omegam = lboxmpch = 0.0;
        return false;
    }
    Vec in[5], o[5];
    for (auto& v : in) v.assign(size, 0.0);
    for (auto& v : o) v.assign(size, 0.0);
    for (unsigned i = 0; i < size; ++i) {
        std::istringstream iss(lines[i]);
        iss >> in[0][i] >> in[1][i] >> in[2][i] >> in[3][i] >> in[4][i];
    }
    // SI units, columns reordered (input.h:1166-1167)
    for (unsigned i = 0; i < size; ++i) {
        o[0][i] = in[4][i] / hubble;
        o[2][i] = in[0][i];
        o[3][i] = in[1][i] * hubble;
    }
    o[0] = shift(o[0], *std::min_element(o[0].begin(), o[0].end()));
    auto g1 = [twob2](double xi, double xj) { return std::exp(-((xi - xj) * (xi - xj)) / twob2); };
    auto g10 = [twob2](double xi, double xj) { return std::exp(-((xi - xj) * (xi - xj) * 10) / twob2); };
    o[3] = smooth(o[2], o[3], g1, size / nsmooth);
    o[1][0] = 0.0;
    // rows whose a does not increase normally are dropped (input.h:1174-1180)
    const double nan = std::numeric_limits<double>::quiet_NaN();
    for (unsigned i = 1; i < size; ++i)
        if (!std::isnormal(std::abs(o[2][i] - o[2][i - 1]))) o[0][i] = nan;
    for (unsigned i = 0; i < size; ++i)
        if (std::isnan(o[0][i] + o[1][i] + o[2][i] + o[3][i] + o[4][i]))
            for (auto& v : o) v[i] = nan;
    for (auto& v : o) v.erase(std::remove_if(v.begin(), v.end(), [](double x) { return std::isnan(x); }), v.end());
    size = (unsigned)o[0].size();
    // conformal time from da / (da/dt), derivative, smoothing, integral
    // (input.h:1183-1194)
    for (unsigned i = 1; i < size; ++i) o[1][i] = o[1][i - 1] + (o[2][i] - o[2][i - 1]) / o[3][i - 1];
    o[4] = derive(o[1], o[3], size / nderive);
    o[4] = smooth(o[2], o[4], g1, size / nsmooth);
    o[4] = smooth(o[2], o[4], g10, size / (nsmooth * 10));
    o[3] = integrate(o[1], o[4], o[3][0]);
    o[1][0] = 0.0;
    for (unsigned i = 1; i < size; ++i) o[1][i] = o[1][i - 1] + (o[2][i] - o[2][i - 1]) / o[3][i - 1];
    // t measured from the observer, increasing into the past (input.h:1196-1201)
    out.t = o[1];
    out.a = o[2];
    out.dadt = o[3];
    out.d2adt2 = o[4];
    out.t = shift(out.t, *std::max_element(out.t.begin(), out.t.end()));
    for (unsigned i = 0; i < size; ++i) {
        out.t[i] = std::abs(out.t[i]);
        out.dadt[i] = -out.dadt[i];
    }
    std::reverse(out.t.begin(), out.t.end());
    std::reverse(out.a.begin(), out.a.end());
    std::reverse(out.dadt.begin(), out.dadt.end());
    std::reverse(out.d2adt2.begin(), out.d2adt2.end());
    return true;
}

void correct(Table& tab, const Vec& traj_t, const Vec& traj_a) {
    const unsigned npasses = 4, nsmooth = 500;
    const unsigned size = (unsigned)tab.t.size();
    const unsigned length = (unsigned)traj_t.size();
    if (length == 0) return;
    const double twob2 = 2.0 * std::pow(traj_t.back() / 1000.0, 2.0);
    auto g1 = [twob2](double xi, double xj) { return std::exp(-((xi - xj) * (xi - xj)) / twob2); };
    auto g10 = [twob2](double xi, double xj) { return std::exp(-((xi - xj) * (xi - xj) * 10) / twob2); };
    auto inside = [&]() {
        for (unsigned i = 0; i < size; ++i)
            if (!((tab.t[i] < traj_t.front()) || (tab.t[i] > traj_t.back())))
                tab.a[i] = interpolate(tab.t[i], traj_t, traj_a);
    };
    inside();
    if (length < size) {
        for (unsigned p = 0; p < npasses; ++p) {
            tab.a = smooth(tab.t, tab.a, g1, size / nsmooth);
            tab.a = smooth(tab.t, tab.a, g10, size / (nsmooth * 10));
            inside();
        }
    }
}

// Integrator::launch from the origin toward +x and the number of steps per
// cell of the homogeneous reference photon (rays.cpp:211-214)
unsigned reference_nsteps(uint64_t table_size, unsigned nreference) {
    return (unsigned)(4.0 * std::pow(2.0, (unsigned)(std::log2((double)table_size / std::pow(2.0, nreference) + 1.0) +
                                                     1.0) + 1.0));
}

}  // namespace cosmo
}  // namespace mpf
