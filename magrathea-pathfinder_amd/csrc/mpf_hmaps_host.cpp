// Host parts of the hmaps program: the homogeneous quantities at the map
// redshifts (hmaps.cpp:297-325) and the HEALPix FITS writer (the
// write_healpix_map of chealpix that hmaps.cpp:510 calls, with the cleaning
// of hmaps.cpp:494-499).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mpf.h"

namespace {

constexpr int ERR_INVAL = -22;
constexpr int ERR_IO = -5;

// one 80-character header card
void card(std::string& h, const char* key, const std::string& value, bool quoted) {
    char buf[81];
    std::string v = value;
    if (quoted) {
        std::string q = "'" + value;
        while (q.size() < 9) q += ' ';
        v = q + "'";
        std::snprintf(buf, sizeof(buf), "%-8.8s= %-70s", key, v.c_str());
    } else {
        std::snprintf(buf, sizeof(buf), "%-8.8s= %20s%50s", key, v.c_str(), "");
    }
    h.append(buf, 80);
}

void end_header(std::string& h) {
    char buf[81];
    std::snprintf(buf, sizeof(buf), "%-80s", "END");
    h.append(buf, 80);
    while (h.size() % 2880) h += ' ';
}

}  // namespace

extern "C" int mpf_hmaps_homogeneous(const double* ref_rows, uint64_t m, const double* z_stop, uint32_t nz,
                                     int stop_ray, double* homo5, double* interp_ref) {
    if (!ref_rows || m < 2 || (nz && (!z_stop || !homo5))) return ERR_INVAL;
    std::vector<double> z(m);
    for (uint64_t k = 0; k < m; ++k) z[k] = ref_rows[(uint64_t)MPF_NOUT * k + MPF_OUT_REDSHIFT];
    for (uint32_t i = 0; i < nz; ++i) {
        const uint64_t marked = std::upper_bound(z.begin(), z.end(), z_stop[i]) - z.begin();
        const uint64_t id = marked - (marked > 0);
        if (id + 1 >= m) return ERR_INVAL;   // beyond the reference ray
        const double* a = ref_rows + (uint64_t)MPF_NOUT * id;
        const double* b = a + MPF_NOUT;
        const double f = (b[MPF_OUT_REDSHIFT] - z_stop[i]) / (b[MPF_OUT_REDSHIFT] - a[MPF_OUT_REDSHIFT]);
        double* h = homo5 + 5 * (uint64_t)i;
        h[0] = a[MPF_OUT_T] * f + b[MPF_OUT_T] * (1 - f);
        h[1] = a[MPF_OUT_CHI] * f + b[MPF_OUT_CHI] * (1 - f);
        h[2] = a[MPF_OUT_LAMBDA] * f + b[MPF_OUT_LAMBDA] * (1 - f);
        h[3] = z_stop[i];
        h[4] = 1. / (1. + z_stop[i]);
        if (interp_ref) {
            switch (stop_ray) {
                case MPF_STOP_T: interp_ref[i] = h[0]; break;
                case MPF_STOP_RADIUS: case MPF_STOP_S: interp_ref[i] = h[1]; break;
                case MPF_STOP_LAMBDA: interp_ref[i] = h[2]; break;
                case MPF_STOP_A: interp_ref[i] = h[4]; break;
                case MPF_STOP_REDSHIFT: interp_ref[i] = h[3]; break;
                default: return ERR_INVAL;
            }
        }
    }
    return 0;
}

extern "C" int mpf_healpix_write_fits(const char* path, const float* map, int64_t nside, int nest, char coordsys,
                                      int clean) {
    if (!path || !map || nside < 1 || nside > (1 << 29)) return ERR_INVAL;
    const uint64_t npix = 12ull * (uint64_t)nside * (uint64_t)nside;
    std::string h;
    card(h, "SIMPLE", "T", false);
    card(h, "BITPIX", "8", false);
    card(h, "NAXIS", "0", false);
    card(h, "EXTEND", "T", false);
    end_header(h);
    card(h, "XTENSION", "BINTABLE", true);
    card(h, "BITPIX", "8", false);
    card(h, "NAXIS", "2", false);
    card(h, "NAXIS1", "4", false);
    card(h, "NAXIS2", std::to_string(npix), false);
    card(h, "PCOUNT", "0", false);
    card(h, "GCOUNT", "1", false);
    card(h, "TFIELDS", "1", false);
    card(h, "TTYPE1", "SIGNAL", true);
    card(h, "TFORM1", "1E", true);
    card(h, "PIXTYPE", "HEALPIX", true);
    card(h, "ORDERING", nest ? "NESTED" : "RING", true);
    card(h, "NSIDE", std::to_string(nside), false);
    card(h, "FIRSTPIX", "0", false);
    card(h, "LASTPIX", std::to_string(npix - 1), false);
    card(h, "COORDSYS", std::string(1, coordsys), true);
    end_header(h);
    FILE* f = std::fopen(path, "wb");
    if (!f) return ERR_IO;
    bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
    // big-endian IEEE single precision, in chunks
    std::vector<unsigned char> buf;
    const uint64_t chunk = 1 << 20;
    const float undef = -1.6375e30f;   // Healpix_undef
    for (uint64_t s0 = 0; s0 < npix && ok; s0 += chunk) {
        const uint64_t m = std::min(chunk, npix - s0);
        buf.resize(4 * m);
        for (uint64_t k = 0; k < m; ++k) {
            float v = map[s0 + k];
            if (clean && !std::isnormal(v)) v = undef;
            uint32_t u;
            std::memcpy(&u, &v, 4);
            buf[4 * k] = (unsigned char)(u >> 24);
            buf[4 * k + 1] = (unsigned char)(u >> 16);
            buf[4 * k + 2] = (unsigned char)(u >> 8);
            buf[4 * k + 3] = (unsigned char)u;
        }
        ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    }
    const uint64_t pad = (2880 - (4 * npix) % 2880) % 2880;
    std::vector<unsigned char> zeros(pad, 0);
    if (ok && pad) ok = std::fwrite(zeros.data(), 1, pad, f) == pad;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? 0 : ERR_IO;
}
