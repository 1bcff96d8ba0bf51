// Catalogue files (host): the output of Catalogues::relCat
// (catalogues.h:862-912) and of relCat_with_previous_cat
// (catalogues.h:1008-1040).  The columns, the routing of each source to the
// catalogue, error or rejection file, and the number formatting (17
// significant digits, or the stream's default 6 for the error file) are the
// reference's, so the files are byte-identical.
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <string>

#include "../../include/mpf.h"

namespace {

constexpr int ERR_INVAL = -22;
constexpr int ERR_IO = -5;

// relCat's catalogue array (catalogues.h:844-859) from a row of MPF_NCAT columns
void catalog16(const double* o, double* c) {
    c[0] = o[MPF_CAT_PHI];
    c[1] = o[MPF_CAT_THETA];
    c[2] = o[MPF_CAT_PHI_OBS];
    c[3] = o[MPF_CAT_THETA_OBS];
    c[4] = o[MPF_CAT_SEP];
    c[5] = o[MPF_CAT_ITER];
    for (int k = 0; k < 6; ++k) c[6 + k] = o[MPF_CAT_Z0 + k];
    c[12] = o[MPF_CAT_A11];
    c[13] = o[MPF_CAT_A12];
    c[14] = o[MPF_CAT_A21];
    c[15] = o[MPF_CAT_A22];
}

void full_row(std::ofstream& f, double id, const double* c, double npart) {
    f << std::setprecision(17) << id;
    for (int k = 0; k < 16; ++k) f << " " << c[k];
    f << " " << npart << std::endl;
}

}  // namespace

extern "C" {

int mpf_catalogue_write(const char* prefix, const double* out, uint64_t n, const double* id_npart) {
    if (!prefix || (n && (!out || !id_npart))) return ERR_INVAL;
    const std::string base(prefix);
    std::ofstream cat(base + ".txt", std::ofstream::out | std::ofstream::trunc);
    std::ofstream err(base + "_err.txt", std::ofstream::out | std::ofstream::trunc);
    std::ofstream rej(base + "_reject.txt", std::ofstream::out | std::ofstream::trunc);
    if (!cat || !err || !rej) return ERR_IO;
    for (uint64_t i = 0; i < n; ++i) {
        double c[16];
        catalog16(out + (uint64_t)MPF_NCAT * i, c);
        const double id = id_npart[2 * i], npart = id_npart[2 * i + 1];
        if (c[2] != 42 && c[3] != 42 && c[5] != 42) {
            full_row(cat, id, c, npart);
        } else if (c[2] == 42 && c[3] == 42) {
            err << id << " " << c[0] << " " << c[1] << " " << c[2] << " " << c[3] << " " << c[4] << " " << c[4]
                << std::endl;
        } else {
            full_row(rej, id, c, npart);
        }
    }
    return (cat && err && rej) ? 0 : ERR_IO;
}

int mpf_catalogue_write_previous(const char* prefix, const double* rows18, uint64_t n) {
    if (!prefix || (n && !rows18)) return ERR_INVAL;
    const std::string base(prefix);
    std::ofstream cat(base + ".txt", std::ofstream::out | std::ofstream::trunc);
    std::ofstream err(base + "_err.txt", std::ofstream::out | std::ofstream::trunc);
    if (!cat || !err) return ERR_IO;
    for (uint64_t i = 0; i < n; ++i) {
        const double* r = rows18 + 18 * i;
        std::ofstream& f = (r[13] != 42 && r[14] != 42) ? cat : err;
        f << std::setprecision(17) << r[0];
        for (int k = 1; k < 18; ++k) f << " " << r[k];
        f << std::endl;
    }
    return (cat && err) ? 0 : ERR_IO;
}

}  // extern "C"
