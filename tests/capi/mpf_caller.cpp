// A compiled C++ caller of libmpf, making the calls that rays.cpp and
// catalogues.cpp would make at the drop-in boundary (include/mpf.h): the
// narrow cones from info_narrow_cone.txt (rays.cpp:186-198), a batch of rays
// integrated in one call (the Integrator::integrate loop, rays.cpp:276-284),
// and the catalogue's Newton root-finding (Catalogues::relCat,
// catalogues.h:803-920).  Inputs and outputs are raw little-endian files, so
// that a test can compare them with the golden vectors.
//
//   mpf_caller narrow <celldir/> <ncones> <out.bin>
//       cones7, cones_rot7 (ncones x 7 each), rotm1 (9) as doubles
//   mpf_caller rays <dir> <order> <precision> <nsteps> <length> <zstop>
//       dir/records.bin (48-byte cone records), dir/table.bin (4 x m: t, a,
//       da/dt, d2a/dt2), dir/in9.bin (n x 9) -> dir/out.bin (n x MPF_NOUT)
//   mpf_caller catalogue <dir> <order> <precision> <nsteps> <length> <accuracy>
//       as rays with dir/src6.bin (n x 6) -> dir/cat.bin (n x MPF_NCAT)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mpf.h"

namespace {

bool read_file(const std::string& path, std::vector<unsigned char>& buf) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    buf.resize(size > 0 ? (size_t)size : 0);
    const size_t got = buf.empty() ? 0 : std::fread(buf.data(), 1, buf.size(), f);
    std::fclose(f);
    return got == buf.size();
}

std::vector<double> read_doubles(const std::string& path) {
    std::vector<unsigned char> b;
    if (!read_file(path, b) || b.size() % sizeof(double)) {
        std::fprintf(stderr, "cannot read %s\n", path.c_str());
        std::exit(2);
    }
    std::vector<double> v(b.size() / sizeof(double));
    if (!v.empty()) std::memcpy(v.data(), b.data(), b.size());
    return v;
}

bool write_doubles(const std::string& path, const double* p, size_t n) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const size_t put = n ? std::fwrite(p, sizeof(double), n, f) : 0;
    std::fclose(f);
    return put == n;
}

int narrow(const std::string& celldir, unsigned ncones, const std::string& out) {
    double ang[4];
    if (mpf_narrow_cone_info((celldir + "info_narrow_cone.txt").c_str(), ang) != 0) {
        std::fprintf(stderr, "%s\n", mpf_cone_last_error());
        return 1;
    }
    std::vector<double> res(14 * (size_t)ncones + 9);
    double* cones = res.data();
    double* rot = cones + 7 * (size_t)ncones;
    double* rotm1 = rot + 7 * (size_t)ncones;
    if (mpf_cones_narrow(ncones, 2.0, ang[0], ang[1], ang[2], ang[3], cones, rot, rotm1, nullptr) != 0) {
        std::fprintf(stderr, "%s\n", mpf_cone_last_error());
        return 1;
    }
    return write_doubles(out, res.data(), res.size()) ? 0 : 1;
}

// Context over the octree of dir/records.bin and the cosmology of dir/table.bin.
mpf_ctx* open_context(const std::string& dir, int order, int precision, unsigned nsteps, double length) {
    std::vector<unsigned char> rec;
    if (!read_file(dir + "/records.bin", rec) || rec.size() % 48) {
        std::fprintf(stderr, "cannot read %s/records.bin\n", dir.c_str());
        return nullptr;
    }
    const std::vector<double> tab = read_doubles(dir + "/table.bin");
    const uint64_t m = tab.size() / 4;
    mpf_config cfg = {/*device*/ 0, order, precision, nsteps, length};
    mpf_ctx* ctx = nullptr;
    if (mpf_create(&ctx, &cfg) != 0) {
        std::fprintf(stderr, "mpf_create: %s\n", mpf_last_error(ctx));
        if (ctx) mpf_destroy(ctx);
        return nullptr;
    }
    if (mpf_load_octree_records(ctx, rec.data(), rec.size() / 48) != 0 ||
        mpf_set_cosmology(ctx, tab.data(), tab.data() + m, tab.data() + 2 * m, m) != 0) {
        std::fprintf(stderr, "%s\n", mpf_last_error(ctx));
        mpf_destroy(ctx);
        return nullptr;
    }
    return ctx;
}

int rays(const std::string& dir, int order, int precision, unsigned nsteps, double length, double zstop) {
    mpf_ctx* ctx = open_context(dir, order, precision, nsteps, length);
    if (!ctx) return 1;
    const std::vector<double> in9 = read_doubles(dir + "/in9.bin");
    const uint64_t n = in9.size() / 9;
    mpf_stop stop = {MPF_STOP_REDSHIFT, zstop, {0, 0, 0}};
    std::vector<double> out(MPF_NOUT * n);
    uint64_t steps = 0;
    const int rc = mpf_integrate(ctx, in9.data(), n, &stop, out.data(), &steps);
    if (rc != 0) std::fprintf(stderr, "mpf_integrate: %s\n", mpf_last_error(ctx));
    mpf_destroy(ctx);
    if (rc != 0 || !write_doubles(dir + "/out.bin", out.data(), out.size())) return 1;
    std::printf("steps %llu\n", (unsigned long long)steps);
    return 0;
}

int catalogue(const std::string& dir, int order, int precision, unsigned nsteps, double length, double accuracy) {
    mpf_ctx* ctx = open_context(dir, order, precision, nsteps, length);
    if (!ctx) return 1;
    const std::vector<double> src6 = read_doubles(dir + "/src6.bin");
    const uint64_t n = src6.size() / 6;
    mpf_newton_cfg ncfg = {};
    ncfg.accuracy = accuracy;
    ncfg.max_iterations = 10;       // catalogues.h:765
    ncfg.h = 0.7;
    ncfg.mpc = 3.08568e22;
    std::vector<double> rows(MPF_NCAT * n);
    uint64_t nrays = 0;
    const int rc = mpf_catalogue_solve(ctx, src6.data(), nullptr, n, &ncfg, rows.data(), &nrays);
    if (rc != 0) std::fprintf(stderr, "mpf_catalogue_solve: %s\n", mpf_last_error(ctx));
    mpf_destroy(ctx);
    if (rc != 0 || !write_doubles(dir + "/cat.bin", rows.data(), rows.size())) return 1;
    std::printf("rays %llu\n", (unsigned long long)nrays);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "narrow" && argc == 5) return narrow(argv[2], (unsigned)std::strtoul(argv[3], nullptr, 10), argv[4]);
    if ((mode == "rays" || mode == "catalogue") && argc == 8) {
        const int order = std::atoi(argv[3]), precision = std::atoi(argv[4]);
        const unsigned nsteps = (unsigned)std::strtoul(argv[5], nullptr, 10);
        const double length = std::strtod(argv[6], nullptr), last = std::strtod(argv[7], nullptr);
        return mode == "rays" ? rays(argv[2], order, precision, nsteps, length, last)
                              : catalogue(argv[2], order, precision, nsteps, length, last);
    }
    std::fprintf(stderr, "usage: mpf_caller narrow <celldir/> <ncones> <out.bin>\n"
                         "       mpf_caller rays|catalogue <dir> <order> <precision> <nsteps> <length> <zstop|acc>\n");
    return 2;
}
