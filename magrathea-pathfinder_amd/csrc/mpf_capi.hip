// C ABI of libmpf.so (declared in include/mpf.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mpf.h"
#include "mpf_kernels.h"
#include "mpf_split.h"
#include "mpf_tree.h"
#include "mpf_healpix.h"
#include "mpf_order.h"
#include "mpf_plane.h"
#include "mpf_propagate.h"
#include "mpf_cone.h"
#include "mpf_mlp.h"
#include "mpf_cosmology.h"

namespace {

constexpr int ERR_INVAL = -22;
constexpr int ERR_NOMEM = -12;
constexpr int ERR_HIP = -5;
constexpr int ERR_STATE = -71;
constexpr int ERR_NOSPC = -28;

__global__ void fill_u64_kernel(unsigned long long* p, uint64_t n, unsigned long long v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(256) copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t ensure(size_t count) {
        if (count <= n && p) return hipSuccess;
        release();
        hipError_t e = hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) n = count;
        else p = nullptr;
        return e;
    }
};

// Workspace of one integration: the queue and step counters, the ray order
// and the work area of its sort, the split tier's launch records.  The
// context's own buffers grow on demand (mpf_integrate_dev); the
// allocation-free entry point (mpf_integrate_dev_static) carves them out of
// the file-scope array below instead.
struct Workspace {
    unsigned long long* counters = nullptr;  // [0] queue, [1] steps, [2] trajectory overflow
    uint32_t* order = nullptr;
    void* order_work = nullptr;
    size_t order_work_bytes = 0;
    double* records = nullptr;
};

// The allocation-free integration's workspace: one array per GPU, held until
// the process ends.  In host code `scratch` names a host-side stand-in, not
// device memory: the device address of each GPU's copy comes from
// hipGetSymbolAddress, taken once per device by mpf_create with that device
// current, and only that address is ever passed to kernels.  Every call on a
// GPU shares its copy, so calls that overlap on one GPU race.
constexpr size_t STATIC_SCRATCH_BYTES = 48ull << 20;
static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};
hipError_t g_scratch_err[MAX_DEVICES] = {};

// the device address of this device's workspace array (device current)
void scratch_address(int device) {
    if (device < 0 || device >= MAX_DEVICES) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) g_scratch_ptr[device] = p;
    else g_scratch_err[device] = (e == hipSuccess) ? hipErrorInvalidSymbol : e;
}

// Layout of n rays in the static workspace: counters (256 B), the order
// (4 B per ray), its sort's work area, the launch records (128 B per ray);
// false when it does not fit.
bool static_layout(void* base, uint64_t n, bool split, Workspace& w) {
    unsigned char* p = static_cast<unsigned char*>(base);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* q = p + off;
        off += (bytes + 255) / 256 * 256;
        return q;
    };
    w.counters = reinterpret_cast<unsigned long long*>(take(256));
    w.order = reinterpret_cast<uint32_t*>(take(4 * std::max<uint64_t>(n, 1)));
    w.order_work_bytes = mpf::ray_order_work_bytes(std::max<uint64_t>(n, 2));
    w.order_work = take(w.order_work_bytes);
    w.records = split ? reinterpret_cast<double*>(take(8ull * mpf::split_launch_doubles() * n)) : nullptr;
    return off <= STATIC_SCRATCH_BYTES;
}

}  // namespace

namespace mpf {
hipError_t launch_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}
}  // namespace mpf

struct mpf_ctx {
    mpf_config cfg{};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t evk0 = nullptr, evk1 = nullptr;  // the persistent kernel alone
    int cus = 256;
    std::string err;
    // octree
    bool have_tree = false;
    uint64_t ncells = 0;
    uint32_t min_level = 0, max_level = 0;
    std::vector<uint32_t> where;  // record -> level-sorted index
    std::vector<uint8_t> lev;     // level-sorted cell levels and coordinates
    std::vector<uint32_t> kx, ky, kz;
    DevBuf<uint32_t> base, child_first, elem;
    DevBuf<uint8_t> child_mask;
    DevBuf<float> f32;
    DevBuf<double> f64;
    mpf::DevTree dt;
    mpf::SplitTree split;
    bool split_ready = false;
    bool split_rho_failed = false;   // the rho blocks did not fit: split surfaces carry NaN densities
    bool split_tidal_failed = false; // the tidal blocks did not fit: the Jacobian samples every neighbour
    bool split_tidal = true;         // tidal blocks allowed (mpf_split_tidal_blocks)
    // cell-major field copies of the lean exact integrator, built at its first TSC run
    DevBuf<float> q32;
    DevBuf<double> q64;
    bool quad_ready = false;
    // velocity field: FP64 sums (n x 4) between begin and finish, the field
    // (3 x n) the tree carries after it, particle staging, parent work array
    DevBuf<double> vacc;
    DevBuf<float> vel, vstage;
    DevBuf<uint32_t> vparent;
    DevBuf<unsigned char> vwork;
    bool vel_open = false;
    double vobs[3] = {0.0, 0.0, 0.0};
    // cosmology
    DevBuf<double> cosmo;
    int ncos = 0;
    // run state
    DevBuf<unsigned long long> counters;  // [0] queue, [1] steps
    DevBuf<double> io_in, io_out, io_surf, io_rv, scratch;
    DevBuf<unsigned long long> map_acc;
    DevBuf<int64_t> io_i64;
    DevBuf<uint32_t> io_u32;
    uint64_t last_steps = 0;
    float last_ms = 0.f, last_kernel_ms = 0.f;
    bool kernel_timed = false;  // the last integration recorded evk0/evk1
    bool surf_jac_chi = false;  // surfaces take the Jacobian at the interpolated chi
    uint64_t bundle_plane_fallback = 0;  // directions of the last plane bundle from whole trajectories
    bool pending = false;
    // learned first guess of the catalogue root-finder
    mpf::MlpModel* mlp = nullptr;
    // launch configuration of the persistent integrators: [0] without
    // surfaces, [1] with surfaces; the ray order they read
    mpf_launch_cfg launch[2] = {{64, 65536, 0, 0}, {64, 65536, 0, 0}};
    DevBuf<uint32_t> order;
    DevBuf<double> homo;   // homogeneous values of the map surfaces
    DevBuf<unsigned char> order_work;
};

static int fail(mpf_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

static int hipfail(mpf_ctx* c, hipError_t e, const char* where) {
    return fail(c, ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIPCHK(c, expr)                                \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return hipfail(c, e_, #expr); \
    } while (0)

// stencil blocks of the split tier, built on first use after a tree load
static int ensure_split(mpf_ctx* c) {
    if (c->split_ready) return 0;
    std::string err;
    if (!mpf::split_build(c->split, c->dt, c->lev.data(), c->kx.data(), c->ky.data(), c->kz.data(), c->cfg.order,
                          c->stream, err))
        return fail(c, ERR_NOMEM, err);
    c->split_ready = true;
    return 0;
}

extern "C" {

int mpf_create(mpf_ctx** out, const mpf_config* cfg) {
    if (!out || !cfg) return ERR_INVAL;
    *out = nullptr;
    if (cfg->order < MPF_HOMOGENEOUS || cfg->order > 2 || cfg->precision < 0 || cfg->precision > 2 || cfg->nsteps == 0 ||
        !(cfg->length > 0))
        return ERR_INVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ERR_HIP;
    if (cfg->device < 0 || cfg->device >= ndev) return ERR_INVAL;
    mpf_ctx* c = new mpf_ctx;
    c->cfg = *cfg;
    if (hipSetDevice(cfg->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->evk0) != hipSuccess || hipEventCreate(&c->evk1) != hipSuccess) {
        delete c;
        return ERR_HIP;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, cfg->device) == hipSuccess && cus > 0)
        c->cus = cus;
    scratch_address(cfg->device);
    if (c->counters.ensure(3) != hipSuccess) {
        delete c;
        return ERR_NOMEM;
    }
    *out = c;
    return 0;
}

int mpf_destroy(mpf_ctx* c) {
    if (!c) return 0;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->base.release(); c->child_first.release(); c->elem.release(); c->child_mask.release();
    c->f32.release(); c->f64.release(); c->cosmo.release(); c->counters.release();
    c->q32.release(); c->q64.release();
    c->io_in.release(); c->io_out.release(); c->io_surf.release(); c->io_rv.release(); c->scratch.release();
    c->order.release(); c->order_work.release(); c->homo.release();
    c->map_acc.release(); c->io_i64.release(); c->io_u32.release();
    c->vacc.release(); c->vel.release(); c->vstage.release(); c->vparent.release(); c->vwork.release();
    mpf::split_release(c->split);
    mpf::mlp_destroy(c->mlp);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->evk0) (void)hipEventDestroy(c->evk0);
    if (c->evk1) (void)hipEventDestroy(c->evk1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

const char* mpf_last_error(const mpf_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mpf_set_precision(mpf_ctx* c, int precision) {
    if (!c) return ERR_INVAL;
    if (precision < 0 || precision > 2) return fail(c, ERR_INVAL, "precision must be 0, 1 or 2");
    c->cfg.precision = precision;
    return 0;
}

int mpf_load_octree_records(mpf_ctx* c, const void* rec48, uint64_t n) {
    if (!c || (!rec48 && n)) return ERR_INVAL;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    mpf::HostTree t;
    std::string err;
    if (!mpf::build_tree(rec48, n, t, err)) return fail(c, ERR_INVAL, err);
    std::vector<float> soa;
    mpf::unpack_fields32(rec48, t, soa);
    c->have_tree = false;
    c->split_ready = false;
    c->split_rho_failed = false;
    c->split_tidal_failed = false;
    c->quad_ready = false;
    c->f64.release();
    if (c->base.ensure(t.base.size()) != hipSuccess || c->child_first.ensure(n) != hipSuccess ||
        c->elem.ensure(n) != hipSuccess || c->child_mask.ensure(n) != hipSuccess || c->f32.ensure(7 * n) != hipSuccess)
        return fail(c, ERR_NOMEM, "device allocation of the octree failed");
    HIPCHK(c, hipMemcpyAsync(c->base.p, t.base.data(), t.base.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->child_first.p, t.child_first.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->elem.p, t.elem.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->child_mask.p, t.child_mask.data(), n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->f32.p, soa.data(), soa.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    mpf::DevTree& d = c->dt;
    d = mpf::DevTree();  // no velocity field
    c->vel_open = false;
    d.B = t.min_level;
    d.maxL = t.max_level;
    d.bx0 = t.bx0; d.by0 = t.by0; d.bz0 = t.bz0;
    d.bnx = t.bnx; d.bny = t.bny; d.bnz = t.bnz;
    d.base = c->base.p;
    d.child_first = c->child_first.p;
    d.child_mask = c->child_mask.p;
    d.elem = c->elem.p;
    d.f32 = c->f32.p;
    d.f64 = nullptr;
    d.n = n;
    c->ncells = n;
    c->min_level = t.min_level;
    c->max_level = t.max_level;
    c->where.swap(t.where);
    c->lev.swap(t.level);
    c->kx.swap(t.kx);
    c->ky.swap(t.ky);
    c->kz.swap(t.kz);
    c->have_tree = true;
    return 0;
}

int mpf_load_octree_fields64(mpf_ctx* c, const double* f7, uint64_t n) {
    if (!c || !f7) return ERR_INVAL;
    if (!c->have_tree) return fail(c, ERR_STATE, "load the octree records first");
    if (n != c->ncells) return fail(c, ERR_INVAL, "fields64 count differs from the octree cell count");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    std::vector<double> soa(7 * n);
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t g = c->where[i];
        for (int k = 0; k < 7; ++k) soa[k * n + g] = f7[7 * i + k];
    }
    if (c->f64.ensure(7 * n) != hipSuccess) return fail(c, ERR_NOMEM, "device allocation of fields64 failed");
    HIPCHK(c, hipMemcpyAsync(c->f64.p, soa.data(), soa.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->dt.f64 = c->f64.p;
    c->split_ready = false;
    c->split_rho_failed = false;
    c->split_tidal_failed = false;
    c->quad_ready = false;
    return 0;
}

int mpf_octree_info(const mpf_ctx* c, uint64_t* ncells, uint32_t* min_level, uint32_t* max_level) {
    if (!c) return ERR_INVAL;
    if (ncells) *ncells = c->have_tree ? c->ncells : 0;
    if (min_level) *min_level = c->min_level;
    if (max_level) *max_level = c->max_level;
    return 0;
}

// ---- velocity field

int mpf_set_observer_velocity(mpf_ctx* c, const double* vobs3) {
    if (!c || !vobs3) return ERR_INVAL;
    for (int k = 0; k < 3; ++k) c->vobs[k] = vobs3[k];
    return 0;
}

int mpf_velocity_begin(mpf_ctx* c) {
    if (!c) return ERR_INVAL;
    if (!c->have_tree) return fail(c, ERR_STATE, "no octree loaded");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->vacc.ensure(4 * c->ncells) != hipSuccess) return fail(c, ERR_NOMEM, "velocity sums");
    HIPCHK(c, hipMemsetAsync(c->vacc.p, 0, 4 * c->ncells * sizeof(double), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->vel_open = true;
    return 0;
}

// particles in chunks of at most 2^27, each put in Morton order (locality of
// the tree descent and of the atomic adds; the sums do not depend on it)
static int velocity_deposit_impl(mpf_ctx* c, const float* dpos, const float* dvel, uint64_t n, int scheme) {
    const uint64_t chunk = std::min<uint64_t>(n, 1ull << 27);
    const size_t wb = mpf::velocity_order_work_bytes(chunk);
    if (c->vwork.ensure(wb) != hipSuccess) return fail(c, ERR_NOMEM, "particle order work area");
    for (uint64_t first = 0; first < n; first += chunk) {
        const uint64_t m = std::min<uint64_t>(chunk, n - first);
        const float *ps = nullptr, *vs = nullptr;
        HIPCHK(c, mpf::launch_velocity_order(dpos + 3 * first, dvel + 3 * first, m, c->vwork.p, &ps, &vs, c->stream));
        for (uint32_t l = c->min_level; l <= c->max_level; ++l)
            HIPCHK(c, mpf::launch_velocity_deposit(c->dt, scheme, ps, vs, m, l, c->vacc.p, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

static int velocity_check(mpf_ctx* c, uint64_t n, const void* pos, const void* vel, int scheme) {
    if (!c->vel_open) return fail(c, ERR_STATE, "call mpf_velocity_begin first");
    if (n && (!pos || !vel)) return fail(c, ERR_INVAL, "null particle arrays");
    if (scheme != MPF_VELOCITY_CIC && scheme != MPF_VELOCITY_TSC)
        return fail(c, ERR_INVAL, "velocity field scheme must be cic (1) or tsc (2)");
    return 0;
}

int mpf_velocity_deposit(mpf_ctx* c, const float* pos3, const float* vel3, uint64_t n, int scheme) {
    if (!c) return ERR_INVAL;
    if (int rc = velocity_check(c, n, pos3, vel3, scheme)) return rc;
    if (n == 0) return 0;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->vstage.ensure(6 * n) != hipSuccess) return fail(c, ERR_NOMEM, "particle staging");
    HIPCHK(c, hipMemcpyAsync(c->vstage.p, pos3, 3 * n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->vstage.p + 3 * n, vel3, 3 * n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    return velocity_deposit_impl(c, c->vstage.p, c->vstage.p + 3 * n, n, scheme);
}

int mpf_velocity_deposit_dev(mpf_ctx* c, const float* dev_pos3, const float* dev_vel3, uint64_t n, int scheme) {
    if (!c) return ERR_INVAL;
    if (int rc = velocity_check(c, n, dev_pos3, dev_vel3, scheme)) return rc;
    if (n == 0) return 0;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    return velocity_deposit_impl(c, dev_pos3, dev_vel3, n, scheme);
}

int mpf_velocity_finish(mpf_ctx* c) {
    if (!c) return ERR_INVAL;
    if (!c->vel_open) return fail(c, ERR_STATE, "call mpf_velocity_begin first");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const uint64_t n = c->ncells;
    if (c->vel.ensure(3 * n) != hipSuccess || c->vparent.ensure(n) != hipSuccess)
        return fail(c, ERR_NOMEM, "velocity field");
    // first level-sorted cell of every level (the levels are stored in order)
    std::vector<uint64_t> first(c->max_level - c->min_level + 2, n);
    for (uint32_t l = c->min_level; l <= c->max_level + 1; ++l)
        first[l - c->min_level] =
            (uint64_t)(std::lower_bound(c->lev.begin(), c->lev.end(), (uint8_t)std::min<uint32_t>(l, 255)) -
                       c->lev.begin());
    first.back() = n;
    HIPCHK(c, hipMemsetAsync(c->vparent.p, 0xFF, n * sizeof(uint32_t), c->stream));
    HIPCHK(c, mpf::launch_velocity_finish(c->dt, c->vacc.p, c->f32.p, c->dt.f64 ? c->f64.p : nullptr, c->vel.p,
                                          c->vparent.p, first.data(), c->min_level, c->max_level, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->dt.vel = c->vel.p;
    c->vel_open = false;
    c->split_ready = false;
    c->split_rho_failed = false;
    c->split_tidal_failed = false;
    c->quad_ready = false;  // rho changed
    return 0;
}

int mpf_velocity_set(mpf_ctx* c, const float* rho, const float* vel3, uint64_t n) {
    if (!c || !vel3) return ERR_INVAL;
    if (!c->have_tree) return fail(c, ERR_STATE, "no octree loaded");
    if (n != c->ncells) return fail(c, ERR_INVAL, "count differs from the octree cell count");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    std::vector<float> soa(3 * n);
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t g = c->where[i];
        for (int k = 0; k < 3; ++k) soa[k * n + g] = vel3[3 * i + k];
    }
    if (c->vel.ensure(3 * n) != hipSuccess) return fail(c, ERR_NOMEM, "velocity field");
    HIPCHK(c, hipMemcpyAsync(c->vel.p, soa.data(), 3 * n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (rho) {
        std::vector<float> r(n);
        std::vector<double> r64(c->dt.f64 ? n : 0);
        for (uint64_t i = 0; i < n; ++i) r[c->where[i]] = rho[i];
        for (uint64_t g = 0; g < r64.size(); ++g) r64[g] = (double)r[g];
        HIPCHK(c, hipMemcpyAsync(c->f32.p, r.data(), n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        if (c->dt.f64)
            HIPCHK(c, hipMemcpyAsync(c->f64.p, r64.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        c->split_ready = false;
    c->split_rho_failed = false;
    c->split_tidal_failed = false;
    c->quad_ready = false;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->dt.vel = c->vel.p;
    c->vel_open = false;
    return 0;
}

int mpf_velocity_get(mpf_ctx* c, float* rho, float* vel3, uint64_t n) {
    if (!c) return ERR_INVAL;
    if (!c->have_tree) return fail(c, ERR_STATE, "no octree loaded");
    if (!c->dt.vel) return fail(c, ERR_STATE, "the tree carries no velocity field");
    if (n != c->ncells) return fail(c, ERR_INVAL, "count differs from the octree cell count");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    std::vector<float> soa(3 * n), r(n);
    HIPCHK(c, hipMemcpyAsync(soa.data(), c->vel.p, 3 * n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(r.data(), c->f32.p, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t g = c->where[i];
        if (rho) rho[i] = r[g];
        if (vel3)
            for (int k = 0; k < 3; ++k) vel3[3 * i + k] = soa[k * n + g];
    }
    return 0;
}

int mpf_velocity_clear(mpf_ctx* c) {
    if (!c) return ERR_INVAL;
    c->dt.vel = nullptr;
    c->vel_open = false;
    return 0;
}

int mpf_validate_octree(const void* rec48, uint64_t n, uint32_t* min_level, uint32_t* max_level, char* msg,
                        uint64_t msg_size) {
    if (!rec48 && n) return ERR_INVAL;
    mpf::HostTree t;
    std::string err;
    const bool ok = mpf::build_tree(rec48, n, t, err);
    if (msg && msg_size) {
        const std::string m = ok ? std::string("ok") : err;
        const size_t k = std::min<size_t>(m.size(), msg_size - 1);
        std::memcpy(msg, m.data(), k);
        msg[k] = 0;
    }
    if (!ok) return ERR_INVAL;
    if (min_level) *min_level = t.min_level;
    if (max_level) *max_level = t.max_level;
    return 0;
}

int mpf_set_cosmology(mpf_ctx* c, const double* t, const double* a, const double* dadt, uint64_t n) {
    if (!c || !t || !a || !dadt) return ERR_INVAL;
    if (n == 0 || n > (1u << 30)) return fail(c, ERR_INVAL, "cosmology table size");
    for (uint64_t i = 1; i < n; ++i)
        if (!(t[i - 1] <= t[i])) return fail(c, ERR_INVAL, "cosmology t column is not sorted");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->cosmo.ensure(4 * n) != hipSuccess) return fail(c, ERR_NOMEM, "cosmology allocation");
    // fourth column: the slope of da/dt on each segment (the split tier's interpolation)
    std::vector<double> slope(n, 0.0);
    for (uint64_t i = 1; i < n; ++i)
        slope[i] = (t[i] > t[i - 1]) ? (dadt[i] - dadt[i - 1]) / (t[i] - t[i - 1]) : 0.0;
    HIPCHK(c, hipMemcpyAsync(c->cosmo.p, t, n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->cosmo.p + n, a, n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->cosmo.p + 2 * n, dadt, n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->cosmo.p + 3 * n, slope.data(), n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->ncos = (int)n;
    return 0;
}

int mpf_locate(mpf_ctx* c, const double* xyz, uint64_t n, int64_t* elem_out, uint32_t* level_out) {
    if (!c || (n && (!xyz || !elem_out || !level_out))) return ERR_INVAL;
    if (!c->have_tree) return fail(c, ERR_STATE, "no octree loaded");
    if (n == 0) return 0;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->io_in.ensure(3 * n) != hipSuccess || c->io_i64.ensure(n) != hipSuccess || c->io_u32.ensure(n) != hipSuccess)
        return fail(c, ERR_NOMEM, "locate buffers");
    HIPCHK(c, hipMemcpyAsync(c->io_in.p, xyz, 3 * n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, mpf::launch_locate(c->dt, c->io_in.p, n, c->io_i64.p, c->io_u32.p, c->stream));
    HIPCHK(c, hipMemcpyAsync(elem_out, c->io_i64.p, n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(level_out, c->io_u32.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int mpf_interpolate(mpf_ctx* c, int order, const double* xyz, uint64_t n, double* out7) {
    if (!c || (n && (!xyz || !out7))) return ERR_INVAL;
    if (order < 0 || order > 3) return fail(c, ERR_INVAL, "order must be 0..3");
    if (!c->have_tree) return fail(c, ERR_STATE, "no octree loaded");
    if (c->cfg.precision == MPF_PREC_SPLIT) return fail(c, ERR_INVAL, "interpolation is offered for the exact tiers");
    if (n == 0) return 0;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->io_in.ensure(3 * n) != hipSuccess || c->io_out.ensure(7 * n) != hipSuccess)
        return fail(c, ERR_NOMEM, "interpolation buffers");
    HIPCHK(c, hipMemcpyAsync(c->io_in.p, xyz, 3 * n * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, mpf::launch_interp(c->cfg.precision, c->dt, order, c->io_in.p, n, c->io_out.p, c->stream));
    HIPCHK(c, hipMemcpyAsync(out7, c->io_out.p, 7 * n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// planes crossed by bundle rays (IntegrateArgs::nplane .. plane_out), device pointers
struct PlaneSet {
    int n = 0;
    unsigned div = 4;
    const double* val = nullptr;
    const double* kit = nullptr;
    double* out = nullptr;
};

static int integrate_impl(mpf_ctx* c, const double* dev_in9, uint64_t n, const mpf_stop* stop, double* dev_out,
                          const mpf_surfaces* surf, double* dev_surf, int sync, double* dev_traj = nullptr,
                          uint64_t traj_cap = 0, const Workspace* ws = nullptr, const PlaneSet* planes = nullptr);
static int finish_run(mpf_ctx* c, const unsigned long long* counters);

int mpf_integrate_dev(mpf_ctx* c, const double* dev_in9, uint64_t n, const mpf_stop* stop, double* dev_out, int sync) {
    return integrate_impl(c, dev_in9, n, stop, dev_out, nullptr, nullptr, sync);
}

int mpf_integrate_surfaces_dev(mpf_ctx* c, const double* dev_in9, uint64_t n, const mpf_surfaces* surf,
                               double* dev_out, double* dev_out_surf, int sync) {
    if (!c || !surf) return ERR_INVAL;
    if (surf->n < 1 || surf->n > MPF_MAX_SURFACES) return fail(c, ERR_INVAL, "surface count must be 1..16");
    if (surf->ray_value && surf->n != 1) return fail(c, ERR_INVAL, "per-ray surface values need exactly one surface");
    if (surf->kind < 0 || surf->kind > 6 || surf->kind == MPF_STOP_PLANE)
        return fail(c, ERR_INVAL, "surface kind must be radius, lambda, redshift, t, a or s");
    for (uint32_t j = 1; j < surf->n; ++j) {
        const bool ok = (surf->kind == MPF_STOP_A) ? (surf->value[j] < surf->value[j - 1]) : (surf->value[j] > surf->value[j - 1]);
        if (!ok) return fail(c, ERR_INVAL, "surface values must be strictly monotonic (decreasing for a)");
    }
    if (n && !dev_out_surf) return ERR_INVAL;
    mpf_stop stop{surf->kind, surf->ray_value ? 0.0 : surf->value[surf->n - 1], {0.0, 0.0, 0.0}};
    return integrate_impl(c, dev_in9, n, &stop, dev_out, surf, dev_out_surf, sync);
}

static int integrate_impl(mpf_ctx* c, const double* dev_in9, uint64_t n, const mpf_stop* stop, double* dev_out,
                          const mpf_surfaces* surf, double* dev_surf, int sync, double* dev_traj, uint64_t traj_cap,
                          const Workspace* ws, const PlaneSet* planes) {
    if (!c || !stop || (n && !dev_in9)) return ERR_INVAL;
    if (ws && (surf || dev_traj || planes || (n && !dev_out))) return ERR_INVAL;
    if (planes && (surf || dev_traj || stop->kind != MPF_STOP_PLANE || planes->n < 1 ||
                   planes->n > MPF_MAX_SURFACES || planes->div < 1 || c->cfg.precision == MPF_PREC_SPLIT))
        return ERR_INVAL;
    if (ws && c->cfg.precision == MPF_PREC_SPLIT && !c->split_ready)
        return fail(c, ERR_STATE, "split-tier stencil blocks not built (mpf_integrate_dev builds them on first use)");
    unsigned long long* counters = ws ? ws->counters : c->counters.p;
    if (n && !dev_out) {
        if (c->io_out.ensure((uint64_t)MPF_NOUT * n) != hipSuccess) return fail(c, ERR_NOMEM, "output buffer");
        dev_out = c->io_out.p;
    }
    if (!c->have_tree) return fail(c, ERR_STATE, "no octree loaded");
    if (c->ncos <= 0) return fail(c, ERR_STATE, "no cosmology table set");
    if (stop->kind < -1 || stop->kind > 6) return fail(c, ERR_INVAL, "unknown stop criterion");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    mpf::IntegrateArgs A;
    A.in9 = dev_in9;
    A.out = dev_out;
    A.n = n;
    A.queue = counters;
    A.steps = counters + 1;
    A.ct = c->cosmo.p;
    A.ca = c->cosmo.p + c->ncos;
    A.cdadt = c->cosmo.p + 2 * c->ncos;
    A.cslope = c->cosmo.p + 3 * c->ncos;
    A.ncos = c->ncos;
    A.order = c->cfg.order;
    A.nsteps = c->cfg.nsteps;
    A.length = c->cfg.length;
    A.stop_kind = stop->kind;
    A.stop_value = stop->value;
    for (int k = 0; k < 3; ++k) A.kit[k] = stop->kitarget[k];
    if (surf) {
        A.surf_kind = surf->kind;
        A.nsurf = (int)surf->n;
        for (uint32_t j = 0; j < surf->n; ++j) A.surf_val[j] = surf->value[j];
        A.surf_out = dev_surf;
        A.jac = surf->jacobian ? 1 : 0;
        A.ray_value = surf->ray_value;
        A.jac_at_chi = c->surf_jac_chi ? 1 : 0;
    }
    if (dev_traj) {
        A.traj = dev_traj;
        A.traj_cap = traj_cap;
        A.traj_over = counters + 2;
    }
    if (planes) {
        A.nplane = planes->n;
        A.plane_div = planes->div;
        A.plane_val = planes->val;
        A.plane_kit = planes->kit;
        A.plane_out = planes->out;
    }
    if (c->cfg.order == MPF_HOMOGENEOUS && c->cfg.precision == MPF_PREC_SPLIT)
        return fail(c, ERR_INVAL, "the homogeneous order runs in the exact tiers (REF, FP64)");
    if (dev_traj && c->cfg.precision == MPF_PREC_SPLIT)
        return fail(c, ERR_INVAL, "trajectories are written by the exact tiers (REF, FP64)");
    if (c->dt.vel && c->cfg.precision == MPF_PREC_SPLIT)
        return fail(c, ERR_INVAL, "a tree with a velocity field is integrated by the exact tiers (REF, FP64)");
    for (int k = 0; k < 3; ++k) A.vobs[k] = c->vobs[k];
    const mpf_launch_cfg& L = c->launch[(surf || planes) ? 1 : 0];
    A.refill = L.refill;
    HIPCHK(c, mpf::launch_fill_u64(counters, 3, 0ull, c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    c->kernel_timed = n > 0;
    if (n) {
        if (c->cfg.precision == MPF_PREC_SPLIT) {
            if (A.nsurf && !mpf::split_supports_surfaces())
                return fail(c, ERR_INVAL, "surfaces are not available in the split tier");
            if (!c->split_ready) {
                if (int rc = ensure_split(c)) return rc;
                HIPCHK(c, hipEventRecord(c->ev0, c->stream));
            }
            // rho blocks for the surfaces' densities, built at the first run with
            // surfaces (without the memory for them the densities stay NaN)
            if (A.nsurf && !c->split.rho && !c->split_rho_failed) {
                std::string err;
                if (!mpf::split_build_rho(c->split, c->dt, c->lev.data(), c->kx.data(), c->ky.data(), c->kz.data(),
                                          c->stream, err)) {
                    c->split_rho_failed = true;
                    (void)hipGetLastError();
                }
                HIPCHK(c, hipEventRecord(c->ev0, c->stream));
            }
            // tidal blocks for the Jacobian, built at the first run with it
            // (without the memory for them every sample is looked up)
            if (A.nsurf && A.jac && c->split_tidal && !c->split.tidal && !c->split_tidal_failed) {
                std::string err;
                if (!mpf::split_build_tidal(c->split, c->dt, c->lev.data(), c->kx.data(), c->ky.data(), c->kz.data(),
                                            c->stream, err)) {
                    c->split_tidal_failed = true;
                    (void)hipGetLastError();
                }
                HIPCHK(c, hipEventRecord(c->ev0, c->stream));
            }
        }
        // the queue order (timed with the integration)
        if (L.sort_bucket && n > 1 && n < (1ull << 31)) {
            uint32_t* order;
            void* work;
            size_t wb;
            if (ws) {
                order = ws->order;
                work = ws->order_work;
                wb = ws->order_work_bytes;
            } else {
                wb = mpf::ray_order_work_bytes(n);
                if (c->order.ensure(n) != hipSuccess || c->order_work.ensure(wb) != hipSuccess)
                    return fail(c, ERR_NOMEM, "ray-order buffers");
                order = c->order.p;
                work = c->order_work.p;
            }
            HIPCHK(c, mpf::launch_ray_order(dev_in9, A.ray_value, stop->kind == MPF_STOP_A ? 1 : 0, n,
                                            L.sort_bucket, order, work, wb, c->stream));
            A.queue_order = order;
        }
        if (c->cfg.precision == MPF_PREC_SPLIT) {
            if (ws) {
                A.scratch = ws->records;
            } else {
                if (c->scratch.ensure((uint64_t)mpf::split_launch_doubles() * n) != hipSuccess)
                    return fail(c, ERR_NOMEM, "launch-record buffer");
                A.scratch = c->scratch.p;
            }
            HIPCHK(c, mpf::launch_integrate_split(A, c->dt, c->split, c->cus, L.blocks_per_cu ? L.blocks_per_cu : 8,
                                                  L.waves_per_simd ? L.waves_per_simd : 2, c->stream, c->evk0,
                                                  c->evk1));
        } else {
            // plain integration runs the lean kernel, two 256-lane blocks per CU (its TSC caches
            // take 55 KB of LDS per block); the kernel with surfaces or trajectories four
            const uint64_t blocks = (n + 255) / 256;
            const bool lean = A.nsurf == 0 && !A.traj && A.nplane == 0;
            if (lean && c->cfg.order == MPF_TSC && !c->quad_ready) {
                // the lean integrator's TSC reads each cell's fields in 16-byte loads
                c->dt.q32 = nullptr;
                c->dt.q64 = nullptr;
                if (c->q32.ensure(8 * c->ncells) != hipSuccess ||
                    (c->dt.f64 && c->q64.ensure(8 * c->ncells) != hipSuccess))
                    return fail(c, ERR_NOMEM, "cell-major field copies");
                HIPCHK(c, mpf::launch_field_quads(c->dt, c->q32.p, c->dt.f64 ? c->q64.p : nullptr, c->stream));
                c->dt.q32 = reinterpret_cast<const float4*>(c->q32.p);
                c->dt.q64 = c->dt.f64 ? reinterpret_cast<const double2*>(c->q64.p) : nullptr;
                c->quad_ready = true;
                HIPCHK(c, hipEventRecord(c->ev0, c->stream));
            }
            if (!c->quad_ready) { c->dt.q32 = nullptr; c->dt.q64 = nullptr; }
            const unsigned dflt = lean ? 2u : 4u;
            const int grid = (int)std::min<uint64_t>(blocks, (uint64_t)c->cus * (L.blocks_per_cu ? L.blocks_per_cu : dflt));
            HIPCHK(c, hipEventRecord(c->evk0, c->stream));
            HIPCHK(c, mpf::launch_integrate_exact(c->cfg.precision, A, c->dt, grid, c->stream));
            HIPCHK(c, hipEventRecord(c->evk1, c->stream));
        }
        if (A.nsurf) HIPCHK(c, mpf::launch_surface_final(A, c->stream));
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->pending = true;
    if (ws) return finish_run(c, counters);
    if (sync) return mpf_synchronize(c);
    return 0;
}

int mpf_integrate_dev_static(mpf_ctx* c, const double* dev_in9, uint64_t n, const mpf_stop* stop, double* dev_out,
                             uint64_t* total_steps) {
    if (!c || !stop || (n && (!dev_in9 || !dev_out))) return ERR_INVAL;
    if (n > MPF_STATIC_MAX_RAYS)
        return fail(c, ERR_INVAL, "mpf_integrate_dev_static takes at most " + std::to_string(MPF_STATIC_MAX_RAYS) +
                                      " rays per call (its workspace is a fixed device array)");
    const int dev = c->cfg.device;
    void* base = (dev >= 0 && dev < MAX_DEVICES) ? g_scratch_ptr[dev] : nullptr;
    if (!base)
        return fail(c, ERR_HIP, std::string("static workspace address of this device unavailable: ") +
                                    hipGetErrorString(dev >= 0 && dev < MAX_DEVICES ? g_scratch_err[dev]
                                                                                    : hipErrorInvalidDevice));
    Workspace w;
    if (!static_layout(base, n, c->cfg.precision == MPF_PREC_SPLIT, w))
        return fail(c, ERR_NOSPC, "static workspace too small for this batch");
    const int rc = integrate_impl(c, dev_in9, n, stop, dev_out, nullptr, nullptr, 1, nullptr, 0, &w);
    if (!rc && total_steps) *total_steps = c->last_steps;
    return rc;
}

int mpf_integrate_surfaces_born_dev(mpf_ctx* c, const double* dev_in9, uint64_t n, const mpf_surfaces* surf,
                                    double* dev_out, double* dev_out_surf, int sync) {
    if (!c) return ERR_INVAL;
    // the path of Integrator::integrate<-1> (homogeneous data) in an exact tier; the Jacobian's samples
    // still come from the octree's data (lensing.h:421-448)
    const int ord = c->cfg.order, prec = c->cfg.precision;
    c->cfg.order = MPF_HOMOGENEOUS;
    if (prec == MPF_PREC_SPLIT) c->cfg.precision = MPF_PREC_FP64;
    const int rc = mpf_integrate_surfaces_dev(c, dev_in9, n, surf, dev_out, dev_out_surf, sync);
    c->cfg.order = ord;
    c->cfg.precision = prec;
    return rc;
}

int mpf_integrate_surfaces_born(mpf_ctx* c, const double* in9, uint64_t n, const mpf_surfaces* surf, double* out,
                                double* out_surf) {
    if (!c) return ERR_INVAL;
    const int ord = c->cfg.order, prec = c->cfg.precision;
    c->cfg.order = MPF_HOMOGENEOUS;
    if (prec == MPF_PREC_SPLIT) c->cfg.precision = MPF_PREC_FP64;
    const int rc = mpf_integrate_surfaces(c, in9, n, surf, out, out_surf);
    c->cfg.order = ord;
    c->cfg.precision = prec;
    return rc;
}

int mpf_propagate(mpf_ctx* c, const double* in9, uint64_t n, uint32_t count, double angle, const double* rotation,
                  int interpolation, double amin, uint64_t capacity, double* traj, double* dist, uint64_t* nelem) {
    if (!c || (n && (!in9 || !traj || !dist || !nelem)) || capacity == 0) return ERR_INVAL;
    if (c->cfg.precision == MPF_PREC_SPLIT)
        return fail(c, ERR_INVAL, "propagate needs whole trajectories: exact tiers (REF, FP64)");
    if (interpolation != MPF_STOP_REDSHIFT && interpolation != MPF_STOP_A && interpolation != MPF_STOP_T &&
        interpolation != MPF_STOP_LAMBDA && interpolation != MPF_STOP_RADIUS)
        return fail(c, ERR_INVAL, "interpolation must be redshift, a, t, lambda or radius");
    if (count > 4096) return fail(c, ERR_INVAL, "at most 4096 bundle rays");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (n == 0) return 0;
    const unsigned nb = count + 1;
    const uint64_t per_ray = capacity * MPF_NOUT * 8;
    if (per_ray / MPF_NOUT / 8 != capacity) return fail(c, ERR_INVAL, "capacity too large");
    // bundles per pass: trajectories of at most 16 GiB on the device
    const uint64_t budget = 16ull << 30;
    const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(n, budget / (per_ray * nb)));
    DevBuf<double> din, dout, dtraj, ddist;
    DevBuf<unsigned> dacc;
    if (din.ensure(9 * nb * chunk) || dout.ensure((uint64_t)MPF_NOUT * nb * chunk) ||
        dtraj.ensure(capacity * MPF_NOUT * nb * chunk) || ddist.ensure(capacity * chunk) || dacc.ensure(chunk)) {
        din.release(); dout.release(); dtraj.release(); ddist.release(); dacc.release();
        return fail(c, ERR_NOMEM, "propagate buffers");
    }
    std::vector<double> bundle(9 * (uint64_t)nb * chunk), last((uint64_t)MPF_NOUT * chunk);
    std::vector<unsigned> acc(chunk);
    const mpf_stop none{MPF_STOP_NONE, 0.0, {0.0, 0.0, 0.0}};
    int rc = 0;
    for (uint64_t s0 = 0; s0 < n && !rc; s0 += chunk) {
        const uint64_t m = std::min(chunk, n - s0);
        for (uint64_t i = 0; i < m; ++i)
            mpf::cone::launch_bundle(in9 + 9 * (s0 + i), count, angle, rotation ? rotation[s0 + i] : 0.0, true,
                                     bundle.data() + 9 * (uint64_t)nb * i);
        if (hipMemcpy(din.p, bundle.data(), 9 * 8 * (uint64_t)nb * m, hipMemcpyHostToDevice) != hipSuccess) {
            rc = fail(c, ERR_HIP, "upload");
            break;
        }
        rc = integrate_impl(c, din.p, m * nb, &none, dout.p, nullptr, nullptr, 1, dtraj.p, capacity);
        if (rc) break;
        unsigned long long over = 0;
        if (hipMemcpy(&over, c->counters.p + 2, 8, hipMemcpyDeviceToHost) != hipSuccess) {
            rc = fail(c, ERR_HIP, "download");
            break;
        }
        if (over) {
            rc = fail(c, -28, std::to_string(over) + " rays have more than capacity elements");
            break;
        }
        mpf::PropagateArgs P;
        P.traj = dtraj.p;
        P.out = dout.p;
        P.n = m;
        P.cap = capacity;
        P.nb = nb;
        P.kind = interpolation;
        P.amin = amin;
        P.length = c->cfg.length;
        P.angle = angle;
        P.nacc = dacc.p;
        P.dist = ddist.p;
        hipError_t e = mpf::launch_bundle_accept(P, c->stream);
        if (e == hipSuccess) e = mpf::launch_bundle_distance(P, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        // central rays: trajectory, last element and distances
        const size_t w = capacity * MPF_NOUT * 8;
        if (e == hipSuccess)
            e = hipMemcpy2D(traj + s0 * capacity * MPF_NOUT, w, dtraj.p, w * nb, w, m, hipMemcpyDeviceToHost);
        if (e == hipSuccess)
            e = hipMemcpy2D(last.data(), MPF_NOUT * 8, dout.p, (size_t)MPF_NOUT * 8 * nb, MPF_NOUT * 8, m,
                            hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(dist + s0 * capacity, ddist.p, capacity * 8 * m, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(acc.data(), dacc.p, 4 * m, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            rc = hipfail(c, e, "propagate");
            break;
        }
        for (uint64_t i = 0; i < m; ++i) {
            const double s = last[(uint64_t)MPF_NOUT * i + MPF_OUT_NSTEPS];
            nelem[s0 + i] = (acc[i] && s >= 0.0) ? (uint64_t)s + 1 : 0;
        }
    }
    din.release(); dout.release(); dtraj.release(); ddist.release(); dacc.release();
    return rc;
}

int mpf_integrate_trajectory(mpf_ctx* c, const double* in9, uint64_t n, const mpf_stop* stop, uint64_t capacity,
                             double* out, double* traj) {
    if (!c || !stop || (n && (!in9 || !out || !traj)) || capacity == 0) return ERR_INVAL;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (n == 0) return 0;
    const uint64_t tn = n * capacity * MPF_NOUT;
    if (tn / capacity / MPF_NOUT != n) return fail(c, ERR_INVAL, "trajectory buffer size overflows");
    DevBuf<double> din, dout, dtraj;
    if (din.ensure(9 * n) || dout.ensure((uint64_t)MPF_NOUT * n) || dtraj.ensure(tn)) {
        din.release(); dout.release(); dtraj.release();
        return fail(c, ERR_NOMEM, "trajectory buffers (" + std::to_string(tn * 8) + " bytes)");
    }
    int rc = 0;
    if (hipMemcpy(din.p, in9, 9 * n * 8, hipMemcpyHostToDevice) != hipSuccess) rc = fail(c, ERR_HIP, "upload");
    if (!rc) rc = integrate_impl(c, din.p, n, stop, dout.p, nullptr, nullptr, 1, dtraj.p, capacity);
    unsigned long long over = 0;
    if (!rc && (hipMemcpy(out, dout.p, (uint64_t)MPF_NOUT * n * 8, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(traj, dtraj.p, tn * 8, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(&over, c->counters.p + 2, 8, hipMemcpyDeviceToHost) != hipSuccess))
        rc = fail(c, ERR_HIP, "download");
    din.release(); dout.release(); dtraj.release();
    if (!rc && over) rc = fail(c, -28, std::to_string(over) + " rays have more than capacity elements");
    return rc;
}

static int copy_table(const mpf::cosmo::Table& tab, double* t, double* a, double* dadt, double* d2adt2,
                      uint64_t capacity, uint64_t* n) {
    const uint64_t m = tab.t.size();
    if (n) *n = m;
    if (capacity >= m && m) {
        if (!t || !a || !dadt || !d2adt2) return ERR_INVAL;
        std::copy(tab.t.begin(), tab.t.end(), t);
        std::copy(tab.a.begin(), tab.a.end(), a);
        std::copy(tab.dadt.begin(), tab.dadt.end(), dadt);
        std::copy(tab.d2adt2.begin(), tab.d2adt2.end(), d2adt2);
    }
    return 0;
}

static thread_local std::string g_cosmo_err;

int mpf_cosmology_acquire(const char* evolution_file, const char* parameter_file, const char* cell_dir, double mpc,
                          double* h, double* omega_m, double* lbox_mpch, double* t, double* a, double* dadt,
                          double* d2adt2, uint64_t capacity, uint64_t* n) {
    if (!evolution_file || !parameter_file || !cell_dir || !h || !omega_m || !lbox_mpch || !n || !(mpc > 0))
        return ERR_INVAL;
    mpf::cosmo::Table tab;
    if (!mpf::cosmo::acquire(evolution_file, parameter_file, cell_dir, mpc, *h, *omega_m, *lbox_mpch, tab,
                             g_cosmo_err)) {
        *n = 0;
        return ERR_INVAL;
    }
    return copy_table(tab, t, a, dadt, d2adt2, capacity, n);
}

int mpf_cosmology_correct(const double* t, double* a, uint64_t n, const double* traj_t, const double* traj_a,
                          uint64_t m) {
    if ((n && (!t || !a)) || (m && (!traj_t || !traj_a))) return ERR_INVAL;
    mpf::cosmo::Table tab;
    tab.t.assign(t, t + n);
    tab.a.assign(a, a + n);
    mpf::cosmo::correct(tab, std::vector<double>(traj_t, traj_t + m), std::vector<double>(traj_a, traj_a + m));
    std::copy(tab.a.begin(), tab.a.end(), a);
    return 0;
}

int mpf_cosmology_prepare(int device, int precision, const char* evolution_file, const char* parameter_file,
                          const char* cell_dir, double mpc, double* h, double* omega_m, double* lbox_mpch, double* t,
                          double* a, double* dadt, double* d2adt2, uint64_t capacity, uint64_t* n) {
    if (!evolution_file || !parameter_file || !cell_dir || !h || !omega_m || !lbox_mpch || !n || !(mpc > 0) ||
        (precision != MPF_PREC_REF && precision != MPF_PREC_FP64))
        return ERR_INVAL;
    mpf::cosmo::Table tab;
    if (!mpf::cosmo::acquire(evolution_file, parameter_file, cell_dir, mpc, *h, *omega_m, *lbox_mpch, tab,
                             g_cosmo_err)) {
        *n = 0;
        return ERR_INVAL;
    }
    // the homogeneous level-5 octree of rays.cpp:211 (assign(5) + homogenize)
    const unsigned nref = 5, side = 1u << nref;
    std::vector<std::pair<unsigned __int128, uint32_t>> keys;
    keys.reserve((size_t)side * side * side);
    for (uint32_t x = 0; x < side; ++x)
        for (uint32_t y = 0; y < side; ++y)
            for (uint32_t z = 0; z < side; ++z) {
                unsigned __int128 k = 0;
                for (unsigned l = 0; l < nref; ++l) {
                    const unsigned b = nref - 1 - l;
                    const unsigned nib = 8u | (((z >> b) & 1u) << 2) | (((y >> b) & 1u) << 1) | ((x >> b) & 1u);
                    k |= (unsigned __int128)nib << (124 - 4 * l);
                }
                keys.push_back({k, 0u});
            }
    std::sort(keys.begin(), keys.end());
    std::vector<unsigned char> rec(keys.size() * 48, 0);
    const float one = 1.0f;
    for (size_t i = 0; i < keys.size(); ++i) {
        const uint64_t lo = (uint64_t)keys[i].first, hi = (uint64_t)(keys[i].first >> 64);
        std::memcpy(&rec[48 * i], &lo, 8);
        std::memcpy(&rec[48 * i + 8], &hi, 8);
        std::memcpy(&rec[48 * i + 20], &one, 4);   // a (homogenize sets a = 1)
    }
    const double length = *lbox_mpch * mpc / *h;   // rays.cpp:208
    mpf_config cfg{device, MPF_HOMOGENEOUS, precision, mpf::cosmo::reference_nsteps(tab.t.size(), nref), length};
    mpf_ctx* c = nullptr;
    int rc = mpf_create(&c, &cfg);
    if (rc) {
        g_cosmo_err = "reference photon: no device context";
        return rc;
    }
    rc = mpf_load_octree_records(c, rec.data(), keys.size());
    if (!rc) rc = mpf_set_cosmology(c, tab.t.data(), tab.a.data(), tab.dadt.data(), tab.t.size());
    // steps to the edge of the box: EXTENT/2 over (cell size / nsteps), plus margin
    const uint64_t cap = 2ull * side * cfg.nsteps + 64;
    std::vector<double> out(MPF_NOUT), traj(cap * MPF_NOUT);
    const double in9[9] = {1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0};   // launch (0,0,0) -> (2,0,0)
    const mpf_stop none{MPF_STOP_NONE, 0.0, {0.0, 0.0, 0.0}};
    if (!rc) rc = mpf_integrate_trajectory(c, in9, 1, &none, cap, out.data(), traj.data());
    if (rc) g_cosmo_err = std::string("reference photon: ") + mpf_last_error(c);
    mpf_destroy(c);
    if (rc) return rc;
    const uint64_t m = out[MPF_OUT_NSTEPS] < 0 ? 0 : (uint64_t)out[MPF_OUT_NSTEPS] + 1;
    std::vector<double> tt(m), aa(m);
    for (uint64_t k = 0; k < m; ++k) {
        tt[k] = traj[k * MPF_NOUT + MPF_OUT_T];
        aa[k] = traj[k * MPF_NOUT + MPF_OUT_A];
    }
    mpf::cosmo::correct(tab, tt, aa);
    return copy_table(tab, t, a, dadt, d2adt2, capacity, n);
}

const char* mpf_cosmology_last_error(void) { return g_cosmo_err.c_str(); }

int mpf_device_memory(int device, uint64_t* free_bytes, uint64_t* total_bytes) {
    if (!free_bytes || !total_bytes) return ERR_INVAL;
    if (hipSetDevice(device) != hipSuccess) return ERR_HIP;
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return ERR_HIP;
    *free_bytes = f;
    *total_bytes = t;
    return 0;
}

int mpf_dev_alloc(mpf_ctx* c, uint64_t bytes, void** dev_ptr) {
    if (!c || !dev_ptr) return ERR_INVAL;
    *dev_ptr = nullptr;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (hipMalloc(dev_ptr, bytes ? bytes : 1) != hipSuccess) {
        *dev_ptr = nullptr;
        return fail(c, ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    return 0;
}

int mpf_dev_free(mpf_ctx* c, void* dev_ptr) {
    if (!c) return ERR_INVAL;
    if (!dev_ptr) return 0;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipFree(dev_ptr));
    return 0;
}

int mpf_dev_upload(mpf_ctx* c, void* dev_dst, const void* host_src, uint64_t bytes) {
    if (!c || (bytes && (!dev_dst || !host_src))) return ERR_INVAL;
    if (!bytes) return 0;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    HIPCHK(c, hipMemcpyAsync(dev_dst, host_src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int mpf_dev_download(mpf_ctx* c, void* host_dst, const void* dev_src, uint64_t bytes) {
    if (!c || (bytes && (!dev_src || !host_dst))) return ERR_INVAL;
    if (!bytes) return 0;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    HIPCHK(c, hipMemcpyAsync(host_dst, dev_src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int mpf_dev_copy(mpf_ctx* c, void* dev_dst, const void* dev_src, uint64_t bytes) {
    if (!c || (bytes && (!dev_dst || !dev_src))) return ERR_INVAL;
    if (bytes % 16) return fail(c, ERR_INVAL, "copy size must be a multiple of 16 bytes");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const uint64_t n = bytes / 16;
    const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, (uint64_t)c->cus * 16);
    HIPCHK(c, mpf::launch_fill_u64(c->counters.p, 2, 0ull, c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    if (n) {
        hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, c->stream, static_cast<const uint4*>(dev_src),
                           static_cast<uint4*>(dev_dst), n);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->pending = true;
    return mpf_synchronize(c);
}

int mpf_catalogue_solve_dev(mpf_ctx* c, const double* dev_src6, const double* dev_init2, uint64_t n,
                            const mpf_newton_cfg* cfg, double* dev_out, uint64_t* total_rays) {
    if (!c || !cfg || (n && (!dev_src6 || !dev_out))) return ERR_INVAL;
    if (!c->have_tree) return fail(c, ERR_STATE, "no octree loaded");
    if (c->ncos <= 0) return fail(c, ERR_STATE, "no cosmology table set");
    if (!(cfg->accuracy > 0) || cfg->max_iterations < 1 || !(cfg->h > 0) || !(cfg->mpc > 0))
        return fail(c, ERR_INVAL, "newton configuration");
    const bool bundle = cfg->jacobian_bundle != 0;
    if (bundle) {
        const int k = cfg->bundle_kind;
        if (k != MPF_STOP_RADIUS && k != MPF_STOP_LAMBDA && k != MPF_STOP_REDSHIFT && k != MPF_STOP_T &&
            k != MPF_STOP_A && k != MPF_STOP_PLANE)
            return fail(c, ERR_INVAL, "stop_bundle must be radius, lambda, redshift, t, a or plane");
        if (cfg->bundle_plane != 0 && cfg->bundle_plane != 1) return fail(c, ERR_INVAL, "plane must be sachs or normal");
        if (!(cfg->bundle_opening > 0) || !(cfg->bundle_opening < 1.5))
            return fail(c, ERR_INVAL, "bundle opening out of range");
        if (cfg->jacobian_born) return fail(c, ERR_INVAL, "jacobiantype is bundle or infinitesimal_born, not both");
    }
    if (cfg->use_previous != 0 && cfg->use_previous != 2)
        return fail(c, ERR_INVAL, "use_previous must be 0 or 2 (mode 1 is mpf_catalogue_previous)");
    if (cfg->jacobian_all_rounds > 2) return fail(c, ERR_INVAL, "jacobian_all_rounds must be 0, 1 or 2");
    const bool rerun = cfg->use_previous == 2;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (total_rays) *total_rays = 0;
    if (n == 0) return 0;
    if (n >= 0xFFFFFFFFull) return fail(c, ERR_INVAL, "too many sources");
    // potential and scale factor at the observer (the launch element, integrator.h:669-683)
    double phi0 = 0.0, a0 = 1.0;
    {
        const double origin[3] = {0.0, 0.0, 0.0};
        double d7[7];
        const int prec = c->cfg.precision;
        if (prec == MPF_PREC_SPLIT) c->cfg.precision = MPF_PREC_FP64;
        const int ord = c->cfg.order == MPF_TSC ? 1 : c->cfg.order;  // tsc at the box centre is cic
        int rc = mpf_interpolate(c, ord, origin, 1, d7);
        c->cfg.precision = prec;
        if (rc) return rc;
        phi0 = d7[1];
        std::vector<double> ct(c->ncos), ca(c->ncos);
        HIPCHK(c, hipMemcpy(ct.data(), c->cosmo.p, c->ncos * 8, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(ca.data(), c->cosmo.p + c->ncos, c->ncos * 8, hipMemcpyDeviceToHost));
        // Utility::interpolate at t = 0
        const long long nn = c->ncos;
        long long i = std::upper_bound(ct.begin(), ct.end(), 0.0) - ct.begin();
        i += (nn > 1) * ((i <= 0) - (i >= nn));
        a0 = (nn > 1) ? (ca[i - 1] + ((ca[i] - ca[i - 1]) * (0.0 - ct[i - 1]) / (ct[i] - ct[i - 1]))) : ca[0];
    }
    DevBuf<double> dist, ang, in9, dact, surf, outb, bctx, jb, gerr, gas, gin9, gdist, gsurf, gout;
    DevBuf<uint32_t> act, next, need, gsrc, gcnt, goff;
    DevBuf<unsigned int> cnt;
    auto release = [&]() {
        dist.release(); ang.release(); in9.release(); dact.release(); surf.release(); act.release(); next.release();
        need.release(); cnt.release(); outb.release(); bctx.release(); jb.release();
        gerr.release(); gas.release(); gin9.release(); gdist.release(); gsurf.release(); gout.release();
        gsrc.release(); gcnt.release(); goff.release();
    };
    if (dist.ensure(n) || ang.ensure(2 * n) || in9.ensure(9 * n) || dact.ensure(n) ||
        surf.ensure((uint64_t)MPF_NSURF * n) || act.ensure(n) || next.ensure(n) || need.ensure(n) || cnt.ensure(3) ||
        outb.ensure((uint64_t)MPF_NOUT * n) ||
        (bundle && (bctx.ensure((uint64_t)mpf::CAT_BCTX * n) || jb.ensure(4 * n))) ||
        (rerun && (gerr.ensure(n) || gas.ensure(n) || gsrc.ensure(n) || gcnt.ensure(n) || goff.ensure(n)))) {
        release();
        return fail(c, ERR_NOMEM, "catalogue buffers");
    }
#define CATCHK(expr)                                  \
    do {                                              \
        hipError_t e_ = (expr);                       \
        if (e_ != hipSuccess) { release(); return hipfail(c, e_, #expr); } \
    } while (0)
    CATCHK(mpf::launch_cat_init(dev_src6, dev_init2, n, dist.p, ang.p, dev_out, c->stream));
    std::vector<uint32_t> ids(n);
    for (uint64_t i = 0; i < n; ++i) ids[i] = (uint32_t)i;
    CATCHK(hipMemcpyAsync(act.p, ids.data(), n * 4, hipMemcpyHostToDevice, c->stream));
    uint64_t nact = n, rays = 0;
    mpf::CatRound R;
    R.src = dev_src6; R.dist = dist.p; R.ang = ang.p; R.out = dev_out; R.surf = surf.p;
    R.next = next.p; R.nnext = cnt.p; R.need_jac = need.p; R.nneed = cnt.p + 1;
    R.max_iter = (int)cfg->max_iterations; R.acc = cfg->accuracy; R.length = c->cfg.length; R.h = cfg->h;
    R.mpc = cfg->mpc; R.a0 = a0; R.phi0 = phi0;
    for (int k = 0; k < 3; ++k) R.vobs[k] = cfg->vobs[k];
    R.bundle = bundle ? 1 : 0; R.bkind = cfg->bundle_kind; R.bplane = cfg->bundle_plane; R.bctx = bctx.p;
    R.born = cfg->jacobian_born ? 1 : 0;
    R.rerun = rerun ? 1 : 0; R.gsrc = gsrc.p; R.ngrid = cnt.p + 2; R.gerr = gerr.p; R.gas = gas.p;
    // the bundle Jacobians (jacobiantype = bundle) of the sources in ids[0, m): their current
    // angles and crossing context read back, mpf_jacobian_bundle on the host, the result in jb
    std::vector<double> h_dist;
    auto bundle_jacobians = [&](const uint32_t* d_ids, uint64_t m) -> int {
        std::vector<uint32_t> hid(m);
        std::vector<double> hctx((uint64_t)mpf::CAT_BCTX * n);
        if (hipMemcpyAsync(hid.data(), d_ids, m * 4, hipMemcpyDeviceToHost, c->stream) ||
            hipMemcpyAsync(hctx.data(), bctx.p, (uint64_t)mpf::CAT_BCTX * n * 8, hipMemcpyDeviceToHost, c->stream) ||
            hipStreamSynchronize(c->stream))
            return fail(c, ERR_HIP, "bundle context download");
        if (h_dist.empty()) {
            h_dist.resize(n);
            if (hipMemcpy(h_dist.data(), dist.p, n * 8, hipMemcpyDeviceToHost)) return fail(c, ERR_HIP, "distances");
        }
        std::vector<double> pt(2 * m), ki(3 * m), ref(m), dd(m), J(4 * m);
        for (uint64_t q = 0; q < m; ++q) {
            const uint64_t i = hid[q];
            const double* b = hctx.data() + (uint64_t)mpf::CAT_BCTX * i;
            pt[2 * q] = b[7]; pt[2 * q + 1] = b[8];   // the direction the bundle is launched around
            for (int k = 0; k < 3; ++k) ki[3 * q + k] = b[3 + k];
            ref[q] = b[6];
            dd[q] = h_dist[i];
        }
        const double origin[3] = {0.0, 0.0, 0.0};
        if (int rc = mpf_jacobian_bundle(c, pt.data(), m, origin, cfg->bundle_kind, ref.data(), 1, ki.data(),
                                         dd.data(), cfg->bundle_opening, J.data()))
            return rc;
        if (hipMemcpyAsync(jb.p, J.data(), 4 * m * 8, hipMemcpyHostToDevice, c->stream))
            return fail(c, ERR_HIP, "bundle Jacobian upload");
        rays += 4 * m;
        return 0;
    };
    // the first round whose rays carry the Jacobian
    const int jac_from = cfg->jacobian_all_rounds == 1 ? 0 : cfg->jacobian_all_rounds == 2 ? 1 : 3;
    unsigned int zero2[2] = {0u, 0u};
    CATCHK(hipMemcpyAsync(cnt.p + 1, zero2, 4, hipMemcpyHostToDevice, c->stream));
    for (int round = 0; nact > 0 && round <= (int)cfg->max_iterations; ++round) {
        // the bundle replaces the ray's own Jacobian from iteration 5 (catalogues.h:509-523); the
        // re-run mode updates without one
        const int jac = (round >= jac_from && !(bundle && round >= 5) && !rerun) ? 1 : 0;
        CATCHK(mpf::launch_cat_launch(ang.p, act.p, nact, in9.p, dact.p, dist.p, c->stream));
        mpf_surfaces sf{};
        sf.kind = MPF_STOP_RADIUS;
        sf.n = 1;
        sf.jacobian = jac;
        sf.ray_value = dact.p;
        int rc = mpf_integrate_surfaces_dev(c, in9.p, nact, &sf, outb.p, surf.p, 1);
        if (rc) { release(); return rc; }
        rays += nact;
        CATCHK(hipMemcpyAsync(cnt.p, zero2, 4, hipMemcpyHostToDevice, c->stream));
        if (rerun) CATCHK(hipMemcpyAsync(cnt.p + 2, zero2, 4, hipMemcpyHostToDevice, c->stream));
        R.act = act.p; R.next = next.p; R.nact = nact; R.round = round; R.jac = jac;
        CATCHK(mpf::launch_newton(R, c->stream));
        if (rerun) {  // the grids of the sources not within the accuracy, in one batch
            unsigned int ng = 0;
            CATCHK(hipMemcpyAsync(&ng, cnt.p + 2, 4, hipMemcpyDeviceToHost, c->stream));
            CATCHK(hipStreamSynchronize(c->stream));
            if (ng) {
                CATCHK(mpf::launch_cat_grid_count(R, ng, gcnt.p, c->stream));
                std::vector<uint32_t> hc(ng), ho(ng);
                CATCHK(hipMemcpyAsync(hc.data(), gcnt.p, 4ull * ng, hipMemcpyDeviceToHost, c->stream));
                CATCHK(hipStreamSynchronize(c->stream));
                uint64_t tot = 0;
                for (unsigned g = 0; g < ng; ++g) {
                    ho[g] = (uint32_t)tot;
                    tot += hc[g];
                }
                if (gin9.ensure(9 * tot) || gdist.ensure(tot) || gsurf.ensure((uint64_t)MPF_NSURF * tot) ||
                    gout.ensure((uint64_t)MPF_NOUT * tot)) {
                    release();
                    return fail(c, ERR_NOMEM, "re-run grid buffers");
                }
                CATCHK(hipMemcpyAsync(goff.p, ho.data(), 4ull * ng, hipMemcpyHostToDevice, c->stream));
                CATCHK(mpf::launch_cat_grid_launch(R, ng, goff.p, gcnt.p, gin9.p, gdist.p, c->stream));
                mpf_surfaces gs{};
                gs.kind = MPF_STOP_RADIUS;
                gs.n = 1;
                gs.jacobian = 0;
                gs.ray_value = gdist.p;
                int rc = mpf_integrate_surfaces_dev(c, gin9.p, tot, &gs, gout.p, gsurf.p, 1);
                if (rc) { release(); return rc; }
                rays += tot;
                CATCHK(mpf::launch_cat_grid_select(R, ng, goff.p, gcnt.p, gsurf.p, c->stream));
            }
        }
        unsigned int nn = 0;
        CATCHK(hipMemcpyAsync(&nn, cnt.p, 4, hipMemcpyDeviceToHost, c->stream));
        CATCHK(hipStreamSynchronize(c->stream));
        std::swap(act.p, next.p);
        nact = nn;
        if (bundle && !rerun && round >= 5 && nact) {  // the pending updates with the bundle Jacobian
            if (int rc = bundle_jacobians(act.p, nact)) { release(); return rc; }
            CATCHK(hipMemcpyAsync(cnt.p + 2, zero2, 4, hipMemcpyHostToDevice, c->stream));
            CATCHK(mpf::launch_newton_bundle(R, act.p, nact, jb.p, next.p, cnt.p + 2, c->stream));
            CATCHK(hipMemcpyAsync(&nn, cnt.p + 2, 4, hipMemcpyDeviceToHost, c->stream));
            CATCHK(hipStreamSynchronize(c->stream));
            std::swap(act.p, next.p);
            nact = nn;
        }
    }
    // Jacobian of the sources that converged before the Jacobian rounds
    unsigned int nneed = 0;
    CATCHK(hipMemcpyAsync(&nneed, cnt.p + 1, 4, hipMemcpyDeviceToHost, c->stream));
    CATCHK(hipStreamSynchronize(c->stream));
    if (bundle) {  // every converged source: the bundle at its converged direction (catalogues.h:666-681)
        if (nneed) {
            if (int rc = bundle_jacobians(need.p, nneed)) { release(); return rc; }
            CATCHK(mpf::launch_cat_jac_bundle(need.p, nneed, jb.p, dev_out, c->stream));
            CATCHK(hipStreamSynchronize(c->stream));
        }
        nneed = 0;
    }
    if (nneed) {
        CATCHK(mpf::launch_cat_launch(ang.p, need.p, nneed, in9.p, dact.p, dist.p, c->stream));
        mpf_surfaces sf{};
        sf.kind = MPF_STOP_RADIUS;
        sf.n = 1;
        sf.jacobian = 1;
        sf.ray_value = dact.p;
        int rc = mpf_integrate_surfaces_dev(c, in9.p, nneed, &sf, outb.p, surf.p, 1);
        if (rc) { release(); return rc; }
        rays += nneed;
        CATCHK(mpf::launch_cat_jac(need.p, nneed, surf.p, dev_out, c->stream));
        CATCHK(hipStreamSynchronize(c->stream));
    }
    // jacobiantype = infinitesimal_born: the final Jacobian from the homogeneous path in the converged
    // direction to the source's distance (catalogues.h:670-674)
    if (cfg->jacobian_born) {
        CATCHK(hipMemcpyAsync(need.p, ids.data(), n * 4, hipMemcpyHostToDevice, c->stream));
        CATCHK(mpf::launch_cat_launch(ang.p, need.p, n, in9.p, dact.p, dist.p, c->stream));
        mpf_surfaces sf{};
        sf.kind = MPF_STOP_RADIUS;
        sf.n = 1;
        sf.jacobian = 1;
        sf.ray_value = dact.p;
        int rc = mpf_integrate_surfaces_born_dev(c, in9.p, n, &sf, outb.p, surf.p, 1);
        if (rc) { release(); return rc; }
        rays += n;
        CATCHK(mpf::launch_cat_jac_born(n, surf.p, dev_out, c->stream));
        CATCHK(hipStreamSynchronize(c->stream));
    }
#undef CATCHK
    release();
    if (total_rays) *total_rays = rays;
    return 0;
}

int mpf_catalogue_solve(mpf_ctx* c, const double* src6, const double* init2, uint64_t n, const mpf_newton_cfg* cfg,
                        double* out, uint64_t* total_rays) {
    if (!c || !cfg || (n && (!src6 || !out))) return ERR_INVAL;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    DevBuf<double> s, i2, o;
    if (s.ensure(6 * std::max<uint64_t>(n, 1)) || o.ensure((uint64_t)MPF_NCAT * std::max<uint64_t>(n, 1)) ||
        (init2 && i2.ensure(2 * std::max<uint64_t>(n, 1)))) {
        s.release(); i2.release(); o.release();
        return fail(c, ERR_NOMEM, "catalogue buffers");
    }
    int rc = 0;
    if (n) {
        if (hipMemcpy(s.p, src6, 6 * n * 8, hipMemcpyHostToDevice) != hipSuccess ||
            (init2 && hipMemcpy(i2.p, init2, 2 * n * 8, hipMemcpyHostToDevice) != hipSuccess))
            rc = fail(c, ERR_HIP, "catalogue upload");
    }
    if (!rc) rc = mpf_catalogue_solve_dev(c, s.p, init2 ? i2.p : nullptr, n, cfg, o.p, total_rays);
    if (!rc && n && hipMemcpy(out, o.p, (uint64_t)MPF_NCAT * n * 8, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(c, ERR_HIP, "catalogue download");
    s.release(); i2.release(); o.release();
    return rc;
}

int mpf_catalogue_previous(mpf_ctx* c, double* rows18, uint64_t n, const mpf_newton_cfg* cfg) {
    if (!c || !cfg || (n && !rows18)) return ERR_INVAL;
    if (!c->have_tree) return fail(c, ERR_STATE, "no octree loaded");
    if (c->ncos <= 0) return fail(c, ERR_STATE, "no cosmology table set");
    const bool bundle = cfg->jacobian_bundle != 0, born = cfg->jacobian_born != 0;
    if (bundle && born) return fail(c, ERR_INVAL, "jacobiantype is bundle or infinitesimal_born, not both");
    if (bundle) {
        const int k = cfg->bundle_kind;
        if (k != MPF_STOP_RADIUS && k != MPF_STOP_LAMBDA && k != MPF_STOP_REDSHIFT && k != MPF_STOP_T &&
            k != MPF_STOP_A && k != MPF_STOP_PLANE)
            return fail(c, ERR_INVAL, "stop_bundle must be radius, lambda, redshift, t, a or plane");
        if (cfg->bundle_plane != 0 && cfg->bundle_plane != 1) return fail(c, ERR_INVAL, "plane must be sachs or normal");
        if (!(cfg->bundle_opening > 0) || !(cfg->bundle_opening < 1.5))
            return fail(c, ERR_INVAL, "bundle opening out of range");
    }
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (n == 0) return 0;
    // the launch from the observed angles (the comoving ones for the Born
    // approximation) to the scale factor of the FLRW redshift (catalogues.h:936-946),
    // launched here on the host as Integrator::launch does
    std::vector<double> h9(9 * n), sf(n), ang(2 * n);
    for (uint64_t i = 0; i < n; ++i) {
        const double* r = rows18 + 18 * i;
        sf[i] = 1. / (1. + r[7]);
        const double phi = born ? r[1] : r[3], theta = born ? r[2] : r[4];
        ang[2 * i] = phi;
        ang[2 * i + 1] = theta;
        double* p = h9.data() + 9 * i;
        p[0] = 1.0; p[1] = 0.0; p[2] = 0.0; p[3] = 0.0; p[4] = 0.0; p[5] = 1.0;
        p[6] = std::cos(phi) * std::sin(theta);
        p[7] = std::sin(phi) * std::sin(theta);
        p[8] = std::cos(theta);
    }
    DevBuf<double> din, dsf, dout, dsurf;
    auto release = [&]() { din.release(); dsf.release(); dout.release(); dsurf.release(); };
    if (din.ensure(9 * n) || dsf.ensure(n) || dout.ensure((uint64_t)MPF_NOUT * n) || dsurf.ensure((uint64_t)MPF_NSURF * n)) {
        release();
        return fail(c, ERR_NOMEM, "previous-catalogue buffers");
    }
    std::vector<double> hs((uint64_t)MPF_NSURF * n);
    int rc = 0;
    if (hipMemcpy(din.p, h9.data(), 9 * n * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dsf.p, sf.data(), n * 8, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(c, ERR_HIP, "previous-catalogue upload");
    if (!rc) {
        mpf_surfaces s{};
        s.kind = MPF_STOP_A;
        s.n = 1;
        s.jacobian = (!bundle && !born) ? 1 : 0;
        s.ray_value = dsf.p;
        c->surf_jac_chi = true;
        const int prec = c->cfg.precision;
        rc = mpf_integrate_surfaces_dev(c, din.p, n, &s, dout.p, dsurf.p, 1);
        c->surf_jac_chi = false;
        c->cfg.precision = prec;
    }
    if (!rc && hipMemcpy(hs.data(), dsurf.p, (uint64_t)MPF_NSURF * n * 8, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(c, ERR_HIP, "previous-catalogue download");
    // rows whose ray reached the scale factor; the others get 42 entries
    std::vector<uint64_t> ok;
    for (uint64_t i = 0; !rc && i < n; ++i) {
        const double* r = hs.data() + (uint64_t)MPF_NSURF * i;
        double* J = rows18 + 18 * i + 13;
        if (r[mpf::SF_CROSSED] == 0.0) {
            J[0] = J[1] = J[2] = J[3] = 42;
            continue;
        }
        ok.push_back(i);
        if (!bundle && !born) { J[0] = r[MPF_SURF_A11]; J[1] = r[MPF_SURF_A12]; J[2] = r[MPF_SURF_A12]; J[3] = r[MPF_SURF_A22]; }
    }
    const uint64_t m = ok.size();
    if (!rc && m && born) {
        // the homogeneous path to the crossing's distance, sampled on the data (catalogues.h:995-999)
        std::vector<double> b9(9 * m), bd(m);
        for (uint64_t q = 0; q < m; ++q) {
            std::copy(h9.begin() + 9 * ok[q], h9.begin() + 9 * ok[q] + 9, b9.begin() + 9 * q);
            bd[q] = hs[(uint64_t)MPF_NSURF * ok[q] + MPF_SURF_CHI];
        }
        if (hipMemcpy(din.p, b9.data(), 9 * m * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dsf.p, bd.data(), m * 8, hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(c, ERR_HIP, "Born upload");
        if (!rc) {
            mpf_surfaces s{};
            s.kind = MPF_STOP_RADIUS;
            s.n = 1;
            s.jacobian = 1;
            s.ray_value = dsf.p;
            rc = mpf_integrate_surfaces_born_dev(c, din.p, m, &s, dout.p, dsurf.p, 1);
        }
        std::vector<double> bs((uint64_t)MPF_NSURF * m);
        if (!rc && hipMemcpy(bs.data(), dsurf.p, (uint64_t)MPF_NSURF * m * 8, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(c, ERR_HIP, "Born download");
        for (uint64_t q = 0; !rc && q < m; ++q) {
            const double* r = bs.data() + (uint64_t)MPF_NSURF * q;
            double* J = rows18 + 18 * ok[q] + 13;
            J[0] = r[MPF_SURF_A11]; J[1] = r[MPF_SURF_A12]; J[2] = r[MPF_SURF_A12]; J[3] = r[MPF_SURF_A22];
        }
    }
    if (!rc && m && bundle) {
        // Lensing::dbetadtheta with kiTarget and interpRef of the crossing (catalogues.h:950-992)
        std::vector<double> pt(2 * m), ki(3 * m), ref(m), dd(m), J(4 * m);
        for (uint64_t q = 0; q < m; ++q) {
            const uint64_t i = ok[q];
            const double* r = hs.data() + (uint64_t)MPF_NSURF * i;
            pt[2 * q] = ang[2 * i];
            pt[2 * q + 1] = ang[2 * i + 1];
            const double fp[3] = {r[MPF_SURF_X], r[MPF_SURF_Y], r[MPF_SURF_Z]};
            for (int k = 0; k < 3; ++k) ki[3 * q + k] = cfg->bundle_plane == 0 ? r[MPF_SURF_KX + k] : fp[k];
            dd[q] = r[MPF_SURF_CHI];
            switch (cfg->bundle_kind) {
                case MPF_STOP_REDSHIFT: ref[q] = r[MPF_SURF_REDSHIFT]; break;
                case MPF_STOP_A: ref[q] = sf[i]; break;
                case MPF_STOP_T: ref[q] = r[MPF_SURF_T]; break;
                case MPF_STOP_LAMBDA: ref[q] = r[MPF_SURF_LAMBDA]; break;
                case MPF_STOP_PLANE: ref[q] = ki[3 * q] * fp[0] + ki[3 * q + 1] * fp[1] + ki[3 * q + 2] * fp[2]; break;
                default: ref[q] = dd[q]; break;
            }
        }
        const double origin[3] = {0.0, 0.0, 0.0};
        rc = mpf_jacobian_bundle(c, pt.data(), m, origin, cfg->bundle_kind, ref.data(), 1, ki.data(), dd.data(),
                                 cfg->bundle_opening, J.data());
        for (uint64_t q = 0; !rc && q < m; ++q)
            for (int k = 0; k < 4; ++k) rows18[18 * ok[q] + 13 + k] = J[4 * q + k];
    }
    release();
    return rc;
}

int mpf_mlp_init(mpf_ctx* c, uint64_t seed) {
    if (!c) return ERR_INVAL;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    mpf::mlp_destroy(c->mlp);
    c->mlp = nullptr;
    std::string err;
    if (!mpf::mlp_create(&c->mlp, c->stream, seed, err)) return fail(c, ERR_HIP, err);
    return 0;
}

static int need_mlp(mpf_ctx* c) {
    if (!c) return ERR_INVAL;
    if (!c->mlp) return fail(c, ERR_STATE, "no network: call mpf_mlp_init first");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    return 0;
}

int mpf_mlp_state(mpf_ctx* c, float* get, const float* set) {
    if (int rc = need_mlp(c)) return rc;
    if (set) {
        for (int i = MPF_MLP_NPARAM + MPF_MLP_NIN; i < MPF_MLP_NPARAM + 2 * MPF_MLP_NIN; ++i)
            if (!(set[i] > 0)) return fail(c, ERR_INVAL, "feature standard deviations must be positive");
        for (int i = MPF_MLP_NPARAM + 2 * MPF_MLP_NIN + MPF_MLP_NOUT; i < MPF_MLP_NSTATE; ++i)
            if (!(set[i] > 0)) return fail(c, ERR_INVAL, "target standard deviations must be positive");
    }
    std::string err;
    if (!mpf::mlp_state(c->mlp, get, set, err)) return fail(c, ERR_HIP, err);
    return 0;
}

static bool finite_all(const float* p, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

int mpf_mlp_train(mpf_ctx* c, const float* feat, const float* target, uint64_t n, uint32_t epochs, uint32_t batch,
                  float lr, uint64_t seed, float* loss_per_epoch) {
    if (int rc = need_mlp(c)) return rc;
    if (n == 0 || !feat || !target || batch == 0 || !(lr > 0)) return fail(c, ERR_INVAL, "training arguments");
    if (!finite_all(feat, n * MPF_MLP_NIN) || !finite_all(target, n * MPF_MLP_NOUT))
        return fail(c, ERR_INVAL, "non-finite training data");
    std::string err;
    if (!mpf::mlp_train(c->mlp, feat, target, n, epochs, batch, lr, seed, loss_per_epoch, err))
        return fail(c, ERR_HIP, err);
    return 0;
}

int mpf_mlp_gradient(mpf_ctx* c, const float* feat, const float* target, uint64_t n, float* grad, float* loss) {
    if (int rc = need_mlp(c)) return rc;
    if (n == 0 || !feat || !target || !grad) return fail(c, ERR_INVAL, "gradient arguments");
    if (!finite_all(feat, n * MPF_MLP_NIN) || !finite_all(target, n * MPF_MLP_NOUT))
        return fail(c, ERR_INVAL, "non-finite batch");
    std::string err;
    if (!mpf::mlp_gradient(c->mlp, feat, target, n, grad, loss, err)) return fail(c, ERR_HIP, err);
    return 0;
}

int mpf_mlp_infer_dev(mpf_ctx* c, const float* dev_feat, uint64_t n, float* dev_out) {
    if (int rc = need_mlp(c)) return rc;
    if (n && (!dev_feat || !dev_out)) return ERR_INVAL;
    std::string err;
    if (!mpf::mlp_infer_dev(c->mlp, dev_feat, n, dev_out, err)) return fail(c, ERR_HIP, err);
    return 0;
}

int mpf_mlp_last_ms(mpf_ctx* c, double* ms) {
    if (!c || !ms) return ERR_INVAL;
    if (int rc = need_mlp(c)) return rc;
    *ms = (double)mpf::mlp_last_ms(c->mlp);
    return 0;
}

int mpf_mlp_infer(mpf_ctx* c, const float* feat, uint64_t n, float* out) {
    if (int rc = need_mlp(c)) return rc;
    if (n && (!feat || !out)) return ERR_INVAL;
    if (n == 0) return 0;
    DevBuf<float> f, o;
    if (f.ensure(n * MPF_MLP_NIN) || o.ensure(n * MPF_MLP_NOUT)) {
        f.release(); o.release();
        return fail(c, ERR_NOMEM, "inference buffers");
    }
    int rc = 0;
    if (hipMemcpy(f.p, feat, n * MPF_MLP_NIN * 4, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(c, ERR_HIP, "inference upload");
    if (!rc) rc = mpf_mlp_infer_dev(c, f.p, n, o.p);
    if (!rc && hipMemcpy(out, o.p, n * MPF_MLP_NOUT * 4, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(c, ERR_HIP, "inference download");
    f.release(); o.release();
    return rc;
}

int mpf_catalogue_features_dev(mpf_ctx* c, const double* dev_src6, uint64_t n, float* dev_feat) {
    if (!c || (n && (!dev_src6 || !dev_feat))) return ERR_INVAL;
    if (!c->have_tree) return fail(c, ERR_STATE, "no octree loaded");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    // sources by direction: a wavefront's lines of sight then share cells
    // (each source's features are its own, whatever the order)
    const uint32_t* order = nullptr;
    if (n > 1 && n < (1ull << 31)) {
        const size_t wb = mpf::ray_order_work_bytes(n);
        if (c->order.ensure(n) != hipSuccess || c->order_work.ensure(wb) != hipSuccess)
            return fail(c, ERR_NOMEM, "source-order buffers");
        HIPCHK(c, mpf::launch_source_order(dev_src6, n, c->order.p, c->order_work.p, wb, c->stream));
        order = c->order.p;
    }
    if (c->cfg.precision == MPF_PREC_SPLIT && c->cfg.order == MPF_TSC) {
        if (int rc = ensure_split(c)) return rc;
        HIPCHK(c, mpf::launch_features_split(c->dt, c->split, dev_src6, n, order, c->cfg.length, dev_feat, c->stream));
    } else {
        HIPCHK(c, mpf::launch_features(c->dt, dev_src6, n, order, c->cfg.length, dev_feat, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int mpf_catalogue_features(mpf_ctx* c, const double* src6, uint64_t n, float* feat) {
    if (!c || (n && (!src6 || !feat))) return ERR_INVAL;
    if (n == 0) return 0;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    DevBuf<double> s;
    DevBuf<float> f;
    if (s.ensure(6 * n) || f.ensure(n * MPF_MLP_NIN)) {
        s.release(); f.release();
        return fail(c, ERR_NOMEM, "feature buffers");
    }
    int rc = 0;
    if (hipMemcpy(s.p, src6, 6 * n * 8, hipMemcpyHostToDevice) != hipSuccess) rc = fail(c, ERR_HIP, "feature upload");
    if (!rc) rc = mpf_catalogue_features_dev(c, s.p, n, f.p);
    if (!rc && hipMemcpy(feat, f.p, n * MPF_MLP_NIN * 4, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(c, ERR_HIP, "feature download");
    s.release(); f.release();
    return rc;
}

int mpf_catalogue_seed_dev(mpf_ctx* c, const double* dev_src6, uint64_t n, int use_network, double* dev_init2) {
    if (!c) return ERR_INVAL;
    if (use_network)
        if (int rc = need_mlp(c)) return rc;
    if (n && (!dev_src6 || !dev_init2)) return ERR_INVAL;
    if (n == 0) return 0;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    DevBuf<float> f, o;
    if (f.ensure(n * MPF_MLP_NIN) || (use_network && o.ensure(n * MPF_MLP_NOUT))) {
        f.release(); o.release();
        return fail(c, ERR_NOMEM, "seed buffers");
    }
    int rc = mpf_catalogue_features_dev(c, dev_src6, n, f.p);
    if (!rc && use_network) rc = mpf_mlp_infer_dev(c, f.p, n, o.p);
    if (!rc) {
        hipError_t e = mpf::launch_seed(dev_src6, f.p, use_network ? o.p : nullptr, n, dev_init2, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hipfail(c, e, "seed");
    }
    f.release(); o.release();
    return rc;
}

int mpf_catalogue_seed(mpf_ctx* c, const double* src6, uint64_t n, int use_network, double* init2) {
    if (!c) return ERR_INVAL;
    if (use_network)
        if (int rc = need_mlp(c)) return rc;
    if (n && (!src6 || !init2)) return ERR_INVAL;
    if (n == 0) return 0;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    DevBuf<double> s, i2;
    if (s.ensure(6 * n) || i2.ensure(2 * n)) {
        s.release(); i2.release();
        return fail(c, ERR_NOMEM, "seed buffers");
    }
    int rc = 0;
    if (hipMemcpy(s.p, src6, 6 * n * 8, hipMemcpyHostToDevice) != hipSuccess) rc = fail(c, ERR_HIP, "seed upload");
    if (!rc) rc = mpf_catalogue_seed_dev(c, s.p, n, use_network, i2.p);
    if (!rc && hipMemcpy(init2, i2.p, 2 * n * 8, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(c, ERR_HIP, "seed download");
    s.release(); i2.release();
    return rc;
}

int mpf_mfma_probe(mpf_ctx* c, const float* a, const float* b, float* out) {
    if (!c || !a || !b || !out) return ERR_INVAL;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    DevBuf<float> d;
    if (d.ensure(3 * 256)) return fail(c, ERR_NOMEM, "probe buffer");
    int rc = 0;
    if (hipMemcpy(d.p, a, 1024, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.p + 256, b, 1024, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(c, ERR_HIP, "probe upload");
    if (!rc) {
        hipError_t e = mpf::launch_mfma_probe(d.p, d.p + 256, d.p + 512, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(out, d.p + 512, 1024, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hipfail(c, e, "probe");
    }
    d.release();
    return rc;
}

int mpf_mfma_probe32(mpf_ctx* c, const float* a, const float* b, int layout, float* out) {
    if (!c || !a || !b || !out || (layout != 0 && layout != 1)) return ERR_INVAL;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    DevBuf<float> d;
    if (d.ensure(512 + 512 + 256)) return fail(c, ERR_NOMEM, "probe buffer");
    int rc = 0;
    if (hipMemcpy(d.p, a, 2048, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d.p + 512, b, 2048, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(c, ERR_HIP, "probe upload");
    if (!rc) {
        hipError_t e = mpf::launch_mfma_probe32(d.p, d.p + 512, d.p + 1024, layout, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(out, d.p + 1024, 1024, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hipfail(c, e, "probe");
    }
    d.release();
    return rc;
}

int mpf_healpix_pix2vec(int64_t nside, const int64_t* pix, uint64_t n, double* vec3) {
    if (nside < 1 || nside > (1 << 29) || (n && (!pix || !vec3))) return ERR_INVAL;
    const int64_t npix = 12 * nside * nside;
    for (uint64_t i = 0; i < n; ++i) {
        if (pix[i] < 0 || pix[i] >= npix) return ERR_INVAL;
        mpf::healpix::pix2vec_ring(nside, pix[i], vec3 + 3 * i);
    }
    return 0;
}

int mpf_healpix_vec2pix(int64_t nside, const double* vec3, uint64_t n, int64_t* pix) {
    if (nside < 1 || nside > (1 << 29) || (n && (!pix || !vec3))) return ERR_INVAL;
    for (uint64_t i = 0; i < n; ++i)
        pix[i] = mpf::healpix::vec2pix_ring(nside, vec3[3 * i], vec3[3 * i + 1], vec3[3 * i + 2]);
    return 0;
}

int mpf_hmaps_launch_dev(mpf_ctx* c, int64_t nside, const int64_t* dev_pix, uint64_t n, double* dev_in9) {
    if (!c || nside < 1 || nside > (1 << 29) || (n && (!dev_pix || !dev_in9))) return ERR_INVAL;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    HIPCHK(c, mpf::launch_pixel_launch(nside, dev_pix, n, dev_in9, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int mpf_hmaps_fill_dev(mpf_ctx* c, int64_t nside, const int64_t* dev_pix, const double* dev_surf, uint64_t n,
                       uint32_t nsurf, double* dev_maps) {
    if (!c || nside < 1 || nside > (1 << 29) || nsurf < 1 || nsurf > MPF_MAX_SURFACES ||
        (n && (!dev_pix || !dev_surf || !dev_maps)))
        return ERR_INVAL;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    HIPCHK(c, mpf::launch_pixel_fill(nside, dev_pix, dev_surf, (int)nsurf, n, dev_maps, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int mpf_hmaps_components_dev(mpf_ctx* c, int64_t nside, const int64_t* dev_pix, const double* dev_surf,
                             const double* dev_surf_born, const double* dev_jac, uint64_t n, uint32_t nz,
                             const int* components, uint32_t ncomp,
                             const double* homo5, float* dev_maps, uint64_t maps_capacity, uint32_t* nmaps_out) {
    if (!c || nside < 1 || nside > (1 << 29) || nz < 1 || nz > MPF_MAX_SURFACES || ncomp < 1 || ncomp > 16 ||
        !components || !homo5 || (n && !dev_pix))
        return ERR_INVAL;
    int nmaps = 0;
    bool lensing = false, born = false, real = false;
    for (uint32_t j = 0; j < ncomp; ++j) {
        const int k = components[j];
        if (k < MPF_MAP_LENSING || k > MPF_MAP_DEFLECTION) return fail(c, ERR_INVAL, "unknown map component");
        lensing |= k == MPF_MAP_LENSING;
        born |= k == MPF_MAP_LENSING_BORN;
        real |= k != MPF_MAP_LENSING_BORN;
        nmaps += (k == MPF_MAP_LENSING) ? (dev_jac ? 5 : 4)
                                         : (k == MPF_MAP_LENSING_BORN ? 4 : (k == MPF_MAP_DEFLECTION ? 2 : 1));
    }
    if (n && ((real && !dev_surf) || (born && !dev_surf_born)))
        return fail(c, ERR_INVAL, "the components need the ray's surfaces and, for lensing_born, the Born surfaces");
    if (nmaps_out) *nmaps_out = (uint32_t)nmaps;
    const uint64_t need = (uint64_t)nz * (uint64_t)nmaps * 12ull * (uint64_t)nside * (uint64_t)nside;
    if (!dev_maps || need > maps_capacity)
        return fail(c, ERR_INVAL, "map buffer holds " + std::to_string(maps_capacity) + " floats, " +
                                      std::to_string(need) + " needed");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->homo.ensure(5 * nz) != hipSuccess) return fail(c, ERR_NOMEM, "homogeneous values");
    HIPCHK(c, hipMemcpy(c->homo.p, homo5, 5 * 8 * (uint64_t)nz, hipMemcpyHostToDevice));
    HIPCHK(c, mpf::launch_component_fill(nside, dev_pix, real ? dev_surf : nullptr, born ? dev_surf_born : nullptr,
                                         lensing ? dev_jac : nullptr, c->homo.p, n, (int)nz, components, (int)ncomp,
                                         nmaps, dev_maps, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int mpf_healpix_bin_dev(mpf_ctx* c, int64_t nside, const double* dev_in9, const double* dev_surf, uint64_t n,
                        uint32_t nsurf, uint32_t isurf, double* dev_maps) {
    if (!c || nside < 1 || nside > (1 << 16) || nsurf < 1 || nsurf > MPF_MAX_SURFACES || isurf >= nsurf ||
        !dev_maps || (n && (!dev_in9 || !dev_surf)))
        return ERR_INVAL;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const uint64_t npix = 12ull * (uint64_t)nside * (uint64_t)nside;
    if (c->map_acc.ensure(4 * npix) != hipSuccess) return fail(c, ERR_NOMEM, "map accumulator");
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    HIPCHK(c, mpf::launch_bin(nside, dev_in9, dev_surf, (int)nsurf, (int)isurf, n, c->map_acc.p, dev_maps, c->stream));
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return 0;
}

int mpf_synchronize(mpf_ctx* c) {
    if (!c) return ERR_INVAL;
    return finish_run(c, c->counters.p);
}

// waits for the context stream; after an integration, reads its step counter
// and timings
static int finish_run(mpf_ctx* c, const unsigned long long* counters) {
    HIPCHK(c, hipSetDevice(c->cfg.device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->pending) {
        unsigned long long h[2] = {0, 0};
        HIPCHK(c, hipMemcpy(h, counters, sizeof(h), hipMemcpyDeviceToHost));
        c->last_steps = h[1];
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->last_ms = ms;
        ms = 0.f;
        if (c->kernel_timed) HIPCHK(c, hipEventElapsedTime(&ms, c->evk0, c->evk1));
        c->last_kernel_ms = ms;
        c->pending = false;
    }
    return 0;
}

static int check_launch_cfg(mpf_ctx* c, const mpf_launch_cfg& L) {
    if (L.refill < 1 || L.refill > 64) return fail(c, ERR_INVAL, "refill must be 1..64");
    if (L.blocks_per_cu > 32) return fail(c, ERR_INVAL, "blocks_per_cu must be 0..32");
    if (L.waves_per_simd > 4) return fail(c, ERR_INVAL, "waves_per_simd must be 0..4");
    return 0;
}

int mpf_set_launch_config(mpf_ctx* c, int kind, const mpf_launch_cfg* cfg) {
    if (!c || !cfg || (kind != 0 && kind != 1)) return ERR_INVAL;
    if (int rc = check_launch_cfg(c, *cfg)) return rc;
    c->launch[kind] = *cfg;
    return 0;
}

int mpf_get_launch_config(mpf_ctx* c, int kind, mpf_launch_cfg* cfg) {
    if (!c || !cfg || (kind != 0 && kind != 1)) return ERR_INVAL;
    *cfg = c->launch[kind];
    return 0;
}

int mpf_autotune(mpf_ctx* c, const double* dev_in9, uint64_t n, const mpf_stop* stop, const mpf_surfaces* surf,
                 double* dev_out, double* dev_out_surf, mpf_launch_cfg* best, double* trials, uint32_t max_trials,
                 uint32_t* ntrials) {
    if (!c || (!stop && !surf) || n == 0 || !dev_in9) return ERR_INVAL;
    const int kind = surf ? 1 : 0;
    const bool split = c->cfg.precision == MPF_PREC_SPLIT;
    uint32_t nt = 0;
    auto run = [&](const mpf_launch_cfg& L, double& ms) -> int {
        c->launch[kind] = L;
        const int rc = surf ? mpf_integrate_surfaces_dev(c, dev_in9, n, surf, dev_out, dev_out_surf, 1)
                            : integrate_impl(c, dev_in9, n, stop, dev_out, nullptr, nullptr, 1);
        if (rc) return rc;
        ms = c->last_ms;
        if (trials && nt < max_trials) {
            double* r = trials + 6 * (uint64_t)nt;
            r[0] = L.refill; r[1] = (double)L.sort_bucket; r[2] = L.blocks_per_cu; r[3] = L.waves_per_simd;
            r[4] = ms; r[5] = (double)c->last_steps;
        }
        ++nt;
        return 0;
    };
    mpf_launch_cfg cur = c->launch[kind];
    if (!cur.blocks_per_cu) cur.blocks_per_cu = split ? 8 : 4;
    if (!cur.waves_per_simd && split) cur.waves_per_simd = 2;
    double ms = 0.0, best_ms = 0.0;
    // the first run also builds what the tier needs once (stencil blocks); time the second
    if (int rc = run(cur, ms)) return rc;
    if (int rc = run(cur, best_ms)) return rc;
    // coordinate descent, one knob at a time; a change is kept when it saves
    // at least 1 % (the run-to-run spread of the kernel time).  The refill
    // first: with refill 1 a wave takes rays as lanes free up, which undoes
    // most of any order the sort buckets give, so buckets compared there
    // are not the ones that win at refill 64; the buckets again last, once
    // the occupancy is set.
    const uint64_t buckets[] = {0, 4096, 16384, 65536, 262144, 1048576, n};
    const uint32_t refills[] = {1, 16, 32, 64};
    const uint32_t waves_plain[] = {1, 2, 3}, waves_surf[] = {1, 2};
    const uint32_t blocks_split[] = {4, 8, 16}, blocks_exact[] = {2, 4, 8};
    auto try_knob = [&](auto setter, const auto& values) -> int {
        mpf_launch_cfg keep = cur;
        for (auto v : values) {
            mpf_launch_cfg L = cur;
            setter(L, v);
            if (std::memcmp(&L, &cur, sizeof(L)) == 0) continue;
            if (int rc = run(L, ms)) return rc;
            if (ms < 0.99 * best_ms) {
                best_ms = ms;
                keep = L;
            }
        }
        cur = keep;
        return 0;
    };
    auto bucket = [](mpf_launch_cfg& L, uint64_t v) { L.sort_bucket = v; };
    int rc = try_knob([](mpf_launch_cfg& L, uint32_t v) { L.refill = v; }, refills);
    if (!rc) rc = try_knob(bucket, buckets);
    if (!rc && split && !surf) rc = try_knob([](mpf_launch_cfg& L, uint32_t v) { L.waves_per_simd = v; }, waves_plain);
    if (!rc && split && surf) rc = try_knob([](mpf_launch_cfg& L, uint32_t v) { L.waves_per_simd = v; }, waves_surf);
    if (!rc && split) rc = try_knob([](mpf_launch_cfg& L, uint32_t v) { L.blocks_per_cu = v; }, blocks_split);
    if (!rc && !split) rc = try_knob([](mpf_launch_cfg& L, uint32_t v) { L.blocks_per_cu = v; }, blocks_exact);
    if (!rc) rc = try_knob(bucket, buckets);
    if (rc) return rc;
    c->launch[kind] = cur;
    if (best) *best = cur;
    if (ntrials) *ntrials = nt;
    return 0;
}

int mpf_last_kernel_ms(mpf_ctx* c, double* ms) {
    if (!c || !ms) return ERR_INVAL;
    int rc = mpf_synchronize(c);
    if (rc) return rc;
    *ms = c->last_kernel_ms;
    return 0;
}

int mpf_last_run(mpf_ctx* c, uint64_t* total_steps, double* kernel_ms) {
    if (!c) return ERR_INVAL;
    int rc = mpf_synchronize(c);
    if (rc) return rc;
    if (total_steps) *total_steps = c->last_steps;
    if (kernel_ms) *kernel_ms = c->last_ms;
    return 0;
}

int mpf_integrate_surfaces(mpf_ctx* c, const double* in9, uint64_t n, const mpf_surfaces* surf, double* out,
                           double* out_surf) {
    if (!c || !surf || (n && (!in9 || !out_surf))) return ERR_INVAL;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const uint64_t ns = surf->n;
    if (ns < 1 || ns > MPF_MAX_SURFACES) return fail(c, ERR_INVAL, "surface count must be 1..16");
    if (c->io_in.ensure(9 * std::max<uint64_t>(n, 1)) != hipSuccess ||
        c->io_out.ensure((uint64_t)MPF_NOUT * std::max<uint64_t>(n, 1)) != hipSuccess ||
        c->io_surf.ensure((uint64_t)MPF_NSURF * ns * std::max<uint64_t>(n, 1)) != hipSuccess)
        return fail(c, ERR_NOMEM, "integration buffers");
    if (n) HIPCHK(c, hipMemcpyAsync(c->io_in.p, in9, 9 * n * 8, hipMemcpyHostToDevice, c->stream));
    mpf_surfaces sd = *surf;
    if (surf->ray_value) {
        if (c->io_rv.ensure(std::max<uint64_t>(n, 1)) != hipSuccess) return fail(c, ERR_NOMEM, "per-ray values");
        if (n) HIPCHK(c, hipMemcpyAsync(c->io_rv.p, surf->ray_value, n * 8, hipMemcpyHostToDevice, c->stream));
        sd.ray_value = c->io_rv.p;
    }
    int rc = mpf_integrate_surfaces_dev(c, c->io_in.p, n, &sd, c->io_out.p, c->io_surf.p, 1);
    if (rc) return rc;
    if (n && out) HIPCHK(c, hipMemcpy(out, c->io_out.p, (uint64_t)MPF_NOUT * n * 8, hipMemcpyDeviceToHost));
    if (n) HIPCHK(c, hipMemcpy(out_surf, c->io_surf.p, (uint64_t)MPF_NSURF * ns * n * 8, hipMemcpyDeviceToHost));
    return 0;
}

// value of the stop parameter in a per-ray output record (lensing.h:205-301)
static double bundle_param(const double* o, int kind) {
    switch (kind) {
        case MPF_STOP_LAMBDA: return o[MPF_OUT_LAMBDA];
        case MPF_STOP_RADIUS: return o[MPF_OUT_CHI];
        case MPF_STOP_A: return o[MPF_OUT_A];
        case MPF_STOP_T: return o[MPF_OUT_T];
        default: return o[MPF_OUT_REDSHIFT];
    }
}

// The four launch photons of a bundle around (phi, theta): +-tan(opening)
// along e1 and e2 (lensing.h:147-195)
static void bundle_photons(const double* phi_theta, const double* observer3, double to, double* in36) {
    double sp, cp, st, ct;
    ::sincos(phi_theta[0], &sp, &cp);
    ::sincos(phi_theta[1], &st, &ct);
    const double t[3] = {cp * st, sp * st, ct};
    const double e1[3] = {-sp, cp, 0}, e2[3] = {-cp * ct, -sp * ct, st};
    const double rini[4][3] = {{t[0] + to * e1[0], t[1] + to * e1[1], t[2]},
                               {t[0] - to * e1[0], t[1] - to * e1[1], t[2]},
                               {t[0] - to * e2[0], t[1] - to * e2[1], t[2] - to * e2[2]},
                               {t[0] + to * e2[0], t[1] + to * e2[1], t[2] + to * e2[2]}};
    for (int k = 0; k < 4; ++k) mpf::cone::launch_towards(observer3, rini[k], in36 + 9 * k);
}

// A direction's Jacobians from its four rays' crossings fp[r][ray][xyz] of
// its nd planes (lensing.h:308-335); a failed direction gets 42, 42, 1, 0
static void bundle_plane_combine(const double* ki_i, const double* dist_i, uint32_t nd, double to,
                                 const double (*fp)[4][3], bool failed, double* J) {
    for (uint64_t q = 0; q < 4ull * nd; ++q) J[q] = 0.0;
    if (failed) {
        for (uint32_t q = 0; q < nd; ++q) {
            J[4 * q] = 42; J[4 * q + 1] = 42; J[4 * q + 2] = 1; J[4 * q + 3] = 0;
        }
        return;
    }
    for (uint32_t r = 0; r < nd; ++r) {
        double rh[3], rvv[3];
        for (int d = 0; d < 3; ++d) {
            rh[d] = fp[r][0][d] - fp[r][1][d];
            rvv[d] = fp[r][2][d] - fp[r][3][d];
        }
        const double* k3 = ki_i + 3 * r;
        const double phik = std::atan2(k3[1], k3[0]);
        const double thetak = std::acos(k3[2] / std::sqrt(k3[0] * k3[0] + k3[1] * k3[1] + k3[2] * k3[2]));
        double spk, cpk, stk, ctk;
        ::sincos(phik, &spk, &cpk);
        ::sincos(thetak, &stk, &ctk);
        const double f1[3] = {-spk, cpk, 0}, f2[3] = {-cpk * ctk, -spk * ctk, stk};
        const double dd = dist_i[r];
        J[4 * r + 0] = (rh[0] * f1[0] + rh[1] * f1[1]) / (2 * dd * to);
        J[4 * r + 1] = (rvv[0] * f1[0] + rvv[1] * f1[1]) / (2 * dd * to);
        J[4 * r + 2] = (rh[0] * f2[0] + rh[1] * f2[1] + rh[2] * f2[2]) / (-2 * dd * to);
        J[4 * r + 3] = (rvv[0] * f2[0] + rvv[1] * f2[1] + rvv[2] * f2[2]) / (-2 * dd * to);
    }
}

// stop_bundle = plane for one direction from whole trajectories (lensing.h:201,
// 281-304): the four rays stopped on the last value along the last ki, each
// crossing found by upper_bound along every ki.  The batched path below
// leaves the directions whose crossings it cannot place to this one.
static int bundle_plane_traj(mpf_ctx* c, const double* phi_theta, const double* observer3, const double* ref,
                             uint32_t nd, const double* ki, const double* dist, double to, double* J) {
    uint64_t cap = 1ull << 14;
    std::vector<double> out((uint64_t)MPF_NOUT * 4), traj;
    double in9[36];
    bundle_photons(phi_theta, observer3, to, in9);
    const double* klast = ki + 3 * (nd - 1);
    mpf_stop stop{MPF_STOP_PLANE, ref[nd - 1], {klast[0], klast[1], klast[2]}};
    int rc = 0;
    for (;;) {
        traj.resize((uint64_t)MPF_NOUT * 4 * cap);
        rc = mpf_integrate_trajectory(c, in9, 4, &stop, cap, out.data(), traj.data());
        if (rc != ERR_NOSPC || cap >= (1ull << 22)) break;
        cap <<= 2;
    }
    if (rc) return rc;
    double fp[MPF_MAX_SURFACES][4][3];
    bool failed = false;
    for (int it = 0; it < 4 && !failed; ++it) {
        const int64_t len = (int64_t)out[(uint64_t)MPF_NOUT * it + MPF_OUT_NSTEPS] + 1;
        const double* tr = traj.data() + (uint64_t)MPF_NOUT * cap * it;
        for (uint32_t r = 0; r < nd && !failed; ++r) {
            const double* kr = ki + 3 * r;
            const double v = ref[r];
            auto along = [&](int64_t e) {
                const double* x = tr + (uint64_t)MPF_NOUT * e;
                return x[MPF_OUT_X] * kr[0] + x[MPF_OUT_Y] * kr[1] + x[MPF_OUT_Z] * kr[2];
            };
            if (len < 2 || along(len - 1) < v) {
                failed = true;
                break;
            }
            int64_t id = 0;
            double f = 0.0;
            mpf::plane_segment(len, v, along, &id, &f);
            const double* a = tr + (uint64_t)MPF_NOUT * id;
            const double* b = a + MPF_NOUT;
            fp[r][it][0] = a[MPF_OUT_X] * f + b[MPF_OUT_X] * (1 - f);
            fp[r][it][1] = a[MPF_OUT_Y] * f + b[MPF_OUT_Y] * (1 - f);
            fp[r][it][2] = a[MPF_OUT_Z] * f + b[MPF_OUT_Z] * (1 - f);
        }
    }
    bundle_plane_combine(ki, dist, nd, to, fp, failed, J);
    return 0;
}

// stop_bundle = plane, batched: the four rays of every direction in one
// exact-tier launch per chunk of directions, each ray recording where it
// crosses its direction's planes while it is integrated (IntegrateArgs::
// nplane); no trajectory is stored.  A ray whose projections on a normal
// decrease somewhere (where upper_bound over the stored trajectory and the
// first crossing could differ) sends its direction to bundle_plane_traj.
static int jacobian_bundle_plane(mpf_ctx* c, const double* phi_theta, uint64_t n, const double* observer3,
                                 const double* ref, uint32_t nd, const double* ki, const double* dist, double opening,
                                 double* out4) {
    const double to = std::tan(opening);
    const int prec = c->cfg.precision;
    if (prec == MPF_PREC_SPLIT) c->cfg.precision = MPF_PREC_FP64;
    const uint64_t chunk = std::min<uint64_t>(n, (1ull << 20) / nd);
    DevBuf<double> din, dval, dkit, dpl;
    std::vector<double> h9(36 * chunk), hpl((uint64_t)4 * chunk * nd * mpf::PL_NREC);
    std::vector<uint64_t> fallback;
    int rc = 0;
    if (din.ensure(36 * chunk) || dval.ensure(chunk * nd) || dkit.ensure(3 * chunk * nd) ||
        dpl.ensure((uint64_t)4 * chunk * nd * mpf::PL_NREC))
        rc = fail(c, ERR_NOMEM, "bundle-plane buffers");
    for (uint64_t i0 = 0; !rc && i0 < n; i0 += chunk) {
        const uint64_t m = std::min(chunk, n - i0);
        for (uint64_t q = 0; q < m; ++q) bundle_photons(phi_theta + 2 * (i0 + q), observer3, to, h9.data() + 36 * q);
        if (hipMemcpy(din.p, h9.data(), 36 * m * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dval.p, ref + i0 * nd, m * nd * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dkit.p, ki + 3 * i0 * nd, 3 * m * nd * 8, hipMemcpyHostToDevice) != hipSuccess) {
            rc = fail(c, ERR_HIP, "bundle-plane upload");
            break;
        }
        PlaneSet P;
        P.n = (int)nd;
        P.div = 4;
        P.val = dval.p;
        P.kit = dkit.p;
        P.out = dpl.p;
        mpf_stop stop{MPF_STOP_PLANE, 0.0, {0.0, 0.0, 0.0}};
        rc = integrate_impl(c, din.p, 4 * m, &stop, nullptr, nullptr, nullptr, 1, nullptr, 0, nullptr, &P);
        if (rc) break;
        if (hipMemcpy(hpl.data(), dpl.p, (uint64_t)4 * m * nd * mpf::PL_NREC * 8, hipMemcpyDeviceToHost) !=
            hipSuccess) {
            rc = fail(c, ERR_HIP, "bundle-plane download");
            break;
        }
        for (uint64_t q = 0; q < m; ++q) {
            double fp[MPF_MAX_SURFACES][4][3];
            bool failed = false, whole = false;
            for (int it = 0; it < 4; ++it)
                for (uint32_t r = 0; r < nd; ++r) {
                    const double* o = hpl.data() + ((4 * q + it) * nd + r) * mpf::PL_NREC;
                    if (o[mpf::PL_STATE] == 0.0) failed = true;
                    else if (o[mpf::PL_STATE] != 1.0) whole = true;
                    else for (int d = 0; d < 3; ++d) fp[r][it][d] = o[mpf::PL_X + d];
                }
            // any ray that never reaches a plane fails the direction, wherever
            // the others cross (lensing.h:281-290 returns at the first)
            if (whole && !failed) {
                fallback.push_back(i0 + q);
                continue;
            }
            bundle_plane_combine(ki + 3 * (i0 + q) * nd, dist + (i0 + q) * nd, nd, to, fp, failed,
                                 out4 + 4 * (uint64_t)nd * (i0 + q));
        }
    }
    din.release(); dval.release(); dkit.release(); dpl.release();
    c->bundle_plane_fallback = fallback.size();
    for (uint64_t i : fallback) {
        if (rc) break;
        rc = bundle_plane_traj(c, phi_theta + 2 * i, observer3, ref + i * nd, nd, ki + 3 * i * nd, dist + i * nd, to,
                               out4 + 4 * (uint64_t)nd * i);
    }
    c->cfg.precision = prec;
    return rc;
}

int mpf_split_tidal_blocks(mpf_ctx* c, int enable, uint64_t* nblocks) {
    if (!c || (enable != 0 && enable != 1 && enable != -1)) return ERR_INVAL;
    if (enable == 0 && c->split.tidal) {
        HIPCHK(c, hipSetDevice(c->cfg.device));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        (void)hipFree(c->split.tidal);
        (void)hipFree(c->split.tidx);
        c->split.tidal = nullptr;
        c->split.tidx = nullptr;
        c->split.ntidal = 0;
    }
    if (enable >= 0) c->split_tidal = enable != 0;
    if (nblocks) *nblocks = c->split.tidal ? c->split.ntidal : 0;
    return 0;
}

int mpf_bundle_plane_info(const mpf_ctx* c, uint64_t* whole_trajectory_directions) {
    if (!c || !whole_trajectory_directions) return ERR_INVAL;
    *whole_trajectory_directions = c->bundle_plane_fallback;
    return 0;
}

int mpf_jacobian_bundle(mpf_ctx* c, const double* phi_theta, uint64_t n, const double* observer3, int kind,
                        const double* ref, uint32_t nd, const double* ki, const double* dist, double opening,
                        double* out4) {
    if (!c || !observer3 || (n && (!phi_theta || !ref || !ki || !dist || !out4))) return ERR_INVAL;
    if (kind != MPF_STOP_RADIUS && kind != MPF_STOP_LAMBDA && kind != MPF_STOP_REDSHIFT && kind != MPF_STOP_T &&
        kind != MPF_STOP_A && kind != MPF_STOP_PLANE)
        return fail(c, ERR_INVAL, "stop_bundle must be radius, lambda, redshift, t, a or plane");
    if (nd < 1 || nd > MPF_MAX_SURFACES) return fail(c, ERR_INVAL, "1..16 values per direction");
    if (n == 0) return 0;
    if (kind == MPF_STOP_PLANE)
        return jacobian_bundle_plane(c, phi_theta, n, observer3, ref, nd, ki, dist, opening, out4);
    // the same values for every direction: surfaces; else one value per direction
    bool same = true;
    for (uint64_t i = 1; i < n && same; ++i)
        for (uint32_t r = 0; r < nd; ++r)
            if (ref[i * nd + r] != ref[r]) { same = false; break; }
    if (!same && nd != 1) return fail(c, ERR_INVAL, "values differing between directions need one value each");
    // four rays per direction at +-tan(opening) along e1, e2 (lensing.h:147-190); sin and cos of one variable
    // come from one sincos call, as the reference build computes them
    const double to = std::tan(opening);
    std::vector<double> in9(36 * n), rv(same ? 0 : 4 * n);
    for (uint64_t i = 0; i < n; ++i) {
        double sp, cp, st, ct;
        ::sincos(phi_theta[2 * i], &sp, &cp);
        ::sincos(phi_theta[2 * i + 1], &st, &ct);
        const double t[3] = {cp * st, sp * st, ct};
        const double e1[3] = {-sp, cp, 0}, e2[3] = {-cp * ct, -sp * ct, st};
        const double rini[4][3] = {{t[0] + to * e1[0], t[1] + to * e1[1], t[2]},
                                   {t[0] - to * e1[0], t[1] - to * e1[1], t[2]},
                                   {t[0] - to * e2[0], t[1] - to * e2[1], t[2] - to * e2[2]},
                                   {t[0] + to * e2[0], t[1] + to * e2[1], t[2] + to * e2[2]}};
        for (int k = 0; k < 4; ++k) {
            mpf::cone::launch_towards(observer3, rini[k], in9.data() + 9 * (4 * i + k));
            if (!same) rv[4 * i + k] = ref[i];
        }
    }
    mpf_surfaces S{};
    S.kind = kind;
    S.n = same ? nd : 1;
    for (uint32_t r = 0; r < S.n; ++r) S.value[r] = same ? ref[r] : 0.0;
    S.jacobian = 0;
    S.ray_value = same ? nullptr : rv.data();
    std::vector<double> out((uint64_t)MPF_NOUT * 4 * n), surf((uint64_t)MPF_NSURF * S.n * 4 * n);
    if (int rc = mpf_integrate_surfaces(c, in9.data(), 4 * n, &S, out.data(), surf.data())) return rc;
    // combine (lensing.h:195-360): failures in the reference's loop order, then finite differences
    for (uint64_t i = 0; i < n; ++i) {
        double* J = out4 + 4 * (uint64_t)nd * i;
        for (uint64_t q = 0; q < 4ull * nd; ++q) J[q] = 0.0;
        double fp[MPF_MAX_SURFACES][4][3];
        bool failed = false;
        for (int it = 0; it < 4 && !failed; ++it) {
            const uint64_t ray = 4 * i + it;
            const double back = bundle_param(out.data() + (uint64_t)MPF_NOUT * ray, kind);
            for (uint32_t r = 0; r < nd && !failed; ++r) {
                const double v = ref[i * nd + r];
                const double* sr = surf.data() + ((uint64_t)S.n * ray + (same ? r : 0)) * MPF_NSURF;
                const bool not_reached = (kind == MPF_STOP_A) ? (back > v) : (back < v);
                if (not_reached || sr[0] == 0.0) {
                    failed = true;
                    if (kind == MPF_STOP_A) {
                        J[4 * r] = 42; J[4 * r + 1] = 42; J[4 * r + 2] = 1; J[4 * r + 3] = 3;
                    } else {
                        for (uint32_t q = 0; q < nd; ++q) {
                            J[4 * q] = 42; J[4 * q + 1] = 42; J[4 * q + 2] = 1; J[4 * q + 3] = 0;
                        }
                    }
                    break;
                }
                fp[r][it][0] = sr[3];
                fp[r][it][1] = sr[4];
                fp[r][it][2] = sr[5];
            }
        }
        if (failed) continue;
        for (uint32_t r = 0; r < nd; ++r) {
            double rh[3], rvv[3];
            for (int d = 0; d < 3; ++d) {
                rh[d] = fp[r][0][d] - fp[r][1][d];
                rvv[d] = fp[r][2][d] - fp[r][3][d];
            }
            const double* k3 = ki + 3 * (i * nd + r);
            const double phik = std::atan2(k3[1], k3[0]);
            const double thetak = std::acos(k3[2] / std::sqrt(k3[0] * k3[0] + k3[1] * k3[1] + k3[2] * k3[2]));
            double spk, cpk, stk, ctk;
            ::sincos(phik, &spk, &cpk);
            ::sincos(thetak, &stk, &ctk);
            const double f1[3] = {-spk, cpk, 0}, f2[3] = {-cpk * ctk, -spk * ctk, stk};
            const double dd = dist[i * nd + r];
            J[4 * r + 0] = (rh[0] * f1[0] + rh[1] * f1[1]) / (2 * dd * to);
            J[4 * r + 1] = (rvv[0] * f1[0] + rvv[1] * f1[1]) / (2 * dd * to);
            J[4 * r + 2] = (rh[0] * f2[0] + rh[1] * f2[1] + rh[2] * f2[2]) / (-2 * dd * to);
            J[4 * r + 3] = (rvv[0] * f2[0] + rvv[1] * f2[1] + rvv[2] * f2[2]) / (-2 * dd * to);
        }
    }
    return 0;
}

int mpf_integrate(mpf_ctx* c, const double* in9, uint64_t n, const mpf_stop* stop, double* out,
                  uint64_t* total_steps) {
    if (!c || !stop || (n && (!in9 || !out))) return ERR_INVAL;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    if (c->io_in.ensure(9 * std::max<uint64_t>(n, 1)) != hipSuccess ||
        c->io_out.ensure((uint64_t)MPF_NOUT * std::max<uint64_t>(n, 1)) != hipSuccess)
        return fail(c, ERR_NOMEM, "integration buffers");
    if (n) HIPCHK(c, hipMemcpyAsync(c->io_in.p, in9, 9 * n * 8, hipMemcpyHostToDevice, c->stream));
    int rc = mpf_integrate_dev(c, c->io_in.p, n, stop, c->io_out.p, 1);
    if (rc) return rc;
    if (n) HIPCHK(c, hipMemcpy(out, c->io_out.p, (uint64_t)MPF_NOUT * n * 8, hipMemcpyDeviceToHost));
    if (total_steps) *total_steps = c->last_steps;
    return 0;
}

}  // extern "C"
