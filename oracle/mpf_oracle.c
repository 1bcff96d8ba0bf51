/* mpf oracle: CPU restatement of the reference photon-propagation path.
 *
 * TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this, as the checker or as the timed CPU
 * baseline.  The product library never links or calls it.
 *
 * Compiled with -ffp-contract=off so that every floating-point expression is
 * evaluated operation by operation, in the order the reference writes it.
 *
 * Pinning (what each function is checked against, tests/test_oracle.py):
 *   key_compute / cipher / level / centre / parent / containing
 *       - the reference's own index class compiled in oracle/_ref
 *         (bit-exact on random and edge points, incl. out-of-box points);
 *       - known-answer keys printed by the reference headers
 *         (cipher(0,0,0) = 0xf888..., compute(9,(0.3,-1.1,1.7)) = 0xdcebec9be0...);
 *       - closed forms: centre = -2 + 4(2k+1)/2^(l+1), extent 4/2^l,
 *         ancestor == key prefix, key order == depth-first pre-order.
 *   locate / ngp / cic / tsc (cached and uncached)
 *       - oracle/_ref on AMR trees with holes and level jumps (bit-exact);
 *       - golden fixtures tests/golden/*.npz written from oracle/_ref;
 *       - closed forms: a field linear in x is reproduced exactly by CIC and
 *         TSC away from level jumps; weights sum to one.
 *   integrate / trajectory (all stop criteria)
 *       - oracle/_ref and golden fixtures (bit-exact per-ray records);
 *       - invariants: phi == 0 gives straight rays and z = a0/a - 1;
 *         null condition error() ~ 0.
 *   jacobian_infinitesimal
 *       - oracle/_ref and golden fixtures (bit-exact: same libm);
 *       - closed form: phi linear in x gives A = identity.
 *   interpolate
 *       - oracle/_ref (bit-exact), linear tables reproduce exactly.
 */
#define _GNU_SOURCE
#include "mpf_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

#define EXTENT 4.0
/* marker bit of every nibble: the index holds 4 bits per level from the MSB,
 * nibble = 8 | z<<2 | y<<1 | x  (simplehyperoctreeindex.h:1167-1189) */
static const u128 MARK = (((u128)0x8888888888888888ULL) << 64) | (u128)0x8888888888888888ULL;
/* invalid index = block(0, Bits-1) (simplehyperoctreeindex.h:1358-1361) */
static const u128 INVALID = (((u128)0x7fffffffffffffffULL) << 64) | (u128)0xffffffffffffffffULL;

static const double C_LIGHT = 299792458.0;          /* constants.h:441-444 */
static const double C2 = 299792458.0 * 299792458.0; /* constants.h:481-484 */

/* rho, phi, dphidx, dphidy, dphidz, a, dphidt (gravity.h:45-87 tuple
 * order), then the velocity of the velocity-field build (gravity2.h:45-90),
 * zero when the octree carries none */
#define NFD 10
typedef struct {
    double v[NFD];
} fdata;
enum { F_RHO, F_PHI, F_DX, F_DY, F_DZ, F_A, F_DT, F_VX, F_VY, F_VZ };

struct mpfo_octree {
    uint64_t n;
    u128* key;
    float* f32;  /* n x 7 */
    double* f64; /* n x 7, optional */
    float* vel;  /* n x 3 (units of c), optional: the velocity-field build */
    double vobs[3];
};

/* ------------------------------------------------------------------ keys */

/* Per-axis integer coordinate at level l: Type((1/E)*2^l*(x - (-E/2))) then
 * glue keeps the low l bits (simplehyperoctreeindex.h:1167-1170).  The
 * conversion behaves as a signed truncation (checked against oracle/_ref). */
static inline uint64_t axis_coord(double x, unsigned l) {
    const double scale = (1.0 / EXTENT) * (double)((u128)1 << l);
    const double v = scale * (x - (-EXTENT / 2.0));
    const int64_t t = (int64_t)v;
    const uint64_t mask = (l >= 64) ? ~0ULL : ((1ULL << l) - 1ULL);
    return (uint64_t)t & mask;
}

static inline u128 key_from_coords(unsigned l, uint64_t kx, uint64_t ky, uint64_t kz) {
    u128 k = 0;
    for (unsigned i = 0; i < l; ++i) {
        const unsigned b = l - 1 - i;
        const unsigned nib = 8u | (unsigned)(((kz >> b) & 1u) << 2) | (unsigned)(((ky >> b) & 1u) << 1) |
                             (unsigned)((kx >> b) & 1u);
        k |= ((u128)nib) << (124 - 4 * i);
    }
    return k;
}

/* compute(l, x, y, z) (simplehyperoctreeindex.h:1167-1170); l = 32 is cipher (:1186-1189) */
static inline u128 key_compute(unsigned l, double x, double y, double z) {
    if (l == 0) return 0;
    return key_from_coords(l, axis_coord(x, l), axis_coord(y, l), axis_coord(z, l));
}

/* level() (simplehyperoctreeindex.h:572-575): trailing zeros over the marker bits */
static inline unsigned key_level(u128 k) {
    const u128 m = k & MARK;
    unsigned tz = 128;
    if (m) {
        const uint64_t lo = (uint64_t)m, hi = (uint64_t)(m >> 64);
        tz = lo ? (unsigned)__builtin_ctzll(lo) : 64u + (unsigned)__builtin_ctzll(hi);
    }
    return (128u - tz) / 4u + (k != 0);
}

/* axis integer coordinate of a key at its level (pext over the axis bits) */
static inline uint64_t key_axis(u128 k, unsigned l, unsigned d) {
    uint64_t r = 0;
    for (unsigned i = 0; i < l; ++i) r = (r << 1) | (uint64_t)((k >> (124 - 4 * i + d)) & 1u);
    return r;
}

/* center(d) (simplehyperoctreeindex.h:890-893): -E/2 + E*((2k+1)/2^(l+1)) */
static inline double key_center(u128 k, unsigned d) {
    const unsigned l = key_level(k);
    const uint64_t a = key_axis(k, l, d);
    return (-EXTENT / 2.0) + (EXTENT * ((double)(2 * (u128)a + 1) / (double)((u128)1 << (l + 1))));
}

/* extent() with the default unit extent (simplehyperoctreeindex.h:952-955) */
static inline double key_extent1(u128 k) { return 1.0 / (1.0 * (double)((u128)1 << key_level(k))); }

/* parent() (simplehyperoctreeindex.h:679-682): clear the last nibble */
static inline u128 key_parent(u128 k) {
    const unsigned l = key_level(k);
    if (l == 0) return k;
    const unsigned sh = 128 - 4 * (l - 1);
    const u128 keep = (sh >= 128) ? 0 : ~((((u128)1) << sh) - 1);
    return k & keep;
}

/* child(isite) (simplehyperoctreeindex.h:690-694) */
static inline u128 key_child(u128 k, unsigned isite) {
    const unsigned l = key_level(k);
    if (l >= 32) return k;
    return k | (((u128)((isite & 7u) | 8u)) << (128 - 4 * (l + 1)));
}

/* containing() (simplehyperoctreeindex.h:998-1001) */
static inline int key_containing(u128 a, u128 b) {
    const u128 x = a ^ b;
    return ((~MARK & x) <= (MARK & x)) && (a <= b);
}

/* ---------------------------------------------------------------- octree */

/* std::upper_bound on the key-sorted container (simplehyperoctree.h:894-899) */
static inline uint64_t upper_bound(const mpfo_octree* o, u128 q) {
    uint64_t lo = 0, n = o->n;
    while (n > 0) {
        const uint64_t half = n / 2;
        if (!(q < o->key[lo + half])) { lo = lo + half + 1; n = n - half - 1; }
        else n = half;
    }
    return lo;
}

static inline void load_data(const mpfo_octree* o, int prec, uint64_t i, fdata* d) {
    if (prec == 1 && o->f64) {
        for (int k = 0; k < 7; ++k) d->v[k] = o->f64[7 * i + k];
    } else {
        for (int k = 0; k < 7; ++k) d->v[k] = (double)o->f32[7 * i + k];
    }
    for (int k = 0; k < 3; ++k) d->v[F_VX + k] = o->vel ? (double)o->vel[3 * i + k] : 0.0;
}

static inline void zero_data(fdata* d) { memset(d, 0, sizeof(*d)); }
static inline int data_empty(const fdata* d) {
    for (int k = 0; k < NFD; ++k) if (d->v[k] != 0.0) return 0;
    return 1;
}

/* mac() (simplehyperoctree.h:1870-1910): accumulator += factor*coefficient.
 * precision 0: accumulator is float, so (float)((double)acc + (double)f*w). */
static inline void mac(fdata* acc, const fdata* f, double w, int prec) {
    if (prec == 0) {
        for (int k = 0; k < NFD; ++k) acc->v[k] = (double)(float)(acc->v[k] + f->v[k] * w);
    } else {
        for (int k = 0; k < NFD; ++k) acc->v[k] = acc->v[k] + f->v[k] * w;
    }
}

/* locate() (simplehyperoctree.h:894-899): returns -1 for end() */
static int64_t locate(const mpfo_octree* o, double x, double y, double z) {
    const u128 q = key_compute(32, x, y, z);
    const uint64_t m = upper_bound(o, q);
    const uint64_t c = m - (m > 0);
    return key_containing(o->key[c], q) ? (int64_t)c : -1;
}

/* ngp() (simplehyperoctree.h:1396-1401) */
static void ngp(const mpfo_octree* o, int prec, double x, double y, double z, fdata* r) {
    const u128 q = key_compute(32, x, y, z);
    const uint64_t m = upper_bound(o, q);
    if (key_containing(o->key[m - (m > 0)], q)) load_data(o, prec, m - 1, r);
    else zero_data(r);
}

/* cic() (simplehyperoctree.h:1412-1469) */
static void cic(const mpfo_octree* o, int prec, double x, double y, double z, fdata* result) {
    const double one = 1.0, two = 2.0;
    const double point[3] = {x, y, z};
    u128 ekey = key_compute(32, x, y, z);
    uint64_t marker = upper_bound(o, ekey);
    u128 idx = key_containing(o->key[marker - (marker > 0)], ekey) ? o->key[marker - 1] : INVALID;
    unsigned ilvl = key_level(idx);
    unsigned lvl = ilvl;
    double half = EXTENT / 1.0 * key_extent1(idx) / two;
    double vertex[3], dist[3];
    unsigned site = 0;
    fdata cell;
    zero_data(result);
    do {
        ilvl = lvl;
        if (idx != INVALID) {
            for (unsigned d = 0; d < 3; ++d) dist[d] = fabs((key_center(idx, d) - point[d]) / (half + half));
            for (unsigned isite = 0; isite < 8; ++isite) {
                if (key_containing(key_child(idx, isite), ekey)) { site = isite; break; }
            }
            for (unsigned iv = 0; iv < 8; ++iv) {
                double weight = one;
                const unsigned bits = iv, diff = site ^ bits;
                for (unsigned d = 0; d < 3; ++d) {
                    const int b = (bits >> d) & 1u, df = (diff >> d) & 1u;
                    vertex[d] = point[d] + (b * half) - ((!b) * half);
                    weight *= ((df) * (one - dist[d])) + ((!df) * dist[d]);
                }
                ekey = key_compute(ilvl, vertex[0], vertex[1], vertex[2]);
                marker = upper_bound(o, ekey);
                if (key_containing(o->key[marker - (marker > 0)], ekey)) {
                    const unsigned fl = key_level(o->key[marker - 1]);
                    lvl = lvl < fl ? lvl : fl;
                    load_data(o, prec, marker - 1, &cell);
                    mac(result, &cell, weight, prec);
                } else {
                    zero_data(result);
                    break;
                }
            }
        }
        if (lvl < ilvl) {
            ekey = key_compute(lvl + 1, x, y, z);
            idx = key_parent(ekey);
            zero_data(result);
            half = EXTENT / 1.0 * key_extent1(idx) / two;
        }
    } while (lvl < ilvl);
}

/* the 27-cell cache of tsc(elemsTsc, ...) (simplehyperoctree.h:1554-1644) */
typedef struct {
    u128 key[27];
    fdata data[27];
} tsc_cache;

static void tsc_cache_init(tsc_cache* c) { memset(c, 0, sizeof(*c)); }

static inline double tsc_weight(int aix, int aiy, int aiz, const double* dist) {
    return (aix * 0.5 * (1.5 - dist[0]) * (1.5 - dist[0]) + (1 - aix) * (0.75 - dist[0] * dist[0])) *
           (aiy * 0.5 * (1.5 - dist[1]) * (1.5 - dist[1]) + (1 - aiy) * (0.75 - dist[1] * dist[1])) *
           (aiz * 0.5 * (1.5 - dist[2]) * (1.5 - dist[2]) + (1 - aiz) * (0.75 - dist[2] * dist[2]));
}

/* tsc() with (cache != NULL, simplehyperoctree.h:1554-1644) or without
 * (cache == NULL, simplehyperoctree.h:1481-1541) the 27-cell cache. */
static void tsc(const mpfo_octree* o, int prec, tsc_cache* cache, double x, double y, double z, fdata* result) {
    const double two = 2.0;
    const double point[3] = {x, y, z};
    if (point[0] == 0.0 && point[1] == 0.0 && point[2] == 0.0) {
        cic(o, prec, 0., 0., 0., result);
        return;
    }
    u128 ekey = key_compute(32, x, y, z);
    uint64_t marker = upper_bound(o, ekey);
    u128 idx = key_containing(o->key[marker - (marker > 0)], ekey) ? o->key[marker - 1] : INVALID;
    unsigned ilvl = key_level(idx);
    unsigned lvl = ilvl;
    double half = EXTENT / 1.0 * key_extent1(idx) / two;
    double dist[3];
    tsc_cache tmp;
    fdata cell;
    zero_data(result);
    do {
        ilvl = lvl;
        const double twohalves = two * half;
        if (cache && cache->key[13] == idx) {
            unsigned ic = 0;
            for (int ix = -1; ix <= 1; ++ix)
                for (int iy = -1; iy <= 1; ++iy)
                    for (int iz = -1; iz <= 1; ++iz) {
                        for (unsigned d = 0; d < 3; ++d)
                            dist[d] = fabs((key_center(cache->key[ic], d) - point[d]) / (twohalves));
                        const double w = tsc_weight(abs(ix), abs(iy), abs(iz), dist);
                        mac(result, &cache->data[ic], w, prec);
                        ic++;
                    }
            return;
        }
        if (idx != INVALID) {
            unsigned ic = 0;
            for (int ix = -1; ix <= 1; ++ix) {
                for (int iy = -1; iy <= 1; ++iy) {
                    for (int iz = -1; iz <= 1; ++iz) {
                        ekey = key_compute(ilvl, point[0] + ix * twohalves, point[1] + iy * twohalves,
                                           point[2] + iz * twohalves);
                        marker = upper_bound(o, ekey);
                        if (key_containing(o->key[marker - (marker > 0)], ekey)) {
                            const uint64_t found = marker - (marker > 0);
                            for (unsigned d = 0; d < 3; ++d)
                                dist[d] = fabs((key_center(ekey, d) - point[d]) / (twohalves));
                            const double w = tsc_weight(abs(ix), abs(iy), abs(iz), dist);
                            const unsigned fl = key_level(o->key[found]);
                            lvl = lvl < fl ? lvl : fl;
                            if (lvl < ilvl) {
                                zero_data(result);
                                goto done;
                            }
                            load_data(o, prec, found, &cell);
                            if (cache) { tmp.key[ic] = o->key[found]; tmp.data[ic] = cell; }
                            mac(result, &cell, w, prec);
                            ic++;
                        } else {
                            zero_data(result);
                            goto done;
                        }
                    }
                }
            }
        done:
            if (cache && ic == 27) *cache = tmp;
        }
        if (lvl < ilvl) {
            ekey = key_compute(lvl + 1, x, y, z);
            idx = key_parent(ekey);
            zero_data(result);
            half = EXTENT / 1.0 * key_extent1(idx) / two;
        }
    } while (lvl < ilvl);
}

/* ------------------------------------------------------------ cosmology */

/* Utility::interpolate (utility.h:442-449): upper_bound, clamp to [1, n-1],
 * linear interpolation/extrapolation. */
static inline double interpolate(double x0, const double* x, const double* y, int64_t n) {
    if (n <= 0) return 0.0;
    if (n == 1) return y[0];
    int64_t lo = 0, cnt = n;
    while (cnt > 0) {
        const int64_t h = cnt / 2;
        if (!(x0 < x[lo + h])) { lo = lo + h + 1; cnt = cnt - h - 1; }
        else cnt = h;
    }
    int64_t i = lo;
    i += (i <= 0) - (i >= n);
    return y[i - 1] + ((y[i] - y[i - 1]) * (x0 - x[i - 1]) / (x[i] - x[i - 1]));
}

/* ----------------------------------------------------------- integration */

typedef struct {
    double core[9]; /* a, t, x, y, z, dtdl, dxdl, dydl, dzdl (photon.h:44-47) */
    double level, ah, rho, phi, dphidx, dphidy, dphidz, dphidl, laplacian, redshift, dsdl2, error, distance;
    double isw, iswold, chi, lambda, dphidt, s;
} photon;
enum { CA, CT, CX, CY, CZ, CKT, CKX, CKY, CKZ };

typedef struct {
    const mpfo_octree* o;
    int order, prec;
    const double *ct, *ca, *cdadt;
    int64_t ncos;
    double length;
    unsigned nsteps;
    tsc_cache cache;
} integ_ctx;

static void interp_order(integ_ctx* g, int order, double x, double y, double z, fdata* d) {
    switch (order) {
        case 0: ngp(g->o, g->prec, x, y, z, d); break;
        case 1: cic(g->o, g->prec, x, y, z, d); break;
        case 2: tsc(g->o, g->prec, &g->cache, x, y, z, d); break;
        default: zero_data(d); break;
    }
}

/* Integrator::dphotondl (integrator.h:409-445) */
static void dphotondl(integ_ctx* g, double* out, const double* in) {
    const double two = 2.0;
    fdata data;
    interp_order(g, g->order, in[CX], in[CY], in[CZ], &data);
    const double dphidl = data.v[F_DT] * in[CKT] + (data.v[F_DX] * in[CKX] + data.v[F_DY] * in[CKY] + data.v[F_DZ] * in[CKZ]);
    const double dadt = interpolate(in[CT], g->ct, g->cdadt, g->ncos);
    const double scale = g->length;
    out[CA] = in[CKT] * dadt;
    out[CT] = in[CKT];
    out[CX] = in[CKX] / scale;
    out[CY] = in[CKY] / scale;
    out[CZ] = in[CKZ] / scale;
    out[CKT] = -(two * dadt / in[CA] * in[CKT] * in[CKT]) - (two / C2 * in[CKT]) * (dphidl - data.v[F_DT] * in[CKT]);
    out[CKX] = -(two * dadt / in[CA] * in[CKT] * in[CKX]) + (two / C2 * dphidl * in[CKX]) - (two * data.v[F_DX] * in[CKT] * in[CKT]);
    out[CKY] = -(two * dadt / in[CA] * in[CKT] * in[CKY]) + (two / C2 * dphidl * in[CKY]) - (two * data.v[F_DY] * in[CKT] * in[CKT]);
    out[CKZ] = -(two * dadt / in[CA] * in[CKT] * in[CKZ]) + (two / C2 * dphidl * in[CKZ]) - (two * data.v[F_DZ] * in[CKT] * in[CKT]);
}

static void photon_to_rec(const photon* p, double* o, double nsteps) {
    for (int k = 0; k < 9; ++k) o[k] = p->core[k];
    o[MPFO_REDSHIFT] = p->redshift; o[MPFO_LAMBDA] = p->lambda; o[MPFO_S] = p->s; o[MPFO_CHI] = p->chi;
    o[MPFO_ISW] = p->isw; o[MPFO_ISWOLD] = p->iswold; o[MPFO_PHI] = p->phi; o[MPFO_DPHIDX] = p->dphidx;
    o[MPFO_DPHIDY] = p->dphidy; o[MPFO_DPHIDZ] = p->dphidz; o[MPFO_DPHIDT] = p->dphidt; o[MPFO_RHO] = p->rho;
    o[MPFO_AH] = p->ah; o[MPFO_DPHIDL] = p->dphidl; o[MPFO_LAPLACIAN] = p->laplacian; o[MPFO_DSDL2] = p->dsdl2;
    o[MPFO_ERROR] = p->error; o[MPFO_LEVEL] = p->level; o[MPFO_NSTEPS] = nsteps;
}

/* growable trajectory (Evolution<Photon>, evolution.h:51, append :908, pop :894) */
typedef struct {
    photon* p;
    uint64_t n, cap;
    int keep; /* keep all elements (1) or only the last two (0) */
} traj_t;

static void traj_append(traj_t* t, const photon* p) {
    if (!t->keep) {
        if (t->n >= 2) { t->p[0] = t->p[1]; t->p[1] = *p; }
        else t->p[t->n] = *p;
        t->n++;
        return;
    }
    if (t->n == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 1024;
        t->p = (photon*)realloc(t->p, t->cap * sizeof(photon));
    }
    t->p[t->n++] = *p;
}
static inline photon* traj_back(traj_t* t) { return t->keep ? &t->p[t->n - 1] : &t->p[t->n >= 2 ? 1 : 0]; }
static inline photon* traj_prev(traj_t* t) { return t->keep ? &t->p[t->n - 2] : &t->p[0]; }

static double locate_level(const mpfo_octree* o, double x, double y, double z) {
    const int64_t e = locate(o, x, y, z);
    return e < 0 ? 0.0 : (double)key_level(o->key[e]);
}
static double locate_extent(const mpfo_octree* o, double x, double y, double z) {
    const int64_t e = locate(o, x, y, z);
    const unsigned l = e < 0 ? 0u : key_level(o->key[e]);
    return EXTENT / (1.0 * (double)((u128)1 << l));
}

/* Integrator::integrate, no stop criterion (integrator.h:477-618) when
 * stop_kind < 0, else with the criterion of integrator.h:648-824:
 * 0 radius, 1 lambda, 2 redshift, 3 t, 4 a, 5 plane, 6 s. */
static void integrate(integ_ctx* g, traj_t* traj, int stop_kind, double ref, const double* kit) {
    const double zero = 0.0, one = 1.0, two = 2.0, six = 6.0;
    const double c = C_LIGHT, c2 = C2;
    const double min = 0.0 - (EXTENT / two), max = 0.0 + (EXTENT / two);
    const double scale = g->length;
    fdata data, empty;
    photon ph;
    double dcoredl[4][9];
    zero_data(&empty);
    memset(&ph, 0, sizeof(ph));
    if (traj->n == 0) return;
    tsc_cache_init(&g->cache);
    {
        photon* b = traj_back(traj);
        if (g->order >= 0) interp_order(g, g->order, b->core[CX], b->core[CY], b->core[CZ], &data);
        else { zero_data(&data); data.v[F_A] = one; }
        const double norm = sqrt((c2 * (one + two * (data.v[F_PHI] / c2)) * b->core[CKT] * b->core[CKT]) /
                                 ((one - two * (data.v[F_PHI] / c2)) *
                                  (b->core[CKX] * b->core[CKX] + b->core[CKY] * b->core[CKY] + b->core[CKZ] * b->core[CKZ])));
        b->core[CKX] *= norm;
        b->core[CKY] *= norm;
        b->core[CKZ] *= norm;
        b->core[CA] = interpolate(b->core[CT], g->ct, g->ca, g->ncos);
        b->level = locate_level(g->o, b->core[CX], b->core[CY], b->core[CZ]);
        b->ah = zero;
        b->rho = data.v[F_RHO];
        b->phi = data.v[F_PHI];
        b->dphidx = data.v[F_DX];
        b->dphidy = data.v[F_DY];
        b->dphidz = data.v[F_DZ];
        b->dphidt = data.v[F_DT];
        b->dphidl = zero;
        b->laplacian = zero;
        b->redshift = zero;
        b->dsdl2 = (b->core[CA] * b->core[CA]) *
                   (-(c2 * (one + two * (b->phi / c2)) * b->core[CKT] * b->core[CKT]) +
                    ((one - two * (b->phi / c2)) *
                     (b->core[CKX] * b->core[CKX] + b->core[CKY] * b->core[CKY] + b->core[CKZ] * b->core[CKZ])));
        b->error = one - ((one - two * (b->phi / c2)) *
                          (b->core[CKX] * b->core[CKX] + b->core[CKY] * b->core[CKY] + b->core[CKZ] * b->core[CKZ])) /
                             (c2 * (one + two * (b->phi / c2)) * b->core[CKT] * b->core[CKT]);
        b->distance = zero;
        b->isw = zero;
        b->chi = zero;
        b->iswold = zero;
        b->lambda = zero;
        b->s = zero;
    }
    photon* b0 = traj_back(traj);
    double ratio = b0->core[CA] * b0->core[CA] * (scale / c) / g->nsteps;
    double dl = locate_extent(g->o, b0->core[CX], b0->core[CY], b0->core[CZ]) * ratio;
    const double a0 = b0->core[CA];
    /* the observer's velocity along the first direction, velocity-field
     * build only (integrator.h:537-539, 707-709) */
    const double* vobs = g->o->vobs;
    const double v0n = (vobs[0] * b0->core[CKX] + vobs[1] * b0->core[CKY] + vobs[2] * b0->core[CKZ]) / (c * b0->core[CKT]);
    const double phi0c2 = b0->phi / c2;
    const int order_step = g->order;
    while (!data_empty(&data)) {
        photon* back = traj_back(traj);
        dphotondl(g, dcoredl[0], back->core);
        for (int i = 0; i < 9; ++i) ph.core[i] = back->core[i] + dl / two * dcoredl[0][i];
        dphotondl(g, dcoredl[1], ph.core);
        for (int i = 0; i < 9; ++i) ph.core[i] = back->core[i] + dl / two * dcoredl[1][i];
        dphotondl(g, dcoredl[2], ph.core);
        for (int i = 0; i < 9; ++i) ph.core[i] = back->core[i] + dl * dcoredl[2][i];
        dphotondl(g, dcoredl[3], ph.core);
        for (int i = 0; i < 9; ++i)
            ph.core[i] = back->core[i] + (dl / six) * (dcoredl[0][i] + two * dcoredl[1][i] + two * dcoredl[2][i] + dcoredl[3][i]);
        if (order_step >= 0) interp_order(g, order_step, ph.core[CX], ph.core[CY], ph.core[CZ], &data);
        else { zero_data(&data); data.v[F_A] = one; }
        if (!(!(ph.core[CA] < zero) && ((ph.core[CX] > min) && (ph.core[CX] < max) && (ph.core[CY] > min) &&
                                        (ph.core[CY] < max) && (ph.core[CZ] > min) && (ph.core[CZ] < max))))
            data = empty;
        ph.level = locate_level(g->o, ph.core[CX], ph.core[CY], ph.core[CZ]);
        ph.ah = data.v[F_A];
        ph.rho = data.v[F_RHO];
        ph.phi = data.v[F_PHI];
        ph.dphidx = data.v[F_DX];
        ph.dphidy = data.v[F_DY];
        ph.dphidz = data.v[F_DZ];
        ph.dphidt = -data.v[F_DT];
        ph.dphidl = data.v[F_DT] * ph.core[CKT] + (ph.dphidx * ph.core[CKX] + ph.dphidy * ph.core[CKY] + ph.dphidz * ph.core[CKZ]);
        ph.laplacian = (data.v[F_PHI] - back->phi) / dl;
        ph.dsdl2 = (ph.core[CA] * ph.core[CA]) *
                   ((-c2 * (one + two * (ph.phi / c2)) * ph.core[CKT] * ph.core[CKT]) +
                    ((one - two * (ph.phi / c2)) *
                     (ph.core[CKX] * ph.core[CKX] + ph.core[CKY] * ph.core[CKY] + ph.core[CKZ] * ph.core[CKZ])));
        ph.error = one - ((one - two * (ph.phi / c2)) *
                          (ph.core[CKX] * ph.core[CKX] + ph.core[CKY] * ph.core[CKY] + ph.core[CKZ] * ph.core[CKZ])) /
                             (c2 * (one + two * (ph.phi / c2)) * ph.core[CKT] * ph.core[CKT]);
        ph.distance = zero;
        ph.isw += (back->dphidt + ph.dphidt) * (ph.core[CT] - back->core[CT]) / c2;
        ph.iswold -= 2 * (ph.phi - back->phi) / c2 -
                     ((back->dphidx + ph.dphidx) * (ph.core[CX] - back->core[CX]) * scale +
                      (back->dphidy + ph.dphidy) * (ph.core[CY] - back->core[CY]) * scale +
                      (back->dphidz + ph.dphidz) * (ph.core[CZ] - back->core[CZ]) * scale) /
                         c2;
        ph.chi = sqrt(ph.core[CX] * ph.core[CX] + ph.core[CY] * ph.core[CY] + ph.core[CZ] * ph.core[CZ]);
        ph.lambda += dl;
        ph.s += sqrt((ph.core[CX] - back->core[CX]) * (ph.core[CX] - back->core[CX]) +
                     (ph.core[CY] - back->core[CY]) * (ph.core[CY] - back->core[CY]) +
                     (ph.core[CZ] - back->core[CZ]) * (ph.core[CZ] - back->core[CZ]));
        if (g->o->vel) /* Doppler and ISW terms (integrator.h:594-595, 765-766) */
            ph.redshift = a0 / ph.core[CA] *
                              (one + phi0c2 - ph.phi / c2 + ph.isw +
                               (data.v[F_VX] * ph.core[CKX] + data.v[F_VY] * ph.core[CKY] + data.v[F_VZ] * ph.core[CKZ]) /
                                   (c * ph.core[CKT]) -
                               v0n) -
                          one;
        else ph.redshift = a0 / ph.core[CA] * (one + phi0c2 - ph.phi / c2) - one;
        if (!data_empty(&data)) {
            ratio = ph.core[CA] * ph.core[CA] * (scale / c) / g->nsteps;
            dl = locate_extent(g->o, ph.core[CX], ph.core[CY], ph.core[CZ]) * ratio;
            traj_append(traj, &ph);
        }
        if (stop_kind >= 0) {
            const photon* bk = traj_back(traj);
            int brk = 0;
            switch (stop_kind) {
                case 0: brk = bk->chi > ref; break;
                case 1: brk = bk->lambda > ref; break;
                case 2: brk = bk->redshift > ref; break;
                case 3: brk = bk->core[CT] > ref; break;
                case 4: brk = bk->core[CA] < ref; break;
                case 5: brk = (ph.core[CX] * kit[0] + ph.core[CY] * kit[1] + ph.core[CZ] * kit[2]) > ref; break;
                case 6: brk = bk->s > ref; break;
                default: brk = 1; break;
            }
            if (brk) break;
        }
    }
    if (traj->n > 0) {
        const photon* bk = traj_back(traj);
        if (signbit(bk->redshift) || signbit(bk->core[CA])) {
            if (traj->keep) traj->n--;
            else {
                /* only the last two are kept: the popped state is the previous one */
                if (traj->n >= 2) traj->p[1] = traj->p[0];
                traj->n--;
            }
        }
    }
}

static void launch(traj_t* t, const double* in9) {
    photon p;
    memset(&p, 0, sizeof(p));
    for (int k = 0; k < 9; ++k) p.core[k] = in9[k];
    traj_append(t, &p);
}

static void ctx_init(integ_ctx* g, const mpfo_octree* o, int order, int prec, const double* tab, uint64_t ncos,
                     double length, unsigned nsteps) {
    g->o = o;
    g->order = order;
    g->prec = prec;
    g->ct = tab;
    g->ca = tab + ncos;
    g->cdadt = tab + 2 * ncos;
    g->ncos = (int64_t)ncos;
    g->length = length;
    g->nsteps = nsteps;
    tsc_cache_init(&g->cache);
}

/* --------------------------------------------------------------- public */

mpfo_octree* mpfo_octree_new(const void* rec48, uint64_t n) {
    mpfo_octree* o = (mpfo_octree*)calloc(1, sizeof(*o));
    o->n = n;
    o->key = (u128*)malloc((n ? n : 1) * sizeof(u128));
    o->f32 = (float*)malloc((n ? n : 1) * 7 * sizeof(float));
    const unsigned char* r = (const unsigned char*)rec48;
    for (uint64_t i = 0; i < n; ++i) {
        const unsigned char* e = r + 48 * i;
        uint64_t lo, hi;
        memcpy(&lo, e, 8);
        memcpy(&hi, e + 8, 8);
        o->key[i] = (((u128)hi) << 64) | lo;
        /* record layout (fact: libstdc++ stores the tuple in reverse):
         * dphidt@16 a@20 dphidx@24 dphidy@28 dphidz@32 phi@36 rho@40 */
        float f[7];
        memcpy(f, e + 16, 28);
        float* d = o->f32 + 7 * i;
        d[F_DT] = f[0]; d[F_A] = f[1]; d[F_DX] = f[2]; d[F_DY] = f[3]; d[F_DZ] = f[4]; d[F_PHI] = f[5]; d[F_RHO] = f[6];
    }
    return o;
}

int mpfo_octree_set_fields64(mpfo_octree* o, const double* f7, uint64_t n) {
    if (n != o->n) return -1;
    free(o->f64);
    o->f64 = (double*)malloc((n ? n : 1) * 7 * sizeof(double));
    memcpy(o->f64, f7, n * 7 * sizeof(double));
    return 0;
}

void mpfo_octree_free(mpfo_octree* o) {
    if (!o) return;
    free(o->key);
    free(o->f32);
    free(o->f64);
    free(o->vel);
    free(o);
}

/* The velocity field of the velocity-field build: the octree's rho replaced
 * (rho, when given) and a velocity per record (vel3, units of c), with the
 * observer's velocity vobs3; vel3 = NULL removes the field. */
int mpfo_octree_set_velocity(mpfo_octree* o, const float* rho, const float* vel3, uint64_t n, const double* vobs3) {
    if (n != o->n) return -1;
    free(o->vel);
    o->vel = NULL;
    o->vobs[0] = o->vobs[1] = o->vobs[2] = 0.0;
    if (rho)
        for (uint64_t i = 0; i < n; ++i) {
            o->f32[7 * i + F_RHO] = rho[i];
            if (o->f64) o->f64[7 * i + F_RHO] = (double)rho[i];
        }
    if (vel3) {
        o->vel = (float*)malloc((n ? n : 1) * 3 * sizeof(float));
        memcpy(o->vel, vel3, n * 3 * sizeof(float));
        if (vobs3) for (int k = 0; k < 3; ++k) o->vobs[k] = vobs3[k];
    }
    return 0;
}

/* Hmaps::CreateOctreeVelocityWithCIC / WithTSC (hmaps.h:439-478, 497-541)
 * for one particle set, then the normalisation by the deposited weight and
 * the parent fill of cells without weight (hmaps.cpp:389-441).  Particles:
 * positions in box units and velocities (units of c), 3 floats each.  rho and vel3
 * (record order) hold the running sums on entry when accumulate != 0 (the
 * sums of earlier particle files, before normalisation) and are cleared
 * otherwise (rho = 0 in every cell, hmaps.cpp:391).  finish != 0 normalises
 * and fills from the parents.  One thread, particles in order within each
 * level and the levels in increasing order: the reference's float
 * accumulation run sequentially.  scheme: 1 CIC, 2 TSC.  Returns -1 on a
 * bad scheme. */
int mpfo_velocity_deposit(const mpfo_octree* o, int scheme, const float* pos, const float* vel, uint64_t np,
                          int accumulate, int finish, float* rho, float* vel3) {
    if (scheme != 1 && scheme != 2) return -1;
    const uint64_t n = o->n;
    if (!accumulate) {
        memset(rho, 0, n * sizeof(float));
        memset(vel3, 0, 3 * n * sizeof(float));
    }
    if (n == 0) return 0;
    unsigned lmin = 64, lmax = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const unsigned l = key_level(o->key[i]);
        if (l < lmin) lmin = l;
        if (l > lmax) lmax = l;
    }
    for (unsigned ilvl = lmin; ilvl <= lmax; ++ilvl) {
        const double half = 0.5 * EXTENT / pow(2, ilvl);
        const double invextension = 1 / (EXTENT / pow(2, ilvl));
        const double twohalves = 2 * half;
        for (uint64_t i = 0; i < np; ++i) {
            const double p[3] = {(double)pos[3 * i], (double)pos[3 * i + 1], (double)pos[3 * i + 2]};
            /* CIC: the 8 cells at +-half a cell; TSC: the 27 at 0, +-one cell */
            const int step = (scheme == 1) ? 2 : 1;
            const double off = (scheme == 1) ? half : twohalves;
            for (int iz = -1; iz <= 1; iz += step)
                for (int iy = -1; iy <= 1; iy += step)
                    for (int ix = -1; ix <= 1; ix += step) {
                        const u128 k = key_compute(ilvl, p[0] + off * ix, p[1] + off * iy, p[2] + off * iz);
                        const uint64_t m = upper_bound(o, k);
                        const uint64_t c = m - (m > 0);
                        if (o->key[c] != k) continue;
                        double vratio;
                        if (scheme == 1) {
                            vratio = (1 - fabs(p[0] - key_center(k, 0)) * invextension) *
                                     (1 - fabs(p[1] - key_center(k, 1)) * invextension) *
                                     (1 - fabs(p[2] - key_center(k, 2)) * invextension);
                        } else {
                            const int ai[3] = {abs(ix), abs(iy), abs(iz)};
                            double w[3];
                            for (int d = 0; d < 3; ++d) {
                                const double dist = fabs((key_center(o->key[c], d) - p[d]) / (twohalves));
                                w[d] = ai[d] * 0.5 * (1.5 - dist) * (1.5 - dist) + (1 - ai[d]) * (0.75 - dist * dist);
                            }
                            vratio = w[0] * w[1] * w[2];
                        }
                        rho[c] += vratio;
                        for (int d = 0; d < 3; ++d) vel3[3 * c + d] += vel[3 * i + d] * vratio;
                    }
        }
    }
    if (!finish) return 0;
    for (uint64_t i = 0; i < n; ++i) {
        double mass = rho[i];
        mass = mass + (mass == 0);
        for (int d = 0; d < 3; ++d) vel3[3 * i + d] /= mass;
    }
    for (unsigned ilvl = lmin + 1; ilvl <= lmax; ++ilvl)
        for (uint64_t i = 0; i < n; ++i) {
            if (key_level(o->key[i]) != ilvl || isnormal(rho[i])) continue;
            const u128 pk = key_parent(o->key[i]);
            const uint64_t m = upper_bound(o, pk);
            const uint64_t c = m - (m > 0);
            if (!key_containing(o->key[c], pk)) continue; /* find() == end(): not on a parent-closed octree */
            for (int d = 0; d < 3; ++d) vel3[3 * i + d] = vel3[3 * c + d];
        }
    return 0;
}

/* order 0 ngp, 1 cic, 2 uncached tsc: 10 values per point (the 7 fields,
 * then the velocity) */
void mpfo_interp10(const mpfo_octree* o, int order, int precision, const double* xyz, uint64_t m, double* out10) {
    for (uint64_t i = 0; i < m; ++i) {
        fdata d;
        const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        if (order == 0) ngp(o, precision, x, y, z, &d);
        else if (order == 1) cic(o, precision, x, y, z, &d);
        else tsc(o, precision, NULL, x, y, z, &d);
        for (int k = 0; k < NFD; ++k) out10[NFD * i + k] = d.v[k];
    }
}

void mpfo_compute_key(unsigned level, const double* xyz, uint64_t m, uint64_t* lohi) {
    for (uint64_t i = 0; i < m; ++i) {
        const u128 k = key_compute(level > 32 ? 32 : level, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        lohi[2 * i] = (uint64_t)k;
        lohi[2 * i + 1] = (uint64_t)(k >> 64);
    }
}

void mpfo_key_info(const uint64_t* lohi, uint64_t m, uint32_t* level, double* center3, double* ext, uint64_t* parent) {
    for (uint64_t i = 0; i < m; ++i) {
        const u128 k = (((u128)lohi[2 * i + 1]) << 64) | lohi[2 * i];
        level[i] = key_level(k);
        for (unsigned d = 0; d < 3; ++d) center3[3 * i + d] = key_center(k, d);
        ext[i] = EXTENT / 1.0 * key_extent1(k);
        const u128 p = key_parent(k);
        parent[2 * i] = (uint64_t)p;
        parent[2 * i + 1] = (uint64_t)(p >> 64);
    }
}

int mpfo_key_containing(const uint64_t* a, const uint64_t* b) {
    const u128 ka = (((u128)a[1]) << 64) | a[0];
    const u128 kb = (((u128)b[1]) << 64) | b[0];
    return key_containing(ka, kb);
}

void mpfo_locate(const mpfo_octree* o, const double* xyz, uint64_t m, int64_t* elem, uint32_t* level) {
    for (uint64_t i = 0; i < m; ++i) {
        const int64_t e = locate(o, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        elem[i] = e;
        level[i] = e < 0 ? 0u : key_level(o->key[e]);
    }
}

void mpfo_interp(const mpfo_octree* o, int order, int precision, const double* xyz, uint64_t m, double* out7) {
    tsc_cache cache;
    tsc_cache_init(&cache);
    for (uint64_t i = 0; i < m; ++i) {
        fdata d;
        const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        if (order == 0) ngp(o, precision, x, y, z, &d);
        else if (order == 1) cic(o, precision, x, y, z, &d);
        else if (order == 2) tsc(o, precision, NULL, x, y, z, &d);
        else tsc(o, precision, &cache, x, y, z, &d);
        for (int k = 0; k < 7; ++k) out7[7 * i + k] = d.v[k];
    }
}

typedef struct {
    const mpfo_octree* o;
    int order, prec, stop_kind;
    const double* tab;
    uint64_t ncos;
    double length, stop_value;
    unsigned nsteps;
    const double* kit;
    const double* in9;
    double* out;
    uint64_t begin, end, steps;
} work_t;

static void* worker(void* arg) {
    work_t* w = (work_t*)arg;
    integ_ctx g;
    photon buf[2];
    ctx_init(&g, w->o, w->order, w->prec, w->tab, w->ncos, w->length, w->nsteps);
    w->steps = 0;
    for (uint64_t i = w->begin; i < w->end; ++i) {
        traj_t t = {buf, 0, 2, 0};
        launch(&t, w->in9 + 9 * i);
        integrate(&g, &t, w->stop_kind, w->stop_value, w->kit);
        double* o = w->out + MPFO_NREC * i;
        if (t.n == 0) {
            memset(o, 0, MPFO_NREC * sizeof(double));
            o[MPFO_NSTEPS] = -1.0;
        } else {
            photon_to_rec(traj_back(&t), o, (double)(t.n - 1));
            w->steps += t.n - 1;
        }
    }
    return NULL;
}

uint64_t mpfo_integrate(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                        double length, unsigned nsteps, int stop_kind, double stop_value, const double* kitarget,
                        const double* in9, uint64_t n, double* out, int nthreads) {
    static const double zk[3] = {0, 0, 0};
    if (nthreads < 1) nthreads = 1;
    if ((uint64_t)nthreads > n) nthreads = n ? (int)n : 1;
    work_t* w = (work_t*)calloc((size_t)nthreads, sizeof(work_t));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    const uint64_t chunk = (n + nthreads - 1) / nthreads;
    for (int k = 0; k < nthreads; ++k) {
        w[k] = (work_t){o, order, precision, stop_kind, cosmo_tab, ncos, length, stop_value, nsteps,
                        kitarget ? kitarget : zk, in9, out, 0, 0, 0};
        w[k].begin = chunk * k < n ? chunk * k : n;
        w[k].end = chunk * (k + 1) < n ? chunk * (k + 1) : n;
    }
    for (int k = 1; k < nthreads; ++k) pthread_create(&th[k], NULL, worker, &w[k]);
    worker(&w[0]);
    uint64_t total = w[0].steps;
    for (int k = 1; k < nthreads; ++k) {
        pthread_join(th[k], NULL);
        total += w[k].steps;
    }
    free(w);
    free(th);
    return total;
}

static void run_full(const mpfo_octree* o, int order, int precision, const double* tab, uint64_t ncos, double length,
                     unsigned nsteps, int stop_kind, double stop_value, const double* kit, const double* in9,
                     traj_t* t) {
    static const double zk[3] = {0, 0, 0};
    integ_ctx g;
    ctx_init(&g, o, order, precision, tab, ncos, length, nsteps);
    t->keep = 1;
    launch(t, in9);
    integrate(&g, t, stop_kind, stop_value, kit ? kit : zk);
}

uint64_t mpfo_trajectory(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                         double length, unsigned nsteps, int stop_kind, double stop_value, const double* kitarget,
                         const double* in9, double* rows, uint64_t maxlen) {
    traj_t t = {NULL, 0, 0, 1};
    run_full(o, order, precision, cosmo_tab, ncos, length, nsteps, stop_kind, stop_value, kitarget, in9, &t);
    for (uint64_t i = 0; i < t.n && i < maxlen; ++i) photon_to_rec(&t.p[i], rows + MPFO_NREC * i, (double)i);
    const uint64_t n = t.n;
    free(t.p);
    return n;
}

/* Lensing::dbetadtheta_infinitesimal (lensing.h:399-480) over the stored
 * trajectory, with its own 27-cell cache. */
uint64_t mpfo_jacobian_infinitesimal(const mpfo_octree* o, int order, int precision, const double* cosmo_tab,
                                     uint64_t ncos, double length, unsigned nsteps, int stop_kind, double stop_value,
                                     const double* kitarget, const double* in9, const double* dist, uint64_t nd,
                                     double* out4) {
    traj_t t = {NULL, 0, 0, 1};
    run_full(o, order, precision, cosmo_tab, ncos, length, nsteps, stop_kind, stop_value, kitarget, in9, &t);
    const double c2 = C2;
    const unsigned levelshift = 2; /* log2(EXTENT) */
    tsc_cache cache;
    tsc_cache_init(&cache);
    double* a11 = (double*)calloc(nd * (t.n ? t.n : 1), sizeof(double));
    double* a12 = (double*)calloc(nd * (t.n ? t.n : 1), sizeof(double));
    double* a22 = (double*)calloc(nd * (t.n ? t.n : 1), sizeof(double));
    for (uint64_t i = 1; i < t.n; ++i) {
        const photon* p = &t.p[i];
        const double x = p->core[CX], y = p->core[CY], z = p->core[CZ], r = p->chi;
        const double phi = atan2(y, x);
        const double theta = acos(z / r);
        const double cp = cos(phi), sp = sin(phi), ct = cos(theta), st = sin(theta);
        const double s2p = sin(2 * phi), s2t = sin(2 * theta);
        const double dl = pow(2, -(p->level - 1 - levelshift));
        fdata xm, xp, ym, yp, zm, zp;
        tsc(o, precision, &cache, x - 0.5 * dl, y, z, &xm);
        tsc(o, precision, &cache, x + 0.5 * dl, y, z, &xp);
        tsc(o, precision, &cache, x, y - 0.5 * dl, z, &ym);
        tsc(o, precision, &cache, x, y + 0.5 * dl, z, &yp);
        tsc(o, precision, &cache, x, y, z - 0.5 * dl, &zm);
        tsc(o, precision, &cache, x, y, z + 0.5 * dl, &zp);
        /* Gravity<float> members: the differences are float subtractions */
#define FDIFF(a, b) (precision == 0 ? (double)((float)(a) - (float)(b)) : ((a) - (b)))
        const double dxxp = FDIFF(xp.v[F_DX], xm.v[F_DX]) / dl;
        const double dxyp = FDIFF(yp.v[F_DX], ym.v[F_DX]) / dl;
        const double dxzp = FDIFF(zp.v[F_DX], zm.v[F_DX]) / dl;
        const double dyyp = FDIFF(yp.v[F_DY], ym.v[F_DY]) / dl;
        const double dyzp = FDIFF(zp.v[F_DY], zm.v[F_DY]) / dl;
        const double dzzp = FDIFF(zp.v[F_DZ], zm.v[F_DZ]) / dl;
#undef FDIFF
        const double d1d1p = sp * sp * dxxp + cp * cp * dyyp - s2p * dxyp;
        const double d2d2p = cp * cp * ct * ct * dxxp + sp * sp * ct * ct * dyyp + st * st * dzzp + s2p * ct * ct * dxyp -
                             sp * s2t * dyzp - cp * s2t * dxzp;
        const double d1d2p = ct * cp * sp * (dyyp - dxxp) + (cp * cp - sp * sp) * ct * dxyp + sp * st * dxzp - cp * st * dyzp;
        const double dr = t.p[i].chi - t.p[i - 1].chi;
        for (uint64_t j = 0; j < nd; ++j) {
            const double gxsx = (dist[j] - p->chi) * p->chi / dist[j];
            a11[j * t.n + i] = a11[j * t.n + i - 1] + gxsx * d1d1p * dr;
            a12[j * t.n + i] = a12[j * t.n + i - 1] + gxsx * d1d2p * dr;
            a22[j * t.n + i] = a22[j * t.n + i - 1] + gxsx * d2d2p * dr;
        }
    }
    for (uint64_t j = 0; j < nd; ++j) {
        /* upper_bound on chi, firstid = marked - (marked > 0) */
        uint64_t lo = 0, cnt = t.n;
        while (cnt > 0) {
            const uint64_t h = cnt / 2;
            if (!(dist[j] < t.p[lo + h].chi)) { lo = lo + h + 1; cnt = cnt - h - 1; }
            else cnt = h;
        }
        const uint64_t first = lo - (lo > 0);
        const double v11 = t.n ? a11[j * t.n + first] : 0.0;
        const double v12 = t.n ? a12[j * t.n + first] : 0.0;
        const double v22 = t.n ? a22[j * t.n + first] : 0.0;
        out4[4 * j + 0] = 1 - 2 * v11 / c2 * length;
        out4[4 * j + 1] = out4[4 * j + 2] = -2 * v12 / c2 * length;
        out4[4 * j + 3] = 1 - 2 * v22 / c2 * length;
    }
    free(a11);
    free(a12);
    free(a22);
    const uint64_t n = t.n;
    free(t.p);
    return n;
}

static double elem_value(const photon* p, int kind) {
    switch (kind) {
        case 0: return p->chi;
        case 1: return p->lambda;
        case 2: return p->redshift;
        case 3: return p->core[CT];
        case 4: return p->core[CA];
        default: return p->s;
    }
}

/* jacobian of lensing.h:399-480 for one distance on a stored trajectory */
static void jacobian_at(const mpfo_octree* o, int precision, const traj_t* t, double dist, double length, double* a3) {
    const double c2 = C2;
    const unsigned levelshift = 2;
    tsc_cache cache;
    tsc_cache_init(&cache);
    double s11 = 0, s12 = 0, s22 = 0;
    /* firstid = upper_bound(chi, dist) - 1 */
    uint64_t lo = 0, cnt = t->n;
    while (cnt > 0) {
        const uint64_t h = cnt / 2;
        if (!(dist < t->p[lo + h].chi)) { lo = lo + h + 1; cnt = cnt - h - 1; }
        else cnt = h;
    }
    const uint64_t first = lo - (lo > 0);
    for (uint64_t i = 1; i <= first && i < t->n; ++i) {
        const photon* p = &t->p[i];
        const double x = p->core[CX], y = p->core[CY], z = p->core[CZ], r = p->chi;
        const double phi = atan2(y, x);
        const double theta = acos(z / r);
        const double cp = cos(phi), sp = sin(phi), ct = cos(theta), st = sin(theta);
        const double s2p = sin(2 * phi), s2t = sin(2 * theta);
        const double dl = pow(2, -(p->level - 1 - levelshift));
        fdata xm, xp, ym, yp, zm, zp;
        tsc(o, precision, &cache, x - 0.5 * dl, y, z, &xm);
        tsc(o, precision, &cache, x + 0.5 * dl, y, z, &xp);
        tsc(o, precision, &cache, x, y - 0.5 * dl, z, &ym);
        tsc(o, precision, &cache, x, y + 0.5 * dl, z, &yp);
        tsc(o, precision, &cache, x, y, z - 0.5 * dl, &zm);
        tsc(o, precision, &cache, x, y, z + 0.5 * dl, &zp);
#define FDIFF(a, b) (precision == 0 ? (double)((float)(a) - (float)(b)) : ((a) - (b)))
        const double dxxp = FDIFF(xp.v[F_DX], xm.v[F_DX]) / dl;
        const double dxyp = FDIFF(yp.v[F_DX], ym.v[F_DX]) / dl;
        const double dxzp = FDIFF(zp.v[F_DX], zm.v[F_DX]) / dl;
        const double dyyp = FDIFF(yp.v[F_DY], ym.v[F_DY]) / dl;
        const double dyzp = FDIFF(zp.v[F_DY], zm.v[F_DY]) / dl;
        const double dzzp = FDIFF(zp.v[F_DZ], zm.v[F_DZ]) / dl;
#undef FDIFF
        const double d1d1p = sp * sp * dxxp + cp * cp * dyyp - s2p * dxyp;
        const double d2d2p = cp * cp * ct * ct * dxxp + sp * sp * ct * ct * dyyp + st * st * dzzp + s2p * ct * ct * dxyp -
                             sp * s2t * dyzp - cp * s2t * dxzp;
        const double d1d2p = ct * cp * sp * (dyyp - dxxp) + (cp * cp - sp * sp) * ct * dxyp + sp * st * dxzp - cp * st * dyzp;
        const double dr = t->p[i].chi - t->p[i - 1].chi;
        const double gxsx = (dist - p->chi) * p->chi / dist;
        s11 = s11 + gxsx * d1d1p * dr;
        s12 = s12 + gxsx * d1d2p * dr;
        s22 = s22 + gxsx * d2d2p * dr;
    }
    a3[0] = 1 - 2 * s11 / c2 * length;
    a3[1] = -2 * s12 / c2 * length;
    a3[2] = 1 - 2 * s22 / c2 * length;
}

uint64_t mpfo_surfaces(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                       double length, unsigned nsteps, int kind, const double* vals, uint32_t nsurf, int jacobian,
                       const double* in9, double* out) {
    traj_t t = {NULL, 0, 0, 1};
    run_full(o, order, precision, cosmo_tab, ncos, length, nsteps, kind, vals[nsurf - 1], NULL, in9, &t);
    for (uint32_t j = 0; j < nsurf; ++j) {
        double* r = out + 32 * j;
        memset(r, 0, 32 * sizeof(double));
        const double ref = vals[j];
        /* hmaps.h:615-660: first element past the surface (find_if for the
         * redshift, upper_bound otherwise: the same on a monotonic sequence) */
        uint64_t marked = t.n;
        for (uint64_t i = 0; i < t.n; ++i) {
            const double v = elem_value(&t.p[i], kind);
            if ((kind == 4) ? (v < ref) : (v > ref)) { marked = i; break; }
        }
        const uint64_t first = marked - (marked > 0);
        r[1] = (double)first;
        if (t.n == 0 || first == t.n - 1 || first == 0) continue; /* rejected (hmaps.h:716) */
        const photon* a = &t.p[first];
        const photon* b = &t.p[first + 1];
        const double f = (elem_value(b, kind) - ref) / (elem_value(b, kind) - elem_value(a, kind));
        r[0] = 1.0;
        r[2] = f;
        r[3] = a->core[CX] * f + b->core[CX] * (1 - f);
        r[4] = a->core[CY] * f + b->core[CY] * (1 - f);
        r[5] = a->core[CZ] * f + b->core[CZ] * (1 - f);
        r[6] = sqrt(r[3] * r[3] + r[4] * r[4] + r[5] * r[5]);
        r[7] = a->chi * f + b->chi * (1 - f);
        r[8] = a->lambda * f + b->lambda * (1 - f);
        r[9] = a->core[CT] * f + b->core[CT] * (1 - f);
        r[10] = a->core[CA] * f + b->core[CA] * (1 - f);
        r[11] = a->redshift * f + b->redshift * (1 - f);
        r[12] = a->s * f + b->s * (1 - f);
        r[13] = a->isw * f + b->isw * (1 - f);
        r[14] = a->rho * f + b->rho * (1 - f);
        r[15] = a->phi * f + b->phi * (1 - f);
        double rmax = t.p[0].rho;   /* max_element over [0, firstid) (hmaps.h:819) */
        for (uint64_t i = 1; i < first; ++i) if (t.p[i].rho > rmax) rmax = t.p[i].rho;
        r[16] = rmax;
        r[17] = a->core[CKT] * f + b->core[CKT] * (1 - f);
        r[18] = a->core[CKX] * f + b->core[CKX] * (1 - f);
        r[19] = a->core[CKY] * f + b->core[CKY] * (1 - f);
        r[20] = a->core[CKZ] * f + b->core[CKZ] * (1 - f);
        r[21] = 1.0; r[22] = 0.0; r[23] = 1.0;
        if (jacobian) jacobian_at(o, precision, &t, r[6], length, r + 21);
    }
    const uint64_t n = t.n;
    free(t.p);
    return n;
}

void mpfo_dbetadtheta(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                      double length, unsigned nsteps, int kind, double opening, const double* observer, double phi,
                      double theta, const double* ref, const double* ki, const double* dist, uint64_t nd,
                      double* out4);

/* ------------------------------------------------- catalogue root-finding */

/* Parameters of the root-finder that change its path: jacobiantype (0
 * infinitesimal, 1 infinitesimal_born, 2 bundle), for the bundle the
 * stop_bundle kind (0 radius, 1 lambda, 2 redshift, 3 t, 4 a, 5 plane), the
 * plane (0 sachs, 1 normal) and the half-opening; use_previous_catalogues
 * (0, or 2: the grid re-run of rejected sources). */
typedef struct {
    int jactype, kind, plane;
    double opening;
    int use_prev;
} cat_cfg;

typedef struct {
    const mpfo_octree* o;
    int order, precision;
    const double* tab;
    uint64_t ncos;
    double length;
    unsigned nsteps;
    double accuracy;
    const double* vobs;
    double h, mpc;
    cat_cfg cfg;
} cat_ctx;

/* What persists across the newtonMethod2d calls of one source in the
 * reference: the jacobian, redshifts and interpRef out-parameters
 * (catalogues.h:811-814), plus counts kept here: rays integrated, and grid
 * rays of the re-run mode that ended before the source radius. */
typedef struct {
    double jac[4];
    double z[6];
    double interp;
    uint64_t rays, short_grid;
} cat_state;

static void launch9(double phi, double theta, double* in9) {
    /* Integrator::launch(0, 0, 0, phi, theta) (integrator.h:304-329) */
    in9[0] = 1.0; in9[1] = 0.0; in9[2] = 0.0; in9[3] = 0.0; in9[4] = 0.0; in9[5] = 1.0;
    in9[6] = cos(phi) * sin(theta);
    in9[7] = sin(phi) * sin(theta);
    in9[8] = cos(theta);
}

/* kiTarget and interpRef of the bundle at a crossing (catalogues.h:458-491):
 * A, B the elements around it, f its weight, fp its position; z1 and as the
 * redshift and scale factor the reference uses there (the central ray's, also
 * when a grid ray is chosen, catalogues.h:586-589) */
static void bundle_ref(const cat_ctx* g, const photon* A, const photon* B, double f, const double* fp,
                       double distTarget, double z1, double as, double* ki, double* ref) {
    if (g->cfg.plane == 0)
        for (int k = 0; k < 3; ++k) ki[k] = A->core[CKX + k] * f + B->core[CKX + k] * (1 - f);
    else
        for (int k = 0; k < 3; ++k) ki[k] = fp[k];
    switch (g->cfg.kind) {
        case 2: *ref = z1; break;
        case 4: *ref = as; break;
        case 3: *ref = A->core[CT] * f + B->core[CT] * (1 - f); break;
        case 1: *ref = A->lambda * f + B->lambda * (1 - f); break;
        case 5: *ref = ki[0] * fp[0] + ki[1] * fp[1] + ki[2] * fp[2]; break;
        default: *ref = distTarget; break;
    }
}

/* Lensing::dbetadtheta from the observer at the origin for one direction */
static void bundle_jac(const cat_ctx* g, const double* ki, double ref, double phi, double theta, double dist,
                       double* jac4) {
    const double observer[3] = {0.0, 0.0, 0.0};
    mpfo_dbetadtheta(g->o, g->order, g->precision, g->tab, g->ncos, g->length, g->nsteps, g->cfg.kind,
                     g->cfg.opening, observer, phi, theta, &ref, ki, &dist, 1, jac4);
}

static void jac_inf(const cat_ctx* g, const traj_t* t, double dist, double* jac4) {
    double j3[3];
    jacobian_at(g->o, g->precision, t, dist, g->length, j3);
    jac4[0] = j3[0]; jac4[1] = j3[1]; jac4[2] = j3[1]; jac4[3] = j3[2];
}

static void invert2(const double* j, double* ji) {   /* Utility::invMatrix2d (utility.h:687-698) */
    const double invdet = 1. / (j[0] * j[3] - j[1] * j[2]);
    ji[0] = j[3] * invdet; ji[1] = -j[1] * invdet; ji[2] = -j[2] * invdet; ji[3] = j[0] * invdet;
}

/* One Catalogues::newtonMethod2d call (catalogues.h:410-754): res receives
 * result[0][0], result[0][1], result[1][0], result[1][1] (the last is left 0
 * where the reference leaves it unset). */
static void newton2d(const cat_ctx* g, const double* tt, const double* vel, double phi, double theta,
                     unsigned iteration, cat_state* st, double* res) {
    const double c = C_LIGHT, c2 = C2;
    const double acc = g->accuracy;
    res[0] = res[1] = res[2] = res[3] = 0.0;
    const double distTarget = sqrt(pow(tt[0], 2) + pow(tt[1], 2) + pow(tt[2], 2));
    double photon9[9];
    launch9(phi, theta, photon9);
    traj_t t = {NULL, 0, 0, 1};
    run_full(g->o, g->order, g->precision, g->tab, g->ncos, g->length, g->nsteps, 0, distTarget, NULL, photon9, &t);
    st->rays++;
    const uint64_t marked = t.n - 1;
    uint64_t firstid = marked - (marked > 0);
    if (marked == 0 || t.p[marked].chi < distTarget) {   /* catalogues.h:433-440 */
        res[0] = 42; res[1] = 42; res[2] = 0; res[3] = (double)firstid;
        free(t.p);
        return;
    }
    double f = (t.p[firstid + 1].chi - distTarget) / (t.p[firstid + 1].chi - t.p[firstid].chi);
    {
        const photon* A = &t.p[firstid];
        const photon* B = &t.p[firstid + 1];
        st->z[1] = A->redshift * f + B->redshift * (1 - f);
    }
    const double as = t.p[firstid].core[CA] * f + t.p[firstid + 1].core[CA] * (1 - f);
    double fp[3];
    for (int k = 0; k < 3; ++k) fp[k] = t.p[firstid].core[CX + k] * f + t.p[firstid + 1].core[CX + k] * (1 - f);
    double ki[3] = {0, 0, 0};
    if (g->cfg.jactype == 2)
        bundle_ref(g, &t.p[firstid], &t.p[firstid + 1], f, fp, distTarget, st->z[1], as, ki, &st->interp);
    res[2] = sqrt(pow(tt[0] - fp[0], 2) + pow(tt[1] - fp[1], 2) + pow(tt[2] - fp[2], 2)) / distTarget;
    if (g->cfg.use_prev == 2) { res[0] = phi; res[1] = theta; }
    traj_t* T = &t;      /* the trajectory the rest reads: the central ray or the chosen grid ray */
    traj_t best = {NULL, 0, 0, 1};
    if (res[2] > acc) {
        double ji[4] = {1, 0, 0, 1};
        if (g->cfg.use_prev == 0) {
            if (iteration < 3) {
            } else if (iteration < 5 || g->cfg.jactype != 2) {
                jac_inf(g, T, distTarget, st->jac);
                invert2(st->jac, ji);
            } else {
                bundle_jac(g, ki, st->interp, phi, theta, distTarget, st->jac);
                if (st->jac[0] == 42 && st->jac[1] == 42) {   /* catalogues.h:518-523 */
                    res[0] = 42; res[1] = 42; res[2] = 22; res[3] = 0;
                    free(t.p);
                    return;
                }
                invert2(st->jac, ji);
            }
        } else {
            /* grid around the last angles, its size and step from the
             * separation (catalogues.h:527-623) */
            const double err = 5 * res[2] / (iteration + 1);
            double phi_tmp = phi, theta_tmp = theta;
            double sep_tmp = res[2];
            for (double i = -err; i <= err; i += 0.2 * err) {
                for (double j = -err; j <= err; j += 0.2 * err) {
                    double in9[9];
                    launch9(phi + i, theta + j, in9);
                    memcpy(photon9, in9, sizeof(in9));
                    traj_t u = {NULL, 0, 0, 1};
                    run_full(g->o, g->order, g->precision, g->tab, g->ncos, g->length, g->nsteps, 0, distTarget, NULL,
                             in9, &u);
                    st->rays++;
                    const uint64_t marked2 = u.n - 1;
                    if (marked2 == 0) {
                        /* the reference reads past this one-element
                         * trajectory: here the ray is no candidate */
                        st->short_grid++;
                        sep_tmp = INFINITY;
                        free(u.p);
                        continue;
                    }
                    if (u.p[marked2].chi < distTarget) st->short_grid++;   /* extrapolated, as the reference does */
                    firstid = marked2 - 1;
                    f = (u.p[firstid + 1].chi - distTarget) / (u.p[firstid + 1].chi - u.p[firstid].chi);
                    for (int k = 0; k < 3; ++k)
                        fp[k] = u.p[firstid].core[CX + k] * f + u.p[firstid + 1].core[CX + k] * (1 - f);
                    sep_tmp = sqrt(pow(tt[0] - fp[0], 2) + pow(tt[1] - fp[1], 2) + pow(tt[2] - fp[2], 2)) / distTarget;
                    if (sep_tmp < acc) {
                        res[2] = sep_tmp;
                        phi_tmp = phi + i;
                        theta_tmp = theta + j;
                        free(best.p);
                        best = u;
                        T = &best;
                        if (g->cfg.jactype == 2)
                            bundle_ref(g, &u.p[firstid], &u.p[firstid + 1], f, fp, distTarget, st->z[1], as, ki,
                                       &st->interp);
                        break;
                    } else if (sep_tmp < res[2]) {
                        res[2] = sep_tmp;
                        phi_tmp = phi + i;
                        theta_tmp = theta + j;
                        free(best.p);
                        best = u;
                        T = &best;
                    } else {
                        free(u.p);
                    }
                }
                if (sep_tmp < acc) break;
            }
            res[0] = phi_tmp;
            res[1] = theta_tmp;
            if (res[2] > acc) {
                free(best.p);
                free(t.p);
                return;
            }
        }
        if (g->cfg.use_prev == 0) {
            /* the update in the (e1, e2) plane at (phi, theta) (catalogues.h:627-660) */
            const double cp = cos(phi), sp = sin(phi), ct = cos(theta), st_ = sin(theta);
            const double e1[3] = {-sp, cp, 0}, e2[3] = {-cp * ct, -sp * ct, st_};
            const double targeted[3] = {distTarget * cp * st_, distTarget * sp * st_, distTarget * ct};
            double sep[3], pos[3];
            for (int k = 0; k < 3; ++k) { sep[k] = fp[k] - tt[k]; pos[k] = targeted[k] - tt[k]; }
            const double sepx = sep[0] * e1[0] + sep[1] * e1[1];
            const double sepy = sep[0] * e2[0] + sep[1] * e2[1] + sep[2] * e2[2];
            const double posx = pos[0] * e1[0] + pos[1] * e1[1];
            const double posy = pos[0] * e2[0] + pos[1] * e2[1] + pos[2] * e2[2];
            const double sepxnew = ji[0] * sepx + ji[1] * sepy;
            const double sepynew = ji[2] * sepx + ji[3] * sepy;
            const double posxnew = posx - sepxnew, posynew = posy - sepynew;
            const double xnew = tt[0] + (posxnew * e1[0] + posynew * e2[0]);
            const double ynew = tt[1] + (posxnew * e1[1] + posynew * e2[1]);
            const double znew = tt[2] + posynew * e2[2];
            res[0] = atan2(ynew, xnew);
            res[1] = acos(znew / sqrt(xnew * xnew + ynew * ynew + znew * znew));
            free(best.p);
            free(t.p);
            return;
        }
    }
    /* the Jacobian of the converged source (catalogues.h:665-681) */
    if (g->cfg.jactype == 2) {
        bundle_jac(g, ki, st->interp, phi, theta, distTarget, st->jac);
    } else if (g->cfg.jactype == 0) {
        jac_inf(g, T, distTarget, st->jac);
    } else {
        /* the last photon launched, on the homogeneous path (Integrator::integrate<-1>) */
        traj_t b = {NULL, 0, 0, 1};
        run_full(g->o, -1, g->precision, g->tab, g->ncos, g->length, g->nsteps, 0, distTarget, NULL, photon9, &b);
        jac_inf(g, &b, distTarget, st->jac);
        free(b.p);
    }
    if (st->jac[0] == 42 && st->jac[1] == 42) {
        res[0] = 42; res[1] = 42; res[2] = 22; res[3] = 0;
        free(best.p);
        free(t.p);
        return;
    }
    /* redshifts (catalogues.h:682-719): potential, ISW and direction on the
     * trajectory read, the scale factor and redshift of the central ray */
    const photon* A = &T->p[firstid];
    const photon* B = &T->p[firstid + 1];
    const double potential = A->phi * f + B->phi * (1 - f);
    const double isw = A->isw * f + B->isw * (1 - f);
    st->z[0] = -1 + 1. / (as);
    const double unit_t = as * as * g->mpc / (g->h * 100 * 1000);
    const double unit_l = as * g->length * 100;
    const double vx = vel[0] * unit_l * 1e-2 / (unit_t * c);
    const double vy = vel[1] * unit_l * 1e-2 / (unit_t * c);
    const double vz = vel[2] * unit_l * 1e-2 / (unit_t * c);
    const double* vobs = g->vobs;
    const photon* F = &T->p[0];
    const double gref = -F->core[CA] * c * F->core[CKT] *
                        (1. + F->phi / c2 + (vobs[0] * F->core[CKX] + vobs[1] * F->core[CKY] + vobs[2] * F->core[CKZ]) / (c * F->core[CKT]) +
                         0.5 * (vobs[0] * vobs[0] + vobs[1] * vobs[1] + vobs[2] * vobs[2]));
    const double dtdl = A->core[CKT] * f + B->core[CKT] * (1 - f);
    const double dxdl = A->core[CKX] * f + B->core[CKX] * (1 - f);
    const double dydl = A->core[CKY] * f + B->core[CKY] * (1 - f);
    const double dzdl = A->core[CKZ] * f + B->core[CKZ] * (1 - f);
    const double gK = -as * c * dtdl * (1. + potential / c2 + (vx * dxdl + vy * dydl + vz * dzdl) / (c * dtdl) +
                                        0.5 * (vx * vx + vy * vy + vz * vz));
    st->z[2] = st->z[1] + (vx * dxdl + vy * dydl + vz * dzdl) / (as * c * dtdl) -
               (vobs[0] * F->core[CKX] + vobs[1] * F->core[CKY] + vobs[2] * F->core[CKZ]) / (as * c * F->core[CKT]);
    st->z[3] = st->z[2] + 0.5 * (vx * vx + vy * vy + vz * vz - vobs[0] * vobs[0] - vobs[1] * vobs[1] - vobs[2] * vobs[2]) / as;
    st->z[4] = st->z[3] + isw / as;
    st->z[5] = gK / gref - 1;
    if (g->cfg.use_prev == 0) { res[0] = phi; res[1] = theta; }
    free(best.p);
    free(t.p);
}

typedef struct {
    const cat_ctx* g;
    const double *src6, *init2;
    unsigned max_it;
    double* out;
    uint64_t begin, end, rays;
} cat_work;

/* Catalogues::iterateNewtonMethod (catalogues.h:757-800) per source, and the
 * routing of relCat (catalogues.h:876-903) as a status */
static void* cat_worker(void* arg) {
    cat_work* w = (cat_work*)arg;
    const cat_ctx* g = w->g;
    w->rays = 0;
    for (uint64_t i = w->begin; i < w->end; ++i) {
        const double* s = w->src6 + 6 * i;
        double* o = w->out + 20 * i;
        memset(o, 0, 20 * sizeof(double));
        const double d = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        o[1] = acos(s[2] / d);
        o[0] = atan2(s[1], s[0]);
        cat_state st;
        memset(&st, 0, sizeof(st));
        double res[4] = {w->init2 ? w->init2[2 * i] : o[0], w->init2 ? w->init2[2 * i + 1] : o[1], 0, 0};
        unsigned iteration = 0;
        int status;
        for (;;) {
            double r[4];
            newton2d(g, s, s + 3, res[0], res[1], iteration, &st, r);
            memcpy(res, r, sizeof(r));
            iteration++;
            if (res[0] == 42 && res[1] == 42) { status = 1; break; }
            if (iteration > w->max_it && res[2] > g->accuracy) { res[3] = 42; status = 2; break; }
            if (!(res[2] > g->accuracy)) { res[3] = iteration; status = 0; break; }
        }
        o[2] = res[0]; o[3] = res[1]; o[4] = res[2]; o[5] = res[3];
        for (int k = 0; k < 6; ++k) o[6 + k] = st.z[k];
        for (int k = 0; k < 4; ++k) o[12 + k] = st.jac[k];
        o[16] = status;
        o[17] = iteration;
        o[18] = (double)st.short_grid;
        w->rays += st.rays;
    }
    return NULL;
}

uint64_t mpfo_catalogue_ex(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                           double length, unsigned nsteps, const double* src6, const double* init2, uint64_t n,
                           double accuracy, unsigned max_iterations, const double* vobs, double h, double mpc,
                           int jactype, int bundle_kind, int plane, double opening, int use_prev, double* out,
                           int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if ((uint64_t)nthreads > n) nthreads = n ? (int)n : 1;
    const cat_ctx g = {o, order, precision, cosmo_tab, ncos, length, nsteps, accuracy, vobs, h, mpc,
                       {jactype, bundle_kind, plane, opening, use_prev}};
    cat_work* w = (cat_work*)calloc((size_t)nthreads, sizeof(cat_work));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    const uint64_t chunk = (n + nthreads - 1) / nthreads;
    for (int k = 0; k < nthreads; ++k) {
        w[k] = (cat_work){&g, src6, init2, max_iterations, out, 0, 0, 0};
        w[k].begin = chunk * k < n ? chunk * k : n;
        w[k].end = chunk * (k + 1) < n ? chunk * (k + 1) : n;
    }
    for (int k = 1; k < nthreads; ++k) pthread_create(&th[k], NULL, cat_worker, &w[k]);
    cat_worker(&w[0]);
    uint64_t rays = w[0].rays;
    for (int k = 1; k < nthreads; ++k) {
        pthread_join(th[k], NULL);
        rays += w[k].rays;
    }
    free(w);
    free(th);
    return rays;
}

uint64_t mpfo_catalogue_bundle(const mpfo_octree* o, int order, int precision, const double* cosmo_tab,
                               uint64_t ncos, double length, unsigned nsteps, const double* src6, const double* init2,
                               uint64_t n, double accuracy, unsigned max_iterations, const double* vobs, double h,
                               double mpc, int bundle_kind, int plane, double opening, double* out, int nthreads) {
    return mpfo_catalogue_ex(o, order, precision, cosmo_tab, ncos, length, nsteps, src6, init2, n, accuracy,
                             max_iterations, vobs, h, mpc, bundle_kind < 0 ? 0 : 2, bundle_kind, plane, opening, 0,
                             out, nthreads);
}

uint64_t mpfo_catalogue(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                        double length, unsigned nsteps, const double* src6, const double* init2, uint64_t n,
                        double accuracy, unsigned max_iterations, const double* vobs, double h, double mpc,
                        double* out, int nthreads) {
    return mpfo_catalogue_ex(o, order, precision, cosmo_tab, ncos, length, nsteps, src6, init2, n, accuracy,
                             max_iterations, vobs, h, mpc, 0, -1, 0, 0.0, 0, out, nthreads);
}

/* Catalogues::relCat_with_previous_cat (catalogues.h:922-1044), per row of a
 * previous catalogue (18 columns): the ray from its observed angles (the
 * comoving ones for infinitesimal_born) to the scale factor of its FLRW
 * redshift, the Jacobian at the interpolated distance; jac4 receives the new
 * columns 13-16.  flag[i] = 1 where the reference reads past a one-element
 * trajectory (jac4 then 42, 42, 42, 42), 2 where the ray ends before the
 * scale factor (the reference extrapolates its last segment; so does this),
 * else 0. */
void mpfo_catalogue_prev(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                         double length, unsigned nsteps, const double* prev18, uint64_t n, int jactype,
                         int bundle_kind, int plane, double opening, double* jac4, int* flag) {
    const double zero3[3] = {0, 0, 0};
    const cat_ctx g = {o, order, precision, cosmo_tab, ncos, length, nsteps, 0.0, zero3, 0.7, 3.08568e22,
                       {jactype, bundle_kind, plane, opening, 1}};
    for (uint64_t i = 0; i < n; ++i) {
        const double* p = prev18 + 18 * i;
        double* J = jac4 + 4 * i;
        flag[i] = 0;
        const double sf = 1. / (1. + p[7]);
        const double phi = (jactype == 1) ? p[1] : p[3];
        const double theta = (jactype == 1) ? p[2] : p[4];
        double in9[9];
        launch9(phi, theta, in9);
        traj_t t = {NULL, 0, 0, 1};
        run_full(o, order, precision, cosmo_tab, ncos, length, nsteps, 4, sf, NULL, in9, &t);
        const uint64_t marked = t.n - 1;
        if (marked == 0) {
            flag[i] = 1;
            J[0] = J[1] = J[2] = J[3] = 42;
            free(t.p);
            continue;
        }
        if (t.p[marked].core[CA] > sf) flag[i] = 2;
        const uint64_t firstid = marked - 1;
        const photon* A = &t.p[firstid];
        const photon* B = &t.p[firstid + 1];
        const double f = (B->core[CA] - sf) / (B->core[CA] - A->core[CA]);
        double fp[3];
        for (int k = 0; k < 3; ++k) fp[k] = A->core[CX + k] * f + B->core[CX + k] * (1 - f);
        const double distTarget = A->chi * f + B->chi * (1 - f);
        if (jactype == 2) {
            double ki[3], ref;
            const double z1 = A->redshift * f + B->redshift * (1 - f);
            bundle_ref(&g, A, B, f, fp, distTarget, z1, sf, ki, &ref);
            bundle_jac(&g, ki, ref, phi, theta, distTarget, J);
        } else if (jactype == 0) {
            jac_inf(&g, &t, distTarget, J);
        } else {
            traj_t b = {NULL, 0, 0, 1};
            run_full(o, -1, precision, cosmo_tab, ncos, length, nsteps, 4, sf, NULL, in9, &b);
            jac_inf(&g, &b, distTarget, J);
            free(b.p);
        }
        free(t.p);
    }
}

void mpfo_interpolate(const double* x, const double* y, uint64_t n, const double* x0, uint64_t m, double* out) {
    for (uint64_t i = 0; i < m; ++i) out[i] = interpolate(x0[i], x, y, (int64_t)n);
}

/* ------------------------------------------------------------------ propagate */

/* Integrator::launch around one photon (integrator.h:345-382): count photons
 * on a circle of half-angle `angle` turned by `rotation`, after the photon
 * itself.  sin and cos of one variable come from one sincos call, as the
 * reference build computes them; the two trigonometric calls of atan2(...)
 * and those of acos(...) stay separate (the two library calls of each are
 * not merged into one value, and glibc's sincos and cos can differ in the
 * last bit, as at 1.8608994686324885). */
void mpfo_launch_bundle(const double* in9, unsigned count, double angle, double rotation, double* out9) {
    const double pi = 4.0 * atan(1.0);
    const double step = (pi + pi) / (double)(count + (!count));
    const double kx = in9[6], ky = in9[7], kz = in9[8];
    const double r = sqrt((kx * kx) + (ky * ky) + (kz * kz));
    double sa, ca, sth, cth, sps, cps;
    sincos(angle, &sa, &ca);
    cth = cos(acos(kz / r));
    sth = sin(acos(kz / r));
    const double cph = cos(atan2(ky, kx));
    const double sph = sin(atan2(ky, kx));
    sincos(rotation, &sps, &cps);
    const double rcos = r * ca, rsin = r * sa;
    for (unsigned i = 0; i <= count; ++i)
        for (int k = 0; k < 9; ++k) out9[9 * i + k] = in9[k];
    if (!(r > 0.0)) return;
    for (unsigned i = 0; i < count; ++i) {
        double si, ci;
        sincos(i * step, &si, &ci);
        const double x = rsin * ci, y = rsin * si, z = rcos;
        double* o = out9 + 9 * (i + 1);
        o[6] = -cph * sps * cth * x - cph * cps * cth * y + cph * sth * z + sph * sps * y - sph * cps * x;
        o[7] = -sph * sps * cth * x - sph * cps * cth * y - cph * sps * y + cph * cps * x + sth * sph * z;
        o[8] = sth * sps * x + sth * cps * y + cth * z;
    }
}

/* interpolation parameter of propagate (integrator.h:916-971): kind 2
 * redshift, 4 a (as 1/a - 1), 3 t, 1 lambda, 0 r (distance from the first
 * element); anything else redshift */
static double prop_ref(const traj_t* t, uint64_t k, int kind) {
    const photon* p = &t->p[k];
    switch (kind) {
        case 4: return 1 / p->core[CA] - 1;
        case 3: return p->core[CT];
        case 1: return p->lambda;
        case 0: {
            const double x = p->core[CX] - t->p[0].core[CX];
            const double y = p->core[CY] - t->p[0].core[CY];
            const double z = p->core[CZ] - t->p[0].core[CZ];
            return sqrt(x * x + y * y + z * z);
        }
        default: return p->redshift;
    }
}

/* Utility::interpolate (utility.h:442-449) of coordinate `dim` of ray t at
 * parameter x0: upper_bound on the parameter sequence (the bisection of
 * std::upper_bound, so a sequence that is not sorted is searched the same
 * way), clamped to the ends */
static double prop_interp(const traj_t* t, int kind, double x0, int dim) {
    const long long n = (long long)t->n;
    long long first = 0, len = n;
    while (len > 0) {
        const long long half = len >> 1;
        if (x0 < prop_ref(t, first + half, kind)) len = half;
        else { first = first + half + 1; len = len - half - 1; }
    }
    long long i = first;
    i += (n > 1) * ((i <= 0) - (i >= n));
    if (n <= 1) return n > 0 ? t->p[0].core[dim] : 0.0;
    const double xa = prop_ref(t, i - 1, kind), xb = prop_ref(t, i, kind);
    const double ya = t->p[i - 1].core[dim], yb = t->p[i].core[dim];
    return ya + ((yb - ya) * (x0 - xa) / (xb - xa));
}

/* Integrator::propagate (integrator.h:826-1015): the bundle of `count` rays
 * around the central photon, every ray integrated until the data ends, the
 * validity checks, and the angular diameter distance along the central ray.
 * rows/dist receive up to maxlen elements of the central trajectory; returns
 * its element count (0 when the bundle is rejected). */
uint64_t mpfo_propagate(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                        double length, unsigned nsteps, const double* in9, unsigned count, double angle,
                        double rotation, int kind, double amin, double* rows, double* dist, uint64_t maxlen) {
    const unsigned nb = count + 1;
    double* init = (double*)malloc(sizeof(double) * 9 * nb);
    traj_t* t = (traj_t*)calloc(nb, sizeof(traj_t));
    double* last = (double*)calloc(nb, sizeof(double));
    mpfo_launch_bundle(in9, count, angle, rotation, init);
    const double quarter = (4.0 / (2.0 * 1.0)) / 2.0;  /* EXTENT / 4 */
    const double limit = 1.0 / (2.0 * 2.0 * 2.0);
    unsigned ntraj = nb;
    for (unsigned j = 0; j < ntraj; ++j) {
        t[j].keep = 1;
        run_full(o, order, precision, cosmo_tab, ncos, length, nsteps, -1, 0.0, NULL, init + 9 * j, &t[j]);
        if (!t[j].n) {
            ntraj = 0;
        } else {
            const photon* b = &t[j].p[t[j].n - 1];
            const photon* f = &t[j].p[0];
            const double cx = b->core[CX] - f->core[CX], cy = b->core[CY] - f->core[CY], cz = b->core[CZ] - f->core[CZ];
            last[j] = sqrt(cx * cx + cy * cy + cz * cz);
            ntraj *= (last[j] > quarter);
            ntraj *= (!(isnormal(amin) && isnormal(b->ah) && (b->ah < 1.0))) || (!(b->ah > amin));
        }
    }
    double mx = last[0], mn = last[0];
    for (unsigned j = 1; j < nb; ++j) {
        if (last[j] > mx) mx = last[j];
        if (last[j] < mn) mn = last[j];
    }
    ntraj *= (fabs(mx - mn) / mx < limit);
    uint64_t size = 0;
    if (ntraj > 0) {
        traj_t* c = &t[0];
        size = c->n;
        if (ntraj > 1) {
            for (uint64_t k = 0; k < size; ++k) {
                const double x0 = prop_ref(c, k, kind);
                double d = c->p[k].distance;
                for (unsigned j = 1; j < ntraj; ++j) {
                    const double dx = prop_interp(&t[j], kind, x0, CX) - c->p[k].core[CX];
                    const double dy = prop_interp(&t[j], kind, x0, CY) - c->p[k].core[CY];
                    const double dz = prop_interp(&t[j], kind, x0, CZ) - c->p[k].core[CZ];
                    d += sqrt(dx * dx + dy * dy + dz * dz);
                }
                d /= (ntraj - 1);
                c->p[k].distance = d;
            }
            for (uint64_t k = 0; k < size; ++k) c->p[k].distance *= (length * c->p[k].core[CA]) / angle;
        }
        for (uint64_t k = 0; k < size && k < maxlen; ++k) {
            photon_to_rec(&c->p[k], rows + MPFO_NREC * k, (double)k);
            dist[k] = c->p[k].distance;
        }
    }
    for (unsigned j = 0; j < nb; ++j) free(t[j].p);
    free(t);
    free(last);
    free(init);
    return size;
}

/* -------------------------------------------------------------- bundle Jacobian */

/* value of the stop parameter of element p (lensing.h:205-301); plane
 * (kind 5): the position along ki (lensing.h:281-293) */
static double bundle_param_ki(const photon* p, int kind, const double* ki) {
    if (kind == 5) return p->core[CX] * ki[0] + p->core[CY] * ki[1] + p->core[CZ] * ki[2];
    switch (kind) {
        case 1: return p->lambda;
        case 0: return p->chi;
        case 4: return p->core[CA];
        case 3: return p->core[CT];
        default: return p->redshift;
    }
}
static double bundle_param(const photon* p, int kind) { return bundle_param_ki(p, kind, NULL); }

/* Lensing::dbetadtheta (lensing.h:140-364) for one direction: four rays at
 * +-tan(opening) along the screen vectors e1, e2 of (phi, theta), integrated
 * to the last value with the stop criterion `kind` (0 radius, 1 lambda,
 * 2 redshift, 3 t, 4 a), each interpolated at every value (upper_bound on the
 * stored trajectory, f = (next - ref) / (next - previous)), then finite
 * differences projected on the screen of each ki.  out4: nd x (a11, a12,
 * a21, a22); on a ray that ends before a value: every entry 42, 42, 1, 0
 * (for `a`: that entry 42, 42, 1, 3 and the others 0). */
void mpfo_dbetadtheta(const mpfo_octree* o, int order, int precision, const double* cosmo_tab, uint64_t ncos,
                      double length, unsigned nsteps, int kind, double opening, const double* observer, double phi,
                      double theta, const double* ref, const double* ki, const double* dist, uint64_t nd,
                      double* out4) {
    double cp, sp, ct, st;
    sincos(phi, &sp, &cp);
    sincos(theta, &st, &ct);
    const double target[3] = {cp * st, sp * st, ct};
    const double e1[3] = {-sp, cp, 0}, e2[3] = {-cp * ct, -sp * ct, st};
    const double to = tan(opening);
    double rini[4][3];
    rini[0][0] = target[0] + to * e1[0]; rini[0][1] = target[1] + to * e1[1]; rini[0][2] = target[2];
    rini[1][0] = target[0] - to * e1[0]; rini[1][1] = target[1] - to * e1[1]; rini[1][2] = target[2];
    rini[2][0] = target[0] - to * e2[0]; rini[2][1] = target[1] - to * e2[1]; rini[2][2] = target[2] - to * e2[2];
    rini[3][0] = target[0] + to * e2[0]; rini[3][1] = target[1] + to * e2[1]; rini[3][2] = target[2] + to * e2[2];
    double* fp = (double*)calloc(nd * 12, sizeof(double));
    for (uint64_t k = 0; k < 4 * nd; ++k) out4[k] = 0.0;
    for (int it = 0; it < 4; ++it) {
        /* Integrator::launch from the observer towards rini (integrator.h:264-290) */
        const double dx = rini[it][0] - observer[0], dy = rini[it][1] - observer[1], dz = rini[it][2] - observer[2];
        const double norm = sqrt(dx * dx + dy * dy + dz * dz);
        const double in9[9] = {1.0, 0.0, observer[0], observer[1], observer[2], 1.0, dx / norm, dy / norm, dz / norm};
        traj_t t = {NULL, 0, 0, 1};
        /* the plane criterion stops at the last value along the last ki (lensing.h:201) */
        run_full(o, order, precision, cosmo_tab, ncos, length, nsteps, kind, ref[nd - 1],
                 kind == 5 ? ki + 3 * (nd - 1) : NULL, in9, &t);
        for (uint64_t r = 0; r < nd; ++r) {
            const double* kr = ki + 3 * r;
            const double back = t.n ? bundle_param_ki(&t.p[t.n - 1], kind, kr) : 0.0;
            const int fail = (kind == 4) ? (back > ref[r]) : (back < ref[r]);
            if (fail) {
                if (kind == 4) {
                    out4[4 * r] = 42; out4[4 * r + 1] = 42; out4[4 * r + 2] = 1; out4[4 * r + 3] = 3;
                } else {
                    for (uint64_t q = 0; q < nd; ++q) {
                        out4[4 * q] = 42; out4[4 * q + 1] = 42; out4[4 * q + 2] = 1; out4[4 * q + 3] = 0;
                    }
                }
                free(t.p);
                free(fp);
                return;
            }
            /* upper_bound with the criterion's order (decreasing for a) */
            uint64_t first = 0, len = t.n;
            while (len > 0) {
                const uint64_t half = len >> 1;
                const double v = bundle_param_ki(&t.p[first + half], kind, kr);
                const int before = (kind == 4) ? (ref[r] > v) : (ref[r] < v);
                if (before) len = half;
                else { first = first + half + 1; len = len - half - 1; }
            }
            uint64_t id = first - (first > 0);
            /* the last element's value equal to ref: upper_bound ends at t.n and the
             * reference reads one element past the trajectory; take the last segment */
            if (id + 1 >= t.n) id = t.n - 2;
            const double next = bundle_param_ki(&t.p[id + 1], kind, kr), prev = bundle_param_ki(&t.p[id], kind, kr);
            const double f = (next - ref[r]) / (next - prev);
            for (int d = 0; d < 3; ++d)
                fp[12 * r + 3 * it + d] = t.p[id].core[CX + d] * f + t.p[id + 1].core[CX + d] * (1 - f);
        }
        free(t.p);
    }
    for (uint64_t r = 0; r < nd; ++r) {
        double rh[3], rv[3];
        for (int d = 0; d < 3; ++d) {
            rh[d] = fp[12 * r + d] - fp[12 * r + 3 + d];
            rv[d] = fp[12 * r + 6 + d] - fp[12 * r + 9 + d];
        }
        const double kx = ki[3 * r], ky = ki[3 * r + 1], kz = ki[3 * r + 2];
        const double phik = atan2(ky, kx);
        const double thetak = acos(kz / sqrt(kx * kx + ky * ky + kz * kz));
        double cpk, spk, ctk, stk;
        sincos(phik, &spk, &cpk);
        sincos(thetak, &stk, &ctk);
        const double f1[3] = {-spk, cpk, 0}, f2[3] = {-cpk * ctk, -spk * ctk, stk};
        out4[4 * r + 0] = (rh[0] * f1[0] + rh[1] * f1[1]) / (2 * dist[r] * to);
        out4[4 * r + 1] = (rv[0] * f1[0] + rv[1] * f1[1]) / (2 * dist[r] * to);
        out4[4 * r + 2] = (rh[0] * f2[0] + rh[1] * f2[1] + rh[2] * f2[2]) / (-2 * dist[r] * to);
        out4[4 * r + 3] = (rv[0] * f2[0] + rv[1] * f2[1] + rv[2] * f2[2]) / (-2 * dist[r] * to);
    }
    free(fp);
}
