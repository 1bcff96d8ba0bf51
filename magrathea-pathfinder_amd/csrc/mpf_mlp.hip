// Learned first guess for the catalogue root-finder: a 10 -> 128 -> 128 -> 2
// GELU network with BF16 operands and FP32 accumulation on the matrix cores,
// trained on this device by its own forward and backward kernels (mixed
// precision: FP32 master weights, biases and Adam moments; BF16 copies of the
// weights, activations and back-propagated deltas).  No counterpart in the
// reference: it seeds the Newton iterations of catalogues.h:757-800, which
// still finish every source to the accuracy.
//
// Products with K = 128 (hidden layers, weight gradients over samples) use
// gfx950's v_mfma_f32_16x16x32_bf16, those with K = 16 (input layer, the
// output delta) v_mfma_f32_16x16x16_bf16.  Operand layouts (wave64, lane l,
// r = l & 15, q = l >> 4):
//   16x16x16: A (M x K) lane holds A[r][4q + j], j < 4; B (K x N) B[4q + j][r]
//   16x16x32: A lane holds A[r][8q + j], j < 8;      B B[8q + j][r]
//   D (M x N), both: lane holds D[4q + j][r], j < 4
// B operands are stored as B^T rows (N x K) so that a lane reads 8 or 16
// contiguous bytes.  mfma_probe_kernel and mfma_probe32_kernel check the
// layouts on the device.
//
// One wave computes 16 samples through all layers: the 16 x 128 hidden
// activations go through LDS (BF16) between layers, the FP32
// pre-activations needed by GELU' stay in registers in the D layout, which is
// also the layout of the back-propagated deltas.  Inference runs a
// persistent kernel that holds the BF16 weights in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mpf_exact_dev.h"
#include "mpf_features.h"
#include "mpf_kernels.h"
#include "mpf_mlp.h"

namespace mpf {

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int KIN = MPF_MLP_KIN;     // 10 input features padded to one MFMA K step
constexpr int NH = MPF_MLP_NHIDDEN;  // 128
constexpr int KOUT = MPF_MLP_KOUT;   // 2 outputs padded to one MFMA tile
constexpr int NT = NH / 16;

constexpr int OFF_W1 = 0;                // [KIN][NH]
constexpr int OFF_B1 = OFF_W1 + KIN * NH;
constexpr int OFF_W2 = OFF_B1 + NH;      // [NH][NH]
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;      // [NH][KOUT]
constexpr int OFF_B3 = OFF_W3 + NH * KOUT;
constexpr int NPARAM = OFF_B3 + KOUT;
static_assert(NPARAM == MPF_MLP_NPARAM, "parameter count");

constexpr int WGRAD_CHUNK = 256;         // samples per weight-gradient split

__host__ __device__ inline unsigned short f2bf(float f) {
    unsigned u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf2f(unsigned short h) {
    const unsigned u = (unsigned)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_d(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

__device__ inline f4 mfma(s4 a, s4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
__device__ inline f4 mfma32(b8 a, b8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

// 8 BF16 values from 16 contiguous, 16-byte aligned bytes
__device__ inline b8 load8(const unsigned short* p) {
    return __builtin_bit_cast(b8, *reinterpret_cast<const uint4*>(p));
}
__device__ inline b8 pack8(const unsigned short* v) {
    uint4 w;
    w.x = (unsigned)v[0] | ((unsigned)v[1] << 16);
    w.y = (unsigned)v[2] | ((unsigned)v[3] << 16);
    w.z = (unsigned)v[4] | ((unsigned)v[5] << 16);
    w.w = (unsigned)v[6] | ((unsigned)v[7] << 16);
    return __builtin_bit_cast(b8, w);
}
constexpr int KT32 = NH / 32;   // K steps of 32 over a hidden layer

__device__ inline s4 load4(const unsigned short* p) {
    const uint2 w = *reinterpret_cast<const uint2*>(p);  // 8-byte aligned by construction
    s4 v;
    v[0] = (short)(w.x & 0xFFFFu); v[1] = (short)(w.x >> 16);
    v[2] = (short)(w.y & 0xFFFFu); v[3] = (short)(w.y >> 16);
    return v;
}

struct Bf16Weights {
    const unsigned short* w1t;  // [NH][KIN]   W1^T
    const unsigned short* w2t;  // [NH][NH]    W2^T
    const unsigned short* w2;   // [NH][NH]    W2
    const unsigned short* w3t;  // [KOUT][NH]  W3^T
    const unsigned short* w3;   // [NH][KOUT]  W3
    const float* b1;
    const float* b2;
    const float* b3;
};

struct Saved {                  // per-sample BF16 tensors of a training batch
    unsigned short* x;          // n x KIN
    unsigned short* h1;         // n x NH
    unsigned short* h2;         // n x NH
    unsigned short* d1;         // n x NH   dL/dZ1
    unsigned short* d2;         // n x NH   dL/dZ2
    unsigned short* d3;         // n x KOUT dL/dZ3
};

// Training step of 16 samples per wave: the forward pass, the loss, the
// backward pass through the activations and the saved tensors.
//   x: normalised features (MPF_MLP_NIN per sample), rows perm[i] (or i)
//   t: normalised targets; loss accumulates sum (y - t)^2 of the outputs y
__global__ void __launch_bounds__(256) mlp_train_tile_kernel(Bf16Weights W, const float* __restrict__ x,
                                                             const uint32_t* __restrict__ perm, uint64_t n,
                                                             const float* __restrict__ t, Saved S, float dscale,
                                                             float* __restrict__ loss) {
    __shared__ __align__(16) unsigned short lds_a[4][16][NH];   // H1
    __shared__ __align__(16) unsigned short lds_b[4][16][NH];   // H2, then dZ2
    __shared__ __align__(16) unsigned short lds_c[4][16][KOUT]; // dZ3
    const int wave = threadIdx.x >> 6;
    const int l = threadIdx.x & 63;
    const int r = l & 15, q = l >> 4;
    const uint64_t base = ((uint64_t)blockIdx.x * 4 + wave) * 16;
    unsigned short(*h1)[NH] = lds_a[wave];
    unsigned short(*h2)[NH] = lds_b[wave];
    // A fragment of the input: sample base + r, features 4q .. 4q+3
    s4 xa;
    {
        const uint64_t si = base + r;
        const bool valid = si < n;
        const uint64_t src = valid ? (perm ? (uint64_t)perm[si] : si) : 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int f = 4 * q + j;
            const float v = (valid && f < MPF_MLP_NIN) ? x[src * MPF_MLP_NIN + f] : 0.0f;
            const unsigned short b = f2bf(v);
            xa[j] = (short)b;
            if (valid) S.x[si * KIN + f] = b;
        }
    }
    float z1[NT][4], z2[NT][4];
    // layer 1: Z1 = X W1 + b1, H1 = gelu(Z1)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(xa, load4(W.w1t + (nt * 16 + r) * KIN + 4 * q), acc);
        const float bias = W.b1[nt * 16 + r];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            z1[nt][j] = acc[j] + bias;
            const unsigned short b = f2bf(gelu(z1[nt][j]));
            h1[4 * q + j][nt * 16 + r] = b;
            if (base + 4 * q + j < n) S.h1[(base + 4 * q + j) * NH + nt * 16 + r] = b;
        }
    }
    __syncthreads();
    // layer 2: Z2 = H1 W2 + b2, H2 = gelu(Z2)
    b8 ha[KT32];
#pragma unroll
    for (int kt = 0; kt < KT32; ++kt) ha[kt] = load8(&h1[r][kt * 32 + 8 * q]);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < KT32; ++kt)
            acc = mfma32(ha[kt], load8(W.w2t + (nt * 16 + r) * NH + kt * 32 + 8 * q), acc);
        const float bias = W.b2[nt * 16 + r];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            z2[nt][j] = acc[j] + bias;
            const unsigned short b = f2bf(gelu(z2[nt][j]));
            h2[4 * q + j][nt * 16 + r] = b;
            if (base + 4 * q + j < n) S.h2[(base + 4 * q + j) * NH + nt * 16 + r] = b;
        }
    }
    __syncthreads();
    // layer 3: Y = H2 W3 + b3 (output r of sample 4q + j)
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < KT32; ++kt)
        acc = mfma32(load8(&h2[r][kt * 32 + 8 * q]), load8(W.w3t + r * NH + kt * 32 + 8 * q), acc);
    const float b3 = W.b3[r];
    // loss = sum over samples and outputs of (y - t)^2; dL/dY scaled by dscale
    float lsum = 0.f;
    unsigned short d3[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint64_t si = base + 4 * q + j;
        float d = 0.f;
        if (si < n && r < MPF_MLP_NOUT) {
            const uint64_t src = perm ? (uint64_t)perm[si] : si;
            const float e = (acc[j] + b3) - t[src * MPF_MLP_NOUT + r];
            lsum += e * e;
            d = 2.0f * e * dscale;
        }
        d3[j] = f2bf(d);
        if (si < n) S.d3[si * KOUT + r] = d3[j];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (l == 0) atomicAdd(loss, lsum);
#pragma unroll
    for (int j = 0; j < 4; ++j) lds_c[wave][4 * q + j][r] = d3[j];
    __syncthreads();   // dZ3 visible; every read of H2 is done
    // dZ2 = (dZ3 W3^T) * gelu'(Z2)
    const s4 da = load4(&lds_c[wave][r][4 * q]);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        f4 g = {0.f, 0.f, 0.f, 0.f};
        g = mfma(da, load4(W.w3 + (nt * 16 + r) * KOUT + 4 * q), g);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const unsigned short b = f2bf(g[j] * gelu_d(z2[nt][j]));
            h2[4 * q + j][nt * 16 + r] = b;
            if (base + 4 * q + j < n) S.d2[(base + 4 * q + j) * NH + nt * 16 + r] = b;
        }
    }
    __syncthreads();
    // dZ1 = (dZ2 W2^T) * gelu'(Z1)
#pragma unroll
    for (int kt = 0; kt < KT32; ++kt) ha[kt] = load8(&h2[r][kt * 32 + 8 * q]);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        f4 g = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < KT32; ++kt) g = mfma32(ha[kt], load8(W.w2 + (nt * 16 + r) * NH + kt * 32 + 8 * q), g);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (base + 4 * q + j < n) S.d1[(base + 4 * q + j) * NH + nt * 16 + r] = f2bf(g[j] * gelu_d(z1[nt][j]));
        }
    }
}

// Weight gradients dW = Act^T dZ over one chunk of WGRAD_CHUNK samples, one
// wave per 16 x 16 tile of dW; bias gradients are the column sums of dZ
// (FP32 sums of the BF16 deltas) from the tiles of the first row.  Each chunk
// writes its own partial vector; reduce_kernel adds the chunks in order, so
// the gradient does not depend on scheduling.
struct GradJob {
    const unsigned short* act;   // n x ka
    const unsigned short* dz;    // n x kz
    int ka, kz;                  // row strides (= the dW dimensions)
    int off_w, off_b;            // offsets in the parameter vector
};

__global__ void __launch_bounds__(64) mlp_wgrad_kernel(GradJob J0, GradJob J1, GradJob J2, uint64_t n,
                                                       float* __restrict__ part) {
    const int t1 = (J0.ka / 16) * (J0.kz / 16);
    const int t2 = t1 + (J1.ka / 16) * (J1.kz / 16);
    const int b = blockIdx.x;
    const GradJob J = (b < t1) ? J0 : ((b < t2) ? J1 : J2);
    const int tile = (b < t1) ? b : ((b < t2) ? b - t1 : b - t2);
    const int ntz = J.kz / 16;
    const int mi = tile / ntz, ni = tile % ntz;
    const int l = threadIdx.x;
    const int r = l & 15, q = l >> 4;
    const uint64_t k_begin = (uint64_t)blockIdx.y * WGRAD_CHUNK;
    const uint64_t k_end = std::min<uint64_t>(n, k_begin + WGRAD_CHUNK);
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;
    // K runs over the chunk's samples, 32 per product (zero rows past its end)
    for (uint64_t k0 = k_begin; k0 < k_end; k0 += 32) {
        unsigned short av[8], bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint64_t s = k0 + 8 * q + j;
            const bool ok = s < k_end;
            av[j] = ok ? J.act[s * J.ka + mi * 16 + r] : (unsigned short)0;
            bv[j] = ok ? J.dz[s * J.kz + ni * 16 + r] : (unsigned short)0;
            bsum += bf2f(bv[j]);
        }
        acc = mfma32(pack8(av), pack8(bv), acc);
    }
    float* out = part + (uint64_t)blockIdx.y * NPARAM;
#pragma unroll
    for (int j = 0; j < 4; ++j) out[J.off_w + (mi * 16 + 4 * q + j) * J.kz + ni * 16 + r] = acc[j];
    if (mi == 0) {
        bsum += __shfl_xor(bsum, 16);
        bsum += __shfl_xor(bsum, 32);
        if (q == 0) out[J.off_b + ni * 16 + r] = bsum;
    }
}

__global__ void reduce_kernel(const float* __restrict__ part, int nchunks, float* __restrict__ g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    float s = 0.f;
    for (int c = 0; c < nchunks; ++c) s += part[(uint64_t)c * NPARAM + i];
    g[i] = s;
}

// Adam (Kingma & Ba 2015, algorithm 1) on the FP32 master parameters
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 operand copies of the master weights
__global__ void pack_kernel(const float* __restrict__ p, unsigned short* __restrict__ w1t,
                            unsigned short* __restrict__ w2t, unsigned short* __restrict__ w2,
                            unsigned short* __restrict__ w3t, unsigned short* __restrict__ w3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < KIN * NH) {
        const int k = i / NH, c = i % NH;
        w1t[c * KIN + k] = f2bf(p[OFF_W1 + i]);
    }
    if (i < NH * NH) {
        const int k = i / NH, c = i % NH;
        const unsigned short b = f2bf(p[OFF_W2 + i]);
        w2[i] = b;
        w2t[c * NH + k] = b;
    }
    if (i < NH * KOUT) {
        const int k = i / KOUT, c = i % KOUT;
        const unsigned short b = f2bf(p[OFF_W3 + i]);
        w3[i] = b;
        w3t[c * NH + k] = b;
    }
}

// 16 x 16 x 16 product through one MFMA: c = bf16(a) bf16(b)
__global__ void mfma_probe_kernel(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ c) {
    const int l = threadIdx.x;
    const int r = l & 15, q = l >> 4;
    s4 av, bv;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        av[j] = (short)f2bf(a[r * 16 + 4 * q + j]);
        bv[j] = (short)f2bf(b[(4 * q + j) * 16 + r]);
    }
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(av, bv, acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) c[(4 * q + j) * 16 + r] = acc[j];
}

// 16 x 16 x 32 product through one MFMA: c = bf16(a) bf16(b), a 16 x 32 and
// b 32 x 16 row-major, with the operand layout the kernels use (split = 0:
// slot j of lane (r, q) holds K = 8q + j in both operands).  A product only
// needs both operands to give slot j the same K; split = 1 gives B's slots
// the K of two halves (4q + j, 16 + 4q + j) while A keeps 8q + j, a
// mismatched pairing that a check of the probe must see fail.
__global__ void mfma_probe32_kernel(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ c,
                                    int split) {
    const int l = threadIdx.x;
    const int r = l & 15, q = l >> 4;
    unsigned short av[8], bv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int ka = 8 * q + j;
        const int kb = split ? ((j < 4) ? 4 * q + j : 16 + 4 * q + (j - 4)) : ka;
        av[j] = f2bf(a[r * 32 + ka]);
        bv[j] = f2bf(b[kb * 16 + r]);
    }
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma32(pack8(av), pack8(bv), acc);
#pragma unroll
    for (int j = 0; j < 4; ++j) c[(4 * q + j) * 16 + r] = acc[j];
}

// Inference: a persistent grid whose blocks hold the BF16 weights and the
// biases in LDS (loaded once) and walk tiles of 64 samples, 16 per wave.
// Features are normalised on the way in (as normalise_kernel does for
// training) and outputs scaled back by the target std and mean on the way
// out, so that a call needs no buffer beyond its input and output.
constexpr int LROW = NH + 8;   // LDS row stride (BF16) of 128-wide rows: spreads 16-byte reads over the banks

__global__ void __launch_bounds__(256) mlp_infer_kernel(Bf16Weights W, const float* __restrict__ norm,
                                                        const float* __restrict__ x, uint64_t n,
                                                        float* __restrict__ y) {
    __shared__ __align__(16) unsigned short w2t[NH][LROW];     // W2^T
    __shared__ __align__(16) unsigned short w3t[KOUT][LROW];   // W3^T
    __shared__ __align__(16) unsigned short w1t[NH][KIN];      // W1^T
    __shared__ __align__(16) unsigned short hb[4][16][LROW];   // per wave: H1, then H2
    __shared__ float bias[2 * NH + KOUT];
    __shared__ float nrm[MPF_MLP_NNORM];
    const int tid = threadIdx.x;
    for (int i = tid; i < NH * NH / 8; i += 256) {
        const int row = i / (NH / 8), c8 = i % (NH / 8);
        *reinterpret_cast<uint4*>(&w2t[row][8 * c8]) = *reinterpret_cast<const uint4*>(W.w2t + row * NH + 8 * c8);
    }
    for (int i = tid; i < KOUT * NH / 8; i += 256) {
        const int row = i / (NH / 8), c8 = i % (NH / 8);
        *reinterpret_cast<uint4*>(&w3t[row][8 * c8]) = *reinterpret_cast<const uint4*>(W.w3t + row * NH + 8 * c8);
    }
    for (int i = tid; i < NH * KIN / 8; i += 256)
        *reinterpret_cast<uint4*>(&w1t[0][0] + 8 * i) = *reinterpret_cast<const uint4*>(W.w1t + 8 * i);
    for (int i = tid; i < NH; i += 256) {
        bias[i] = W.b1[i];
        bias[NH + i] = W.b2[i];
    }
    if (tid < KOUT) bias[2 * NH + tid] = W.b3[tid];
    if (tid < MPF_MLP_NNORM) nrm[tid] = norm[tid];
    __syncthreads();
    const int wave = tid >> 6;
    const int l = tid & 63;
    const int r = l & 15, q = l >> 4;
    unsigned short(*h)[LROW] = hb[wave];
    const uint64_t ngroups = (n + 63) / 64;
    for (uint64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {   // the same trip count for the block's waves
        const uint64_t base = g * 64 + (uint64_t)wave * 16;
        // layer 1 (K = 16): sample base + r, features 4q .. 4q + 3
        s4 xa;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int f = 4 * q + j;
            const uint64_t si = base + r;
            const float v = (si < n && f < MPF_MLP_NIN)
                                ? (x[si * MPF_MLP_NIN + f] - nrm[f]) / nrm[MPF_MLP_NIN + f] : 0.0f;
            xa[j] = (short)f2bf(v);
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(xa, load4(&w1t[nt * 16 + r][4 * q]), acc);
            const float b = bias[nt * 16 + r];
#pragma unroll
            for (int j = 0; j < 4; ++j) h[4 * q + j][nt * 16 + r] = f2bf(gelu(acc[j] + b));
        }
        __syncthreads();
        // layer 2 (K = 128): the whole H1 row of sample r into registers, then H2 over it
        b8 ha[KT32];
#pragma unroll
        for (int kt = 0; kt < KT32; ++kt) ha[kt] = load8(&h[r][kt * 32 + 8 * q]);
        __syncthreads();
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < KT32; ++kt) acc = mfma32(ha[kt], load8(&w2t[nt * 16 + r][kt * 32 + 8 * q]), acc);
            const float b = bias[NH + nt * 16 + r];
#pragma unroll
            for (int j = 0; j < 4; ++j) h[4 * q + j][nt * 16 + r] = f2bf(gelu(acc[j] + b));
        }
        __syncthreads();
        // layer 3 (K = 128): output r of sample 4q + j
        f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < KT32; ++kt) acc = mfma32(load8(&h[r][kt * 32 + 8 * q]), load8(&w3t[r][kt * 32 + 8 * q]), acc);
        if (r < MPF_MLP_NOUT) {
            const float b = bias[2 * NH + r];
            const float sd = nrm[2 * MPF_MLP_NIN + MPF_MLP_NOUT + r], mu = nrm[2 * MPF_MLP_NIN + r];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint64_t si = base + 4 * q + j;
                if (si < n) y[si * MPF_MLP_NOUT + r] = (acc[j] + b) * sd + mu;
            }
        }
        __syncthreads();   // every read of H2 is done before the next tile writes H1
    }
}

// Features of catalogue sources with exact FP64 TSC: one 27-cell stencil
// cache per source, carried outward along the line of sight (the cached tsc()
// of simplehyperoctree.h:1554-1644 that the integrator uses).
struct ExactSampler {
    const DevTree& t;
    exact::TscCache cache;
    __device__ explicit ExactSampler(const DevTree& tree) : t(tree) { exact::cache_init(cache); }
    __device__ void restart() { exact::cache_init(cache); }
    __device__ void operator()(double x, double y, double z, double* g) {
        double d[7];
        int loc;
        exact::tsc<1, true>(t, cache, x, y, z, d, &loc);
        g[0] = d[F_PHI]; g[1] = d[F_DX]; g[2] = d[F_DY]; g[3] = d[F_DZ];
    }
};

__global__ void features_kernel(DevTree t, const double* __restrict__ src, uint64_t n,
                                const uint32_t* __restrict__ order, double length, float* __restrict__ feat) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint64_t i = order ? (uint64_t)order[j] : j;
    ExactSampler sample(t);
    source_features(src + 6 * i, length, sample, feat + (uint64_t)MPF_MLP_NIN * i);
}

__global__ void normalise_kernel(const float* __restrict__ in, uint64_t n, int dim, const float* __restrict__ mean,
                                 const float* __restrict__ stdv, float* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * (uint64_t)dim) return;
    const int d = (int)(i % (uint64_t)dim);
    out[i] = (in[i] - mean[d]) / stdv[d];
}

// first guess = unlensed direction + Born deflection (features 1, 2) + the
// network's correction (arcsec along e_phi and e_theta; none when pred is null)
__global__ void seed_kernel(const double* __restrict__ src, const float* __restrict__ feat,
                            const float* __restrict__ pred, uint64_t n, double* __restrict__ init) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* s = src + 6 * i;
    const double D = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    const float* f = feat + (uint64_t)MPF_MLP_NIN * i;
    const double theta = acos(s[2] / D);
    const double y0 = pred ? (double)pred[2 * i] : 0.0, y1 = pred ? (double)pred[2 * i + 1] : 0.0;
    init[2 * i] = atan2(s[1], s[0]) + ((double)f[1] + y0) / (FEAT_ARCSEC * sin(theta));
    init[2 * i + 1] = theta + ((double)f[2] + y1) / FEAT_ARCSEC;
}

__global__ void zero_f32(float* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

inline unsigned nblk(uint64_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

template <class T>
struct Tmp {  // device scratch owned by one call
    T* p = nullptr;
    Tmp() = default;
    Tmp(const Tmp&) = delete;
    Tmp& operator=(const Tmp&) = delete;
    ~Tmp() { if (p) (void)hipFree(p); }
    bool alloc(uint64_t n) { return hipMalloc((void**)&p, std::max<uint64_t>(n, 1) * sizeof(T)) == hipSuccess; }
};

// buffers of one training or gradient call
struct Batch {
    Tmp<float> x, t, xn, tn, loss, part;
    Tmp<uint32_t> perm;
    Tmp<unsigned short> saved;
    Saved S{};
    bool alloc(uint64_t n, uint64_t batch) {
        const uint64_t per = KIN + 4 * NH + KOUT;
        if (!x.alloc(n * MPF_MLP_NIN) || !t.alloc(n * MPF_MLP_NOUT) || !xn.alloc(n * MPF_MLP_NIN) ||
            !tn.alloc(n * MPF_MLP_NOUT) || !loss.alloc(1) || !perm.alloc(n) || !saved.alloc(batch * per) ||
            !part.alloc(((batch + WGRAD_CHUNK - 1) / WGRAD_CHUNK) * (uint64_t)NPARAM))
            return false;
        S.x = saved.p;
        S.h1 = S.x + batch * KIN;
        S.h2 = S.h1 + batch * NH;
        S.d1 = S.h2 + batch * NH;
        S.d2 = S.d1 + batch * NH;
        S.d3 = S.d2 + batch * NH;
        return true;
    }
};

}  // namespace

// ------------------------------------------------------------------ model

struct MlpModel {
    hipStream_t stream = nullptr;
    float* p = nullptr;      // NPARAM FP32 master parameters
    float* g = nullptr;      // gradient
    float* m = nullptr;      // Adam moments
    float* v = nullptr;
    unsigned short* wb = nullptr;   // BF16 operand copies
    float* norm = nullptr;   // in mean[10], in std[10], out mean[2], out std[2]
    float hnorm[MPF_MLP_NNORM];
    bool norm_set = false;
    long step = 0;
    unsigned short *w1t = nullptr, *w2t = nullptr, *w2 = nullptr, *w3t = nullptr, *w3 = nullptr;
    int cus = 0;                          // compute units of the device (inference grid)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = -1.f;                 // duration of the last inference kernel
};

static Bf16Weights weights_of(const MlpModel& M) {
    Bf16Weights W;
    W.w1t = M.w1t; W.w2t = M.w2t; W.w2 = M.w2; W.w3t = M.w3t; W.w3 = M.w3;
    W.b1 = M.p + OFF_B1; W.b2 = M.p + OFF_B2; W.b3 = M.p + OFF_B3;
    return W;
}

static void pack(MlpModel* M) {
    hipLaunchKernelGGL(pack_kernel, dim3(nblk(NH * NH)), dim3(256), 0, M->stream, M->p, M->w1t, M->w2t, M->w2,
                       M->w3t, M->w3);
}

void mlp_destroy(MlpModel* M) {
    if (!M) return;
    for (void* q : {(void*)M->p, (void*)M->g, (void*)M->m, (void*)M->v, (void*)M->wb, (void*)M->norm})
        if (q) (void)hipFree(q);
    if (M->ev0) (void)hipEventDestroy(M->ev0);
    if (M->ev1) (void)hipEventDestroy(M->ev1);
    delete M;
}

bool mlp_create(MlpModel** out, hipStream_t st, uint64_t seed, std::string& err) {
    MlpModel* M = new MlpModel;
    M->stream = st;
    const size_t nbw = (size_t)KIN * NH + 2 * NH * NH + 2 * NH * KOUT;
    int dev = 0;
    if (hipMalloc((void**)&M->p, NPARAM * 4) || hipMalloc((void**)&M->g, NPARAM * 4) ||
        hipMalloc((void**)&M->m, NPARAM * 4) || hipMalloc((void**)&M->v, NPARAM * 4) ||
        hipMalloc((void**)&M->wb, nbw * 2) || hipMalloc((void**)&M->norm, MPF_MLP_NNORM * 4) ||
        hipEventCreate(&M->ev0) || hipEventCreate(&M->ev1) || hipGetDevice(&dev) ||
        hipDeviceGetAttribute(&M->cus, hipDeviceAttributeMultiprocessorCount, dev)) {
        err = "MLP allocation failed";
        mlp_destroy(M);
        return false;
    }
    M->w1t = M->wb;
    M->w2t = M->w1t + KIN * NH;
    M->w2 = M->w2t + NH * NH;
    M->w3t = M->w2 + NH * NH;
    M->w3 = M->w3t + NH * KOUT;
    // Glorot-uniform hidden weights over the real fan-in/fan-out; zero output
    // layer (the untrained network adds the target mean: a residual start),
    // zero biases and pads
    std::vector<float> h(NPARAM, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int rows, int cols, int real_rows, int real_cols) {
        const float a = std::sqrt(6.0f / (float)(real_rows + real_cols));
        std::uniform_real_distribution<float> u(-a, a);
        for (int i = 0; i < rows; ++i)
            for (int j = 0; j < cols; ++j) h[off + i * cols + j] = (i < real_rows && j < real_cols) ? u(rng) : 0.f;
    };
    fill(OFF_W1, KIN, NH, MPF_MLP_NIN, NH);
    fill(OFF_W2, NH, NH, NH, NH);
    for (int i = 0; i < MPF_MLP_NNORM; ++i) M->hnorm[i] = 0.f;
    for (int i = 0; i < MPF_MLP_NIN; ++i) M->hnorm[MPF_MLP_NIN + i] = 1.f;
    for (int i = 0; i < MPF_MLP_NOUT; ++i) M->hnorm[2 * MPF_MLP_NIN + MPF_MLP_NOUT + i] = 1.f;
    std::vector<float> zeros(NPARAM, 0.f);
    if (hipMemcpy(M->p, h.data(), NPARAM * 4, hipMemcpyHostToDevice) ||
        hipMemcpy(M->m, zeros.data(), NPARAM * 4, hipMemcpyHostToDevice) ||
        hipMemcpy(M->v, zeros.data(), NPARAM * 4, hipMemcpyHostToDevice) ||
        hipMemcpy(M->g, zeros.data(), NPARAM * 4, hipMemcpyHostToDevice) ||
        hipMemcpy(M->norm, M->hnorm, MPF_MLP_NNORM * 4, hipMemcpyHostToDevice)) {
        err = "MLP initialisation failed";
        mlp_destroy(M);
        return false;
    }
    pack(M);
    if (hipStreamSynchronize(st) != hipSuccess) {
        err = "MLP weight packing failed";
        mlp_destroy(M);
        return false;
    }
    *out = M;
    return true;
}

bool mlp_state(MlpModel* M, float* get, const float* set, std::string& err) {
    if (set) {
        if (hipStreamSynchronize(M->stream) || hipMemcpy(M->p, set, NPARAM * 4, hipMemcpyHostToDevice) ||
            hipMemcpy(M->norm, set + NPARAM, MPF_MLP_NNORM * 4, hipMemcpyHostToDevice)) {
            err = "MLP state upload failed";
            return false;
        }
        memcpy(M->hnorm, set + NPARAM, MPF_MLP_NNORM * 4);
        M->norm_set = true;
        pack(M);
        if (hipStreamSynchronize(M->stream) != hipSuccess) { err = "MLP weight packing failed"; return false; }
    }
    if (get) {
        if (hipStreamSynchronize(M->stream) || hipMemcpy(get, M->p, NPARAM * 4, hipMemcpyDeviceToHost)) {
            err = "MLP state download failed";
            return false;
        }
        memcpy(get + NPARAM, M->hnorm, MPF_MLP_NNORM * 4);
    }
    return true;
}

// per-dimension mean and standard deviation of the training set
static bool set_norm_from(MlpModel* M, const float* feat, const float* target, uint64_t n) {
    for (int d = 0; d < MPF_MLP_NIN + MPF_MLP_NOUT; ++d) {
        const bool in = d < MPF_MLP_NIN;
        const float* a = in ? feat : target;
        const int dim = in ? MPF_MLP_NIN : MPF_MLP_NOUT;
        const int k = in ? d : d - MPF_MLP_NIN;
        double s = 0, s2 = 0;
        for (uint64_t i = 0; i < n; ++i) {
            const double x = a[i * dim + k];
            s += x;
            s2 += x * x;
        }
        const double mu = s / (double)n;
        double sd = std::sqrt(std::max(s2 / (double)n - mu * mu, 0.0));
        if (!(sd > 1e-30)) sd = 1.0;
        if (in) {
            M->hnorm[k] = (float)mu;
            M->hnorm[MPF_MLP_NIN + k] = (float)sd;
        } else {
            M->hnorm[2 * MPF_MLP_NIN + k] = (float)mu;
            M->hnorm[2 * MPF_MLP_NIN + MPF_MLP_NOUT + k] = (float)sd;
        }
    }
    M->norm_set = true;
    return hipMemcpy(M->norm, M->hnorm, MPF_MLP_NNORM * 4, hipMemcpyHostToDevice) == hipSuccess;
}

// loss sum and gradient (into M->g) of rows perm[0 .. nb) of the normalised set
// (the loss is added to B.loss, which the caller clears)
static void batch_gradient(MlpModel* M, Batch& B, const uint32_t* perm, uint64_t nb, float dscale) {
    hipStream_t st = M->stream;
    const Bf16Weights W = weights_of(*M);
    hipLaunchKernelGGL(mlp_train_tile_kernel, dim3(nblk(nb, 64)), dim3(256), 0, st, W, B.xn.p, perm, nb, B.tn.p, B.S,
                       dscale, B.loss.p);
    const GradJob j1{B.S.x, B.S.d1, KIN, NH, OFF_W1, OFF_B1};
    const GradJob j2{B.S.h1, B.S.d2, NH, NH, OFF_W2, OFF_B2};
    const GradJob j3{B.S.h2, B.S.d3, NH, KOUT, OFF_W3, OFF_B3};
    const int tiles = (KIN / 16) * (NH / 16) + (NH / 16) * (NH / 16) + (NH / 16) * (KOUT / 16);
    const int chunks = (int)((nb + WGRAD_CHUNK - 1) / WGRAD_CHUNK);
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3(tiles, chunks), dim3(64), 0, st, j1, j2, j3, nb, B.part.p);
    hipLaunchKernelGGL(reduce_kernel, dim3(nblk(NPARAM)), dim3(256), 0, st, B.part.p, chunks, M->g);
}

static bool upload_set(MlpModel* M, Batch& B, const float* feat, const float* target, uint64_t n) {
    hipStream_t st = M->stream;
    if (hipMemcpyAsync(B.x.p, feat, n * MPF_MLP_NIN * 4, hipMemcpyHostToDevice, st) ||
        hipMemcpyAsync(B.t.p, target, n * MPF_MLP_NOUT * 4, hipMemcpyHostToDevice, st))
        return false;
    hipLaunchKernelGGL(normalise_kernel, dim3(nblk(n * MPF_MLP_NIN)), dim3(256), 0, st, B.x.p, n, MPF_MLP_NIN,
                       M->norm, M->norm + MPF_MLP_NIN, B.xn.p);
    hipLaunchKernelGGL(normalise_kernel, dim3(nblk(n * MPF_MLP_NOUT)), dim3(256), 0, st, B.t.p, n, MPF_MLP_NOUT,
                       M->norm + 2 * MPF_MLP_NIN, M->norm + 2 * MPF_MLP_NIN + MPF_MLP_NOUT, B.tn.p);
    return hipGetLastError() == hipSuccess;
}

bool mlp_train(MlpModel* M, const float* feat, const float* target, uint64_t n, uint32_t epochs, uint32_t batch,
               float lr, uint64_t seed, float* loss_per_epoch, std::string& err) {
    if (n == 0 || batch == 0) { err = "empty training set or batch"; return false; }
    if (n > 0xFFFFFFFFull) { err = "training set larger than 2^32 samples"; return false; }
    if (!M->norm_set && !set_norm_from(M, feat, target, n)) { err = "MLP normalisation upload failed"; return false; }
    hipStream_t st = M->stream;
    const uint64_t nbm = std::min<uint64_t>(batch, n);
    Batch B;
    if (!B.alloc(n, nbm)) { err = "MLP training buffers: allocation failed"; return false; }
    if (!upload_set(M, B, feat, target, n)) { err = "MLP training set upload failed"; return false; }
    std::vector<uint32_t> perm(n);
    for (uint64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
    std::mt19937_64 rng(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), rng);
        // the host vector is reshuffled next epoch: wait for the copy
        if (hipMemcpyAsync(B.perm.p, perm.data(), n * 4, hipMemcpyHostToDevice, st) || hipStreamSynchronize(st)) {
            err = "MLP permutation upload failed";
            return false;
        }
        // the epoch's loss sums on the device and is read once
        hipLaunchKernelGGL(zero_f32, dim3(1), dim3(64), 0, st, B.loss.p, 1);
        for (uint64_t b0 = 0; b0 < n; b0 += nbm) {
            const uint64_t nb = std::min<uint64_t>(nbm, n - b0);
            batch_gradient(M, B, B.perm.p + b0, nb, 1.0f / (float)(nb * MPF_MLP_NOUT));
            M->step++;
            const float c1 = 1.0f - std::pow(b1, (float)M->step), c2 = 1.0f - std::pow(b2, (float)M->step);
            hipLaunchKernelGGL(adam_kernel, dim3(nblk(NPARAM)), dim3(256), 0, st, M->p, M->g, M->m, M->v, lr, b1, b2,
                               eps, c1, c2);
            pack(M);
        }
        if (loss_per_epoch) {
            float l = 0.f;
            if (hipMemcpyAsync(&l, B.loss.p, 4, hipMemcpyDeviceToHost, st) || hipStreamSynchronize(st)) {
                err = "MLP loss download failed";
                return false;
            }
            loss_per_epoch[ep] = (float)((double)l / (double)(n * MPF_MLP_NOUT));
        }
    }
    if (hipStreamSynchronize(st) != hipSuccess) { err = "MLP training failed"; return false; }
    return true;
}

bool mlp_gradient(MlpModel* M, const float* feat, const float* target, uint64_t n, float* grad, float* loss,
                  std::string& err) {
    if (n == 0) { err = "empty batch"; return false; }
    if (!M->norm_set && !set_norm_from(M, feat, target, n)) { err = "MLP normalisation upload failed"; return false; }
    Batch B;
    if (!B.alloc(n, n)) { err = "MLP gradient buffers: allocation failed"; return false; }
    if (!upload_set(M, B, feat, target, n)) { err = "MLP batch upload failed"; return false; }
    hipLaunchKernelGGL(zero_f32, dim3(1), dim3(64), 0, M->stream, B.loss.p, 1);
    batch_gradient(M, B, nullptr, n, 1.0f / (float)(n * MPF_MLP_NOUT));
    float l = 0.f;
    if (hipMemcpyAsync(grad, M->g, NPARAM * 4, hipMemcpyDeviceToHost, M->stream) ||
        hipMemcpyAsync(&l, B.loss.p, 4, hipMemcpyDeviceToHost, M->stream) || hipStreamSynchronize(M->stream)) {
        err = "MLP gradient failed";
        return false;
    }
    if (loss) *loss = l / (float)(n * MPF_MLP_NOUT);
    return true;
}

bool mlp_infer_dev(MlpModel* M, const float* dfeat, uint64_t n, float* dout, std::string& err) {
    if (n == 0) return true;
    hipStream_t st = M->stream;
    // two 256-lane blocks per compute unit (60 KB of LDS each), or fewer for a small batch
    const uint64_t groups = (n + 63) / 64;
    const unsigned grid = (unsigned)std::min<uint64_t>(groups, 2ull * (uint64_t)std::max(M->cus, 1));
    if (hipEventRecord(M->ev0, st) != hipSuccess) { err = "MLP inference: event"; return false; }
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(grid), dim3(256), 0, st, weights_of(*M), M->norm, dfeat, n, dout);
    if (hipGetLastError() != hipSuccess || hipEventRecord(M->ev1, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        err = "MLP inference failed";
        return false;
    }
    float ms = -1.f;
    if (hipEventElapsedTime(&ms, M->ev0, M->ev1) == hipSuccess) M->last_ms = ms;
    return true;
}

float mlp_last_ms(const MlpModel* M) { return M ? M->last_ms : -1.f; }

hipError_t launch_features(const DevTree& t, const double* src, uint64_t n, const uint32_t* order, double length,
                           float* feat, hipStream_t st) {
    if (n) hipLaunchKernelGGL(features_kernel, dim3(nblk(n)), dim3(256), 0, st, t, src, n, order, length, feat);
    return hipGetLastError();
}

hipError_t launch_seed(const double* src, const float* feat, const float* pred, uint64_t n, double* init,
                       hipStream_t st) {
    if (n) hipLaunchKernelGGL(seed_kernel, dim3(nblk(n)), dim3(256), 0, st, src, feat, pred, n, init);
    return hipGetLastError();
}

hipError_t launch_mfma_probe(const float* a, const float* b, float* c, hipStream_t st) {
    hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, st, a, b, c);
    return hipGetLastError();
}

hipError_t launch_mfma_probe32(const float* a, const float* b, float* c, int split, hipStream_t st) {
    hipLaunchKernelGGL(mfma_probe32_kernel, dim3(1), dim3(64), 0, st, a, b, c, split);
    return hipGetLastError();
}

}  // namespace mpf
