// Irradiance surrogate kernels (gfx950, BF16 MFMA 32x32x8 with FP32
// accumulation).  No reference counterpart: the network stands in for
// RadiosityFunction::GatherLight (core/lighting/radiosity.cpp:509-743) at
// cache misses, trained on the gathers the plain path makes.
//
// Layout trick: every layer is computed transposed, Y^T = W^T X^T, with the
// 32 points of a tile as the MFMA N dimension.  The 32x32 accumulator of a
// lane then holds, for its point (lane % 32), neurons 8q + 4(lane/32) + r,
// which is exactly the B operand the next layer's K step q needs — so the
// activations stay in registers from layer to layer, forward and backward.
// Weight gradients (sums over points) are separate GEMMs over the stored
// activations, split over chunks of 128 points and reduced in fixed order
// (deterministic), followed by Adam on the FP32 master weights, which also
// rewrites the BF16 copies the kernels read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"
#include "mlp_internal.h"

using namespace pvx;

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// BF16 weight copies, element offsets (rows padded so that 4-element loads
// stay 8-byte aligned and rows start on different LDS banks).
constexpr int kP1 = kMlpIn + 4;                   // W1T row pitch
constexpr int kP = kMlpHid + 4;                   // 64-wide row pitch
constexpr int kOffP_W1T = 0;                      // [64 out][192 in]
constexpr int kOffP_W2T = kOffP_W1T + 64 * kP1;   // [64 out][64 in]
constexpr int kOffP_W3T = kOffP_W2T + 64 * kP;
constexpr int kOffP_W4T = kOffP_W3T + 64 * kP;    // [32 out (3 used)][64 in]
constexpr int kFwdElems = kOffP_W4T + 32 * kP;    // forward pass needs these
constexpr int kOffP_W2N = kFwdElems;              // [64 in][64 out]
constexpr int kOffP_W3N = kOffP_W2N + 64 * kP;
constexpr int kOffP_W4N = kOffP_W3N + 64 * kP;    // [64 in][8 out (3 used)]
constexpr int kAllElems = kOffP_W4N + 64 * 8;
constexpr int kBiasFloats = 3 * kMlpHid + 4;      // b1 b2 b3 b4(+pad)
constexpr size_t kPackBytes = (size_t)kAllElems * 2 + kBiasFloats * 4;
static_assert((kFwdElems * 2) % 16 == 0 && (kAllElems * 2) % 16 == 0, "16-byte LDS copies");

// Stored activations per batch slot (rows of `batch` BF16 values each).
constexpr int kRowX0 = 0;                 // 192 feature rows
constexpr int kRowX1 = kRowX0 + kMlpIn;   // h1
constexpr int kRowX2 = kRowX1 + kMlpHid;  // h2
constexpr int kRowX3 = kRowX2 + kMlpHid;  // h3
constexpr int kRowZ1 = kRowX3 + kMlpHid;  // dz1
constexpr int kRowZ2 = kRowZ1 + kMlpHid;
constexpr int kRowZ3 = kRowZ2 + kMlpHid;
constexpr int kRowDY = kRowZ3 + kMlpHid;  // dy (3 rows used, 4 reserved)
constexpr int kActRows = kRowDY + 4;

constexpr int kWgradItems = 12 + 4 + 4 + 2; // 32x32 gradient tiles of W1..W4

__device__ inline uint16_t f2bf(float f) { // round to nearest even (finite inputs)
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ inline f32x16 mfma(bf16x4 a, bf16x4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

__device__ inline f32x16 zero16() {
    f32x16 z;
    for (int g = 0; g < 16; g++) z[g] = 0.0f;
    return z;
}

__device__ inline bf16x4 lds4(const uint16_t *base, int off) {
    return *reinterpret_cast<const bf16x4 *>(base + off);
}

// Frequency k of the encoding: 2^(k/3), k = 0..15 (1 .. 32 periods over
// the unit range): the FP32 value of 2^((k mod 3)/3) scaled exactly by
// 2^(k div 3).
__device__ inline float freq(int k) {
    const int q = k / 3, r = k - 3 * q;
    const float base = r == 0 ? 1.0f : (r == 1 ? 1.25992104989487316f : 1.58740105196819947f);
    return base * (float)(1u << q);
}

// Features 8s + 4h .. 8s + 4h + 3 of one point: dimension s / 4, frequencies
// k0 and k0 + 1 (k0 = 4 (s % 4) + 2h), each as (sin, cos) of pi f x.  The
// normal's dimensions (3..5) use a quarter of the position frequencies:
// irradiance varies smoothly with the normal.
__device__ inline bf16x4 feat4(const float *x, int s, int h) {
    const int j = s >> 2, k0 = 4 * (s & 3) + 2 * h;
    const float sc = j >= 3 ? 0.25f : 1.0f;
    float s0, c0, s1, c1;
    sincospif((freq(k0) * sc) * x[j], &s0, &c0);
    sincospif((freq(k0 + 1) * sc) * x[j], &s1, &c1);
    bf16x4 r;
    r[0] = (short)f2bf(s0);
    r[1] = (short)f2bf(c0);
    r[2] = (short)f2bf(s1);
    r[3] = (short)f2bf(c1);
    return r;
}

__device__ inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Cooperative 16-byte copy of the packed weights (and biases) into LDS.
__device__ inline void load_pack(const uint8_t *__restrict__ pack, uint16_t *sw, int welems, float *sb) {
    const uint4 *src = reinterpret_cast<const uint4 *>(pack);
    uint4 *dst = reinterpret_cast<uint4 *>(sw);
    for (int i = threadIdx.x; i < welems * 2 / 16; i += blockDim.x) dst[i] = src[i];
    const float *bsrc = reinterpret_cast<const float *>(pack + (size_t)kAllElems * 2);
    for (int i = threadIdx.x; i < kBiasFloats; i += blockDim.x) sb[i] = bsrc[i];
}

// One hidden layer forward: acc_mb = W^T[mb*32 + p][.] * B, 64 -> 64.
__device__ inline void hidden_fwd(const uint16_t *sw, int offT, const bf16x4 *Bin, int p, int h,
                                  f32x16 &a0, f32x16 &a1) {
    a0 = zero16();
    a1 = zero16();
    for (int s = 0; s < 8; s++) {
        a0 = mfma(lds4(sw, offT + p * kP + 8 * s + 4 * h), Bin[s], a0);
        a1 = mfma(lds4(sw, offT + (32 + p) * kP + 8 * s + 4 * h), Bin[s], a1);
    }
}

// Bias + ReLU; returns the B operands of the next layer in `out` and stores
// the activations (row = neuron) for the weight-gradient GEMM when `st` is set.
__device__ inline void bias_relu(const f32x16 &a0, const f32x16 &a1, const float *bias, int h, bf16x4 *out,
                                 uint16_t *st, uint32_t batch, uint32_t slot) {
    for (int g = 0; g < 16; g++) {
        const int i = 8 * (g >> 2) + 4 * h + (g & 3);
        uint16_t v0 = f2bf(fmaxf(a0[g] + bias[i], 0.0f));
        uint16_t v1 = f2bf(fmaxf(a1[g] + bias[32 + i], 0.0f));
        out[g >> 2][g & 3] = (short)v0;
        out[4 + (g >> 2)][g & 3] = (short)v1;
        if (st) {
            st[(size_t)i * batch + slot] = v0;
            st[(size_t)(32 + i) * batch + slot] = v1;
        }
    }
}

// dz = dh * [h > 0], as next B operands, stored as rows for the gradient GEMM.
__device__ inline void relu_back(const f32x16 &a0, const f32x16 &a1, const bf16x4 *H, int h, bf16x4 *out,
                                 uint16_t *st, uint32_t batch, uint32_t slot) {
    for (int g = 0; g < 16; g++) {
        const int i = 8 * (g >> 2) + 4 * h + (g & 3);
        uint16_t v0 = H[g >> 2][g & 3] > 0 ? f2bf(a0[g]) : (uint16_t)0; // h >= 0 (post ReLU, BF16 bits)
        uint16_t v1 = H[4 + (g >> 2)][g & 3] > 0 ? f2bf(a1[g]) : (uint16_t)0;
        out[g >> 2][g & 3] = (short)v0;
        out[4 + (g >> 2)][g & 3] = (short)v1;
        st[(size_t)i * batch + slot] = v0;
        st[(size_t)(32 + i) * batch + slot] = v1;
    }
}

// Features of n points, [point][192] BF16 (the B-operand order of feat4),
// computed once per training set instead of once per step.
__global__ void k_mlp_features(const float *__restrict__ x, uint32_t n, uint16_t *__restrict__ F) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * 48u) return;
    const uint32_t pt = i / 48u, sh = i % 48u, s = sh >> 1, h = sh & 1u;
    float xv[6];
    for (int c = 0; c < 6; c++) xv[c] = x[6 * (size_t)pt + c];
    *reinterpret_cast<bf16x4 *>(F + (size_t)pt * kMlpIn + 8 * s + 4 * h) = feat4(xv, (int)s, (int)h);
}

struct TrainArgs {
    const uint8_t *pack;
    const float *x; // n x 6
    const uint16_t *F; // n x 192 precomputed features (or null: computed from x)
    const float *t; // n x 3
    uint32_t n;
    uint32_t batch;  // multiple of kMlpChunk
    uint32_t sample; // 1: slot -> hashed point index; 0: slot = point
    uint32_t seed;
    float dscale;    // 2 / (points in the loss)
    uint16_t *acts;  // kActRows x batch
    float *loss_part;// per tile
};

// Forward + backward of one 32-point tile per wavefront.
__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(TrainArgs a) {
    __shared__ __align__(16) uint16_t sw[kAllElems];
    __shared__ float sb[kBiasFloats];
    load_pack(a.pack, sw, kAllElems, sb);
    __syncthreads();
    const int lane = threadIdx.x & 63, p = lane & 31, h = lane >> 5;
    const uint32_t tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile * 32 >= a.batch) return;
    const uint32_t slot = tile * 32 + p;
    uint32_t src = slot;
    bool valid = slot < a.n;
    if (a.sample) {
        src = mix32(a.seed * 0x9E3779B9u + slot) % a.n;
        valid = true;
    }
    float x[6];
    if (!a.F)
        for (int c = 0; c < 6; c++) x[c] = valid ? a.x[6 * (size_t)src + c] : 0.0f;
    const uint16_t *Fp = a.F ? a.F + (size_t)(valid ? src : 0) * kMlpIn : nullptr;
    uint16_t *A = a.acts;
    const uint32_t B = a.batch;
    // layer 1 (features loaded, or computed in place)
    f32x16 a0 = zero16(), a1 = zero16();
    for (int s = 0; s < 24; s++) {
        bf16x4 b;
        if (Fp) {
            b = *reinterpret_cast<const bf16x4 *>(Fp + 8 * s + 4 * h);
            if (!valid)
                for (int t = 0; t < 4; t++) b[t] = 0;
        } else {
            b = feat4(x, s, h);
        }
        for (int t = 0; t < 4; t++) A[(size_t)(kRowX0 + 8 * s + 4 * h + t) * B + slot] = (uint16_t)b[t];
        a0 = mfma(lds4(sw, kOffP_W1T + p * kP1 + 8 * s + 4 * h), b, a0);
        a1 = mfma(lds4(sw, kOffP_W1T + (32 + p) * kP1 + 8 * s + 4 * h), b, a1);
    }
    bf16x4 H1[8], H2[8], H3[8];
    bias_relu(a0, a1, sb, h, H1, A + (size_t)kRowX1 * B, B, slot);
    hidden_fwd(sw, kOffP_W2T, H1, p, h, a0, a1);
    bias_relu(a0, a1, sb + 64, h, H2, A + (size_t)kRowX2 * B, B, slot);
    hidden_fwd(sw, kOffP_W3T, H2, p, h, a0, a1);
    bias_relu(a0, a1, sb + 128, h, H3, A + (size_t)kRowX3 * B, B, slot);
    f32x16 ao = zero16();
    for (int s = 0; s < 8; s++) ao = mfma(lds4(sw, kOffP_W4T + p * kP + 8 * s + 4 * h), H3[s], ao);
    // loss and output gradient (rows 0..2 of the output tile live in lanes 0..31)
    float l2 = 0.0f;
    bf16x4 DY;
    for (int t = 0; t < 4; t++) DY[t] = 0;
    if (h == 0) {
        for (int c = 0; c < 3; c++) {
            float y = ao[c] + sb[192 + c];
            float e = valid ? y - a.t[3 * (size_t)src + c] : 0.0f;
            l2 += e * e;
            uint16_t d = f2bf(a.dscale * e);
            DY[c] = (short)d;
            A[(size_t)(kRowDY + c) * B + slot] = d;
        }
    }
    for (int off = 32; off > 0; off >>= 1) l2 += __shfl_xor(l2, off, 64);
    if (lane == 0) a.loss_part[tile] = l2;
    // backward: dh3 = W4 dy
    a0 = mfma(lds4(sw, kOffP_W4N + p * 8 + 4 * h), DY, zero16());
    a1 = mfma(lds4(sw, kOffP_W4N + (32 + p) * 8 + 4 * h), DY, zero16());
    bf16x4 Z[8];
    relu_back(a0, a1, H3, h, Z, A + (size_t)kRowZ3 * B, B, slot);
    // dh2 = W3 dz3
    a0 = zero16();
    a1 = zero16();
    for (int s = 0; s < 8; s++) {
        a0 = mfma(lds4(sw, kOffP_W3N + p * kP + 8 * s + 4 * h), Z[s], a0);
        a1 = mfma(lds4(sw, kOffP_W3N + (32 + p) * kP + 8 * s + 4 * h), Z[s], a1);
    }
    relu_back(a0, a1, H2, h, Z, A + (size_t)kRowZ2 * B, B, slot);
    // dh1 = W2 dz2
    a0 = zero16();
    a1 = zero16();
    for (int s = 0; s < 8; s++) {
        a0 = mfma(lds4(sw, kOffP_W2N + p * kP + 8 * s + 4 * h), Z[s], a0);
        a1 = mfma(lds4(sw, kOffP_W2N + (32 + p) * kP + 8 * s + 4 * h), Z[s], a1);
    }
    relu_back(a0, a1, H1, h, Z, A + (size_t)kRowZ1 * B, B, slot);
}

// Weight-gradient GEMMs: dW[in][out] = sum_p X[in][p] dZ[out][p] over one
// chunk of 128 points per wavefront; bias gradients ride along as a
// ones-row A operand on the m-block-0 tiles.  Partial sums per chunk.
__global__ __launch_bounds__(256) void k_mlp_wgrad(const uint16_t *__restrict__ acts, uint32_t batch,
                                                   uint32_t nchunks, float *__restrict__ partial) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (wave >= kWgradItems * nchunks) return;
    const int lane = threadIdx.x & 63, p = lane & 31, h = lane >> 5;
    const uint32_t item = wave % kWgradItems, chunk = wave / kWgradItems;
    int xrow, zrow, mb, nb, outdim, offW, offB;
    if (item < 12) {
        xrow = kRowX0; zrow = kRowZ1; mb = item / 2; nb = item % 2; outdim = 64; offW = kOffW1; offB = kOffB1;
    } else if (item < 16) {
        xrow = kRowX1; zrow = kRowZ2; mb = (item - 12) / 2; nb = (item - 12) % 2; outdim = 64;
        offW = kOffW2; offB = kOffB2;
    } else if (item < 20) {
        xrow = kRowX2; zrow = kRowZ3; mb = (item - 16) / 2; nb = (item - 16) % 2; outdim = 64;
        offW = kOffW3; offB = kOffB3;
    } else {
        xrow = kRowX3; zrow = kRowDY; mb = item - 20; nb = 0; outdim = 3; offW = kOffW4; offB = kOffB4;
    }
    const uint16_t *X = acts + (size_t)(xrow + mb * 32 + p) * batch;
    const int j = nb * 32 + p;
    const bool jv = j < outdim;
    const uint16_t *Z = acts + (size_t)(zrow + (jv ? j : 0)) * batch;
    bf16x4 ones, zero;
    for (int t = 0; t < 4; t++) {
        ones[t] = (short)0x3F80;
        zero[t] = 0;
    }
    f32x16 acc = zero16(), accb = zero16();
    const uint32_t p0 = chunk * kMlpChunk;
    for (int s = 0; s < kMlpChunk / 8; s++) {
        const uint32_t kp = p0 + 8 * s + 4 * h;
        bf16x4 av = *reinterpret_cast<const bf16x4 *>(X + kp);
        bf16x4 bv = jv ? *reinterpret_cast<const bf16x4 *>(Z + kp) : zero;
        acc = mfma(av, bv, acc);
        if (mb == 0) accb = mfma(ones, bv, accb);
    }
    float *out = partial + (size_t)chunk * kMlpParams;
    if (jv) {
        for (int g = 0; g < 16; g++) {
            const int i = mb * 32 + 8 * (g >> 2) + 4 * h + (g & 3);
            out[offW + i * outdim + j] = acc[g];
        }
        if (mb == 0 && h == 0) out[offB + j] = accb[0];
    }
}

// Writes the BF16 / bias copies of parameter i.
__device__ inline void pack_param(uint8_t *pack, int i, float w) {
    uint16_t *pw = reinterpret_cast<uint16_t *>(pack);
    float *pb = reinterpret_cast<float *>(pack + (size_t)kAllElems * 2);
    const uint16_t b = f2bf(w);
    if (i < kOffB1) {
        int in = i / 64, out = i % 64;
        pw[kOffP_W1T + out * kP1 + in] = b;
    } else if (i < kOffW2) {
        pb[i - kOffB1] = w;
    } else if (i < kOffB2) {
        int in = (i - kOffW2) / 64, out = (i - kOffW2) % 64;
        pw[kOffP_W2T + out * kP + in] = b;
        pw[kOffP_W2N + in * kP + out] = b;
    } else if (i < kOffW3) {
        pb[64 + i - kOffB2] = w;
    } else if (i < kOffB3) {
        int in = (i - kOffW3) / 64, out = (i - kOffW3) % 64;
        pw[kOffP_W3T + out * kP + in] = b;
        pw[kOffP_W3N + in * kP + out] = b;
    } else if (i < kOffW4) {
        pb[128 + i - kOffB3] = w;
    } else if (i < kOffB4) {
        int in = (i - kOffW4) / 3, out = (i - kOffW4) % 3;
        pw[kOffP_W4T + out * kP + in] = b;
        pw[kOffP_W4N + in * 8 + out] = b;
    } else {
        pb[192 + i - kOffB4] = w;
    }
}

enum { kUpdPack = 0, kUpdGrad = 1, kUpdAdam = 2 };

// Reduces the chunk partials of every parameter (fixed order) and then packs
// (mode 0), writes the gradient (mode 1) or takes one Adam step and packs
// (mode 2).
__global__ void k_mlp_update(int mode, const float *__restrict__ partial, uint32_t nchunks, float *__restrict__ w,
                             float *__restrict__ m, float *__restrict__ v, float *__restrict__ grad,
                             uint8_t *__restrict__ pack, float lr, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    float wi = w[i];
    if (mode != kUpdPack) {
        float g = 0.0f;
        for (uint32_t c = 0; c < nchunks; c++) g += partial[(size_t)c * kMlpParams + i];
        if (mode == kUpdGrad) {
            grad[i] = g;
            return;
        }
        const float b1 = 0.9f, b2 = 0.99f, eps = 1e-8f;
        float mi = b1 * m[i] + (1.0f - b1) * g;
        float vi = b2 * v[i] + (1.0f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        wi -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
        w[i] = wi;
    }
    pack_param(pack, i, wi);
}

__global__ void k_fill_zero(uint32_t *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Forward only: y = net(x) for n points, one 32-point tile per wavefront
// (grid-stride over tiles).
__global__ __launch_bounds__(256) void k_mlp_infer(const uint8_t *__restrict__ pack, const float *__restrict__ x,
                                                   uint32_t n, float *__restrict__ y) {
    __shared__ __align__(16) uint16_t sw[kFwdElems];
    __shared__ float sb[kBiasFloats];
    load_pack(pack, sw, kFwdElems, sb);
    __syncthreads();
    const int lane = threadIdx.x & 63, p = lane & 31, h = lane >> 5;
    const uint32_t ntiles = (n + 31) / 32;
    const uint32_t nw = gridDim.x * 4;
    for (uint32_t tile = blockIdx.x * 4 + (threadIdx.x >> 6); tile < ntiles; tile += nw) {
        const uint32_t pt = tile * 32 + p;
        const bool valid = pt < n;
        float xv[6];
        for (int c = 0; c < 6; c++) xv[c] = valid ? x[6 * (size_t)pt + c] : 0.0f;
        f32x16 a0 = zero16(), a1 = zero16();
        for (int s = 0; s < 24; s++) {
            bf16x4 b = feat4(xv, s, h);
            a0 = mfma(lds4(sw, kOffP_W1T + p * kP1 + 8 * s + 4 * h), b, a0);
            a1 = mfma(lds4(sw, kOffP_W1T + (32 + p) * kP1 + 8 * s + 4 * h), b, a1);
        }
        bf16x4 H[8], G[8];
        bias_relu(a0, a1, sb, h, H, nullptr, 0, 0);
        hidden_fwd(sw, kOffP_W2T, H, p, h, a0, a1);
        bias_relu(a0, a1, sb + 64, h, G, nullptr, 0, 0);
        hidden_fwd(sw, kOffP_W3T, G, p, h, a0, a1);
        bias_relu(a0, a1, sb + 128, h, H, nullptr, 0, 0);
        f32x16 ao = zero16();
        for (int s = 0; s < 8; s++) ao = mfma(lds4(sw, kOffP_W4T + p * kP + 8 * s + 4 * h), H[s], ao);
        if (h == 0 && valid)
            for (int c = 0; c < 3; c++) y[3 * (size_t)pt + c] = ao[c] + sb[192 + c];
    }
}

inline uint32_t nblocks(size_t n, uint32_t b = 256) { return (uint32_t)((n + b - 1) / b); }

} // namespace

// ----------------------------------------------------------------------------
namespace pvx {

int Surrogate::init(const float *host_params, uint64_t sd, hipStream_t st) {
    seed = sd;
    step = 0;
    int r;
    if ((r = w.reserve(kMlpParams * 4)) || (r = m.reserve(kMlpParams * 4)) || (r = v.reserve(kMlpParams * 4)) ||
        (r = pack.reserve(kPackBytes)) || (r = grad.reserve(kMlpParams * 4)))
        return r;
    std::vector<float> hw(kMlpParams, 0.0f);
    if (host_params) {
        std::memcpy(hw.data(), host_params, kMlpParams * 4);
    } else { // He-uniform weights from a splitmix64 stream, zero biases
        uint64_t s = sd;
        auto next = [&s]() {
            uint64_t z = (s += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            return (double)(z >> 11) * (1.0 / 9007199254740992.0);
        };
        auto fill = [&](int off, int fan_in, int count, double gain) {
            double lim = gain * std::sqrt(6.0 / fan_in);
            for (int k = 0; k < count; k++) hw[off + k] = (float)((2.0 * next() - 1.0) * lim);
        };
        fill(kOffW1, kMlpIn, kMlpIn * kMlpHid, 1.0);
        fill(kOffW2, kMlpHid, kMlpHid * kMlpHid, 1.0);
        fill(kOffW3, kMlpHid, kMlpHid * kMlpHid, 1.0);
        fill(kOffW4, kMlpHid, kMlpHid * kMlpOut, 0.5);
    }
    HIP_TRY(hipMemcpyAsync(w.p, hw.data(), kMlpParams * 4, hipMemcpyHostToDevice, st));
    k_fill_zero<<<nblocks(kMlpParams), 256, 0, st>>>(static_cast<uint32_t *>(m.p), kMlpParams);
    k_fill_zero<<<nblocks(kMlpParams), 256, 0, st>>>(static_cast<uint32_t *>(v.p), kMlpParams);
    k_fill_zero<<<nblocks(kPackBytes / 4), 256, 0, st>>>(static_cast<uint32_t *>(pack.p), kPackBytes / 4);
    k_mlp_update<<<nblocks(kMlpParams), 256, 0, st>>>(kUpdPack, nullptr, 0, static_cast<float *>(w.p), nullptr,
                                                      nullptr, nullptr, static_cast<uint8_t *>(pack.p), 0, 1, 1);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    return PVX_OK;
}

int Surrogate::ensure_acts(size_t batch) {
    int r;
    if ((r = acts.reserve((size_t)kActRows * batch * 2)) ||
        (r = partial.reserve((batch / kMlpChunk) * (size_t)kMlpParams * 4)) || (r = loss_part.reserve(batch / 32 * 4)))
        return r;
    acts_batch = batch;
    return PVX_OK;
}

int Surrogate::fwd_bwd(const float *x, const uint16_t *F, const float *t, uint32_t n, uint32_t batch, bool sample,
                       uint32_t sseed, hipStream_t st) {
    int r;
    if ((r = ensure_acts(batch))) return r;
    TrainArgs a;
    a.pack = static_cast<const uint8_t *>(pack.p);
    a.x = x;
    a.F = F;
    a.t = t;
    a.n = n;
    a.batch = batch;
    a.sample = sample ? 1 : 0;
    a.seed = sseed;
    a.dscale = 2.0f / (float)(sample ? batch : n);
    a.acts = static_cast<uint16_t *>(acts.p);
    a.loss_part = static_cast<float *>(loss_part.p);
    k_mlp_fwd_bwd<<<batch / 128, 256, 0, st>>>(a);
    const uint32_t nchunks = batch / kMlpChunk;
    k_mlp_wgrad<<<nblocks((size_t)kWgradItems * nchunks * 64), 256, 0, st>>>(
        static_cast<const uint16_t *>(acts.p), batch, nchunks, static_cast<float *>(partial.p));
    HIP_TRY(hipGetLastError());
    return PVX_OK;
}

int Surrogate::grad_batch(const float *x, const float *t, uint32_t n, float *loss, float *grad_host,
                          hipStream_t st) {
    const uint32_t batch = (uint32_t)(((size_t)n + kMlpChunk - 1) / kMlpChunk * kMlpChunk);
    int r;
    if ((r = fwd_bwd(x, nullptr, t, n, batch, false, 0, st))) return r;
    k_mlp_update<<<nblocks(kMlpParams), 256, 0, st>>>(kUpdGrad, static_cast<const float *>(partial.p),
                                                      batch / kMlpChunk, static_cast<float *>(w.p), nullptr,
                                                      nullptr, static_cast<float *>(grad.p), nullptr, 0, 1, 1);
    HIP_TRY(hipGetLastError());
    std::vector<float> lp(batch / 32);
    HIP_TRY(hipMemcpyAsync(lp.data(), loss_part.p, lp.size() * 4, hipMemcpyDeviceToHost, st));
    if (grad_host) HIP_TRY(hipMemcpyAsync(grad_host, grad.p, kMlpParams * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    double sum = 0;
    for (float f : lp) sum += f;
    if (loss) *loss = (float)(sum / n);
    return PVX_OK;
}

int Surrogate::train(const float *x, const float *t, uint32_t n, uint32_t steps, uint32_t batch, float lr,
                     float *final_loss, hipStream_t st) {
    int r;
    if ((r = feats.reserve((size_t)n * kMlpIn * 2))) return r;
    k_mlp_features<<<nblocks((size_t)n * 48), 256, 0, st>>>(x, n, static_cast<uint16_t *>(feats.p));
    HIP_TRY(hipGetLastError());
    for (uint32_t k = 0; k < steps; k++) {
        const uint32_t sseed = (uint32_t)(seed * 0x9E3779B97F4A7C15ull >> 32) + 7919u * (step + 1);
        if ((r = fwd_bwd(x, static_cast<const uint16_t *>(feats.p), t, n, batch, true, sseed, st))) return r;
        step++;
        // exponential decay to a tenth of the rate over the run
        float lr_k = lr * std::pow(0.1f, (float)k / (float)std::max<uint32_t>(1, steps));
        float bc1 = 1.0f - std::pow(0.9f, (float)step), bc2 = 1.0f - std::pow(0.99f, (float)step);
        k_mlp_update<<<nblocks(kMlpParams), 256, 0, st>>>(kUpdAdam, static_cast<const float *>(partial.p),
                                                          batch / kMlpChunk, static_cast<float *>(w.p),
                                                          static_cast<float *>(m.p), static_cast<float *>(v.p),
                                                          nullptr, static_cast<uint8_t *>(pack.p), lr_k, bc1, bc2);
        HIP_TRY(hipGetLastError());
    }
    if (final_loss) {
        std::vector<float> lp(batch / 32);
        HIP_TRY(hipMemcpyAsync(lp.data(), loss_part.p, lp.size() * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        double sum = 0;
        for (float f : lp) sum += f;
        *final_loss = (float)(sum / batch);
    }
    return PVX_OK;
}

int Surrogate::infer(const float *x, uint32_t n, float *y, hipStream_t st) {
    if (n == 0) return PVX_OK;
    const uint32_t tiles = (n + 31) / 32;
    const uint32_t grid = std::min<uint32_t>((tiles + 3) / 4, 2048);
    k_mlp_infer<<<grid, 256, 0, st>>>(static_cast<const uint8_t *>(pack.p), x, n, y);
    HIP_TRY(hipGetLastError());
    return PVX_OK;
}

int Surrogate::get_params(float *host, hipStream_t st) {
    HIP_TRY(hipMemcpyAsync(host, w.p, kMlpParams * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return PVX_OK;
}

} // namespace pvx

// ----------------------------------------------------------------------------
// C-ABI
// ----------------------------------------------------------------------------
struct pvx_surrogate {
    pvx::Surrogate net;
};

extern "C" int pvx_surrogate_create(const float *params, uint64_t seed, pvx_surrogate **out) {
    if (!out) return fail(PVX_EINVAL, "null output handle");
    *out = nullptr;
    pvx_surrogate *s = new (std::nothrow) pvx_surrogate;
    if (!s) return fail(PVX_ENOMEM, "host allocation failed");
    int r = s->net.init(params, seed, nullptr);
    if (r) {
        delete s;
        return r;
    }
    *out = s;
    return PVX_OK;
}

extern "C" void pvx_surrogate_destroy(pvx_surrogate *s) { delete s; }

extern "C" int pvx_surrogate_get_params(pvx_surrogate *s, float *params) {
    if (!s || !params) return fail(PVX_EINVAL, "null argument");
    return s->net.get_params(params, nullptr);
}

extern "C" int pvx_surrogate_grad(pvx_surrogate *s, const float *x, const float *target, size_t n, float *loss,
                                  float *grad, void *stream) {
    if (!s || !x || !target || !loss || !grad) return fail(PVX_EINVAL, "null argument");
    if (n == 0 || n > (1u << 24)) return fail(PVX_EINVAL, "n must be in [1, 2^24]");
    return s->net.grad_batch(x, target, (uint32_t)n, loss, grad, static_cast<hipStream_t>(stream));
}

extern "C" int pvx_surrogate_train(pvx_surrogate *s, const float *x, const float *target, size_t n,
                                   uint32_t steps, uint32_t batch, float lr, float *final_loss, void *stream) {
    if (!s || !x || !target) return fail(PVX_EINVAL, "null argument");
    if (n == 0 || n > (1u << 30)) return fail(PVX_EINVAL, "n must be in [1, 2^30]");
    if (batch == 0 || batch % 256 || batch > (1u << 20))
        return fail(PVX_EINVAL, "batch must be a multiple of 256 up to 2^20");
    if (!(lr > 0.0f)) return fail(PVX_EINVAL, "lr must be positive");
    return s->net.train(x, target, (uint32_t)n, steps, batch, lr, final_loss, static_cast<hipStream_t>(stream));
}

extern "C" int pvx_surrogate_infer(pvx_surrogate *s, const float *x, size_t n, float *out, void *stream) {
    if (!s || (n && (!x || !out))) return fail(PVX_EINVAL, "null argument");
    if (n > (1u << 30)) return fail(PVX_EINVAL, "n too large");
    return s->net.infer(x, (uint32_t)n, out, static_cast<hipStream_t>(stream));
}
