// Irradiance surrogate kernels (gfx950: BF16 MFMA v_mfma_f32_32x32x16_bf16
// with FP32 accumulation).  No reference counterpart: the network stands in
// for RadiosityFunction::GatherLight (core/lighting/radiosity.cpp:509-743) at
// cache misses, trained on the gathers the plain path makes.
//
// Layout trick: every layer is computed transposed, Y^T = W^T X^T, with the
// 32 points of a tile as the MFMA N dimension.  The 32x32 accumulator of a
// lane holds, for its point (lane % 32), neurons 8q + 4h + r (h = lane / 32,
// q = 0..3, r = 0..3).  A K step of 16 neurons (16s .. 16s + 15) takes from
// lane half h the eight neurons 16s + 4h + {0..3} and 16s + 8 + 4h + {0..3},
// i.e. the lane's own accumulator groups q = 2s and 2s + 1: the activations
// stay in registers from layer to layer, forward and backward.  The weight
// copies in LDS store their K columns in the same order (bits 2 and 3 of the
// neuron index swapped), so each lane's eight A values are one 16-byte load.
// Weight gradients (sums over points) are GEMMs over the stored activations.
// Training runs as one cooperative launch per network: each workgroup takes
// 64 points of the step's batch (activations in LDS), writes its partial
// gradient, and after a grid barrier reduces a slice of the parameters in
// fixed order (deterministic) and takes the Adam step on the FP32 master
// weights, rewriting the BF16 copies; a second barrier ends the step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "internal.h"
#include "mlp_internal.h"

using namespace pvx;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// K-column order of the packed weights: bits 2 and 3 of the index swapped.
__host__ __device__ constexpr int kperm(int i) { return (i & ~12) | ((i & 4) << 1) | ((i & 8) >> 1); }

// BF16 weight copies, element offsets; row pitches keep 16-byte loads
// aligned and put consecutive rows on different LDS banks.
constexpr int kP1 = kMlpIn + 8;                   // W1T row pitch (200)
constexpr int kP = kMlpHid + 8;                   // 64-wide row pitch (72)
constexpr int kP4 = 24;                           // W4N row pitch
constexpr int kOffP_W1T = 0;                      // [64 out][192 in]
constexpr int kOffP_W2T = kOffP_W1T + 64 * kP1;   // [64 out][64 in]
constexpr int kOffP_W3T = kOffP_W2T + 64 * kP;
constexpr int kOffP_W4T = kOffP_W3T + 64 * kP;    // [32 out (3 used)][64 in]
constexpr int kFwdElems = kOffP_W4T + 32 * kP;    // forward pass needs these
constexpr int kOffP_W2N = kFwdElems;              // [64 in][64 out]
constexpr int kOffP_W3N = kOffP_W2N + 64 * kP;
constexpr int kOffP_W4N = kOffP_W3N + 64 * kP;    // [64 in][16 out (3 used)]
constexpr int kAllElems = kOffP_W4N + 64 * kP4;
constexpr int kBiasFloats = 3 * kMlpHid + 4;      // b1 b2 b3 b4(+pad)
constexpr size_t kPackBytes = (size_t)kAllElems * 2 + kBiasFloats * 4;
static_assert((kFwdElems * 2) % 16 == 0 && (kAllElems * 2) % 16 == 0, "16-byte LDS copies");

// Stored activations (rows of points, BF16).
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

// fused trainer: points per workgroup, LDS activation row pitch
constexpr int kTrainPts = 64;
constexpr int kTrainPitch = kTrainPts + 8;
constexpr int kTrainThreads = 256;
constexpr uint32_t kSpinLimit = 1u << 24; // grid-barrier wait bound (about a second)

__device__ inline uint16_t f2bf(float f) { // round to nearest even (finite inputs)
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline f32x16 mfma(s16x8 a, s16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

__device__ inline f32x16 zero16() {
    f32x16 z;
    for (int g = 0; g < 16; g++) z[g] = 0.0f;
    return z;
}

__device__ inline s16x8 ld8(const uint16_t *base, int off) {
    return *reinterpret_cast<const s16x8 *>(base + off);
}

__device__ inline s16x8 cat(s16x4 a, s16x4 b) {
    s16x8 r;
    for (int t = 0; t < 4; t++) {
        r[t] = a[t];
        r[4 + t] = b[t];
    }
    return r;
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
__device__ inline s16x4 feat4(const float *x, int s, int h) {
    const int j = s >> 2, k0 = 4 * (s & 3) + 2 * h;
    const float sc = j >= 3 ? 0.25f : 1.0f;
    float s0, c0, s1, c1;
    sincospif((freq(k0) * sc) * x[j], &s0, &c0);
    sincospif((freq(k0 + 1) * sc) * x[j], &s1, &c1);
    s16x4 r;
    r[0] = (short)f2bf(s0);
    r[1] = (short)f2bf(c0);
    r[2] = (short)f2bf(s1);
    r[3] = (short)f2bf(c1);
    return r;
}

// The B operand of K step s of layer 1 for lane half h: features
// 16s + 4h + {0..3} and 16s + 8 + 4h + {0..3}.
__device__ inline s16x8 feat8(const float *x, int s, int h) { return cat(feat4(x, 2 * s, h), feat4(x, 2 * s + 1, h)); }

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
    // eight 16-byte loads in flight per lane before the LDS stores: the pack
    // was just rewritten by other workgroups, so each load waits the full
    // memory latency, once per batch instead of once per 16 bytes
    constexpr int kInFlight = 8;
    const int n16 = welems * 2 / 16;
    const int step = kInFlight * (int)blockDim.x;
    int i0 = threadIdx.x;
    for (; i0 + step - (int)blockDim.x < n16; i0 += step) { // whole batches
        uint4 v0 = src[i0], v1 = src[i0 + blockDim.x], v2 = src[i0 + 2 * blockDim.x], v3 = src[i0 + 3 * blockDim.x];
        uint4 v4 = src[i0 + 4 * blockDim.x], v5 = src[i0 + 5 * blockDim.x], v6 = src[i0 + 6 * blockDim.x],
              v7 = src[i0 + 7 * blockDim.x];
        dst[i0] = v0;
        dst[i0 + blockDim.x] = v1;
        dst[i0 + 2 * blockDim.x] = v2;
        dst[i0 + 3 * blockDim.x] = v3;
        dst[i0 + 4 * blockDim.x] = v4;
        dst[i0 + 5 * blockDim.x] = v5;
        dst[i0 + 6 * blockDim.x] = v6;
        dst[i0 + 7 * blockDim.x] = v7;
    }
    for (; i0 < n16; i0 += blockDim.x) dst[i0] = src[i0]; // the rest
    const float *bsrc = reinterpret_cast<const float *>(pack + (size_t)kAllElems * 2);
    for (int i = threadIdx.x; i < kBiasFloats; i += blockDim.x) sb[i] = bsrc[i];
}

// One 64 -> 64 layer: acc = W^T[mb*32 + p][.] * B over 4 K steps.
__device__ inline void hidden_mm(const uint16_t *sw, int off, const s16x4 *Bin, int p, int h, f32x16 &a0,
                                 f32x16 &a1) {
    a0 = zero16();
    a1 = zero16();
    for (int s = 0; s < 4; s++) {
        const s16x8 b = cat(Bin[2 * s], Bin[2 * s + 1]);
        a0 = mfma(ld8(sw, off + p * kP + 16 * s + 8 * h), b, a0);
        a1 = mfma(ld8(sw, off + (32 + p) * kP + 16 * s + 8 * h), b, a1);
    }
}

// Bias + ReLU; returns the lane's activation groups in `out` and stores the
// activations (row = neuron) for the weight-gradient GEMM when `st` is set.
__device__ inline void bias_relu(const f32x16 &a0, const f32x16 &a1, const float *bias, int h, s16x4 *out,
                                 uint16_t *st, uint32_t pitch, uint32_t slot) {
    for (int g = 0; g < 16; g++) {
        const int i = 8 * (g >> 2) + 4 * h + (g & 3);
        uint16_t v0 = f2bf(fmaxf(a0[g] + bias[i], 0.0f));
        uint16_t v1 = f2bf(fmaxf(a1[g] + bias[32 + i], 0.0f));
        out[g >> 2][g & 3] = (short)v0;
        out[4 + (g >> 2)][g & 3] = (short)v1;
        if (st) {
            st[(size_t)i * pitch + slot] = v0;
            st[(size_t)(32 + i) * pitch + slot] = v1;
        }
    }
}

// dz = dh * [h > 0] as the lane's groups, stored as rows for the gradient GEMM.
__device__ inline void relu_back(const f32x16 &a0, const f32x16 &a1, const s16x4 *H, int h, s16x4 *out,
                                 uint16_t *st, uint32_t pitch, uint32_t slot) {
    for (int g = 0; g < 16; g++) {
        const int i = 8 * (g >> 2) + 4 * h + (g & 3);
        uint16_t v0 = H[g >> 2][g & 3] > 0 ? f2bf(a0[g]) : (uint16_t)0; // h >= 0 (post ReLU, BF16 bits)
        uint16_t v1 = H[4 + (g >> 2)][g & 3] > 0 ? f2bf(a1[g]) : (uint16_t)0;
        out[g >> 2][g & 3] = (short)v0;
        out[4 + (g >> 2)][g & 3] = (short)v1;
        st[(size_t)i * pitch + slot] = v0;
        st[(size_t)(32 + i) * pitch + slot] = v1;
    }
}

// Forward + backward of one 32-point tile (one wavefront).  Features come
// from F (per point, in K-step order) or are computed from x; activations
// go to rows of A (element (row, slot) at A[row * pitch + slot]).  Returns
// the tile's squared error (lane 0).
__device__ inline float tile_fwd_bwd(const uint16_t *sw, const float *sb, const uint16_t *Fp, const float *x,
                                     bool valid, const float *target, float dscale, uint16_t *A, uint32_t pitch,
                                     uint32_t slot) {
    const int lane = threadIdx.x & 63, p = lane & 31, h = lane >> 5;
    f32x16 a0 = zero16(), a1 = zero16();
    for (int s = 0; s < kMlpIn / 16; s++) {
        s16x8 b;
        if (Fp) {
            b = *reinterpret_cast<const s16x8 *>(Fp + 16 * s + 8 * h);
            if (!valid)
                for (int t = 0; t < 8; t++) b[t] = 0;
        } else {
            b = feat8(x, s, h);
        }
        for (int t = 0; t < 8; t++) {
            const int f = 16 * s + 4 * h + (t & 3) + 8 * (t >> 2);
            A[(size_t)(kRowX0 + f) * pitch + slot] = (uint16_t)b[t];
        }
        a0 = mfma(ld8(sw, kOffP_W1T + p * kP1 + 16 * s + 8 * h), b, a0);
        a1 = mfma(ld8(sw, kOffP_W1T + (32 + p) * kP1 + 16 * s + 8 * h), b, a1);
    }
    s16x4 H1[8], H2[8], H3[8];
    bias_relu(a0, a1, sb, h, H1, A + (size_t)kRowX1 * pitch, pitch, slot);
    hidden_mm(sw, kOffP_W2T, H1, p, h, a0, a1);
    bias_relu(a0, a1, sb + 64, h, H2, A + (size_t)kRowX2 * pitch, pitch, slot);
    hidden_mm(sw, kOffP_W3T, H2, p, h, a0, a1);
    bias_relu(a0, a1, sb + 128, h, H3, A + (size_t)kRowX3 * pitch, pitch, slot);
    f32x16 ao = zero16();
    for (int s = 0; s < 4; s++) ao = mfma(ld8(sw, kOffP_W4T + p * kP + 16 * s + 8 * h), cat(H3[2 * s], H3[2 * s + 1]), ao);
    // loss and output gradient (rows 0..2 of the output tile live in lanes 0..31)
    float l2 = 0.0f;
    s16x8 DY;
    for (int t = 0; t < 8; t++) DY[t] = 0;
    if (h == 0) {
        for (int c = 0; c < 3; c++) {
            float y = ao[c] + sb[192 + c];
            float e = valid ? y - target[c] : 0.0f;
            l2 += e * e;
            uint16_t d = f2bf(dscale * e);
            DY[c] = (short)d;
            A[(size_t)(kRowDY + c) * pitch + slot] = d;
        }
        A[(size_t)(kRowDY + 3) * pitch + slot] = 0;
    }
    for (int off = 32; off > 0; off >>= 1) l2 += __shfl_xor(l2, off, 64);
    // backward: dh3 = W4 dy
    a0 = mfma(ld8(sw, kOffP_W4N + p * kP4 + 8 * h), DY, zero16());
    a1 = mfma(ld8(sw, kOffP_W4N + (32 + p) * kP4 + 8 * h), DY, zero16());
    s16x4 Z[8];
    relu_back(a0, a1, H3, h, Z, A + (size_t)kRowZ3 * pitch, pitch, slot);
    hidden_mm(sw, kOffP_W3N, Z, p, h, a0, a1); // dh2 = W3 dz3
    relu_back(a0, a1, H2, h, Z, A + (size_t)kRowZ2 * pitch, pitch, slot);
    hidden_mm(sw, kOffP_W2N, Z, p, h, a0, a1); // dh1 = W2 dz2
    relu_back(a0, a1, H1, h, Z, A + (size_t)kRowZ1 * pitch, pitch, slot);
    return l2;
}

// Where each of the 22 gradient tiles reads and writes.
struct WgradItem {
    int xrow, zrow, mb, nb, outdim, offW, offB;
};
__device__ inline WgradItem wgrad_item(uint32_t item) {
    WgradItem w;
    if (item < 12) {
        w = {kRowX0, kRowZ1, (int)item / 2, (int)item % 2, 64, kOffW1, kOffB1};
    } else if (item < 16) {
        w = {kRowX1, kRowZ2, (int)(item - 12) / 2, (int)(item - 12) % 2, 64, kOffW2, kOffB2};
    } else if (item < 20) {
        w = {kRowX2, kRowZ3, (int)(item - 16) / 2, (int)(item - 16) % 2, 64, kOffW3, kOffB3};
    } else {
        w = {kRowX3, kRowDY, (int)item - 20, 0, 3, kOffW4, kOffB4};
    }
    return w;
}

// dW[in][out] = sum_p X[in][p] dZ[out][p] over points [p0, p0 + npts) of
// the activation rows (pitch elements apart) for one 32x32 tile; bias
// gradients ride along as a ones-row A operand on the m-block-0 tiles.
// Writes the sums to out (a parameter vector).
__device__ inline void wgrad_tile(const uint16_t *acts, uint32_t pitch, uint32_t p0, int npts, uint32_t item,
                                  float *out) {
    const int lane = threadIdx.x & 63, p = lane & 31, h = lane >> 5;
    const WgradItem w = wgrad_item(item);
    const uint16_t *X = acts + (size_t)(w.xrow + w.mb * 32 + p) * pitch;
    const int j = w.nb * 32 + p;
    const bool jv = j < w.outdim;
    const uint16_t *Zr = acts + (size_t)(w.zrow + (jv ? j : 0)) * pitch;
    s16x8 ones, zero;
    for (int t = 0; t < 8; t++) {
        ones[t] = (short)0x3F80;
        zero[t] = 0;
    }
    f32x16 acc = zero16(), accb = zero16();
    for (int s = 0; s < npts / 16; s++) {
        const uint32_t kp = p0 + 16 * s + 8 * h;
        const s16x8 av = *reinterpret_cast<const s16x8 *>(X + kp);
        const s16x8 bv = jv ? *reinterpret_cast<const s16x8 *>(Zr + kp) : zero;
        acc = mfma(av, bv, acc);
        if (w.mb == 0) accb = mfma(ones, bv, accb);
    }
    if (jv) {
        for (int g = 0; g < 16; g++) {
            const int i = w.mb * 32 + 8 * (g >> 2) + 4 * h + (g & 3);
            out[w.offW + i * w.outdim + j] = acc[g];
        }
        if (w.mb == 0 && h == 0) out[w.offB + j] = accb[0];
    }
}

// Features of n points, [point][192] BF16 in K-step order (16s + 8h + t is
// feature 16s + 4h + (t & 3) + 8 (t >> 2)), computed once per training set.
__global__ void k_mlp_features(const float *__restrict__ x, uint32_t n, uint16_t *__restrict__ F) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * 24u) return;
    const uint32_t pt = i / 24u, sh = i % 24u, s = sh >> 1, h = sh & 1u;
    float xv[6];
    for (int c = 0; c < 6; c++) xv[c] = x[6 * (size_t)pt + c];
    *reinterpret_cast<s16x8 *>(F + (size_t)pt * kMlpIn + 16 * s + 8 * h) = feat8(xv, (int)s, (int)h);
}

struct TrainArgs {
    const uint8_t *pack;
    const float *x; // n x 6
    const float *t; // n x 3
    uint32_t n;
    uint32_t batch;  // multiple of kMlpChunk
    float dscale;    // 2 / (points in the loss)
    uint16_t *acts;  // kActRows x batch
    float *loss_part;// per tile
};

// Forward + backward over all n points, one 32-point tile per wavefront
// (the full-gradient path of pvx_surrogate_grad).
__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(TrainArgs a) {
    __shared__ __align__(16) uint16_t sw[kAllElems];
    __shared__ float sb[kBiasFloats];
    load_pack(a.pack, sw, kAllElems, sb);
    __syncthreads();
    const int lane = threadIdx.x & 63, p = lane & 31;
    const uint32_t tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile * 32 >= a.batch) return;
    const uint32_t slot = tile * 32 + p;
    const bool valid = slot < a.n;
    float x[6];
    for (int c = 0; c < 6; c++) x[c] = valid ? a.x[6 * (size_t)slot + c] : 0.0f;
    const float *tg = a.t + 3 * (size_t)(valid ? slot : 0);
    const float l2 = tile_fwd_bwd(sw, sb, nullptr, x, valid, tg, a.dscale, a.acts, a.batch, slot);
    if (lane == 0) a.loss_part[tile] = l2;
}

// Weight-gradient GEMMs over chunks of kMlpChunk points, one tile per
// wavefront; partial sums per chunk.
__global__ __launch_bounds__(256) void k_mlp_wgrad(const uint16_t *__restrict__ acts, uint32_t batch,
                                                   uint32_t nchunks, float *__restrict__ partial) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (wave >= kWgradItems * nchunks) return;
    const uint32_t item = wave % kWgradItems, chunk = wave / kWgradItems;
    wgrad_tile(acts, batch, chunk * kMlpChunk, kMlpChunk, item, partial + (size_t)chunk * kMlpParams);
}

// Writes the BF16 / bias copies of parameter i.
__device__ inline void pack_param(uint8_t *pack, int i, float w) {
    uint16_t *pw = reinterpret_cast<uint16_t *>(pack);
    float *pb = reinterpret_cast<float *>(pack + (size_t)kAllElems * 2);
    const uint16_t b = f2bf(w);
    if (i < kOffB1) {
        int in = i / 64, out = i % 64;
        pw[kOffP_W1T + out * kP1 + kperm(in)] = b;
    } else if (i < kOffW2) {
        pb[i - kOffB1] = w;
    } else if (i < kOffB2) {
        int in = (i - kOffW2) / 64, out = (i - kOffW2) % 64;
        pw[kOffP_W2T + out * kP + kperm(in)] = b;
        pw[kOffP_W2N + in * kP + kperm(out)] = b;
    } else if (i < kOffW3) {
        pb[64 + i - kOffB2] = w;
    } else if (i < kOffB3) {
        int in = (i - kOffW3) / 64, out = (i - kOffW3) % 64;
        pw[kOffP_W3T + out * kP + kperm(in)] = b;
        pw[kOffP_W3N + in * kP + kperm(out)] = b;
    } else if (i < kOffW4) {
        pb[128 + i - kOffB3] = w;
    } else if (i < kOffB4) {
        int in = (i - kOffW4) / 3, out = (i - kOffW4) % 3;
        pw[kOffP_W4T + out * kP + kperm(in)] = b;
        pw[kOffP_W4N + in * kP4 + kperm(out)] = b;
    } else {
        pb[192 + i - kOffB4] = w;
    }
}

// One Adam step on parameter i with gradient g (beta 0.9 / 0.99).
__device__ inline float adam(int i, float g, float *w, float *m, float *v, float lr, float bc1, float bc2) {
    const float b1 = 0.9f, b2 = 0.99f, eps = 1e-8f;
    float mi = b1 * m[i] + (1.0f - b1) * g;
    float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    w[i] = wi;
    return wi;
}

enum { kUpdPack = 0, kUpdGrad = 1 };

// Reduces the chunk partials of every parameter (fixed order) and writes the
// gradient (mode 1), or packs the weights (mode 0).
__global__ void k_mlp_update(int mode, const float *__restrict__ partial, uint32_t nchunks, float *__restrict__ w,
                             float *__restrict__ grad, uint8_t *__restrict__ pack) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    if (mode == kUpdGrad) {
        float g = 0.0f;
        for (uint32_t c = 0; c < nchunks; c++) g += partial[(size_t)c * kMlpParams + i];
        grad[i] = g;
        return;
    }
    pack_param(pack, i, w[i]);
}

__global__ void k_fill_zero(uint32_t *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Grid barrier of the cooperative trainer: every thread releases its writes,
// one thread per block arrives and waits for the generation to move on
// (bounded: after kSpinLimit polls the error flag is raised and every block
// leaves).  Returns false once the flag is up.
struct GridBar {
    unsigned count, gen, err, pad;
};
__device__ inline bool grid_sync(GridBar *b, unsigned nblocks) {
    __shared__ int s_ok;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned g = __hip_atomic_load(&b->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned arrived = __hip_atomic_fetch_add(&b->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        if (arrived == nblocks - 1) {
            __hip_atomic_store(&b->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&b->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            uint32_t spins = 0;
            while (__hip_atomic_load(&b->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > kSpinLimit) {
                    __hip_atomic_store(&b->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }
        if (__hip_atomic_load(&b->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ok = 0;
        s_ok = ok;
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return s_ok != 0;
}

struct FusedArgs {
    uint8_t *pack;
    const uint16_t *F; // n x 192 features (K-step order)
    const float *t;    // n x 3
    uint32_t n, steps, step0;
    uint32_t seed;     // sample stream seed
    const float *sched;// per step: lr, bc1, bc2
    float dscale;
    float *w, *m, *v;
    float *partial;    // gridDim.x x kMlpParams
    float *loss_part;  // gridDim.x x 2 (last step's per-tile squared errors)
    GridBar *bar;
};

// `steps` Adam steps of batch = 64 x gridDim.x points in one cooperative
// launch (waves 0 and 1 of a block take 32 points each; all four share the
// gradient tiles, the reduction and the weight copies).
__global__ __launch_bounds__(kTrainThreads) void k_mlp_train(FusedArgs a) {
    __shared__ __align__(16) uint16_t sw[kAllElems];
    __shared__ __align__(16) uint16_t acts[kActRows * kTrainPitch];
    __shared__ float sb[kBiasFloats];
    const int lane = threadIdx.x & 63, p = lane & 31, wv = threadIdx.x >> 6;
    const uint32_t G = gridDim.x, wg = blockIdx.x;
    const uint32_t per = (kMlpParams + G - 1) / G;
    const uint32_t lo = wg * per, hi = min((uint32_t)kMlpParams, lo + per);
    for (uint32_t k = 0; k < a.steps; k++) {
        load_pack(a.pack, sw, kAllElems, sb);
        __syncthreads();
        // forward + backward of this block's 64 samples
        float l2 = 0.0f;
        if (wv < 2) {
            const uint32_t slot = (uint32_t)wv * 32 + p;
            const uint32_t sseed = a.seed + 7919u * (a.step0 + k + 1);
            const uint32_t src = mix32(sseed * 0x9E3779B9u + wg * kTrainPts + slot) % a.n;
            l2 = tile_fwd_bwd(sw, sb, a.F + (size_t)src * kMlpIn, nullptr, true, a.t + 3 * (size_t)src, a.dscale,
                              acts, kTrainPitch, slot);
        }
        if (lane == 0 && wv < 2) a.loss_part[2 * wg + wv] = l2;
        __syncthreads();
        // this block's partial gradient
        for (uint32_t item = wv; item < (uint32_t)kWgradItems; item += kTrainThreads / 64)
            wgrad_tile(acts, kTrainPitch, 0, kTrainPts, item, a.partial + (size_t)wg * kMlpParams);
        if (!grid_sync(a.bar, G)) return;
        // reduce this block's slice of the parameters in block order, Adam
        const float lr = a.sched[3 * k], bc1 = a.sched[3 * k + 1], bc2 = a.sched[3 * k + 2];
        for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
            // block partials summed in block order, sixteen loads in flight
            float g = 0.0f;
            for (uint32_t c0 = 0; c0 < G; c0 += 16) {
                float pv[16];
#pragma unroll
                for (int u = 0; u < 16; u++)
                    pv[u] = c0 + u < G ? a.partial[(size_t)(c0 + u) * kMlpParams + i] : 0.0f;
#pragma unroll
                for (int u = 0; u < 16; u++)
                    if (c0 + u < G) g += pv[u];
            }
            pack_param(a.pack, (int)i, adam((int)i, g, a.w, a.m, a.v, lr, bc1, bc2));
        }
        if (!grid_sync(a.bar, G)) return;
    }
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
        for (int s = 0; s < kMlpIn / 16; s++) {
            const s16x8 b = feat8(xv, s, h);
            a0 = mfma(ld8(sw, kOffP_W1T + p * kP1 + 16 * s + 8 * h), b, a0);
            a1 = mfma(ld8(sw, kOffP_W1T + (32 + p) * kP1 + 16 * s + 8 * h), b, a1);
        }
        s16x4 H[8], G[8];
        bias_relu(a0, a1, sb, h, H, nullptr, 0, 0);
        hidden_mm(sw, kOffP_W2T, H, p, h, a0, a1);
        bias_relu(a0, a1, sb + 64, h, G, nullptr, 0, 0);
        hidden_mm(sw, kOffP_W3T, G, p, h, a0, a1);
        bias_relu(a0, a1, sb + 128, h, H, nullptr, 0, 0);
        f32x16 ao = zero16();
        for (int s = 0; s < 4; s++) ao = mfma(ld8(sw, kOffP_W4T + p * kP + 16 * s + 8 * h), cat(H[2 * s], H[2 * s + 1]), ao);
        if (h == 0 && valid)
            for (int c = 0; c < 3; c++) y[3 * (size_t)pt + c] = ao[c] + sb[192 + c];
    }
}

inline uint32_t nblocks(size_t n, uint32_t b = 256) { return (uint32_t)((n + b - 1) / b); }

std::mutex g_train_mu; // serialises cooperative trainer launches (see Surrogate::train)

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
                                                      static_cast<uint8_t *>(pack.p));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    return PVX_OK;
}

int Surrogate::fwd_bwd(const float *x, const float *t, uint32_t n, uint32_t batch, hipStream_t st) {
    int r;
    if ((r = acts.reserve((size_t)kActRows * batch * 2)) ||
        (r = partial.reserve((batch / kMlpChunk) * (size_t)kMlpParams * 4)) || (r = loss_part.reserve(batch / 32 * 4)))
        return r;
    TrainArgs a;
    a.pack = static_cast<const uint8_t *>(pack.p);
    a.x = x;
    a.t = t;
    a.n = n;
    a.batch = batch;
    a.dscale = 2.0f / (float)n;
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
    if ((r = fwd_bwd(x, t, n, batch, st))) return r;
    k_mlp_update<<<nblocks(kMlpParams), 256, 0, st>>>(kUpdGrad, static_cast<const float *>(partial.p),
                                                      batch / kMlpChunk, static_cast<float *>(w.p),
                                                      static_cast<float *>(grad.p), nullptr);
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

int Surrogate::max_train_batch() {
    int per_cu = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k_mlp_train), kTrainThreads, 0) !=
            hipSuccess)
        return 0;
    return per_cu * cus * kTrainPts;
}

int Surrogate::train(const float *x, const float *t, uint32_t n, uint32_t steps, uint32_t batch, float lr,
                     float *final_loss, hipStream_t st) {
    if (batch % kTrainPts) return fail(PVX_EINVAL, "training batch must be a multiple of 64");
    const uint32_t G = batch / kTrainPts;
    const int maxb = max_train_batch();
    if (maxb <= 0) return fail(PVX_EDEVICE, "occupancy query for the training kernel failed");
    if (batch > (uint32_t)maxb)
        return fail(PVX_EINVAL, "training batch above " + std::to_string(maxb) +
                                    " points (64 per co-resident block of the cooperative trainer)");
    int r;
    if ((r = feats.reserve((size_t)n * kMlpIn * 2)) || (r = partial.reserve((size_t)G * kMlpParams * 4)) ||
        (r = loss_part.reserve((size_t)2 * G * 4)) || (r = bar.reserve(16)) ||
        (r = sched.reserve((size_t)std::max<uint32_t>(steps, 1) * 12)))
        return r;
    k_mlp_features<<<nblocks((size_t)n * 24), 256, 0, st>>>(x, n, static_cast<uint16_t *>(feats.p));
    k_fill_zero<<<1, 64, 0, st>>>(static_cast<uint32_t *>(bar.p), 4);
    HIP_TRY(hipGetLastError());
    if (steps > 0) {
        // exponential decay to a tenth of the rate over the run; Adam's bias
        // corrections for the running step count
        hsched.resize((size_t)steps * 3);
        for (uint32_t k = 0; k < steps; k++) {
            const float sk = (float)(step + k + 1);
            hsched[3 * k] = lr * std::pow(0.1f, (float)k / (float)steps);
            hsched[3 * k + 1] = 1.0f - std::pow(0.9f, sk);
            hsched[3 * k + 2] = 1.0f - std::pow(0.99f, sk);
        }
        HIP_TRY(hipMemcpyAsync(sched.p, hsched.data(), hsched.size() * 4, hipMemcpyHostToDevice, st));
        FusedArgs a;
        a.pack = static_cast<uint8_t *>(pack.p);
        a.F = static_cast<const uint16_t *>(feats.p);
        a.t = t;
        a.n = n;
        a.steps = steps;
        a.step0 = step;
        a.seed = (uint32_t)(seed * 0x9E3779B97F4A7C15ull >> 32);
        a.sched = static_cast<const float *>(sched.p);
        a.dscale = 2.0f / (float)batch;
        a.w = static_cast<float *>(w.p);
        a.m = static_cast<float *>(m.p);
        a.v = static_cast<float *>(v.p);
        a.partial = static_cast<float *>(partial.p);
        a.loss_part = static_cast<float *>(loss_part.p);
        a.bar = static_cast<GridBar *>(bar.p);
        void *args[] = {&a};
        // one cooperative trainer on the device at a time: two of them,
        // each partly resident and spinning at its barrier, could hold the
        // CUs the other's missing blocks need
        std::lock_guard<std::mutex> lock(g_train_mu);
        HIP_TRY(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_mlp_train), dim3(G), dim3(kTrainThreads),
                                           args, 0, st));
        step += steps;
        HIP_TRY(hipStreamSynchronize(st));
    }
    std::vector<float> lp(2 * G);
    uint32_t hb[4] = {0, 0, 0, 0};
    HIP_TRY(hipMemcpyAsync(hb, bar.p, 16, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(lp.data(), loss_part.p, lp.size() * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (hb[2]) return fail(PVX_EDEVICE, "training grid barrier timed out");
    if (final_loss) {
        double sum = 0;
        for (float f : lp) sum += f;
        *final_loss = steps ? (float)(sum / batch) : 0.0f;
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
