// Irradiance surrogate: Fourier features of (position, normal) -> three
// ReLU layers of width 64 -> RGB irradiance.  BF16 MFMA (32x32x16) with FP32
// accumulation for the forward pass, the backward pass and the weight
// gradients; FP32 master weights with Adam.
#ifndef PVX_MLP_INTERNAL_H
#define PVX_MLP_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "internal.h"

namespace pvx {

constexpr int kMlpFreq = 16;                   // frequencies per input dimension
constexpr int kMlpDims = 6;                    // position (normalised) + normal
constexpr int kMlpIn = 2 * kMlpFreq * kMlpDims; // 192 features
constexpr int kMlpHid = 64;
constexpr int kMlpOut = 3;

// FP32 parameter vector: W1[in][out], b1, W2, b2, W3, b3, W4[64][3], b4.
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpIn * kMlpHid;
constexpr int kOffW2 = kOffB1 + kMlpHid;
constexpr int kOffB2 = kOffW2 + kMlpHid * kMlpHid;
constexpr int kOffW3 = kOffB2 + kMlpHid;
constexpr int kOffB3 = kOffW3 + kMlpHid * kMlpHid;
constexpr int kOffW4 = kOffB3 + kMlpHid;
constexpr int kOffB4 = kOffW4 + kMlpHid * kMlpOut;
constexpr int kMlpParams = kOffB4 + kMlpOut; // 20867

constexpr int kMlpChunk = 128; // points per weight-gradient partial sum

struct Surrogate {
    uint64_t seed = 0;
    uint32_t step = 0;        // Adam steps taken
    DevBuf w, m, v;           // FP32 master weights and Adam moments
    DevBuf pack;              // BF16 weight copies laid out for the kernels + FP32 biases
    DevBuf acts, partial, loss_part, grad, feats;
    DevBuf bar, sched;          // cooperative trainer: grid barrier, per-step lr / bias corrections
    std::vector<float> hsched;  // host copy of sched (kept until the copy has run)

    int init(const float *host_params, uint64_t seed, hipStream_t st);
    // Loss and full gradient over n points (x: n*6 floats, t: n*3 floats, device).
    int grad_batch(const float *x, const float *t, uint32_t n, float *loss, float *grad_host, hipStream_t st);
    // `steps` Adam steps on mini-batches of `batch` points drawn from n, in
    // one cooperative launch; batch is a multiple of 64 up to max_train_batch().
    int train(const float *x, const float *t, uint32_t n, uint32_t steps, uint32_t batch, float lr,
              float *final_loss, hipStream_t st);
    int infer(const float *x, uint32_t n, float *y, hipStream_t st);
    int get_params(float *host, hipStream_t st);

    static int max_train_batch(); // 64 x co-resident trainer blocks on the current device

  private:
    int fwd_bwd(const float *x, const float *t, uint32_t n, uint32_t batch, hipStream_t st);
};

} // namespace pvx

#endif
