// t2x_cnn.h — pose-proposal network state held by a device handle.
#pragma once
#include "t2x_internal.h"

namespace t2x {
namespace cnn {

constexpr int kIn = 64;                                  // input image size
constexpr int kChannels[5] = {1, 32, 64, 128, 128};      // input + four conv layers
constexpr int kFeat = (kIn / 16) * (kIn / 16) * 128;     // flattened features (2048)

struct Net {
    int n_bins = 0;
    long long n_params = 0;
    long long w_off[4] = {}, b_off[4] = {}, wh_off = 0, bh_off = 0;
    DevBuf params, grad, m1, m2;        // FP32 master weights, gradient, Adam moments
    DevBuf act[5], dact[5], dcols, dz, y, dy, part, loss, correct, rank, label, shift_t, splitk;
    int cap = 0;                        // batch capacity of the activation buffers
    long long adam_t = 0;
    bool trained = false;
    // training-set geometry the network was trained for (checked at inference)
    int train_ori = 0, train_K = 0, train_order = 0;
    double train_max_shift = 0.;

    void layout(int n_bins_);
    int init(int n_bins_, uint64_t seed, hipStream_t st);
    int ensure_batch(int B);
    int forward(int B, bool bf16, hipStream_t st);           // act[0] -> y
    int backward(int B, bool bf16, hipStream_t st);          // dy -> grad
    int loss_grad(int B, bool bf16, float shift_weight, hipStream_t st);
    int adam_step(float lr, hipStream_t st);
};

}  // namespace cnn
}  // namespace t2x
