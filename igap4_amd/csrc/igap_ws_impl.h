// igap_ws_impl.h -- internal state and launchers of the warm-start network.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace igap {
namespace ws {

constexpr int NL = 4;                  // convolution layers
constexpr int CH = 16;                 // channels between layers (and padded in/out)
constexpr int LAYER_W = 27 * CH * CH;  // weights of one layer
constexpr int LAYER_P = LAYER_W + CH;  // weights + bias
constexpr int CHUNK = 512;             // voxels per weight-gradient partial
constexpr float ADAM_B1 = 0.9f, ADAM_B2 = 0.999f, ADAM_EPS = 1e-8f;

struct Net {
    int nx = 0, ny = 0, nz = 0, cin = 0, cout = 0;
    int dil[NL] = {1, 1, 1, 1};
    int nchunk = 0;
    int64_t step = 0;                   // Adam step count
    float *P = nullptr, *G = nullptr, *M = nullptr, *S = nullptr;   // FP32 params, grads, moments
    uint16_t *Wb = nullptr, *WTb = nullptr;                          // BF16 weight copies
    uint16_t *a[NL] = {};               // layer inputs (a[0] = packed input)
    uint16_t *dz[NL] = {};              // gradients w.r.t. each layer's pre-activation output
    float *y = nullptr, *tg = nullptr, *x = nullptr;
    float *part = nullptr, *bpart = nullptr, *lpart = nullptr;
};

hipError_t conv_forward(const Net &n, int layer, const uint16_t *ain, uint16_t *aout, float *yout, hipStream_t s);
hipError_t conv_input_grad(const Net &n, int layer, const uint16_t *dz, const uint16_t *mask, uint16_t *dzprev, hipStream_t s);
hipError_t weight_grad(const Net &n, int layer, const uint16_t *ain, const uint16_t *dz, hipStream_t s);
hipError_t loss_grad(const Net &n, const float *y, const float *tg, uint16_t *dz, hipStream_t s);
hipError_t pack_input(const Net &n, const float *x, uint16_t *a, hipStream_t s);
hipError_t pack_weights(const Net &n, hipStream_t s);
hipError_t adam(const Net &n, float lr, float c1, float c2, hipStream_t s);
hipError_t selftest_bf16(float *out, hipStream_t s);

}  // namespace ws
}  // namespace igap
