// zh_mlp.h -- the learned colour surrogate: a 3 -> 64 -> 64 -> 3 network with
// SiLU activations standing in for a pointwise colour chain (R^3 -> R^3 on
// sRGB-encoded values).
//
// Parameter vector (4,611 FP32 master values):
//   W1[3][64] @0, b1[64] @192, W2[64][64] @256, b2[64] @4352, W3[64][3] @4416,
//   b3[3] @4608;   W[i][o] multiplies input i into output o.
//
// Inference numerics: weights, biases and both hidden activations are
// bfloat16 (round to nearest even), products accumulate in FP32 on the
// matrix cores.  The input stays exact: each component is split into a
// bfloat16 high part and a bfloat16 remainder, both multiplied by W1.
// Training runs the same forward rounding in FP32 and passes gradients
// straight through the roundings.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zh {
namespace mlp {

constexpr int kIn = 3, kHid = 64, kOut = 3;
constexpr int kOffW1 = 0, kOffB1 = 192, kOffW2 = 256, kOffB2 = 4352, kOffW3 = 4416, kOffB3 = 4608;
constexpr int kParams = 4611;

// Matrix-core operands of the inference kernel, bfloat16 bits, one 4-value
// group per lane of a wave (v_mfma_f32_32x32x8_bf16 A operand: lane l holds
// row l % 32, k = 4 * (l / 32) + 0..3):
//   A1[m][lane][4]      m = 0..1 (features 32m..32m+31), k = (x, y, z, b1 | x, y, z, 0)
//   A2[m][kb][lane][4]  kb = 0..7 inputs 8kb..8kb+7, kb = 8: bias column
//   A3[kb][lane][4]     rows = output channels (0..2 used)
constexpr int kA1 = 2 * 64 * 4;
constexpr int kA2 = 2 * 9 * 64 * 4;
constexpr int kA3 = 9 * 64 * 4;
constexpr int kOps = kA1 + kA2 + kA3;  // uint16 values

}  // namespace mlp
struct StageDesc;
}  // namespace zh

extern "C" {
// Training kernels on a batch of n samples (device pointers; see zh_mlp.hip).
struct zh_mlp_batch {
  const float* x;      // inputs, rows of `x_stride` floats (first 3 used)
  const float* t;      // targets, rows of `x_stride` floats
  const uint32_t* idx; // sample index of each batch row (n entries)
  uint32_t x_stride;
  uint32_t n;
  float* a1;  // n x 64 pre-activations
  float* h1;  // n x 64 rounded activations
  float* a2;
  float* h2;
  float* da1;
  float* da2;
  float* dy;   // n x 3
  float* loss; // one float, accumulated
};
// Scratch floats after the gradient for a batch of n samples.
size_t zh_mlp_part_floats(uint32_t n);
hipError_t zh_mlp_quantise(const float* p, float* q, hipStream_t s);
hipError_t zh_mlp_loss_grad(const float* q, const zh_mlp_batch* b, float* grad, hipStream_t s);
hipError_t zh_mlp_adam(float* p, const float* g, float* m, float* v, float lr, float bc1, float bc2,
                       hipStream_t s);
hipError_t zh_mlp_pack(const float* p, uint16_t* ops, hipStream_t s);
// fmt_in/fmt_out: 1 = RGBA8 sRGB codes (encoded value = code / 255, alpha
// passed through), 6 = RGBA32F (encoded values, alpha passed through).
// Surrogate of a flat RGBA8 sRGB stage [load, colour ops, store] (n pixels).
hipError_t zh_mlp_infer_stage(const uint16_t* ops, const zh::StageDesc* d_dev, uint64_t n, uint32_t cus,
                              hipStream_t s);
hipError_t zh_mlp_infer(const uint16_t* ops, const void* src, void* dst, uint64_t n, int fmt_in, int fmt_out,
                        uint32_t cus, hipStream_t s);
}
