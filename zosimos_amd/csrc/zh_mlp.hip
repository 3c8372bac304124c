// zh_mlp.hip -- training and inference kernels of the learned colour
// surrogate (zh_mlp.h).
//
// Training (FP32 vector ALUs): per step, forward through both hidden layers
// with the inference kernel's bfloat16 roundings, the mean-squared-error
// gradient at the output, backward through both layers (straight-through the
// roundings), weight gradients as chunked products reduced in a fixed order
// (deterministic), and an Adam update of the FP32 master weights.
//
// Inference (matrix cores): a wave evaluates 32 pixels per step with
// v_mfma_f32_32x32x8_bf16.  Features run along the matrix rows and pixels
// along the columns (H^T = W^T X^T), so each layer's accumulator -- lane l
// holds column l % 32, rows 8g + 4 (l / 32) + r in register 4g + r -- is
// already the next layer's B operand for the input block 8g..8g+7: the
// activations never leave the registers.  Biases of the second and third
// layers enter as a ninth K block against a constant-one input.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "zh_mlp.h"
#include "zh_device.h"

namespace zh {
namespace mlp {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int kBlock = 256;
constexpr int kChunk = 64;  // samples per weight-gradient partial (33 KB of LDS)

__device__ __forceinline__ float bf16r(float x) {
  uint32_t u = __float_as_uint(x);
  u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
  return __uint_as_float(u);
}
__device__ __forceinline__ short bf16_bits(float x) {
  const uint32_t u = __float_as_uint(x);
  return (short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float sigmoid_p(float a) { return 1.0f / (1.0f + expf(-a)); }
__device__ __forceinline__ float silu_p(float a) { return a * sigmoid_p(a); }
__device__ __forceinline__ float silu_grad(float a) {
  const float s = sigmoid_p(a);
  return s * (1.0f + a * (1.0f - s));
}
// Inference SiLU: hardware exp2 and reciprocal (x -> -inf gives 0, x -> +inf gives x).
__device__ __forceinline__ float silu_f(float a) {
  return a * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-a * 1.4426950408889634f));
}

// ---- training -----------------------------------------------------------------
__global__ void k_quant(const float* p, float* q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kParams) q[i] = bf16r(p[i]);
}

__device__ __forceinline__ uint32_t row_of(const zh_mlp_batch& b, uint32_t r) { return b.idx ? b.idx[r] : r; }

// Hidden layer forward, one thread per (sample, feature).  Layer 1 reads the
// batch inputs, layer 2 the previous activations.
__global__ __launch_bounds__(kBlock) void k_fwd_hidden(zh_mlp_batch b, int layer, const float* q) {
  const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint32_t r = (uint32_t)(t / kHid), f = (uint32_t)(t % kHid);
  if (r >= b.n) return;
  float a;
  if (layer == 1) {
    const float* x = b.x + (uint64_t)row_of(b, r) * b.x_stride;
    a = q[kOffB1 + f];
#pragma unroll
    for (int k = 0; k < kIn; ++k) a = fmaf(x[k], q[kOffW1 + k * kHid + f], a);
    b.a1[(uint64_t)r * kHid + f] = a;
    b.h1[(uint64_t)r * kHid + f] = bf16r(silu_p(a));
  } else {
    const float* h = b.h1 + (uint64_t)r * kHid;
    a = q[kOffB2 + f];
#pragma unroll 16
    for (int k = 0; k < kHid; ++k) a = fmaf(h[k], q[kOffW2 + k * kHid + f], a);
    b.a2[(uint64_t)r * kHid + f] = a;
    b.h2[(uint64_t)r * kHid + f] = bf16r(silu_p(a));
  }
}

// Output layer and the loss gradient, one thread per sample.
__global__ __launch_bounds__(kBlock) void k_fwd_out(zh_mlp_batch b, const float* q) {
  __shared__ float red[kBlock];
  const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
  float l = 0.0f;
  if (r < b.n) {
    const float* h = b.h2 + (uint64_t)r * kHid;
    const float* t = b.t + (uint64_t)row_of(b, r) * b.x_stride;
    const float scale = 2.0f / (3.0f * (float)b.n);
#pragma unroll
    for (int c = 0; c < kOut; ++c) {
      float y = q[kOffB3 + c];
      for (int k = 0; k < kHid; ++k) y = fmaf(h[k], q[kOffW3 + k * kOut + c], y);
      const float e = y - t[c];
      b.dy[(uint64_t)r * kOut + c] = scale * e;
      l += e * e;
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(b.loss, red[0] / (3.0f * (float)b.n));
}

// Hidden layer backward: dA = (dOut . W^T) * silu'(A), one thread per (sample, feature).
__global__ __launch_bounds__(kBlock) void k_bwd_hidden(zh_mlp_batch b, int layer, const float* q) {
  const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint32_t r = (uint32_t)(t / kHid), i = (uint32_t)(t % kHid);
  if (r >= b.n) return;
  if (layer == 2) {
    const float* d = b.dy + (uint64_t)r * kOut;
    float g = 0.0f;
#pragma unroll
    for (int o = 0; o < kOut; ++o) g = fmaf(d[o], q[kOffW3 + i * kOut + o], g);
    b.da2[(uint64_t)r * kHid + i] = g * silu_grad(b.a2[(uint64_t)r * kHid + i]);
  } else {
    const float* d = b.da2 + (uint64_t)r * kHid;
    float g = 0.0f;
#pragma unroll 16
    for (int o = 0; o < kHid; ++o) g = fmaf(d[o], q[kOffW2 + i * kHid + o], g);
    b.da1[(uint64_t)r * kHid + i] = g * silu_grad(b.a1[(uint64_t)r * kHid + i]);
  }
}

// Weight-gradient partials of one chunk of kChunk samples:
//   part[chunk][i * out + o] = sum_r in[r][i] * d[r][o]   (i = n_in: bias, in = 1)
// staged through LDS; the chunks are summed in order by k_grad_reduce.
__global__ __launch_bounds__(kBlock) void k_grad_part(zh_mlp_batch b, int layer, float* part) {
  __shared__ float sin_[kChunk * (kHid + 1)];
  __shared__ float sd[kChunk * kHid];
  const int n_in = layer == 1 ? kIn : kHid;
  const int n_out = layer == 3 ? kOut : kHid;
  const uint32_t r0 = blockIdx.x * kChunk;
  const uint32_t rows = min((uint32_t)kChunk, b.n - r0);
  for (int e = threadIdx.x; e < kChunk * (n_in + 1); e += kBlock) {
    const uint32_t r = e / (n_in + 1), i = e % (n_in + 1);
    float v = 0.0f;
    if (r < rows) {
      const uint32_t gr = r0 + r;
      if (i == (uint32_t)n_in) v = 1.0f;
      else if (layer == 1) v = b.x[(uint64_t)row_of(b, gr) * b.x_stride + i];
      else if (layer == 2) v = b.h1[(uint64_t)gr * kHid + i];
      else v = b.h2[(uint64_t)gr * kHid + i];
    }
    sin_[r * (n_in + 1) + i] = v;
  }
  for (int e = threadIdx.x; e < kChunk * n_out; e += kBlock) {
    const uint32_t r = e / n_out, o = e % n_out;
    float v = 0.0f;
    if (r < rows) {
      const uint32_t gr = r0 + r;
      if (layer == 1) v = b.da1[(uint64_t)gr * kHid + o];
      else if (layer == 2) v = b.da2[(uint64_t)gr * kHid + o];
      else v = b.dy[(uint64_t)gr * kOut + o];
    }
    sd[r * n_out + o] = v;
  }
  __syncthreads();
  const int entries = (n_in + 1) * n_out;
  float* out = part + (uint64_t)blockIdx.x * entries;
  for (int e = threadIdx.x; e < entries; e += kBlock) {
    const int i = e / n_out, o = e % n_out;
    float s = 0.0f;
    for (int r = 0; r < kChunk; ++r) s = fmaf(sin_[r * (n_in + 1) + i], sd[r * n_out + o], s);
    out[e] = s;
  }
}

// Sum the chunk partials into the gradient vector in a fixed two-level order
// (deterministic, and kReduceGroups times the parallelism of one serial sum
// per entry): group g sums chunks [g per, (g + 1) per) of every entry, then
// each entry sums its kReduceGroups group sums in order.
constexpr uint32_t kReduceGroups = 16;

__global__ __launch_bounds__(kBlock) void k_grad_reduce_groups(const float* part, uint32_t chunks, int entries,
                                                               float* sums) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= entries) return;
  const uint32_t g = blockIdx.y;
  const uint32_t per = (chunks + kReduceGroups - 1) / kReduceGroups;
  const uint32_t c0 = g * per, c1 = min(chunks, c0 + per);
  float s = 0.0f;
  for (uint32_t c = c0; c < c1; ++c) s += part[(uint64_t)c * entries + e];
  sums[(uint64_t)g * entries + e] = s;
}

__global__ __launch_bounds__(kBlock) void k_grad_reduce(const float* sums, int layer, float* grad) {
  const int n_in = layer == 1 ? kIn : kHid;
  const int n_out = layer == 3 ? kOut : kHid;
  const int entries = (n_in + 1) * n_out;
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= entries) return;
  float s = 0.0f;
  for (uint32_t g = 0; g < kReduceGroups; ++g) s += sums[(uint64_t)g * entries + e];
  const int i = e / n_out, o = e % n_out;
  const int wo = layer == 1 ? kOffW1 : (layer == 2 ? kOffW2 : kOffW3);
  const int bo = layer == 1 ? kOffB1 : (layer == 2 ? kOffB2 : kOffB3);
  if (i == n_in) grad[bo + o] = s;
  else grad[wo + i * n_out + o] = s;
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, float lr, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
}

// ---- inference operands ------------------------------------------------------------
__global__ void k_pack(const float* p, uint16_t* ops) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kOps) return;
  const int r = e & 3;
  float v = 0.0f;
  if (e < kA1) {
    const int lane = (e >> 2) & 63, m = e >> 8;
    const int f = 32 * m + (lane & 31), k = 4 * (lane >> 5) + r;
    if (k < 3) v = p[kOffW1 + k * kHid + f];
    else if (k == 3) v = p[kOffB1 + f];
    else if (k < 7) v = p[kOffW1 + (k - 4) * kHid + f];
  } else if (e < kA1 + kA2) {
    const int x = e - kA1;
    const int lane = (x >> 2) & 63, kb = (x >> 8) % 9, m = (x >> 8) / 9;
    const int f = 32 * m + (lane & 31), k = 4 * (lane >> 5) + r;
    if (kb < 8) v = p[kOffW2 + (8 * kb + k) * kHid + f];
    else if (k == 0) v = p[kOffB2 + f];
  } else {
    const int x = e - kA1 - kA2;
    const int lane = (x >> 2) & 63, kb = x >> 8;
    const int c = lane & 31, k = 4 * (lane >> 5) + r;
    if (c < kOut) {
      if (kb < 8) v = p[kOffW3 + (8 * kb + k) * kOut + c];
      else if (k == 0) v = p[kOffB3 + c];
    }
  }
  ops[e] = (uint16_t)bf16_bits(v);
}

// Four floats as bfloat16, rounded to nearest even: gfx950 converts a pair
// per instruction (v_cvt_pk_bf16_f32), the same rounding as bf16_bits.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack2(float a, float b) {
  bf16x2 r;
  r.x = (__bf16)a;
  r.y = (__bf16)b;
  return __builtin_bit_cast(uint32_t, r);
}
__device__ __forceinline__ s4 pack4(float a, float b, float c, float d) {
  const uint2 u = make_uint2(pack2(a, b), pack2(c, d));
  return __builtin_bit_cast(s4, u);
}

// Activations of accumulator group g (rows 8g + 4(l/32) + 0..3) as the next
// layer's B operand.
__device__ __forceinline__ s4 act_block(const f16v& acc, int g) {
  return pack4(silu_f(acc[4 * g]), silu_f(acc[4 * g + 1]), silu_f(acc[4 * g + 2]), silu_f(acc[4 * g + 3]));
}

// FIN / FOUT: 1 = RGBA8 sRGB codes, 6 = RGBA32F.
template <int FIN, int FOUT>
__device__ __forceinline__ void infer_body(const uint16_t* ops, const void* src, void* dst, uint64_t n) {
  const int lane = (int)(threadIdx.x & 63);
  const bool upper = lane >= 32;
  const uint64_t wave = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const uint64_t nwaves = (uint64_t)gridDim.x * (kBlock / 64);
  const s4* o = reinterpret_cast<const s4*>(ops);
  s4 a1[2], a2[2][9], a3[9];
#pragma unroll
  for (int m = 0; m < 2; ++m) a1[m] = o[m * 64 + lane];
  o += kA1 / 4;
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int kb = 0; kb < 9; ++kb) a2[m][kb] = o[(m * 9 + kb) * 64 + lane];
  o += kA2 / 4;
#pragma unroll
  for (int kb = 0; kb < 9; ++kb) a3[kb] = o[kb * 64 + lane];
  s4 one;
  one.x = upper ? (short)0 : (short)0x3f80;  // bfloat16 1.0
  one.y = one.z = one.w = 0;
  const f16v zero = {};

  for (uint64_t tile = wave; tile * 32 < n; tile += nwaves) {
    const uint64_t p = tile * 32 + (uint64_t)(lane & 31);
    const uint64_t pc = p < n ? p : n - 1;
    float x, y, z;
    uint32_t araw = 0;
    float af = 0.0f;
    if constexpr (FIN == 1) {
      const uint32_t w = reinterpret_cast<const uint32_t*>(src)[pc];
      x = (float)(w & 255u) / 255.0f;
      y = (float)((w >> 8) & 255u) / 255.0f;
      z = (float)((w >> 16) & 255u) / 255.0f;
      araw = w >> 24;
    } else {
      const float4 v = reinterpret_cast<const float4*>(src)[pc];
      x = v.x;
      y = v.y;
      z = v.z;
      af = v.w;
    }
    // Layer 1: (x_hi, y_hi, z_hi, 1 | x_lo, y_lo, z_lo, 0) against (W1, b1 | W1, 0).
    const float xh = bf16r(x), yh = bf16r(y), zh_ = bf16r(z);
    const s4 b1 = upper ? pack4(x - xh, y - yh, z - zh_, 0.0f) : pack4(xh, yh, zh_, 1.0f);
    f16v acc0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1[0], b1, zero, 0, 0, 0);
    f16v acc1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a1[1], b1, zero, 0, 0, 0);
    // Layer 2.
    s4 b2[8];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      b2[g] = act_block(acc0, g);
      b2[4 + g] = act_block(acc1, g);
    }
    acc0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a2[0][8], one, zero, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a2[1][8], one, zero, 0, 0, 0);
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) {
      acc0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a2[0][kb], b2[kb], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a2[1][kb], b2[kb], acc1, 0, 0, 0);
    }
    // Layer 3 (rows 0..2 of the single output tile).
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      b2[g] = act_block(acc0, g);
      b2[4 + g] = act_block(acc1, g);
    }
    f16v out = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a3[8], one, zero, 0, 0, 0);
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) out = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a3[kb], b2[kb], out, 0, 0, 0);
    if (!upper && p < n) {
      if constexpr (FOUT == 1) {
        const uint32_t r = (uint32_t)rintf(fminf(fmaxf(out[0], 0.0f), 1.0f) * 255.0f);
        const uint32_t g = (uint32_t)rintf(fminf(fmaxf(out[1], 0.0f), 1.0f) * 255.0f);
        const uint32_t bb = (uint32_t)rintf(fminf(fmaxf(out[2], 0.0f), 1.0f) * 255.0f);
        const uint32_t a = FIN == 1 ? araw : (uint32_t)rintf(fminf(fmaxf(af, 0.0f), 1.0f) * 255.0f);
        reinterpret_cast<uint32_t*>(dst)[p] = r | (g << 8) | (bb << 16) | (a << 24);
      } else {
        const float a = FIN == 1 ? (float)araw / 255.0f : af;
        reinterpret_cast<float4*>(dst)[p] = make_float4(out[0], out[1], out[2], a);
      }
    }
  }
}

template <int FIN, int FOUT>
__global__ __launch_bounds__(kBlock) void k_infer(const uint16_t* ops, const void* src, void* dst, uint64_t n) {
  infer_body<FIN, FOUT>(ops, src, dst, n);
}

// A flat pointwise stage [load, colour ops, store] replaced by the surrogate:
// the image addresses come from the stage descriptor, which a rebind updates
// without touching a captured graph.
__global__ __launch_bounds__(kBlock) void k_infer_stage(const uint16_t* ops, const zh::StageDesc* d) {
  const uint64_t n = (uint64_t)d->width * d->height;
  const zh::CDesc c = zh::as_const(d);
  const void* src = (const void*)c->slot[zh::instr_slot(c, 0)].ptr;
  void* dst = (void*)c->slot[zh::instr_slot(c, c->n_instr - 1)].ptr;
  infer_body<1, 1>(ops, src, dst, n);
}

}  // namespace mlp
}  // namespace zh

extern "C" {

using namespace zh::mlp;

size_t zh_mlp_part_floats(uint32_t n) {
  return ((size_t)((n + kChunk - 1) / kChunk) + kReduceGroups) * (kHid + 1) * kHid;
}

hipError_t zh_mlp_quantise(const float* p, float* q, hipStream_t s) {
  hipLaunchKernelGGL(k_quant, dim3((kParams + 255) / 256), dim3(256), 0, s, p, q);
  return hipGetLastError();
}

// Loss gradient of one batch into grad[0, kParams); `b->loss` is incremented
// by the batch's mean squared error.  The buffer continues with the chunk
// partials: it must hold kParams + zh_mlp_part_floats(n) floats.
hipError_t zh_mlp_loss_grad(const float* q, const zh_mlp_batch* b, float* grad, hipStream_t s) {
  const uint32_t n = b->n;
  if (n == 0) return hipSuccess;
  const uint64_t hid = (uint64_t)n * kHid;
  const dim3 gh((uint32_t)((hid + kBlock - 1) / kBlock));
  hipLaunchKernelGGL(k_fwd_hidden, gh, dim3(kBlock), 0, s, *b, 1, q);
  hipLaunchKernelGGL(k_fwd_hidden, gh, dim3(kBlock), 0, s, *b, 2, q);
  hipLaunchKernelGGL(k_fwd_out, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, *b, q);
  hipLaunchKernelGGL(k_bwd_hidden, gh, dim3(kBlock), 0, s, *b, 2, q);
  hipLaunchKernelGGL(k_bwd_hidden, gh, dim3(kBlock), 0, s, *b, 1, q);
  const uint32_t chunks = (n + kChunk - 1) / kChunk;
  float* part = grad + kParams;                        // chunks * (kHid + 1) * kHid floats
  float* sums = part + (size_t)chunks * (kHid + 1) * kHid;  // kReduceGroups * (kHid + 1) * kHid
  for (int layer = 3; layer >= 1; --layer) {
    const int entries = ((layer == 1 ? kIn : kHid) + 1) * (layer == 3 ? kOut : kHid);
    const uint32_t eb = (entries + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_grad_part, dim3(chunks), dim3(kBlock), 0, s, *b, layer, part);
    hipLaunchKernelGGL(k_grad_reduce_groups, dim3(eb, kReduceGroups), dim3(kBlock), 0, s, part, chunks, entries,
                       sums);
    hipLaunchKernelGGL(k_grad_reduce, dim3(eb), dim3(kBlock), 0, s, sums, layer, grad);
  }
  return hipGetLastError();
}

hipError_t zh_mlp_adam(float* p, const float* g, float* m, float* v, float lr, float bc1, float bc2,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, p, g, m, v, lr, bc1, bc2);
  return hipGetLastError();
}

hipError_t zh_mlp_pack(const float* p, uint16_t* ops, hipStream_t s) {
  hipLaunchKernelGGL(k_pack, dim3((kOps + 255) / 256), dim3(256), 0, s, p, ops);
  return hipGetLastError();
}

hipError_t zh_mlp_infer(const uint16_t* ops, const void* src, void* dst, uint64_t n, int fmt_in, int fmt_out,
                        uint32_t cus, hipStream_t s) {
  if (n == 0) return hipSuccess;
  static uint32_t resident = 0;
  if (!resident) resident = zh::resident_blocks((const void*)k_infer<1, 1>, 1);
  const uint64_t tiles = (n + 31) / 32;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((tiles + 3) / 4, (uint64_t)resident * cus);
  if (fmt_in == 1 && fmt_out == 1)
    hipLaunchKernelGGL((k_infer<1, 1>), dim3(blocks), dim3(kBlock), 0, s, ops, src, dst, n);
  else if (fmt_in == 6 && fmt_out == 6)
    hipLaunchKernelGGL((k_infer<6, 6>), dim3(blocks), dim3(kBlock), 0, s, ops, src, dst, n);
  else if (fmt_in == 1 && fmt_out == 6)
    hipLaunchKernelGGL((k_infer<1, 6>), dim3(blocks), dim3(kBlock), 0, s, ops, src, dst, n);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t zh_mlp_infer_stage(const uint16_t* ops, const zh::StageDesc* d_dev, uint64_t n, uint32_t cus,
                              hipStream_t s) {
  if (n == 0) return hipSuccess;
  static uint32_t resident = 0;
  if (!resident) resident = zh::resident_blocks((const void*)k_infer_stage, 1);
  const uint64_t tiles = (n + 31) / 32;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((tiles + 3) / 4, (uint64_t)resident * cus);
  hipLaunchKernelGGL(k_infer_stage, dim3(blocks), dim3(kBlock), 0, s, ops, d_dev);
  return hipGetLastError();
}

}  // extern "C"
