// flat_field.hip -- a user stage kernel (binary_dynamic): flat-field
// correction of the reference's extension test (lib/std/src/flat_field.frag,
// driven by lib/zosimos/tests/custom.rs:187-285).  Shader data: {mean} f32.
// Both operands are sampled at the texel's texture coordinate (nearest,
// clamp-to-edge); the colour of the first is scaled by mean / field.x of the
// second, alpha kept.
#include <hip/hip_runtime.h>

#include "zosimos_hip.h"

__device__ static inline float4 nearest(const zh_dynamic_args& a, int k, float u, float v) {
  const uint32_t w = a.in_width[k], h = a.in_height[k];
  const uint32_t tx = min((uint32_t)fmaxf(floorf(u * (float)w), 0.0f), w - 1);
  const uint32_t ty = min((uint32_t)fmaxf(floorf(v * (float)h), 0.0f), h - 1);
  return reinterpret_cast<const float4*>(a.in[k])[(size_t)ty * w + tx];
}

extern "C" __global__ void __launch_bounds__(256) flat_field(zh_dynamic_args a) {
  const uint32_t x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
  if (x >= a.width || y >= a.height) return;
  const float mean = *static_cast<const float*>(a.data);
  const float u = ((float)x + 0.5f) / (float)a.width, v = ((float)y + 0.5f) / (float)a.height;
  const float4 c = nearest(a, 0, u, v);
  const float4 f = nearest(a, 1, u, v);
  const float eff = mean / f.x;
  reinterpret_cast<float4*>(a.out)[(size_t)y * a.width + x] = make_float4(c.x * eff, c.y * eff, c.z * eff, c.w);
}
