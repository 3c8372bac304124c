// crt.hip -- a user stage kernel (unary_dynamic): the shadow-mask pattern of
// the reference's extension test (lib/std/src/crt.frag, driven by
// lib/zosimos/tests/custom.rs:98-185).  Shader data: the result's size
// {w, h} as u32.  Each result texel samples its operand at the same texture
// coordinate with a nearest, clamp-to-edge lookup, keeps alpha and one colour
// channel (x mod 3), scaled by 0.16 c (5 - c) for the cell c = (y + 3 ((x / 3)
// mod 2)) mod 6 of a staggered 3 x 6 grid.
#include <hip/hip_runtime.h>

#include "zosimos_hip.h"

extern "C" __global__ void __launch_bounds__(256) crt(zh_dynamic_args a) {
  const uint32_t x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
  if (x >= a.width || y >= a.height) return;
  const uint32_t* scale = static_cast<const uint32_t*>(a.data);
  const float u = ((float)x + 0.5f) / (float)a.width, v = ((float)y + 0.5f) / (float)a.height;
  const uint32_t sx = (uint32_t)(u * (float)scale[0]), sy = (uint32_t)(v * (float)scale[1]);
  const uint32_t w = a.in_width[0], h = a.in_height[0];
  const uint32_t tx = min((uint32_t)fmaxf(floorf(u * (float)w), 0.0f), w - 1);
  const uint32_t ty = min((uint32_t)fmaxf(floorf(v * (float)h), 0.0f), h - 1);
  const float4 c = reinterpret_cast<const float4*>(a.in[0])[(size_t)ty * w + tx];
  const uint32_t bias = ((sx / 3) % 2) * 3;
  const uint32_t cell = (sy + bias) % 6;
  const float k = 0.16f * ((float)cell * (float)(5 - cell));
  const uint32_t ch = sx % 3;
  float4 r = make_float4(0.0f, 0.0f, 0.0f, c.w);
  if (ch == 0) r.x = c.x * k;
  if (ch == 1) r.y = c.y * k;
  if (ch == 2) r.z = c.z * k;
  reinterpret_cast<float4*>(a.out)[(size_t)y * a.width + x] = r;
}
