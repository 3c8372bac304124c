// mandelbrot.hip -- a user stage kernel (construct_dynamic, no operands):
// the Mandelbrot colouring of the reference's extension test
// (lib/std/src/mandelbrot.frag, driven by lib/zosimos/tests/custom.rs:14-96).
// Shader data: {scale.x, scale.y, position.x, position.y} f32; the point of
// texel (x, y) is c = (uv - position) * scale.  After 2048 steps of
// z <- z^2 + c the result is (L, a, b, 1) = (clamp(2 - |z|, 0, 0.7),
// mean(z) - c halved): lightness inside the set, chroma from the mean of the
// orbit.
#include <hip/hip_runtime.h>

#include "zosimos_hip.h"

// Every product and sum rounds on its own (fma only where the shader writes
// it), so the result is fixed by IEEE f32 arithmetic.
#pragma clang fp contract(off)

extern "C" __global__ void __launch_bounds__(256) mandelbrot(zh_dynamic_args a) {
  const uint32_t x = blockIdx.x * 16 + threadIdx.x, y = blockIdx.y * 16 + threadIdx.y;
  if (x >= a.width || y >= a.height) return;
  const float* p = static_cast<const float*>(a.data);
  const float u = ((float)x + 0.5f) / (float)a.width, v = ((float)y + 0.5f) / (float)a.height;
  const float cx = (u - p[2]) * p[0], cy = (v - p[3]) * p[1];
  float zx = 0.0f, zy = 0.0f, sx = 0.0f, sy = 0.0f;
  constexpr int kSteps = 2048;
  for (int i = 0; i < kSteps; ++i) {
    sx += zx;
    sy += zy;
    const float re = zx * zx - zy * zy;
    const float im = fmaf(2.0f * zx, zy, cy);
    zx = re + cx;
    zy = im;
  }
  sx = sx / (float)kSteps - cx;
  sy = sy / (float)kSteps - cy;
  const float len = sqrtf(zx * zx + zy * zy);
  const float light = fminf(fmaxf(2.0f - len, 0.0f), 0.7f);
  float4* out = reinterpret_cast<float4*>(a.out);
  out[(size_t)y * a.width + x] = make_float4(light, sx * 0.5f, sy * 0.5f, 1.0f);
}
