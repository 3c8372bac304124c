// zh_device.h -- device helpers shared by the interpreter kernels
// (zh_kernels.hip) and the specialised pointwise kernels (zh_spec.hip):
// constant-space descriptor access, global-space 128-bit streaming accesses
// and the in-line codecs of the common formats.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zh_codec.h"
#include "zh_stage.h"

namespace zh {

typedef const __attribute__((address_space(4))) StageDesc* CDesc;
typedef const __attribute__((address_space(4))) BatchDesc* CBatch;

__device__ __forceinline__ CDesc as_const(const StageDesc* p) { return (CDesc)(uintptr_t)p; }

// Copy a descriptor record out of the constant address space (scalar loads).
template <typename T>
__device__ __forceinline__ T ldc(const __attribute__((address_space(4))) T* p) {
  static_assert(sizeof(T) % 4 == 0, "records are dword multiples");
  T r;
  const __attribute__((address_space(4))) uint32_t* src = (const __attribute__((address_space(4))) uint32_t*)p;
  uint32_t* dst = reinterpret_cast<uint32_t*>(&r);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = src[i];
  return r;
}

// Word loads through the scalar unit.  The descriptor tables are uniform and
// read-only; reading them as dwords from the constant address space keeps
// them on s_load (a byte field or a generic pointer would become a vector
// load, a full memory round trip on each tile's critical path).
typedef const __attribute__((address_space(4))) uint32_t* CWords;
__device__ __forceinline__ uint32_t cword(const uint32_t* p, uint32_t i) { return ((CWords)(uintptr_t)p)[i]; }
// Slot operand (the `a` byte) of instruction i of a stage.
__device__ __forceinline__ uint32_t instr_slot(CDesc d, uint32_t i) {
  return (((CWords)&d->code[i])[0] >> 8) & 255u;
}

typedef float4 Px[kPix];

// Image memory is addressed in the global address space (addrspace 1): the
// uniform slot base stays in SGPRs and loads/stores are global_* instructions
// counted by vmcnt alone, so they overlap with LDS traffic (generic FLAT
// accesses would also hold lgkmcnt and serialise with every LDS read).
typedef __attribute__((address_space(1))) uint8_t gbyte;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ T gld(const gbyte* p) {
  return *reinterpret_cast<const __attribute__((address_space(1))) T*>(p);
}
template <typename T>
__device__ __forceinline__ void gst(gbyte* p, T v) {
  *reinterpret_cast<__attribute__((address_space(1))) T*>(p) = v;
}

// Streaming (non-temporal) 128-bit accesses: every texel is read and written
// exactly once, so they do not claim cache residency.
__device__ __forceinline__ uint4 nt_load(const gbyte* p) {
  const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const __attribute__((address_space(1))) u32x4*>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void nt_store(uint4 v, gbyte* p) {
  const u32x4 w = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(w, reinterpret_cast<__attribute__((address_space(1))) u32x4*>(p));
}
__device__ __forceinline__ uint4 ld16(const gbyte* p) {
  const u32x4 v = gld<u32x4>(p);
  return make_uint4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ gbyte* texel_addr(const Slot& s, int x, int y) {
  return (gbyte*)s.ptr + ((uint64_t)y * s.row_stride + (uint64_t)x * s.bpp);
}

// ---- in-line codecs of the FmtCode formats -----------------------------------
template <uint32_t F>
__device__ __forceinline__ float4 dec_u8(uint32_t w, const Luts* L) {
  constexpr bool bgra = F == FMT_U8_SRGB_BGRA || F == FMT_U8_LIN_BGRA;
  constexpr bool srgb = F == FMT_U8_SRGB_RGBA || F == FMT_U8_SRGB_BGRA;
  const uint32_t b0 = w & 255u, b1 = (w >> 8) & 255u, b2 = (w >> 16) & 255u, b3 = w >> 24;
  const uint32_t r = bgra ? b2 : b0, b = bgra ? b0 : b2;
  const float a = (float)b3 * (1.0f / 255.0f);
  if (srgb) return make_float4(L->dec[r], L->dec[b1], L->dec[b], a);
  return make_float4((float)r * (1.0f / 255.0f), (float)b1 * (1.0f / 255.0f), (float)b * (1.0f / 255.0f), a);
}

// The colour bytes of an 8-bit texel (alpha byte 0).
template <uint32_t F>
__device__ __forceinline__ uint32_t enc_u8_rgb(float4 v, const Luts* L) {
  constexpr bool bgra = F == FMT_U8_SRGB_BGRA || F == FMT_U8_LIN_BGRA;
  constexpr bool srgb = F == FMT_U8_SRGB_RGBA || F == FMT_U8_SRGB_BGRA;
  const float lo = bgra ? v.z : v.x, hi = bgra ? v.x : v.z;
  if (srgb) return srgb_encode8(lo, L) | (srgb_encode8(v.y, L) << 8) | (srgb_encode8(hi, L) << 16);
  return put_unorm8(hi, 2u, put_unorm8(v.y, 1u, put_unorm8(lo, 0u, 0u)));
}

template <uint32_t F>
__device__ __forceinline__ uint32_t enc_u8(float4 v, const Luts* L) {
  return put_unorm8(v.w, 3u, enc_u8_rgb<F>(v, L));
}

__device__ __forceinline__ float4 dec_f16(uint32_t lo, uint32_t hi) {
  return make_float4(h2f(lo & 0xffffu), h2f(lo >> 16), h2f(hi & 0xffffu), h2f(hi >> 16));
}
__device__ __forceinline__ uint32_t pack_f16(float x, float y) { return f2h(x) | (f2h(y) << 16); }

__device__ __forceinline__ float3 mat3(const __attribute__((address_space(4))) float* m, float3 v) {
  return make_float3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
                     m[6] * v.x + m[7] * v.y + m[8] * v.z);
}

// Signed cube root: sign(x) 2^(log2|x| / 3) through the hardware log2/exp2
// (relative error ~1e-7 over the colour range; cbrt(0) = 0 since
// exp2(-inf) = 0).
__device__ __forceinline__ float cbrt_s(float x) { return copysignf(pow_pos(fabsf(x), 1.0f / 3.0f), x); }
// SrLab2's cone non-linearity and its inverse (srlab2.frag:87-120).  The
// shader's pow of a negative base is undefined; below -216/24389 the cube
// root is taken with its sign.
__device__ __forceinline__ float srlab_nl(float v) {
  return fabsf(v) < 216.0f / 24389.0f ? v * 24389.0f / 2700.0f : 1.16f * cbrt_s(v) - 0.16f;
}
__device__ __forceinline__ float srlab_nl_inv(float v) {
  if (fabsf(v) < 0.08f) return v * 2700.0f / 24389.0f;
  const float p = (v + 0.16f) / 1.16f;
  return p * p * p;
}

#ifndef __HIPCC_RTC__  // host code: kernels compiled at run time (zh_jit.cpp) see device code only
// Workgroups of `kernel` (kBlock lanes) resident on a device of `cus` compute
// units: the grid of a persistent kernel, which then has no second partial
// round of workgroups.
inline uint32_t resident_blocks(const void* kernel, uint32_t cus, size_t dyn_lds = 0) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, kBlock, dyn_lds) != hipSuccess || n <= 0) n = 1;
  return (uint32_t)n * cus;
}
#endif

// Copy the context's sRGB tables (zh_codec.h Luts) into LDS.
__device__ __forceinline__ void load_lut(CDesc d, Luts* lut) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(d->srgb_lut);
  uint32_t* dst = reinterpret_cast<uint32_t*>(lut);
  for (uint32_t i = threadIdx.x; i < sizeof(Luts) / 4; i += kBlock) dst[i] = src[i];
}

// ---- affine layers -------------------------------------------------------------
// Map a pixel centre through a layer's inverse transform (FP64).
__device__ __forceinline__ double2 map_pt(const double* m, double px, double py) {
  const double q = m[6] * px + m[7] * py + m[8];
  return make_double2((m[0] * px + m[1] * py + m[2]) / q, (m[3] * px + m[4] * py + m[5]) / q);
}

// Split a mapped coordinate into the texel left of the sample and the
// bilinear weight of the right texel (texel centres at i + 0.5).
__device__ __forceinline__ void split_tap(double u, int& i, float& f) {
  const double t = u - 0.5;
  const double fl = floor(t);
  i = (int)fl;
  f = (float)(t - fl);
}

// Source box of every bilinear tap of the tile's pixel centres (one texel of
// margin against rounding), clipped to the source grid.  Returns false when
// the tile's footprint misses the source.
__device__ __forceinline__ bool tap_box(const Layer& L, double ex0, double ey0, double ex1, double ey1, int& bx0,
                                        int& by0, int& bw, int& bh) {
  const double2 p0 = map_pt(L.inv, ex0, ey0), p1 = map_pt(L.inv, ex1, ey0);
  const double2 p2 = map_pt(L.inv, ex0, ey1), p3 = map_pt(L.inv, ex1, ey1);
  const double umin = fmin(fmin(p0.x, p1.x), fmin(p2.x, p3.x)) - 0.5;
  const double umax = fmax(fmax(p0.x, p1.x), fmax(p2.x, p3.x)) - 0.5;
  const double vmin = fmin(fmin(p0.y, p1.y), fmin(p2.y, p3.y)) - 0.5;
  const double vmax = fmax(fmax(p0.y, p1.y), fmax(p2.y, p3.y)) - 0.5;
  const int sw = (int)L.src_w, sh = (int)L.src_h;
  if (!(umax > -1.0 && vmax > -1.0 && umin < (double)sw && vmin < (double)sh)) return false;
  bx0 = max(0, (int)floor(umin) - 1);
  by0 = max(0, (int)floor(vmin) - 1);
  const int bx1 = min(sw - 1, (int)floor(umax) + 2), by1 = min(sh - 1, (int)floor(vmax) + 2);
  bw = bx1 - bx0 + 1;
  bh = by1 - by0 + 1;
  return bw > 0 && bh > 0;
}

}  // namespace zh
