// zh_codec.h -- gfx950 texel unpack/pack and transfer functions.
//
// Semantics follow the reference's staging shader lib/std/src/stage.frag:
//   transfers              stage.frag:279-425 (parts_(un)transfer 750-809)
//   demux / mux            stage.frag:533-641 (division by BITSn, truncating
//                          uint(c * BITSn) after clamp(0, 1))
//   parts (de)normalize    stage.frag:654-747
// and, for the formats the hardware samples natively (RGBA8/BGRA8 with sRGB or
// linear transfer, program.rs:794-838), the unorm rules of the texture unit:
// decode x / 255 (sRGB through the exact 8-bit EOTF table), encode with
// round-to-nearest after clamp(0, 1).  RGBA16F / RGBA32F (absent in the
// reference, shaders/stage.rs:200-218) are IEEE values, f32 -> f16 rounds to
// nearest even, no clamp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zh_stage.h"

namespace zh {

// ---------------------------------------------------------------------------
// 8-bit sRGB tables, built on the host in FP64 and staged in LDS per block.
//   dec[k]   linear value of code k (IEC 61966-2-1, rounded once to f32)
//   idx[i]   code of the linear value i / 4096 (0 <= i <= 4096)
//   thr[k]   smallest f32 x whose code is >= k (thr[256] = +inf)
// Adjacent codes are at least 1 / (255 * 12.92) = 3.0e-4 apart in linear
// light, more than a 1/4096 bucket, so a bucket holds at most one threshold:
// code(x) = idx[i] + (x >= thr[idx[i] + 1]), exact for every f32 x.
// ---------------------------------------------------------------------------
// (struct Luts is declared in zh_stage.h, shared with the host builder.)

__device__ __forceinline__ float clamp01_(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

__device__ __forceinline__ uint32_t srgb_encode8(float v, const Luts* L) {
  const float c = clamp01_(v);  // NaN -> 0
  const uint32_t k = L->idx[(uint32_t)(c * 4096.0f)];  // bucket 4096 holds 1.0 alone
  return k + (c >= L->thr[k + 1] ? 1u : 0u);
}

// Round-to-nearest-even 8-bit unorm of v * 255 written into byte `pos` of w
// (v_cvt_pk_u8_f32: saturates to [0, 255] and maps NaN to 0, so no clamp is
// needed first; equal to rint(clamp01(v) * 255) for every f32 v).
__device__ __forceinline__ uint32_t put_unorm8(float v, uint32_t pos, uint32_t w) {
  return __builtin_amdgcn_cvt_pk_u8_f32(v * 255.0f, pos, w);
}

// ---------------------------------------------------------------------------
// transfer functions
// ---------------------------------------------------------------------------
// x^e for x > 0 through the hardware log2/exp2 (v_log_f32 / v_exp_f32, ~1 ulp).
__device__ __forceinline__ float pow_pos(float x, float e) {
  return __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(x));
}

__device__ __forceinline__ float oe_srgb(float v) {
  const float a = fabsf(v);
  const float r = a <= 0.0031308f ? a * 12.92f : 1.055f * pow_pos(a, 1.0f / 2.4f) - 0.055f;
  return copysignf(r, v);
}
__device__ __forceinline__ float eo_srgb(float v) {
  const float a = fabsf(v);
  const float r = a <= 0.04045f ? a * (1.0f / 12.92f) : pow_pos((a + 0.055f) * (1.0f / 1.055f), 2.4f);
  return copysignf(r, v);
}
__device__ __forceinline__ float oe_bt709(float v) {
  return v >= 0.018f ? 1.099f * powf(v, 0.45f) - 0.099f : 4.5f * v;
}
__device__ __forceinline__ float eo_bt709(float v) {
  // Threshold oe_bt709(0.018) as stage.frag:291 computes it.
  return v >= 1.099f * powf(0.018f, 0.45f) - 0.099f ? powf((v + 0.099f) / 1.099f, 1.0f / 0.45f) : v / 4.5f;
}
__device__ __forceinline__ float oe_smpte240(float v) {
  return v < 0.0228f ? 4.0f * v : 1.1115f * powf(v, 0.45f) - 0.1115f;
}
__device__ __forceinline__ float eo_smpte240(float v) {
  // stage.frag:330 subtracts 0.1115 (kept for parity with the reference).
  return v < 0.0913f ? v / 4.0f : powf((v - 0.1115f) / 1.1115f, 1.0f / 0.45f);
}
#define ZH_PQ_M1 (2610.0f / 16384.0f)
#define ZH_PQ_M2 (2523.0f / 4096.0f)
#define ZH_PQ_C1 (3424.0f / 4096.0f)
#define ZH_PQ_C2 (2413.0f / 128.0f)
#define ZH_PQ_C3 (2392.0f / 128.0f)
__device__ __forceinline__ float eo_pq(float v) {
  const float n = powf(v, 1.0f / ZH_PQ_M2);
  const float nom = fmaxf(n - ZH_PQ_C1, 0.0f);
  return powf(nom / (ZH_PQ_C2 - ZH_PQ_C3 * n), 1.0f / ZH_PQ_M1);
}
__device__ __forceinline__ float oe_pq(float v) {
  const float display = powf(oe_bt709(59.5208f * v), 2.4f) / 100.0f;
  const float y = powf(display, ZH_PQ_M1);
  return powf((ZH_PQ_C1 + ZH_PQ_C2 * y) / (ZH_PQ_C3 * y + 1.0f), ZH_PQ_M2);
}

// parts_untransfer (stage.frag:781-809): electrical -> linear on rgb.
__device__ __noinline__ float4 untransfer_generic(float4 c, uint32_t t) {
  switch (t) {
    case 5: return make_float4(eo_srgb(c.x), eo_srgb(c.y), eo_srgb(c.z), c.w);
    case 4: case 9: case 10: case 11: return c;
    case 0: case 6: case 7: return make_float4(eo_bt709(c.x), eo_bt709(c.y), eo_bt709(c.z), c.w);
    case 1: return make_float4(powf(c.x, 2.2f), powf(c.y, 2.2f), powf(c.z, 2.2f), c.w);
    case 2: return make_float4(oe_bt709(c.x), oe_bt709(c.y), oe_bt709(c.z), c.w);
    case 3: return make_float4(eo_smpte240(c.x), eo_smpte240(c.y), eo_smpte240(c.z), c.w);
    case 8: return make_float4(eo_pq(c.x), eo_pq(c.y), eo_pq(c.z), c.w);
    case 0x100: {  // transfer_lch_to_lab, stage.frag:422-425
      const float ang = (360.0f * (c.z - 0.5f)) * (3.14159265358979f / 180.0f);
      return make_float4(c.x, c.y * cosf(ang), c.y * sinf(ang), c.w);
    }
  }
  return c;
}

// parts_transfer (stage.frag:750-778): linear -> electrical on rgb.
__device__ __noinline__ float4 transfer_generic(float4 c, uint32_t t) {
  switch (t) {
    case 5: return make_float4(oe_srgb(c.x), oe_srgb(c.y), oe_srgb(c.z), c.w);
    case 4: case 9: case 10: case 11: return c;
    case 0: case 6: case 7: return make_float4(oe_bt709(c.x), oe_bt709(c.y), oe_bt709(c.z), c.w);
    case 1: return make_float4(powf(c.x, 1.0f / 2.2f), powf(c.y, 1.0f / 2.2f), powf(c.z, 1.0f / 2.2f), c.w);
    case 2: return make_float4(eo_bt709(c.x), eo_bt709(c.y), eo_bt709(c.z), c.w);
    case 3: return make_float4(oe_smpte240(c.x), oe_smpte240(c.y), oe_smpte240(c.z), c.w);
    case 8: return make_float4(oe_pq(c.x), oe_pq(c.y), oe_pq(c.z), c.w);
    case 0x100: {  // transfer_lab_to_lch, stage.frag:415-420
      const float ch = sqrtf(c.y * c.y + c.z * c.z);
      const float h = (atan2f(c.z, c.y) * (180.0f / 3.14159265358979f)) / 360.0f + 0.5f;
      return make_float4(c.x, ch, h, c.w);
    }
  }
  return c;
}

// Inline front ends: linear and sRGB in line, every other transfer out of line.
__device__ __forceinline__ float4 untransfer(float4 c, uint32_t t) {
  if (t == 4) return c;
  if (t == 5) return make_float4(eo_srgb(c.x), eo_srgb(c.y), eo_srgb(c.z), c.w);
  return untransfer_generic(c, t);
}
__device__ __forceinline__ float4 transfer(float4 c, uint32_t t) {
  if (t == 4) return c;
  if (t == 5) return make_float4(oe_srgb(c.x), oe_srgb(c.y), oe_srgb(c.z), c.w);
  return transfer_generic(c, t);
}

// ---------------------------------------------------------------------------
// demux / mux and parts permutations (stage.frag:533-747)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float fd(uint32_t v, float bits) { return (float)v / bits; }

__device__ __forceinline__ float4 demux(uint32_t n, uint32_t kind) {
  switch (kind) {
    case 0: { const float f = fd(n, 255.0f); return make_float4(f, f, f, f); }
    case 1: return make_float4(fd(n & 3u, 3.0f), fd((n >> 2) & 7u, 7.0f), fd(n >> 5, 7.0f), 1.0f);
    case 2: return make_float4(fd(n & 7u, 7.0f), fd((n >> 3) & 7u, 7.0f), fd(n >> 6, 3.0f), 1.0f);
    case 3: { const float f = fd(n, 65535.0f); return make_float4(f, f, f, f); }
    case 4: return make_float4(fd(n & 15u, 15.0f), fd((n >> 4) & 15u, 15.0f), fd((n >> 8) & 15u, 15.0f), fd(n >> 12, 15.0f));
    case 5: return make_float4(fd(n & 15u, 15.0f), fd((n >> 4) & 15u, 15.0f), fd((n >> 8) & 15u, 15.0f), 1.0f);
    case 6: return make_float4(fd((n >> 4) & 15u, 15.0f), fd((n >> 9) & 15u, 15.0f), fd((n >> 12) & 15u, 15.0f), 1.0f);
    case 7: return make_float4(fd(n & 31u, 31.0f), fd((n >> 5) & 63u, 63.0f), fd(n >> 11, 31.0f), 1.0f);
    case 8: return make_float4(fd(n & 255u, 255.0f), 0.0f, 0.0f, fd((n >> 8) & 255u, 255.0f));
    case 9: return make_float4(fd(n & 255u, 255.0f), fd((n >> 8) & 255u, 255.0f), fd((n >> 16) & 255u, 255.0f), 1.0f);
    case 10: return make_float4(fd(n & 255u, 255.0f), fd((n >> 8) & 255u, 255.0f), fd((n >> 16) & 255u, 255.0f), fd(n >> 24, 255.0f));
    case 11: return make_float4(fd(n & 65535u, 65535.0f), fd(n >> 16, 65535.0f), 0.0f, 1.0f);
    case 14: return make_float4(fd(n & 3u, 3.0f), fd((n >> 2) & 1023u, 1023.0f), fd((n >> 12) & 1023u, 1023.0f), fd(n >> 22, 1023.0f));
    case 15: return make_float4(fd(n & 1023u, 1023.0f), fd((n >> 10) & 1023u, 1023.0f), fd((n >> 20) & 1023u, 1023.0f), fd(n >> 30, 3.0f));
    case 16: return make_float4(fd((n >> 2) & 1023u, 1023.0f), fd((n >> 12) & 1023u, 1023.0f), fd(n >> 22, 1023.0f), 1.0f);
    case 17: return make_float4(fd(n & 1023u, 1023.0f), fd((n >> 10) & 1023u, 1023.0f), fd((n >> 20) & 1023u, 1023.0f), 1.0f);
  }
  return make_float4(1.0f, 0.0f, 0.0f, 1.0f);  // BIT_DECODE_FAIL
}

// uint(c * BITSn), c already clamped to [0, 1].
__device__ __forceinline__ uint32_t qt(float c, uint32_t nbits) {
  return (uint32_t)(c * (float)((1u << nbits) - 1u));
}

__device__ __forceinline__ uint32_t mux(float4 c, uint32_t kind) {
  switch (kind) {
    case 0: return qt(c.x, 8);
    case 1: return qt(c.x, 2) + (qt(c.y, 3) << 2) + (qt(c.z, 3) << 5);
    case 2: return qt(c.x, 3) + (qt(c.y, 3) << 3) + (qt(c.z, 2) << 6);
    case 3: return qt(c.x, 16);
    case 4: return qt(c.x, 4) + (qt(c.y, 4) << 4) + (qt(c.z, 4) << 8) + (qt(c.w, 4) << 12);
    case 5: return qt(c.x, 4) + (qt(c.y, 4) << 4) + (qt(c.z, 4) << 8);
    case 6: return (qt(c.x, 4) << 4) + (qt(c.y, 4) << 8) + (qt(c.z, 4) << 12);
    case 7: return qt(c.x, 5) + (qt(c.y, 6) << 5) + (qt(c.z, 5) << 11);
    case 8: return qt(c.x, 8) + (qt(c.w, 8) << 8);
    case 9: return qt(c.x, 8) + (qt(c.y, 8) << 8) + (qt(c.z, 8) << 16);
    case 10: return qt(c.x, 8) + (qt(c.y, 8) << 8) + (qt(c.z, 8) << 16) + (qt(c.w, 8) << 24);
    case 11: return qt(c.x, 16) + (qt(c.w, 16) << 16);
    case 14: return qt(c.x, 2) + (qt(c.y, 10) << 2) + (qt(c.z, 10) << 12) + (qt(c.w, 10) << 22);
    case 15: return qt(c.x, 10) + (qt(c.y, 10) << 10) + (qt(c.z, 10) << 20) + (qt(c.w, 2) << 30);
    case 16: return (qt(c.x, 10) << 2) + (qt(c.y, 10) << 12) + (qt(c.z, 10) << 22);
    case 17: return qt(c.x, 10) + (qt(c.y, 10) << 10) + (qt(c.z, 10) << 20);
  }
  return 0x55445544u;  // BIT_ENCODE_FAIL
}

__device__ __forceinline__ float4 parts_normalize(float4 c, uint32_t parts) {
  switch (parts) {
    case 0: return make_float4(0.0f, 0.0f, 0.0f, c.x);
    case 1: return make_float4(c.x, 0.0f, 0.0f, 1.0f);
    case 2: return make_float4(0.0f, c.x, 0.0f, 1.0f);
    case 3: return make_float4(0.0f, 0.0f, c.x, 1.0f);
    case 4: return make_float4(c.x, c.x, c.x, 1.0f);
    case 5: return make_float4(c.x, c.x, c.x, c.w);
    case 6: case 9: case 17: case 19: return make_float4(c.x, c.y, c.z, 1.0f);
    case 7: case 11: return make_float4(c.z, c.y, c.x, 1.0f);
    case 8: case 10: case 18: case 20: return c;  // Bgra unswizzled: stage.frag:676-677
    case 12: return make_float4(c.y, c.z, c.w, c.x);
    case 13: return make_float4(c.y, c.z, c.w, 1.0f);
    case 14: return make_float4(c.w, c.z, c.y, c.x);
    case 15: return make_float4(c.w, c.z, c.y, 1.0f);
  }
  return make_float4(1.0f, 0.0f, 0.0f, 1.0f);
}

__device__ __forceinline__ float4 parts_denormalize(float4 c, uint32_t parts) {
  switch (parts) {
    case 8: case 18: case 20: return c;
    case 0: return make_float4(c.w, 0.0f, 0.0f, 1.0f);
    case 1: return make_float4(c.x, 0.0f, 0.0f, 1.0f);
    case 2: return make_float4(c.y, 0.0f, 0.0f, 1.0f);
    case 3: return make_float4(c.z, 0.0f, 0.0f, 1.0f);
    case 4: return make_float4(c.x, c.x, c.x, 1.0f);
    case 5: return make_float4(c.x, c.x, c.x, c.w);
    case 6: case 9: case 17: case 19: return make_float4(c.x, c.y, c.z, 1.0f);
    case 7: case 11: return make_float4(c.z, c.y, c.x, 1.0f);
    case 10: return make_float4(c.z, c.y, c.x, c.w);
    case 12: return make_float4(c.w, c.x, c.y, c.z);
    case 13: return make_float4(1.0f, c.x, c.y, c.z);
    case 14: return make_float4(c.w, c.z, c.y, c.x);
    case 15: return make_float4(1.0f, c.z, c.y, c.x);
  }
  return c;
}

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

__device__ __forceinline__ float h2f(uint32_t bits16) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)bits16);
}
__device__ __forceinline__ uint32_t f2h(float v) {
  return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)v);
}

// ---------------------------------------------------------------------------
// whole-texel decode / encode.  A raw texel is up to 16 bytes in a uint4.
// ---------------------------------------------------------------------------
// Formats other than the fast ones below: staged codecs of stage.frag and
// 16-bit unorm / float texels with any parts and transfer.
__device__ __noinline__ float4 decode_generic(uint32_t bits, uint32_t parts, uint32_t tr, uint4 raw) {
  float4 comp;
  if (bits == 18) {
    comp = make_float4(h2f(raw.x & 0xffffu), h2f(raw.x >> 16), h2f(raw.y & 0xffffu), h2f(raw.y >> 16));
  } else if (bits == 19) {
    comp = make_float4(__uint_as_float(raw.x), __uint_as_float(raw.y), __uint_as_float(raw.z), __uint_as_float(raw.w));
  } else if (bits == 13) {
    comp = make_float4(fd(raw.x & 0xffffu, 65535.0f), fd(raw.x >> 16, 65535.0f), fd(raw.y & 0xffffu, 65535.0f),
                       fd(raw.y >> 16, 65535.0f));
  } else if (bits == 12) {
    comp = make_float4(fd(raw.x & 0xffffu, 65535.0f), fd(raw.x >> 16, 65535.0f), fd(raw.y & 0xffffu, 65535.0f), 1.0f);
  } else {
    comp = demux(raw.x, bits);
  }
  return untransfer(parts_normalize(comp, parts), tr);
}

__device__ __noinline__ uint4 encode_generic(uint32_t bits, uint32_t parts, uint32_t tr, float4 v) {
  const float4 c = parts_denormalize(transfer(v, tr), parts);
  if (bits == 18) return make_uint4(f2h(c.x) | (f2h(c.y) << 16), f2h(c.z) | (f2h(c.w) << 16), 0u, 0u);
  if (bits == 19)
    return make_uint4(__float_as_uint(c.x), __float_as_uint(c.y), __float_as_uint(c.z), __float_as_uint(c.w));
  const float4 k = make_float4(clamp01(c.x), clamp01(c.y), clamp01(c.z), clamp01(c.w));
  if (bits == 13) return make_uint4(qt(k.x, 16) | (qt(k.y, 16) << 16), qt(k.z, 16) | (qt(k.w, 16) << 16), 0u, 0u);
  if (bits == 12) return make_uint4(qt(k.x, 16) | (qt(k.y, 16) << 16), qt(k.z, 16), 0u, 0u);
  return make_uint4(mux(k, bits), 0u, 0u, 0u);
}

// Parts whose normalisation is the identity on rgba.
__device__ __forceinline__ bool parts_identity(uint32_t p) { return p == 8 || p == 18 || p == 20; }

// FAST instantiations only handle the formats with an in-line path (native
// RGBA8/BGRA8, linear RGBA16F/RGBA32F); the planner routes a stage there only
// when every slot has one of them.  Leaving the out-of-line generic codecs
// out keeps their register budget out of the kernel.
template <bool FAST = false>
__device__ __forceinline__ float4 decode_texel(const Slot& s, uint4 raw, const Luts* lut) {
  const uint32_t bits = s.bits;
  if (bits == 10 && s.native) {
    const uint32_t w = raw.x;
    const uint32_t b0 = w & 255u, b1 = (w >> 8) & 255u, b2 = (w >> 16) & 255u, b3 = w >> 24;
    const uint32_t r = s.parts == 10 ? b2 : b0;
    const uint32_t b = s.parts == 10 ? b0 : b2;
    const float a = (float)b3 * (1.0f / 255.0f);
    if (s.transfer == 5) return make_float4(lut->dec[r], lut->dec[b1], lut->dec[b], a);
    return make_float4((float)r * (1.0f / 255.0f), (float)b1 * (1.0f / 255.0f), (float)b * (1.0f / 255.0f), a);
  }
  if (bits == 18 && parts_identity(s.parts) && s.transfer == 4)
    return make_float4(h2f(raw.x & 0xffffu), h2f(raw.x >> 16), h2f(raw.y & 0xffffu), h2f(raw.y >> 16));
  if (bits == 19 && parts_identity(s.parts) && s.transfer == 4)
    return make_float4(__uint_as_float(raw.x), __uint_as_float(raw.y), __uint_as_float(raw.z), __uint_as_float(raw.w));
  if (FAST) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  return decode_generic(bits, s.parts, s.transfer, raw);
}

template <bool FAST = false>
__device__ __forceinline__ uint4 encode_texel(const Slot& s, float4 v, const Luts* lut) {
  const uint32_t bits = s.bits;
  if (bits == 10 && s.native) {
    // clamp01 maps NaN to 0 (fmaxf returns the non-NaN operand).
    const float lo_v = s.parts == 10 ? v.z : v.x, hi_v = s.parts == 10 ? v.x : v.z;
    uint32_t w;
    if (s.transfer == 5) {
      w = srgb_encode8(lo_v, lut) | (srgb_encode8(v.y, lut) << 8) | (srgb_encode8(hi_v, lut) << 16);
    } else {
      w = put_unorm8(hi_v, 2u, put_unorm8(v.y, 1u, put_unorm8(lo_v, 0u, 0u)));
    }
    return make_uint4(put_unorm8(v.w, 3u, w), 0u, 0u, 0u);
  }
  if (bits == 18 && parts_identity(s.parts) && s.transfer == 4)
    return make_uint4(f2h(v.x) | (f2h(v.y) << 16), f2h(v.z) | (f2h(v.w) << 16), 0u, 0u);
  if (bits == 19 && parts_identity(s.parts) && s.transfer == 4)
    return make_uint4(__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z), __float_as_uint(v.w));
  if (FAST) return make_uint4(0u, 0u, 0u, 0u);
  return encode_generic(bits, s.parts, s.transfer, v);
}

}  // namespace zh
