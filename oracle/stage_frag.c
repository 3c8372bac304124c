/* FP32, bit-level C restatement of the reference's staging codec
 * (lib/std/src/stage.frag).  TEST INFRASTRUCTURE ONLY: linked by tests/ through
 * ctypes as a checker; the product library never links it.
 *
 * Functions follow stage.frag line by line:
 *   demux_uint          stage.frag:533-582
 *   mux_uint            stage.frag:590-641   (truncating uint(c * BITS))
 *   parts_normalize     stage.frag:654-699
 *   parts_denormalize   stage.frag:703-747
 *   parts_transfer      stage.frag:750-778
 *   parts_untransfer    stage.frag:781-809
 *   transfer_*          stage.frag:279-425
 * The decode entry (stage.frag:439-450 / 471-482 / 503-513) is
 * demux -> normalize -> untransfer; the encode entry (452-469 / 484-501 /
 * 515-525) is transfer -> denormalize -> clamp(0, 1) -> mux.
 * GLSL float ops map to C float ops; GLSL pow maps to powf (the driver's pow
 * accuracy is not pinned by anything in the reference).
 */
#include <math.h>
#include <stdint.h>

typedef struct { float x, y, z, w; } vec4;

static vec4 v4(float x, float y, float z, float w) { vec4 r = {x, y, z, w}; return r; }
static const vec4 DECODE_FAIL = {1.0f, 0.0f, 0.0f, 1.0f};
#define ENCODE_FAIL 0x55445544u

/* ---- transfer functions ------------------------------------------------ */
static float oe_bt709(float v) { return v >= 0.018f ? 1.099f * powf(v, 0.45f) - 0.099f : 4.500f * v; }
static float eo_bt709(float v) {
  return v >= oe_bt709(0.018f) ? powf((v + 0.099f) / 1.099f, 1.0f / 0.45f) : v / 4.500f;
}
static float oe_bt470m(float v) { return powf(v, 1.0f / 2.200f); }
static float eo_bt470m(float v) { return powf(v, 2.200f); }
static float oe_bt601(float v) { return eo_bt709(v); }
static float eo_bt601(float v) { return oe_bt709(v); }
static float oe_smpte240(float v) { return v < 0.0228f ? 4.0f * v : 1.1115f * powf(v, 0.45f) - 0.1115f; }
static float eo_smpte240(float v) { return v < 0.0913f ? v / 4.0f : powf((v - 0.1115f) / 1.1115f, 1.0f / 0.45f); }
static float oe_srgb(float v) {
  if (v < -0.0031308f) return -1.055f * powf(-v, 1.0f / 2.4f) + 0.055f;
  if (v <= 0.0031308f) return v * 12.92f;
  return 1.055f * powf(v, 1.0f / 2.4f) - 0.055f;
}
static float eo_srgb(float v) {
  if (v < -0.04045f) return -powf((-v + 0.055f) / 1.055f, 2.4f);
  if (v <= 0.04045f) return v / 12.92f;
  return powf((v + 0.055f) / 1.055f, 2.4f);
}
#define PQ_M1 (2610.0f / 16384.0f)
#define PQ_M2 (2523.0f / 4096.0f)
#define PQ_C1 (3424.0f / 4096.0f)
#define PQ_C2 (2413.0f / 128.0f)
#define PQ_C3 (2392.0f / 128.0f)
static float eo_smpte2084(float v) {
  float n = powf(v, 1.0f / PQ_M2);
  float nom = fmaxf(n - PQ_C1, 0.0f);
  float den = PQ_C2 - PQ_C3 * n;
  return powf(nom / den, 1.0f / PQ_M1);
}
static float eo_inv_smpte2084(float v) {
  float y = powf(v, PQ_M1);
  return powf((PQ_C1 + PQ_C2 * y) / (PQ_C3 * y + 1.0f), PQ_M2);
}
static float oe_smpte2084(float v) {
  float e = oe_bt709(59.5208f * v);
  return eo_inv_smpte2084(powf(e, 2.4f) / 100.0f);
}

static vec4 lab_to_lch(vec4 c) {
  float ch = sqrtf(c.y * c.y + c.z * c.z);
  float h = (atan2f(c.z, c.y) * (180.0f / 3.14159265358979f)) / 360.0f + 0.5f;
  return v4(c.x, ch, h, c.w);
}
static vec4 lch_to_lab(vec4 c) {
  float a = (360.0f * (c.z - 0.5f)) * (3.14159265358979f / 180.0f);
  return v4(c.x, c.y * cosf(a), c.y * sinf(a), c.w);
}

#define XYZ(E, FN) v4(FN(E.x), FN(E.y), FN(E.z), E.w)
static vec4 parts_transfer(vec4 l, uint32_t t) {
  switch (t) {
    case 0: return XYZ(l, oe_bt709);
    case 1: return XYZ(l, oe_bt470m);
    case 2: return XYZ(l, oe_bt601);
    case 3: return XYZ(l, oe_smpte240);
    case 4: return l;
    case 5: return XYZ(l, oe_srgb);
    case 6: case 7: return XYZ(l, oe_bt709);
    case 8: return XYZ(l, oe_smpte2084);
    case 10: return l;
    case 0x100: return lab_to_lch(l);
  }
  return l;
}
static vec4 parts_untransfer(vec4 n, uint32_t t) {
  switch (t) {
    case 0: return XYZ(n, eo_bt709);
    case 1: return XYZ(n, eo_bt470m);
    case 2: return XYZ(n, eo_bt601);
    case 3: return XYZ(n, eo_smpte240);
    case 4: return n;
    case 5: return XYZ(n, eo_srgb);
    case 6: case 7: return XYZ(n, eo_bt709);
    case 8: return XYZ(n, eo_smpte2084);
    case 10: return n;
    case 0x100: return lch_to_lab(n);
  }
  return n;
}

/* ---- bits -------------------------------------------------------------- */
static vec4 demux_uint(uint32_t num, uint32_t kind) {
  switch (kind) {
    case 0: { float f = (float)num / 255.0f; return v4(f, f, f, f); }
    case 1: return v4((float)(num & 3u) / 3.0f, (float)((num >> 2) & 7u) / 7.0f, (float)(num >> 5) / 7.0f, 1.0f);
    case 2: return v4((float)(num & 7u) / 7.0f, (float)((num >> 3) & 7u) / 7.0f, (float)(num >> 6) / 3.0f, 1.0f);
    case 3: { float f = (float)num / 65535.0f; return v4(f, f, f, f); }
    case 4: return v4((float)(num & 15u) / 15.0f, (float)((num >> 4) & 15u) / 15.0f, (float)((num >> 8) & 15u) / 15.0f, (float)(num >> 12) / 15.0f);
    case 5: return v4((float)(num & 15u) / 15.0f, (float)((num >> 4) & 15u) / 15.0f, (float)((num >> 8) & 15u) / 15.0f, 15.0f / 15.0f);
    case 6: return v4((float)((num >> 4) & 15u) / 15.0f, (float)((num >> 9) & 15u) / 15.0f, (float)((num >> 12) & 15u) / 15.0f, 15.0f / 15.0f);
    case 7: return v4((float)(num & 31u) / 31.0f, (float)((num >> 5) & 63u) / 63.0f, (float)(num >> 11) / 31.0f, 1.0f);
    case 8: return v4((float)(num & 255u) / 255.0f, 0.0f / 255.0f, 0.0f / 255.0f, (float)((num >> 8) & 255u) / 255.0f);
    case 9: return v4((float)(num & 255u) / 255.0f, (float)((num >> 8) & 255u) / 255.0f, (float)((num >> 16) & 255u) / 255.0f, 255.0f / 255.0f);
    case 10: return v4((float)(num & 255u) / 255.0f, (float)((num >> 8) & 255u) / 255.0f, (float)((num >> 16) & 255u) / 255.0f, (float)(num >> 24) / 255.0f);
    case 11: return v4((float)(num & 65535u) / 65535.0f, (float)((num >> 16) & 65535u) / 65535.0f, 0.0f / 65535.0f, 65535.0f / 65535.0f);
    case 14: return v4((float)(num & 3u) / 3.0f, (float)((num >> 2) & 1023u) / 1023.0f, (float)((num >> 12) & 1023u) / 1023.0f, (float)(num >> 22) / 1023.0f);
    case 15: return v4((float)(num & 1023u) / 1023.0f, (float)((num >> 10) & 1023u) / 1023.0f, (float)((num >> 20) & 1023u) / 1023.0f, (float)(num >> 30) / 3.0f);
    case 16: return v4((float)((num >> 2) & 1023u) / 1023.0f, (float)((num >> 12) & 1023u) / 1023.0f, (float)(num >> 22) / 1023.0f, 1.0f);
    case 17: return v4((float)(num & 1023u) / 1023.0f, (float)((num >> 10) & 1023u) / 1023.0f, (float)((num >> 20) & 1023u) / 1023.0f, 1.0f);
  }
  return DECODE_FAIL;
}

/* GLSL uint(float) truncates toward zero; inputs are clamped to [0, 1].
 * BITSn = 2^n - 1 exactly as stage.frag:175-190 defines them. */
static uint32_t q(float c, uint32_t n) { return (uint32_t)(c * (float)((1u << n) - 1u)); }

/* STEALTH_MUX_UINTk(s0, b0, s1, b1, ...) of stage.frag:591-594: each field is
 * skipped by s_i bits and quantised to b_i bits; fields are added in order. */
static uint32_t mux_uint(vec4 c, uint32_t kind) {
#define M1(s0, b0) q(c.x, b0)
#define M2(s0, b0, s1, b1) ((q(c.x, b0) << (s0)) + (q(c.w, b1) << ((s0) + (b0) + (s1))))
#define M3(s0, b0, s1, b1, s2, b2) ((q(c.x, b0) << (s0)) + (q(c.y, b1) << ((s0) + (b0) + (s1))) + (q(c.z, b2) << ((s0) + (b0) + (s1) + (b1) + (s2))))
#define M4(s0, b0, s1, b1, s2, b2, s3, b3) (M3(s0, b0, s1, b1, s2, b2) + (q(c.w, b3) << ((s0) + (b0) + (s1) + (b1) + (s2) + (b2) + (s3))))
  switch (kind) {
    case 0: return M1(0, 8);
    case 1: return M3(0, 2, 0, 3, 0, 3);
    case 2: return M3(0, 3, 0, 3, 0, 2);
    case 3: return M1(0, 16);
    case 4: return M4(0, 4, 0, 4, 0, 4, 0, 4);
    case 5: return M3(0, 4, 0, 4, 0, 4);
    case 6: return M3(4, 4, 0, 4, 0, 4);
    case 7: return M3(0, 5, 0, 6, 0, 5);
    case 8: return M2(0, 8, 0, 8);
    case 9: return M3(0, 8, 0, 8, 0, 8);
    case 10: return M4(0, 8, 0, 8, 0, 8, 0, 8);
    case 11: return M2(0, 16, 0, 16);
    case 14: return M4(0, 2, 0, 10, 0, 10, 0, 10);
    case 15: return M4(0, 10, 0, 10, 0, 10, 0, 2);
    case 16: return M3(2, 10, 0, 10, 0, 10);
    case 17: return M3(0, 10, 0, 10, 0, 10);
  }
  return ENCODE_FAIL;
#undef M1
#undef M2
#undef M3
#undef M4
}

static vec4 parts_normalize(vec4 c, uint32_t parts) {
  switch (parts) {
    case 0: return v4(0, 0, 0, c.x);
    case 1: return v4(c.x, 0, 0, 1);
    case 2: return v4(0, c.x, 0, 1);
    case 3: return v4(0, 0, c.x, 1);
    case 4: return v4(c.x, c.x, c.x, 1);
    case 5: return v4(c.x, c.x, c.x, c.w);
    case 6: case 9: return v4(c.x, c.y, c.z, 1);
    case 7: case 11: return v4(c.z, c.y, c.x, 1);
    case 8: return c;
    case 10: return c;
    case 12: return v4(c.y, c.z, c.w, c.x);
    case 14: return v4(c.w, c.z, c.y, c.x);
    case 13: return v4(c.y, c.z, c.w, 1);
    case 15: return v4(c.w, c.z, c.y, 1);
    case 17: case 19: return v4(c.x, c.y, c.z, 1);
    case 18: case 20: return c;
  }
  return DECODE_FAIL;
}

static vec4 parts_denormalize(vec4 c, uint32_t parts) {
  switch (parts) {
    case 8: return c;
    case 0: return v4(c.w, 0, 0, 1);
    case 1: return v4(c.x, 0, 0, 1);
    case 2: return v4(c.y, 0, 0, 1);
    case 3: return v4(c.z, 0, 0, 1);
    case 4: return v4(c.x, c.x, c.x, 1);
    case 5: return v4(c.x, c.x, c.x, c.w);
    case 6: case 9: return v4(c.x, c.y, c.z, 1);
    case 7: case 11: return v4(c.z, c.y, c.x, 1);
    case 10: return v4(c.z, c.y, c.x, c.w);
    case 12: return v4(c.w, c.x, c.y, c.z);
    case 14: return v4(c.w, c.z, c.y, c.x);
    case 13: return v4(1, c.x, c.y, c.z);
    case 15: return v4(1, c.z, c.y, c.x);
    case 17: case 19: return v4(c.x, c.y, c.z, 1);
    case 18: case 20: return c;
  }
  return c;
}

static float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }

/* Decode one staged texel value (already extracted from its u32 stage word by
 * the caller: stage.frag:441 / 473 / 505). */
void zso_stage_decode(uint32_t num, uint32_t bits, uint32_t parts, uint32_t transfer, float out[4]) {
  vec4 c = demux_uint(num, bits);
  vec4 e = parts_normalize(c, parts);
  vec4 p = parts_untransfer(e, transfer);
  out[0] = p.x; out[1] = p.y; out[2] = p.z; out[3] = p.w;
}

/* Encode one linear RGBA value into its texel bits. */
uint32_t zso_stage_encode(const float in[4], uint32_t bits, uint32_t parts, uint32_t transfer) {
  vec4 p = v4(in[0], in[1], in[2], in[3]);
  vec4 e = parts_transfer(p, transfer);
  vec4 c = parts_denormalize(e, parts);
  c = v4(clamp01(c.x), clamp01(c.y), clamp01(c.z), clamp01(c.w));
  return mux_uint(c, bits);
}

/* Batched helpers for numpy callers. */
void zso_stage_decode_n(const uint32_t* nums, int64_t n, uint32_t bits, uint32_t parts, uint32_t transfer, float* out) {
  for (int64_t i = 0; i < n; i++) zso_stage_decode(nums[i], bits, parts, transfer, out + 4 * i);
}
void zso_stage_encode_n(const float* in, int64_t n, uint32_t bits, uint32_t parts, uint32_t transfer, uint32_t* out) {
  for (int64_t i = 0; i < n; i++) out[i] = zso_stage_encode(in + 4 * i, bits, parts, transfer);
}

/* Single-channel transfer access for tests of the closed forms. */
float zso_transfer_decode1(float v, uint32_t transfer) {
  vec4 r = parts_untransfer(v4(v, v, v, 1.0f), transfer);
  return r.x;
}
float zso_transfer_encode1(float v, uint32_t transfer) {
  vec4 r = parts_transfer(v4(v, v, v, 1.0f), transfer);
  return r.x;
}
