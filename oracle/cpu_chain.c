/* All-core FP32 CPU restatement of the five benchmark workloads (OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY: the benchmark's CPU-baseline leg times it on the
 * host cores beside the GPU, and tests/ check it against the FP64 oracle
 * (oracle/color.py, oracle/pipelines.py); the product library never links
 * it.  It stands in for the reference's own CPU path (its wgpu pipeline on
 * a software Vulkan adapter), which cannot be built here.
 *
 * What it computes follows the reference's per-pixel definitions:
 *   sRGB EOTF / OETF                  lib/std/src/stage.frag:335-352
 *   3x3 primaries matrices            lib/std/src/linear.frag:12-17
 *   Oklab (M1, cube root, M2, back)   lib/std/src/oklab.frag:14-64 (decode clamp :63)
 *   premultiplied 'over'              Porter & Duff 1984 (absent in the reference)
 *   bilinear affine, transparent edge the build's definition of BiLinear
 *                                     (command.rs:1659-1665 leaves it unimplemented)
 * with the colour matrices folded on the host (the caller passes them), as a
 * CPU implementation would.  8-bit sRGB decode reads a 256-entry table, encode
 * rounds to nearest; RGBA16F rounds to nearest even.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static float g_srgb_dec[256];
static float g_half[65536];
static int g_ready = 0;

static double eotf64(double v) { return v <= 0.04045 ? v / 12.92 : pow((v + 0.055) / 1.055, 2.4); }

static float half_to_float_bits(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  float f;
  if (e == 0) {
    f = (float)m * (1.0f / 16777216.0f); /* m * 2^-24, exact */
    uint32_t u;
    memcpy(&u, &f, 4);
    u |= sign;
    memcpy(&f, &u, 4);
    return f;
  }
  uint32_t u;
  if (e == 31) u = sign | 0x7f800000u | (m << 13);
  else u = sign | ((e + 112u) << 23) | (m << 13);
  memcpy(&f, &u, 4);
  return f;
}

/* IEEE binary16, round to nearest even. */
static uint16_t float_to_half(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u, a = x & 0x7fffffffu;
  if (a >= 0x7f800000u) return (uint16_t)(sign | (a > 0x7f800000u ? 0x7e00u : 0x7c00u));
  if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 rounds to infinity */
  if (a < 0x38800000u) {                                     /* below 2^-14: subnormal or zero */
    float v;
    memcpy(&v, &a, 4);
    return (uint16_t)(sign | (uint32_t)lrintf(v * 16777216.0f)); /* exact scaling, then nearest even */
  }
  uint32_t h = (((a >> 23) - 112u) << 10) | ((a & 0x7fffffu) >> 13);
  const uint32_t rem = a & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return (uint16_t)(sign | h);
}

void zsc_init(void) {
  if (g_ready) return;
  for (int k = 0; k < 256; ++k) g_srgb_dec[k] = (float)eotf64(k / 255.0);
  for (int k = 0; k < 65536; ++k) g_half[k] = half_to_float_bits((uint16_t)k);
  g_ready = 1;
}

int zsc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

uint16_t zsc_f2h(float f) { return float_to_half(f); }

static inline float clamp01f(float v) { return v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f; } /* NaN -> 0 */

static inline uint8_t srgb_enc8(float l) {
  const float c = clamp01f(l);
  const float e = c <= 0.0031308f ? c * 12.92f : 1.055f * powf(c, 1.0f / 2.4f) - 0.055f;
  return (uint8_t)lrintf(clamp01f(e) * 255.0f);
}
static inline uint8_t unorm8(float v) { return (uint8_t)lrintf(clamp01f(v) * 255.0f); }

static inline void mat3(const float* m, float* v) {
  const float x = v[0], y = v[1], z = v[2];
  v[0] = m[0] * x + m[1] * y + m[2] * z;
  v[1] = m[3] * x + m[4] * y + m[5] * z;
  v[2] = m[6] * x + m[7] * y + m[8] * z;
}

/* Oklab round trip with the adjustment folded into `mid`: lms = A rgb,
 * cube root, mid (M2^-1 adjust M2), cube, rgb = C lms, clamp (oklab.frag:63). */
static inline void oklab_chain(const float* A, const float* mid, const float* C, float* v) {
  mat3(A, v);
  for (int k = 0; k < 3; ++k) v[k] = cbrtf(v[k]);
  mat3(mid, v);
  for (int k = 0; k < 3; ++k) v[k] = v[k] * v[k] * v[k];
  mat3(C, v);
  for (int k = 0; k < 3; ++k) v[k] = clamp01f(v[k]);
}

static inline void load_px(const void* src, int fmt, int64_t i, float* v) {
  if (fmt == 0) {
    const uint8_t* p = (const uint8_t*)src + 4 * i;
    v[0] = g_srgb_dec[p[0]];
    v[1] = g_srgb_dec[p[1]];
    v[2] = g_srgb_dec[p[2]];
    v[3] = p[3] * (1.0f / 255.0f);
  } else {
    const uint16_t* p = (const uint16_t*)src + 4 * i;
    for (int k = 0; k < 4; ++k) v[k] = g_half[p[k]];
  }
}

static inline void store_px(void* dst, int fmt, int64_t i, const float* v) {
  if (fmt == 0) {
    uint8_t* p = (uint8_t*)dst + 4 * i;
    p[0] = srgb_enc8(v[0]);
    p[1] = srgb_enc8(v[1]);
    p[2] = srgb_enc8(v[2]);
    p[3] = unorm8(v[3]);
  } else {
    uint16_t* p = (uint16_t*)dst + 4 * i;
    for (int k = 0; k < 4; ++k) p[k] = float_to_half(v[k]);
  }
}

static inline void unpremul(float* v) {
  if (v[3] > 0.0f) {
    const float r = 1.0f / v[3];
    v[0] *= r;
    v[1] *= r;
    v[2] *= r;
  } else {
    v[0] = v[1] = v[2] = 0.0f;
  }
}

/* Premultiplied bilinear sample of a (w, h) source at output pixel (x, y)
 * through the inverse affine map inv (2x3 row-major), transparent outside.
 * `tap` fills the premultiplied value of texel (i, j). */
#define BILINEAR(OUT, INV, X, Y, W, H, TAP)                                            \
  do {                                                                                \
    const double px_ = (X) + 0.5, py_ = (Y) + 0.5;                                   \
    const double u_ = (INV)[0] * px_ + (INV)[1] * py_ + (INV)[2] - 0.5;               \
    const double v_ = (INV)[3] * px_ + (INV)[4] * py_ + (INV)[5] - 0.5;               \
    const double fx0_ = floor(u_), fy0_ = floor(v_);                                  \
    const float fx_ = (float)(u_ - fx0_), fy_ = (float)(v_ - fy0_);                   \
    const int64_t x0_ = (int64_t)fx0_, y0_ = (int64_t)fy0_;                           \
    float t_[4];                                                                      \
    const float w_[4] = {(1 - fx_) * (1 - fy_), fx_ * (1 - fy_), (1 - fx_) * fy_, fx_ * fy_}; \
    for (int c_ = 0; c_ < 4; ++c_) (OUT)[c_] = 0.0f;                                  \
    for (int k_ = 0; k_ < 4; ++k_) {                                                  \
      const int64_t xi_ = x0_ + (k_ & 1), yi_ = y0_ + (k_ >> 1);                       \
      if (xi_ < 0 || yi_ < 0 || xi_ >= (W) || yi_ >= (H)) continue;                    \
      TAP(t_, xi_, yi_);                                                              \
      for (int c_ = 0; c_ < 4; ++c_) (OUT)[c_] += w_[k_] * t_[c_];                    \
    }                                                                                 \
  } while (0)

/* C3, one image: decode -> 3x3 matrix -> bilinear resample by an axis map
 * (inv) -> encode in the input's format (fmt 0 RGBA8 sRGB, 1 RGBA16F). */
void zsc_c3_image(const void* src, int fmt, int w, int h, void* dst, int ow, int oh, const float* m3,
                  const double* inv) {
  zsc_init();
#define C3_TAP(T, XI, YI)                                   \
  do {                                                      \
    load_px(src, fmt, (int64_t)(YI) * w + (XI), T);         \
    mat3(m3, T);                                            \
    T[0] *= T[3];                                           \
    T[1] *= T[3];                                           \
    T[2] *= T[3];                                           \
  } while (0)
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < oh; ++y)
    for (int x = 0; x < ow; ++x) {
      float v[4];
      BILINEAR(v, inv, x, y, w, h, C3_TAP);
      unpremul(v);
      store_px(dst, fmt, (int64_t)y * ow + x, v);
    }
#undef C3_TAP
}

/* C2: RGBA8 sRGB -> (A, cube root, mid, cube, C) -> 'over' an RGBA16F layer ->
 * RGBA8 sRGB, n pixels. */
void zsc_c2(const uint8_t* src, const uint16_t* layer, uint8_t* dst, int64_t n, const float* A, const float* mid,
            const float* C) {
  zsc_init();
#pragma omp parallel for schedule(static, 65536)
  for (int64_t i = 0; i < n; ++i) {
    float v[4], u[4];
    load_px(src, 0, i, v);
    oklab_chain(A, mid, C, v);
    load_px(layer, 1, i, u);
    /* straight up over straight v */
    const float k = 1.0f - u[3], a = u[3] + v[3] * k;
    float o[4] = {u[0] * u[3] + v[0] * v[3] * k, u[1] * u[3] + v[1] * v[3] * k, u[2] * u[3] + v[2] * v[3] * k, a};
    unpremul(o);
    store_px(dst, 0, i, o);
  }
}

/* C4: RGBA8 sRGB through the Oklab chain, n pixels. */
void zsc_c4(const uint8_t* src, uint8_t* dst, int64_t n, const float* A, const float* mid, const float* C) {
  zsc_init();
#pragma omp parallel for schedule(static, 65536)
  for (int64_t i = 0; i < n; ++i) {
    float v[4];
    load_px(src, 0, i, v);
    oklab_chain(A, mid, C, v);
    store_px(dst, 0, i, v);
  }
}

/* C5: `layers` RGBA16F layers of (w, h), each resampled through its inverse
 * map inv[6 l ..] and composited 'over' in order onto transparency, then
 * un-premultiplied, the Oklab chain, RGBA8 sRGB (ow, oh). */
void zsc_c5(const uint16_t* const* src, int layers, int w, int h, const double* inv, uint8_t* dst, int ow, int oh,
            const float* A, const float* mid, const float* C) {
  zsc_init();
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < oh; ++y)
    for (int x = 0; x < ow; ++x) {
      float acc[4] = {0, 0, 0, 0};
      for (int l = 0; l < layers; ++l) {
        const uint16_t* s = src[l];
#define C5_TAP(T, XI, YI)                                  \
  do {                                                     \
    load_px(s, 1, (int64_t)(YI) * w + (XI), T);            \
    T[0] *= T[3];                                          \
    T[1] *= T[3];                                          \
    T[2] *= T[3];                                          \
  } while (0)
        float v[4];
        BILINEAR(v, inv + 6 * l, x, y, w, h, C5_TAP);
#undef C5_TAP
        const float k = 1.0f - v[3];
        for (int c = 0; c < 4; ++c) acc[c] = v[c] + acc[c] * k;
      }
      unpremul(acc);
      oklab_chain(A, mid, C, acc);
      store_px(dst, 0, (int64_t)y * ow + x, acc);
    }
}

/* C1: two RGBA8 sRGB images, premultiply, B 'over' A, bilinear affine onto
 * transparency (inv), un-premultiply, RGBA8 sRGB. */
void zsc_c1(const uint8_t* a, const uint8_t* b, int w, int h, const double* inv, uint8_t* dst, int ow, int oh) {
  zsc_init();
#define C1_TAP(T, XI, YI)                                                      \
  do {                                                                         \
    float p_[4], q_[4];                                                        \
    load_px(a, 0, (int64_t)(YI) * w + (XI), p_);                               \
    load_px(b, 0, (int64_t)(YI) * w + (XI), q_);                               \
    const float k_ = 1.0f - q_[3];                                             \
    for (int c_ = 0; c_ < 3; ++c_) T[c_] = q_[c_] * q_[3] + p_[c_] * p_[3] * k_; \
    T[3] = q_[3] + p_[3] * k_;                                                 \
  } while (0)
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < oh; ++y)
    for (int x = 0; x < ow; ++x) {
      float v[4];
      BILINEAR(v, inv, x, y, w, h, C1_TAP);
      unpremul(v);
      store_px(dst, 0, (int64_t)y * ow + x, v);
    }
#undef C1_TAP
}
