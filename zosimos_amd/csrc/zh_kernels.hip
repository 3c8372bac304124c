// zh_kernels.hip -- gfx950 kernels of the fused compositing stages.
//
// One lane owns kPix = 4 horizontally adjacent pixels.  A stage's micro-op
// program (zh_stage.h) is read through the constant address space, so every
// instruction fetch and dispatch is scalar; the per-pixel colour work runs on
// the vector ALUs.  Contiguous, 16-byte aligned groups move through 128-bit
// global loads/stores (RGBA8: one dwordx4 per lane, RGBA16F: two, RGBA32F:
// four), so each wave touches 1-4 KiB of consecutive HBM per instruction.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>

#include "../../include/zosimos_hip.h"
#include "zh_device.h"

namespace zh {

// The pixels a lane works on: either kPix contiguous pixels of one row
// starting at (x0, y), of which the first n exist, or kPix gathered texel
// coordinates with their own validity.
struct Ctx {
  int x0, y, n;
  int dx, dy;       // OP_OFFSET displacement for loads
  bool gather;
  int gx[kPix], gy[kPix];
  bool gv[kPix];
};

// Fetch one texel's bytes (bpp in {1, 2, 3, 4, 6, 8, 16}).
__device__ __forceinline__ uint4 fetch_one(const Slot& s, int x, int y) {
  const gbyte* p = texel_addr(s, x, y);
  switch (s.bpp) {
    case 1: return make_uint4(gld<uint8_t>(p), 0u, 0u, 0u);
    case 2: return make_uint4(gld<uint16_t>(p), 0u, 0u, 0u);
    case 3: return make_uint4((uint32_t)gld<uint8_t>(p) | ((uint32_t)gld<uint8_t>(p + 1) << 8) |
                                  ((uint32_t)gld<uint8_t>(p + 2) << 16), 0u, 0u, 0u);
    case 4: return make_uint4(gld<uint32_t>(p), 0u, 0u, 0u);
    case 6:
      return make_uint4((uint32_t)gld<uint16_t>(p) | ((uint32_t)gld<uint16_t>(p + 2) << 16), gld<uint16_t>(p + 4),
                        0u, 0u);
    case 8: {
      const u32x2 v = gld<u32x2>(p);
      return make_uint4(v.x, v.y, 0u, 0u);
    }
    default: return ld16(p);
  }
}

__device__ __forceinline__ void store_one(const Slot& s, int x, int y, uint4 v) {
  gbyte* p = texel_addr(s, x, y);
  switch (s.bpp) {
    case 1: gst<uint8_t>(p, (uint8_t)v.x); break;
    case 2: gst<uint16_t>(p, (uint16_t)v.x); break;
    case 3:
      gst<uint8_t>(p, (uint8_t)v.x);
      gst<uint8_t>(p + 1, (uint8_t)(v.x >> 8));
      gst<uint8_t>(p + 2, (uint8_t)(v.x >> 16));
      break;
    case 4: gst<uint32_t>(p, v.x); break;
    case 6:
      gst<uint16_t>(p, (uint16_t)v.x);
      gst<uint16_t>(p + 2, (uint16_t)(v.x >> 16));
      gst<uint16_t>(p + 4, (uint16_t)v.y);
      break;
    case 8: gst<u32x2>(p, u32x2{v.x, v.y}); break;
    default: gst<u32x4>(p, u32x4{v.x, v.y, v.z, v.w}); break;
  }
}

__device__ __forceinline__ bool in_slot(const Slot& s, int x, int y) {
  return x >= 0 && y >= 0 && x < (int)s.width && y < (int)s.height;
}

// Decode a full group of kPix texels of 4 or 8 bytes from its raw 16-byte
// words (bpp 4: a; bpp 8: a, b).
template <bool FAST = false>
__device__ __forceinline__ void decode_group(const Slot& s, uint4 a, uint4 b, Px& X, const Luts* lut) {
  if (s.bpp == 4) {
    X[0] = decode_texel<FAST>(s, make_uint4(a.x, 0u, 0u, 0u), lut);
    X[1] = decode_texel<FAST>(s, make_uint4(a.y, 0u, 0u, 0u), lut);
    X[2] = decode_texel<FAST>(s, make_uint4(a.z, 0u, 0u, 0u), lut);
    X[3] = decode_texel<FAST>(s, make_uint4(a.w, 0u, 0u, 0u), lut);
  } else {
    X[0] = decode_texel<FAST>(s, make_uint4(a.x, a.y, 0u, 0u), lut);
    X[1] = decode_texel<FAST>(s, make_uint4(a.z, a.w, 0u, 0u), lut);
    X[2] = decode_texel<FAST>(s, make_uint4(b.x, b.y, 0u, 0u), lut);
    X[3] = decode_texel<FAST>(s, make_uint4(b.z, b.w, 0u, 0u), lut);
  }
}

// Decode the lane's full group of slot s from raw words a, b (4-byte texels:
// a; 8-byte: a, b); 16-byte texels are read here.
__device__ __forceinline__ void dec_group_fmt(const Slot& s, int x0, uint4 a, uint4 b, Px& X, const Luts* L) {
  switch (s.fmt) {
#define ZH_U8(F)                        \
  case F:                               \
    X[0] = dec_u8<F>(a.x, L);           \
    X[1] = dec_u8<F>(a.y, L);           \
    X[2] = dec_u8<F>(a.z, L);           \
    X[3] = dec_u8<F>(a.w, L);           \
    break;
    ZH_U8(FMT_U8_SRGB_RGBA)
    ZH_U8(FMT_U8_SRGB_BGRA)
    ZH_U8(FMT_U8_LIN_RGBA)
    ZH_U8(FMT_U8_LIN_BGRA)
#undef ZH_U8
    case FMT_F16:
      X[0] = dec_f16(a.x, a.y);
      X[1] = dec_f16(a.z, a.w);
      X[2] = dec_f16(b.x, b.y);
      X[3] = dec_f16(b.z, b.w);
      break;
    default: {  // FMT_F32
      const gbyte* q = texel_addr(s, x0, 0);
#pragma unroll
      for (int i = 0; i < kPix; ++i) {
        const uint4 v = nt_load(q + 16 * i);
        X[i] = make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
      }
      break;
    }
  }
}

__device__ __forceinline__ void enc_group_fmt(const Slot& s, int x0, const Px& X, const Luts* L) {
  gbyte* q = texel_addr(s, x0, 0);
  switch (s.fmt) {
#define ZH_U8(F)                                                                                         \
  case F:                                                                                                \
    nt_store(make_uint4(enc_u8<F>(X[0], L), enc_u8<F>(X[1], L), enc_u8<F>(X[2], L), enc_u8<F>(X[3], L)), q); \
    break;
    ZH_U8(FMT_U8_SRGB_RGBA)
    ZH_U8(FMT_U8_SRGB_BGRA)
    ZH_U8(FMT_U8_LIN_RGBA)
    ZH_U8(FMT_U8_LIN_BGRA)
#undef ZH_U8
    case FMT_F16:
      nt_store(make_uint4(pack_f16(X[0].x, X[0].y), pack_f16(X[0].z, X[0].w), pack_f16(X[1].x, X[1].y),
                          pack_f16(X[1].z, X[1].w)), q);
      nt_store(make_uint4(pack_f16(X[2].x, X[2].y), pack_f16(X[2].z, X[2].w), pack_f16(X[3].x, X[3].y),
                          pack_f16(X[3].z, X[3].w)), q + 16);
      break;
    default:  // FMT_F32
#pragma unroll
      for (int i = 0; i < kPix; ++i)
        nt_store(make_uint4(__float_as_uint(X[i].x), __float_as_uint(X[i].y), __float_as_uint(X[i].z),
                            __float_as_uint(X[i].w)), q + 16 * i);
      break;
  }
}

// X[i] <- decode(slot) at the lane's pixels; missing pixels read zero.
template <bool FAST = false>
__device__ __forceinline__ void load_px(const Slot& s, const Ctx& c, Px& X, const Luts* lut) {
  if (!c.gather) {
    const int x = c.x0 + c.dx, y = c.y + c.dy;
    const bool full = c.n == kPix && s.vec_ok && (((uint32_t)x * s.bpp) & 15u) == 0 && in_slot(s, x, y) &&
                      in_slot(s, x + kPix - 1, y);
    if (full && (s.bpp == 4 || s.bpp == 8)) {
      const gbyte* p = texel_addr(s, x, y);
      const uint4 a = nt_load(p);
      const uint4 b = s.bpp == 8 ? nt_load(p + 16) : make_uint4(0u, 0u, 0u, 0u);
      decode_group<FAST>(s, a, b, X, lut);
      return;
    }
    if (full && s.bpp == 16) {
      const gbyte* p = texel_addr(s, x, y);
#pragma unroll
      for (int i = 0; i < kPix; ++i) X[i] = decode_texel<FAST>(s, nt_load(p + 16 * i), lut);
      return;
    }
#pragma unroll
    for (int i = 0; i < kPix; ++i) {
      const bool ok = i < c.n && in_slot(s, x + i, y);
      X[i] = ok ? decode_texel<FAST>(s, fetch_one(s, x + i, y), lut) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < kPix; ++i) {
    const int x = c.gx[i] + c.dx, y = c.gy[i] + c.dy;
    const bool ok = c.gv[i] && in_slot(s, x, y);
    X[i] = ok ? decode_texel<FAST>(s, fetch_one(s, x, y), lut) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

template <bool FAST = false>
__device__ __forceinline__ void store_px(const Slot& s, const Ctx& c, const Px& X, const Luts* lut) {
  const int x = c.x0, y = c.y;
  if (c.n == kPix && s.vec_ok && (((uint32_t)x * s.bpp) & 15u) == 0) {
    if (s.bpp == 4) {
      const uint4 w = make_uint4(encode_texel<FAST>(s, X[0], lut).x, encode_texel<FAST>(s, X[1], lut).x,
                                 encode_texel<FAST>(s, X[2], lut).x, encode_texel<FAST>(s, X[3], lut).x);
      nt_store(w, texel_addr(s, x, y));
      return;
    }
    if (s.bpp == 8) {
      const uint4 a0 = encode_texel<FAST>(s, X[0], lut), a1 = encode_texel<FAST>(s, X[1], lut);
      const uint4 a2 = encode_texel<FAST>(s, X[2], lut), a3 = encode_texel<FAST>(s, X[3], lut);
      gbyte* p = texel_addr(s, x, y);
      nt_store(make_uint4(a0.x, a0.y, a1.x, a1.y), p);
      nt_store(make_uint4(a2.x, a2.y, a3.x, a3.y), p + 16);
      return;
    }
    if (s.bpp == 16) {
      gbyte* p = texel_addr(s, x, y);
#pragma unroll
      for (int i = 0; i < kPix; ++i) nt_store(encode_texel<FAST>(s, X[i], lut), p + 16 * i);
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < kPix; ++i)
    if (i < c.n) store_one(s, x + i, y, encode_texel<FAST>(s, X[i], lut));
}

// Raw texels of a lane's group for up to two contiguous slots, loaded one
// grid-stride iteration ahead of their use so the loads of the next group
// are in flight while this group's colour math runs.
struct Pre {
  uint4 a0, b0;  // slot pre_slot[0]
  uint4 a1, b1;  // slot pre_slot[1]
};

// Evaluate instructions [begin, end) of the stage program.  FAST programs
// (see decode_texel) take their contiguous loads from the prefetched words.
// ---- fractal_noise.frag ---------------------------------------------------------
// pcg4d (Jarzynski and Olano, "Hash Functions for GPU Rendering", JCGT 9(3),
// 2020): an LCG step per lane, then two rounds of cross-lane multiply-adds
// around a 16-bit xorshift; wrapping 32-bit arithmetic.
__device__ __forceinline__ uint4 pcg4d(uint4 v) {
  v.x = v.x * 1664525u + 1013904223u;
  v.y = v.y * 1664525u + 1013904223u;
  v.z = v.z * 1664525u + 1013904223u;
  v.w = v.w * 1664525u + 1013904223u;
  v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
  v.x ^= v.x >> 16; v.y ^= v.y >> 16; v.z ^= v.z >> 16; v.w ^= v.w >> 16;
  v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
  return v;
}

// The corner value of a cell: pcg4d of (cell, 0, 0) over 2^32 - 1 (as f32,
// that is 2^32), each lane a channel.
__device__ __forceinline__ float4 cell_value(uint32_t sx, uint32_t sy) {
  const uint4 h = pcg4d(make_uint4(sx, sy, 0u, 0u));
  const float k = 4294967295.0f;
  return make_float4((float)h.x / k, (float)h.y / k, (float)h.z / k, (float)h.w / k);
}

// f32 floor as the shader's ivec2(floor(.)) cell index, saturated to int32.
__device__ __forceinline__ uint32_t cell_index(float f) {
  return (uint32_t)(int32_t)fminf(fmaxf(f, -2147483648.0f), 2147483520.0f);
}

// Octave-summed value noise at (x, y) (uv of the pixel centre); every
// product and sum rounded on its own (no contraction), in the shader's order,
// so the result is a fixed function of the parameters.
__device__ __forceinline__ float4 fractal_noise(float x, float y, const __attribute__((address_space(4))) float* q) {
#pragma clang fp contract(off)
  const float sx = q[0], sy = q[1], damping = q[3];
  const uint32_t octaves = (uint32_t)q[4];
  // 2 mat3(cos .5, sin .5, 1, -sin .5, cos .5, 1, 0, 0, 1): only x and y
  // reach the noise.
  const float kc = 1.75516512378075f, ks = 0.958851077208406f;
  float amp = q[2];
  float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (uint32_t o = 0; o < octaves; ++o) {  // uniform
    const float px = x * sx, py = y * sy;
    const float flx = floorf(px), fly = floorf(py);
    const float fx = px - flx, fy = py - fly;
    const uint32_t cx = cell_index(flx), cy = cell_index(fly);
    const float4 a = cell_value(cx, cy), b = cell_value(cx + 1u, cy);
    const float4 c = cell_value(cx, cy + 1u), e = cell_value(cx + 1u, cy + 1u);
    const float ux = fx * fx * (3.0f - 2.0f * fx), uy = fy * fy * (3.0f - 2.0f * fy);
    const float vx = 1.0f - ux;
    // mix(a, b, ux) + (c - a) uy (1 - ux) + (d - b) ux uy, mix as a (1 - t) + b t.
    v.x = v.x + amp * (a.x * vx + b.x * ux + (c.x - a.x) * uy * vx + (e.x - b.x) * ux * uy);
    v.y = v.y + amp * (a.y * vx + b.y * ux + (c.y - a.y) * uy * vx + (e.y - b.y) * ux * uy);
    v.z = v.z + amp * (a.z * vx + b.z * ux + (c.z - a.z) * uy * vx + (e.z - b.z) * ux * uy);
    v.w = v.w + amp * (a.w * vx + b.w * ux + (c.w - a.w) * uy * vx + (e.w - b.w) * ux * uy);
    const float nx = kc * x + -ks * y, ny = ks * x + kc * y;
    x = nx;
    y = ny;
    amp = amp * damping;
  }
  return v;
}

template <bool FAST = false>
__device__ __forceinline__ void run(CDesc d, uint32_t begin, uint32_t end, Ctx& c, Px& X, Px& Y, Px& Z,
                                    const Luts* lut, Pre pre = Pre{}, bool has_pre = false, bool full = false) {
  for (uint32_t pc = begin; pc < end; ++pc) {
    const Instr in = ldc(&d->code[pc]);
    switch (in.op) {
      case OP_LOAD:
      case OP_SOLID: {
        const Slot sl = ldc(&d->slot[in.a]);
        if (in.b) {
#pragma unroll
          for (int i = 0; i < kPix; ++i) { Z[i] = Y[i]; Y[i] = X[i]; }
        }
        if (in.op == OP_LOAD) {
          if (FAST && full) {
            // Full aligned group of a FmtCode format: the slot tests are
            // uniform; take the prefetched words or load them, then one
            // specialised decode.
            const bool p0 = has_pre && (int)in.a == d->pre_slot[0];
            const bool p1 = has_pre && (int)in.a == d->pre_slot[1];
            uint4 a = make_uint4(0u, 0u, 0u, 0u), b = a;
            if (p0 || p1) {
              a = p0 ? pre.a0 : pre.a1;
              b = p0 ? pre.b0 : pre.b1;
            } else if (sl.fmt != FMT_F32) {
              const gbyte* q = texel_addr(sl, c.x0, 0);
              a = nt_load(q);
              if (sl.fmt == FMT_F16) b = nt_load(q + 16);
            }
            dec_group_fmt(sl, c.x0, a, b, X, lut);
          } else {
            load_px<FAST>(sl, c, X, lut);
          }
        } else {
          const float4 v = make_float4(d->params[in.arg], d->params[in.arg + 1], d->params[in.arg + 2],
                                       d->params[in.arg + 3]);
#pragma unroll
          for (int i = 0; i < kPix; ++i) X[i] = v;
        }
        break;
      }
      case OP_GRID: {
        // bilinear.frag at the pixel centres: the stage runs 2-D (not flat),
        // so (x0 + i, y) are image coordinates.  mix(a, b, t) is evaluated as
        // a + (b - a) t, exact where a = b.
        if (in.b) {
#pragma unroll
          for (int i = 0; i < kPix; ++i) { Z[i] = Y[i]; Y[i] = X[i]; }
        }
        const __attribute__((address_space(4))) float* q = d->params + in.arg;
        const float v = ((float)c.y + 0.5f) / (float)d->height;
#pragma unroll
        for (int i = 0; i < kPix; ++i) {
          const float u = ((float)(c.x0 + i) + 0.5f) / (float)d->width, uv = u * v;
          float r[4];
#pragma unroll
          for (int k = 0; k < 4; ++k)
            r[k] = fmaf(q[4 + k] - q[k], u, q[k]) + fmaf(q[12 + k] - q[8 + k], v, q[8 + k]) +
                   fmaf(q[20 + k] - q[16 + k], uv, q[16 + k]);
          X[i] = make_float4(r[0], r[1], r[2], r[3]);
        }
        break;
      }
      case OP_NORMAL2D: {
        // distribution_normal2d.frag at the pixel centres (2-D stage): the
        // quadratic form of the inverse covariance, read as the shader's
        // column-major mat2 of the row-major blob (the form is symmetric in
        // the off-diagonal pair).
        if (in.b) {
#pragma unroll
          for (int i = 0; i < kPix; ++i) { Z[i] = Y[i]; Y[i] = X[i]; }
        }
        const __attribute__((address_space(4))) float* q = d->params + in.arg;
        const float py = 2.0f * (((float)c.y + 0.5f) / (float)d->height - 0.5f) - q[1];
        const float root = sqrtf(q[6]);
#pragma unroll
        for (int i = 0; i < kPix; ++i) {
          const float px = 2.0f * (((float)(c.x0 + i) + 0.5f) / (float)d->width - 0.5f) - q[0];
          const float e = 0.5f * (px * (q[2] * px + q[4] * py) + py * (q[3] * px + q[5] * py));
          const float g = expf(-e) / root;
          X[i] = make_float4(g, g, g, 1.0f);
        }
        break;
      }
      case OP_PALETTE: {
        // palette.frag: the texel of the slot nearest to the coordinates the
        // pixel's colour selects, clamped to the edge (the index is clamped
        // before the conversion, so any value, NaN included, stays inside).
        const Slot sl = ldc(&d->slot[in.a]);
        const __attribute__((address_space(4))) float* q = d->params + in.arg;
        const float w = (float)sl.width, h = (float)sl.height;
#pragma unroll
        for (int i = 0; i < kPix; ++i) {
          const float4 v = X[i];
          const float u = q[0] * v.x + q[1] * v.y + q[2] * v.z + q[3] * v.w + q[8];
          const float t = q[4] * v.x + q[5] * v.y + q[6] * v.z + q[7] * v.w + q[9];
          const int tx = (int)fminf(fmaxf(floorf(u * w), 0.0f), w - 1.0f);
          const int ty = (int)fminf(fmaxf(floorf(t * h), 0.0f), h - 1.0f);
          X[i] = decode_texel<FAST>(sl, fetch_one(sl, tx, ty), lut);
        }
        break;
      }
      case OP_NOISE: {
        // fractal_noise.frag at the pixel centres (2-D stage).
        if (in.b) {
#pragma unroll
          for (int i = 0; i < kPix; ++i) { Z[i] = Y[i]; Y[i] = X[i]; }
        }
        const __attribute__((address_space(4))) float* q = d->params + in.arg;
        const float v = ((float)c.y + 0.5f) / (float)d->height;
#pragma unroll
        for (int i = 0; i < kPix; ++i) X[i] = fractal_noise(((float)(c.x0 + i) + 0.5f) / (float)d->width, v, q);
        break;
      }
      case OP_BOX3: {
        // box3.frag: the weighted 3x3 neighbourhood of the slot's texels,
        // coordinates clamped to the slot (clamp-to-edge sampling); alpha 1.
        // The stage runs 2-D, so (x0 + i, y) are image coordinates.
        const Slot sl = ldc(&d->slot[in.a]);
        if (in.b) {
#pragma unroll
          for (int i = 0; i < kPix; ++i) { Z[i] = Y[i]; Y[i] = X[i]; }
        }
        const __attribute__((address_space(4))) float* w = d->params + in.arg;
        const int W = (int)sl.width - 1, H = (int)sl.height - 1;
        for (int i = 0; i < kPix; ++i) {
          float r = 0.0f, g = 0.0f, b = 0.0f;
          if (i < c.n) {
            for (int dy = 0; dy < 3; ++dy) {
              const int yy = min(max(c.y + dy - 1, 0), H);
#pragma unroll
              for (int dx = 0; dx < 3; ++dx) {
                const int xx = min(max(c.x0 + i + dx - 1, 0), W);
                const float4 t = decode_texel<FAST>(sl, fetch_one(sl, xx, yy), lut);
                const float k = w[dy * 3 + dx];
                r = fmaf(k, t.x, r);
                g = fmaf(k, t.y, g);
                b = fmaf(k, t.z, b);
              }
            }
          }
          X[i] = make_float4(r, g, b, 1.0f);
        }
        break;
      }
      case OP_MAT3: {
        const __attribute__((address_space(4))) float* m = d->params + in.arg;
#pragma unroll
        for (int i = 0; i < kPix; ++i) {
          const float3 r = mat3(m, make_float3(X[i].x, X[i].y, X[i].z));
          X[i] = make_float4(r.x, r.y, r.z, X[i].w);
        }
        break;
      }
      case OP_CBRT:
        if (in.a) {  // input proved non-negative
#pragma unroll
          for (int i = 0; i < kPix; ++i)
            X[i] = make_float4(pow_pos(X[i].x, 1.0f / 3.0f), pow_pos(X[i].y, 1.0f / 3.0f), pow_pos(X[i].z, 1.0f / 3.0f),
                               X[i].w);
        } else {
#pragma unroll
          for (int i = 0; i < kPix; ++i) X[i] = make_float4(cbrt_s(X[i].x), cbrt_s(X[i].y), cbrt_s(X[i].z), X[i].w);
        }
        break;
      case OP_CUBE:
#pragma unroll
        for (int i = 0; i < kPix; ++i)
          X[i] = make_float4(X[i].x * X[i].x * X[i].x, X[i].y * X[i].y * X[i].y, X[i].z * X[i].z * X[i].z, X[i].w);
        break;
      case OP_SRLAB:
#pragma unroll
        for (int i = 0; i < kPix; ++i) X[i] = make_float4(srlab_nl(X[i].x), srlab_nl(X[i].y), srlab_nl(X[i].z), X[i].w);
        break;
      case OP_SRLAB_INV:
#pragma unroll
        for (int i = 0; i < kPix; ++i)
          X[i] = make_float4(srlab_nl_inv(X[i].x), srlab_nl_inv(X[i].y), srlab_nl_inv(X[i].z), X[i].w);
        break;
      case OP_CLAMP01:
#pragma unroll
        for (int i = 0; i < kPix; ++i) X[i] = make_float4(clamp01(X[i].x), clamp01(X[i].y), clamp01(X[i].z), X[i].w);
        break;
      case OP_PREMUL:
#pragma unroll
        for (int i = 0; i < kPix; ++i) X[i] = make_float4(X[i].x * X[i].w, X[i].y * X[i].w, X[i].z * X[i].w, X[i].w);
        break;
      case OP_UNPREMUL:
#pragma unroll
        for (int i = 0; i < kPix; ++i) {
          const float a = X[i].w;
          const float r = a > 0.0f ? __builtin_amdgcn_rcpf(a) : 0.0f;
          X[i] = make_float4(X[i].x * r, X[i].y * r, X[i].z * r, a);
        }
        break;
      case OP_OVER:
#pragma unroll
        for (int i = 0; i < kPix; ++i) {
          const float k = 1.0f - X[i].w;
          X[i] = make_float4(fmaf(Y[i].x, k, X[i].x), fmaf(Y[i].y, k, X[i].y), fmaf(Y[i].z, k, X[i].z),
                             fmaf(Y[i].w, k, X[i].w));
          Y[i] = Z[i];
        }
        break;
      case OP_SWAP:
#pragma unroll
        for (int i = 0; i < kPix; ++i) { const float4 t = X[i]; X[i] = Y[i]; Y[i] = t; }
        break;
      case OP_CHANNEL: {
        const __attribute__((address_space(4))) float* q = d->params + in.arg;
#pragma unroll
        for (int i = 0; i < kPix; ++i)
          X[i] = make_float4(fmaf(X[i].x, q[0], q[4]), fmaf(X[i].y, q[1], q[5]), fmaf(X[i].z, q[2], q[6]),
                             fmaf(X[i].w, q[3], q[7]));
        break;
      }
      case OP_INJECT: {
        // inject.frag: mix(below, vec4(dot(above, color)), select) with
        // below in Y and above in X; GLSL mix as a (1 - t) + b t.
        const __attribute__((address_space(4))) float* q = d->params + in.arg;
#pragma unroll
        for (int i = 0; i < kPix; ++i) {
          const float v = q[4] * X[i].x + q[5] * X[i].y + q[6] * X[i].z + q[7] * X[i].w;
          X[i] = make_float4(Y[i].x * (1.0f - q[0]) + v * q[0], Y[i].y * (1.0f - q[1]) + v * q[1],
                             Y[i].z * (1.0f - q[2]) + v * q[2], Y[i].w * (1.0f - q[3]) + v * q[3]);
          Y[i] = Z[i];
        }
        break;
      }
      case OP_EOTF:
        if constexpr (!FAST) {
#pragma unroll
          for (int i = 0; i < kPix; ++i) X[i] = untransfer(X[i], in.arg);
        }
        break;
      case OP_OETF:
        if constexpr (!FAST) {
#pragma unroll
          for (int i = 0; i < kPix; ++i) X[i] = transfer(X[i], in.arg);
        }
        break;
      case OP_QUANT: {
        if constexpr (!FAST) {
          const Slot s = ldc(&d->slot[in.a]);
#pragma unroll
          for (int i = 0; i < kPix; ++i) X[i] = decode_texel<FAST>(s, encode_texel<FAST>(s, X[i], lut), lut);
        }
        break;
      }
      case OP_STORE: {
        const Slot s = ldc(&d->slot[in.a]);
        if (FAST && full)
          enc_group_fmt(s, c.x0, X, lut);
        else
          store_px<FAST>(s, c, X, lut);
        break;
      }
      case OP_OFFSET:
        c.dx = in.a ? (int)d->params[in.arg] : 0;
        c.dy = in.a ? (int)d->params[in.arg + 1] : 0;
        break;
      default:
        break;
    }
  }
}

// Raw loads of the full group g for the prefetched slots (flat stages).
__device__ __forceinline__ void issue(CDesc d, uint64_t g, Pre& p) {
  const uint64_t x0 = g * kPix;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int sl = d->pre_slot[k];
    if (sl < 0) continue;
    const uint64_t ptr = d->slot[sl].ptr;
    const uint32_t bpp = d->slot[sl].bpp;
    const gbyte* q = (const gbyte*)ptr + x0 * bpp;
    const uint4 a = nt_load(q);
    const uint4 b = bpp == 8 ? nt_load(q + 16) : make_uint4(0u, 0u, 0u, 0u);
    if (k == 0) { p.a0 = a; p.b0 = b; } else { p.a1 = a; p.b1 = b; }
  }
}

__device__ __forceinline__ Ctx flat_ctx(uint64_t g, int n) {
  Ctx c;
  c.dx = c.dy = 0;
  c.gather = false;
  c.x0 = (int)(g * kPix);
  c.y = 0;
  c.n = n;
  return c;
}

// ---------------------------------------------------------------------------
// pointwise stage: a persistent grid-stride loop; each lane owns kPix pixels
// per step.
//   FAST (flat, in-line codecs): full groups with the raw texels of the next
//   step already in flight (the address of the last step is reused instead
//   of branching, so the loop body has no divergent control flow), then the
//   one partial group of the image tail.
//   Otherwise: the general interpreter, one step at a time.
// ---------------------------------------------------------------------------
template <bool FAST>
__global__ __launch_bounds__(kBlock) void k_pointwise(const StageDesc* dd) {
  const CDesc d = as_const(dd);
  __shared__ Luts luts;
  load_lut(d, &luts);
  __syncthreads();
  const bool flat = d->flat != 0;
  const uint64_t total = (uint64_t)d->width * d->height;
  const uint32_t gpr = (d->width + kPix - 1) / kPix;
  const uint64_t ngroups = flat ? (total + kPix - 1) / kPix : (uint64_t)gpr * d->height;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if constexpr (FAST) {
    const uint64_t nfull = total / kPix;
    const bool prefetch = d->pre_slot[0] >= 0;
    // Two register buffers in ping-pong: the step on buffer A runs while B's
    // loads are in flight and vice versa (no copy between them, which would
    // make the step wait for the newest loads).
    Pre A, B;
    if (prefetch && g < nfull) issue(d, g, A);
    // (Programs write a register before reading it, so the stack needs no
    // initialisation here.)
    while (g < nfull) {
      if (prefetch) issue(d, min(g + stride, nfull - 1), B);
      {
        Ctx c = flat_ctx(g, kPix);
        Px X, Y, Z;
        run<true>(d, 0, d->n_instr, c, X, Y, Z, &luts, A, prefetch, true);
      }
      g += stride;
      if (g >= nfull) break;
      if (prefetch) issue(d, min(g + stride, nfull - 1), A);
      {
        Ctx c = flat_ctx(g, kPix);
        Px X, Y, Z;
        run<true>(d, 0, d->n_instr, c, X, Y, Z, &luts, B, prefetch, true);
      }
      g += stride;
    }
    if (g == nfull && nfull * kPix < total) {
      Ctx c = flat_ctx(g, (int)(total - nfull * kPix));
      Px X, Y, Z;
#pragma unroll
      for (int i = 0; i < kPix; ++i) X[i] = Y[i] = Z[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      run<true>(d, 0, d->n_instr, c, X, Y, Z, &luts);
    }
    return;
  }
  for (; g < ngroups; g += stride) {
    Ctx c;
    c.dx = c.dy = 0;
    c.gather = false;
    if (flat) {
      const uint64_t x0 = g * kPix;
      c.x0 = (int)x0;
      c.y = 0;
      c.n = (int)min((uint64_t)kPix, total - x0);
    } else {
      const uint32_t y = (uint32_t)(g / gpr);
      c.y = (int)y;
      c.x0 = (int)((g - (uint64_t)y * gpr) * kPix);
      c.n = min(kPix, (int)d->width - c.x0);
    }
    Px X, Y, Z;
#pragma unroll
    for (int i = 0; i < kPix; ++i) X[i] = Y[i] = Z[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    run<false>(d, 0, d->n_instr, c, X, Y, Z, &luts);
  }
}

// ---------------------------------------------------------------------------
// resample stage
// ---------------------------------------------------------------------------
__device__ __forceinline__ float4 f4_fma(float4 a, float k, float4 b) {
  return make_float4(fmaf(a.x, k, b.x), fmaf(a.y, k, b.y), fmaf(a.z, k, b.z), fmaf(a.w, k, b.w));
}
__device__ __forceinline__ float4 f4_scale(float4 a, float k) {
  return make_float4(a.x * k, a.y * k, a.z * k, a.w * k);
}
__device__ __forceinline__ float4 over4(float4 top, float4 below) {
  return f4_fma(below, 1.0f - top.w, top);
}

enum TileMode { TILE_STAGED = 0, TILE_DIRECT = 1, TILE_NEAREST = 2 };

// One output tile of a resample stage.  All lanes of the workgroup take the
// same branches at every barrier (tile and layer decisions are uniform).
template <int MODE>
__device__ __forceinline__ void process_tile(CDesc d, int tx, int ty, float4* tile, const Luts* lut) {
  const int TW = (int)d->tile_w, TH = (int)d->tile_h;
  const int tpr = TW / kPix;
  const int lx = (int)threadIdx.x % tpr, ly = (int)threadIdx.x / tpr;
  const int W = (int)d->width, H = (int)d->height;
  Ctx c;
  c.dx = c.dy = 0;
  c.gather = false;
  c.x0 = tx * TW + lx * kPix;
  c.y = ty * TH + ly;
  c.n = (ly < TH && c.y < H) ? max(0, min(kPix, W - c.x0)) : 0;

  Px X, Y, Z;
#pragma unroll
  for (int i = 0; i < kPix; ++i) X[i] = Y[i] = Z[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (d->base_end > d->base_begin && c.n > 0) run(d, d->base_begin, d->base_end, c, X, Y, Z, lut);

  // Pixel-centre extent of the tile (clipped to the image).
  const double ex0 = tx * TW + 0.5, ey0 = ty * TH + 0.5;
  const double ex1 = min(tx * TW + TW, W) - 0.5, ey1 = min(ty * TH + TH, H) - 0.5;

  for (uint32_t l = 0; l < d->n_layers; ++l) {
    const Layer L = ldc(&d->layer[l]);
    if (MODE == TILE_STAGED) {
     // The tile is processed in bands of band_h rows whose source boxes fit
     // the LDS budget (strong downscales and rotations widen the box).
     const int BH = (int)d->band_h;
     for (int band = 0; band < TH; band += BH) {
      const double by_lo = ty * TH + band + 0.5;
      const double by_hi = min(ty * TH + band + BH, H) - 0.5;
      if (by_lo > by_hi) break;
      int bx0, by0, bw, bh;
      // A band outside the layer contributes nothing: 'over' a transparent
      // sample is the identity.
      if (!tap_box(L, ex0, by_lo, ex1, by_hi, bx0, by0, bw, bh)) continue;
      if ((uint32_t)(bw * bh) > d->lds_texels) {
        // The planner sizes bands so this cannot happen; flag it loudly.
        if (threadIdx.x == 0 && d->error_flag) *d->error_flag = 1u;
        bh = (int)d->lds_texels / bw;
      }
      // Cooperative fill: the layer's source program runs once per texel of
      // the box, kPix texels per lane, and lands premultiplied in LDS.
      const int gprow = (bw + kPix - 1) / kPix;
      const int ngroups = gprow * bh;
      for (int g = (int)threadIdx.x; g < ngroups; g += kBlock) {
        const int r = g / gprow;
        const int cx = (g - r * gprow) * kPix;
        Ctx s;
        s.dx = s.dy = 0;
        s.gather = false;
        s.x0 = bx0 + cx;
        s.y = by0 + r;
        s.n = min(kPix, bw - cx);
        Px V, V1, V2;
        run(d, L.prog_begin, L.prog_end, s, V, V1, V2, lut);
#pragma unroll
        for (int i = 0; i < kPix; ++i)
          if (i < s.n) tile[r * bw + cx + i] = V[i];
      }
      __syncthreads();
      if (c.n > 0 && ly >= band && ly < band + BH) {
        const int sw = (int)L.src_w, sh = (int)L.src_h;
#pragma unroll
        for (int i = 0; i < kPix; ++i) {
          if (i >= c.n) continue;
          const double2 uv = map_pt(L.inv, c.x0 + i + 0.5, c.y + 0.5);
          int iu, iv;
          float ax, ay;
          split_tap(uv.x, iu, ax);
          split_tap(uv.y, iv, ay);
          float4 sm = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          // Taps outside the source are transparent.
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const int xx = iu + (t & 1), yy = iv + (t >> 1);
            const float w = ((t & 1) ? ax : 1.0f - ax) * ((t >> 1) ? ay : 1.0f - ay);
            const int ty_ = yy - by0, tx_ = xx - bx0;
            if (xx >= 0 && yy >= 0 && xx < sw && yy < sh && tx_ >= 0 && tx_ < bw && ty_ >= 0 && ty_ < bh)
              sm = f4_fma(tile[ty_ * bw + tx_], w, sm);
          }
          X[i] = over4(sm, X[i]);
        }
      }
      __syncthreads();
     }
    } else if (MODE == TILE_DIRECT) {
      // Four gathered taps per pixel, each running the source program at its
      // texel (served by L2); used when a tile's source box exceeds LDS.
      if (c.n > 0) {
        const int sw = (int)L.src_w, sh = (int)L.src_h;
        float ax[kPix], ay[kPix];
        int iu[kPix], iv[kPix];
#pragma unroll
        for (int i = 0; i < kPix; ++i) {
          const double2 uv = map_pt(L.inv, c.x0 + i + 0.5, c.y + 0.5);
          split_tap(uv.x, iu[i], ax[i]);
          split_tap(uv.y, iv[i], ay[i]);
          Z[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        for (int t = 0; t < 4; ++t) {
          Ctx s;
          s.dx = s.dy = 0;
          s.gather = true;
          s.x0 = s.y = 0;
          s.n = kPix;
#pragma unroll
          for (int i = 0; i < kPix; ++i) {
            s.gx[i] = iu[i] + (t & 1);
            s.gy[i] = iv[i] + (t >> 1);
            s.gv[i] = i < c.n && s.gx[i] >= 0 && s.gy[i] >= 0 && s.gx[i] < sw && s.gy[i] < sh;
          }
          Px V, V1, V2;
          run(d, L.prog_begin, L.prog_end, s, V, V1, V2, lut);
#pragma unroll
          for (int i = 0; i < kPix; ++i) {
            const float w = ((t & 1) ? ax[i] : 1.0f - ax[i]) * ((t >> 1) ? ay[i] : 1.0f - ay[i]);
            if (s.gv[i]) Z[i] = f4_fma(V[i], w, Z[i]);
          }
        }
#pragma unroll
        for (int i = 0; i < kPix; ++i)
          if (i < c.n) X[i] = over4(Z[i], X[i]);
      }
    } else {
      // Nearest modes: the pixel is covered when its centre maps inside the
      // selection (the rasteriser's coverage of the painted quad); the
      // sampled texel is the floor of the mapped coordinate.
      if (c.n > 0) {
        Ctx s;
        s.dx = s.dy = 0;
        s.gather = true;
        s.x0 = s.y = 0;
        s.n = kPix;
#pragma unroll
        for (int i = 0; i < kPix; ++i) {
          const double2 uv = map_pt(L.inv, c.x0 + i + 0.5, c.y + 0.5);
          s.gv[i] = i < c.n && uv.x >= L.cover[0] && uv.x < L.cover[2] && uv.y >= L.cover[1] && uv.y < L.cover[3];
          s.gx[i] = (int)floor(uv.x);
          s.gy[i] = (int)floor(uv.y);
        }
        Px V, V1, V2;
        run(d, L.prog_begin, L.prog_end, s, V, V1, V2, lut);
#pragma unroll
        for (int i = 0; i < kPix; ++i) {
          if (!s.gv[i]) continue;
          X[i] = L.mode == LAYER_NEAREST_REPLACE ? V[i] : over4(V[i], X[i]);
        }
      }
    }
  }

  if (c.n > 0) run(d, d->post_begin, d->post_end, c, X, Y, Z, lut);
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_resample(const StageDesc* dd) {
  const CDesc d = as_const(dd);
  extern __shared__ float4 tile[];
  __shared__ Luts lut[1];
  load_lut(d, lut);
  __syncthreads();
  process_tile<MODE>(d, (int)blockIdx.x, (int)blockIdx.y, tile, lut);
}

// Persistent tile queue over many LDS-staged resample stages (one launch for a
// whole batch of images of uneven sizes and formats).  Every workgroup takes
// tiles from a global counter until it passes the total, so the grid drains.
__global__ __launch_bounds__(kBlock) void k_batch(const BatchDesc* bb) {
  const CBatch b = (CBatch)(uintptr_t)bb;
  extern __shared__ float4 tile[];
  __shared__ Luts lut[1];
  __shared__ uint32_t job;
  load_lut(as_const(b->stages), lut);
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) job = atomicAdd(b->counter, 1u);
    __syncthreads();
    const uint32_t j = __builtin_amdgcn_readfirstlane(job);
    if (j >= b->total_tiles) break;
    // Stage of tile j: the last s with prefix[s] <= j (uniform binary search).
    uint32_t lo = 0, hi = b->n_stages;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (cword(b->tile_prefix, mid) <= j) lo = mid; else hi = mid;
    }
    const CDesc d = as_const(b->stages + cword(b->stage_ids, lo));
    const uint32_t t = j - cword(b->tile_prefix, lo);
    process_tile<TILE_STAGED>(d, (int)(t % d->tiles_x), (int)(t / d->tiles_x), tile, lut);
  }
}

__global__ void k_reset_counter(uint32_t* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0u;
}

// ---------------------------------------------------------------------------
// byte copy of rows (High::Copy, same-format inscribe): 16 bytes per lane.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_copy(const StageDesc* dd) {
  const CDesc d = as_const(dd);
  const uint64_t chunks_per_row = (d->copy_row_bytes + 15) / 16;
  const uint64_t gid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t row = gid / chunks_per_row;
  if (row >= d->copy_rows) return;
  const uint64_t off = (gid - row * chunks_per_row) * 16;
  const uint64_t s_off = row * d->copy_src_stride + off, d_off = row * d->copy_dst_stride + off;
  const gbyte* src = (const gbyte*)d->copy_src + s_off;
  gbyte* dst = (gbyte*)d->copy_dst + d_off;
  const uint64_t n = min((uint64_t)16, d->copy_row_bytes - off);
  const bool aligned = (((d->copy_src + s_off) | (d->copy_dst + d_off)) & 15) == 0;
  if (n == 16 && aligned) {
    nt_store(nt_load(src), dst);
  } else {
    for (uint64_t i = 0; i < n; ++i) gst<uint8_t>(dst + i, gld<uint8_t>(src + i));
  }
}

}  // namespace zh

// ---------------------------------------------------------------------------
// host-side launch wrappers (C linkage inside the library)
// ---------------------------------------------------------------------------
extern "C" {

hipError_t zh_launch_pointwise(const zh::StageDesc* d_dev, uint32_t width, uint32_t height, int flat, int fast,
                               uint32_t cus, hipStream_t stream) {
  uint64_t groups;
  if (flat) {
    groups = ((uint64_t)width * height + zh::kPix - 1) / zh::kPix;
  } else {
    groups = (uint64_t)((width + zh::kPix - 1) / zh::kPix) * height;
  }
  static uint32_t resident[2] = {};
  const void* fn = fast ? (const void*)zh::k_pointwise<true> : (const void*)zh::k_pointwise<false>;
  if (!resident[fast ? 1 : 0]) resident[fast ? 1 : 0] = zh::resident_blocks(fn, 1);
  // Persistent grid: the resident workgroups stride over the groups.
  const uint64_t blocks =
      std::min<uint64_t>((groups + zh::kBlock - 1) / zh::kBlock, (uint64_t)resident[fast ? 1 : 0] * cus);
  if (blocks == 0) return hipSuccess;
  if (fast)
    hipLaunchKernelGGL(zh::k_pointwise<true>, dim3((uint32_t)blocks), dim3(zh::kBlock), 0, stream, d_dev);
  else
    hipLaunchKernelGGL(zh::k_pointwise<false>, dim3((uint32_t)blocks), dim3(zh::kBlock), 0, stream, d_dev);
  return hipGetLastError();
}

hipError_t zh_launch_resample(const zh::StageDesc* d_dev, int mode, uint32_t tiles_x, uint32_t tiles_y,
                              uint32_t lds_bytes, hipStream_t stream) {
  if (tiles_x == 0 || tiles_y == 0) return hipSuccess;
  const dim3 grid(tiles_x, tiles_y), block(zh::kBlock);
  switch (mode) {
    case zh::TILE_STAGED: hipLaunchKernelGGL(zh::k_resample<zh::TILE_STAGED>, grid, block, lds_bytes, stream, d_dev); break;
    case zh::TILE_DIRECT: hipLaunchKernelGGL(zh::k_resample<zh::TILE_DIRECT>, grid, block, 0, stream, d_dev); break;
    default: hipLaunchKernelGGL(zh::k_resample<zh::TILE_NEAREST>, grid, block, 0, stream, d_dev); break;
  }
  return hipGetLastError();
}

hipError_t zh_launch_batch(const zh::BatchDesc* b_dev, uint32_t* counter, uint32_t grid, uint32_t lds_bytes,
                           hipStream_t stream) {
  hipLaunchKernelGGL(zh::k_reset_counter, dim3(1), dim3(64), 0, stream, counter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(zh::k_batch, dim3(grid), dim3(zh::kBlock), lds_bytes, stream, b_dev);
  return hipGetLastError();
}

hipError_t zh_launch_copy(const zh::StageDesc* d_dev, uint64_t row_bytes, uint32_t rows, hipStream_t stream) {
  const uint64_t chunks = ((row_bytes + 15) / 16) * rows;
  const uint64_t blocks = (chunks + zh::kBlock - 1) / zh::kBlock;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(zh::k_copy, dim3((uint32_t)blocks), dim3(zh::kBlock), 0, stream, d_dev);
  return hipGetLastError();
}

// The static scratch of zh_program_launch_static: device tables of a
// program launched without allocating (descriptors, tile queues and their
// counters, tile-layer masks, buffer registers, intermediates).  Each device
// holds its own copy from the library's load to the end of the process.
static __device__ __align__(256) uint8_t scratch[ZH_STATIC_SCRATCH_BYTES];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};
}  // namespace

// The device address of `scratch` on `device`, which must be current.  In
// host code the array's name is only a stand-in for the symbol, so its
// address is resolved once per device here and kept; callers pass this
// pointer, never the name, to kernels and copies.
hipError_t zh_scratch_setup(int device, void** ptr, size_t* bytes) {
  if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
  }
  *ptr = g_scratch_ptr[device];
  *bytes = sizeof(scratch);
  return hipSuccess;
}

hipError_t zh_set_resample_lds(uint32_t bytes) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(zh::k_resample<zh::TILE_STAGED>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(zh::k_batch), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes);
}

}  // extern "C"
