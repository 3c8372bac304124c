// zh_warp_kernels.h -- the bodies of the specialised kernels of affine layer
// stacks, device code only: included by zh_warp.hip for the registered forms
// and by the kernels zh_jit.cpp compiles at run time for the others.
//
// A resample stage composites bilinear samples of affine-mapped layers 'over'
// a solid base and finishes with a per-pixel post program (zh_stage.h).  Its
// registered forms (zh_spec.h) run here as straight-line code:
//
//   * the output is cut into tiles of tile_w x tile_h pixels (one workgroup
//     of 256 lanes, kPix = 4 horizontally adjacent pixels per lane); a
//     persistent grid walks the tiles with a stride;
//   * per tile and layer, the layer is skipped when it misses the tile (the
//     host's per-tile mask, or the tile's footprint mapped in FP64); each
//     lane's first pixel centre and the step between the lane's pixels are
//     held in 32.32 fixed point, computed with integers from the layer's map
//     in fixed point (the tile's part uniform), so every pixel's map is one
//     64-bit add: the integer texel is the high word and the bilinear
//     fraction the low word (about 1e-9 px error at any image size; FP32
//     steps would err by ~1e-5 px at 8K, which un-premultiplying a small
//     alpha turns into several f16 ulps);
//   * the four taps are read straight from memory (the L1/L2 caches serve
//     the overlap between neighbouring pixels) through buffer loads whose
//     range is the source image, decoded in-line, premultiplied and
//     weighted; taps outside the source get weight 0 (transparent border),
//     and a tap offset past the end of the source reads 0 in hardware, so no
//     coordinate is clamped;
//   * a 3x3 matrix in the source program is linear, so it is applied once to
//     the interpolated premultiplied value rather than to every tap.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zh_prog.h"

namespace zh {
namespace spec {

// Raw texel of format F as loaded (4, 8 or 16 bytes).
template <uint32_t F> struct RawT { typedef uint32_t type; };
template <> struct RawT<FMT_F16> { typedef u32x2 type; };
template <> struct RawT<FMT_F32> { typedef u32x4 type; };

template <uint32_t F>
__device__ __forceinline__ float4 raw_dec(typename RawT<F>::type r, const Luts* L) {
  if constexpr (F == FMT_F16) {
    return dec_f16(r.x, r.y);
  } else if constexpr (F == FMT_F32) {
    return make_float4(__uint_as_float(r.x), __uint_as_float(r.y), __uint_as_float(r.z), __uint_as_float(r.w));
  } else {
    return dec_u8<F>(r, L);
  }
}

__device__ __forceinline__ float4 premul4(float4 c) { return make_float4(c.x * c.w, c.y * c.w, c.z * c.w, c.w); }

// Uniform data of one layer, read per tile.  Both sources of WSRC_OVER2 have
// the same row stride and size (the planner checks), so one offset serves
// both.  Each source is a buffer resource whose range is its rows (stride x
// height bytes): a load at an offset past the range returns 0 without a
// memory access.  Row offsets use signed 24-bit multiplies: the planner
// admits only sources under 4 GiB with fewer than 2^23 rows of fewer than
// 2^23 bytes, so the offset of every row from -1 to the last is exact, and
// the next row's offset is the row's plus the stride modulo 2^32 (row -1's
// is negative: out of range).  Offsets of taps farther outside the source
// (weight 0) may wrap and land anywhere, in range or not.
struct LayerArgs {
  __amdgpu_buffer_rsrc_t src0;
  __amdgpu_buffer_rsrc_t src1;  // WSRC_OVER2: the upper source
  uint32_t stride;
  int sw, sh;
  float m[9];  // WSRC_ONE with mat3
};

// Fourth word of a raw buffer resource on gfx9-family GPUs (32-bit data
// format; offsets in bytes, range check against the byte count).
constexpr int kBufferFlags = 0x00020000;

// A tap offset kept in its own register.  Left visible, the constant step
// between taps would be folded into the load's immediate offset; a tap at
// texel column 0 would then be addressed as a negative register offset (the
// column -1 tap's) plus the step, and whether the hardware's range check
// sees that sum wrapped to 32 bits is not something to rely on.
__device__ __forceinline__ uint32_t opaque(uint32_t v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <class R>
__device__ __forceinline__ R bload(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <>
__device__ __forceinline__ uint32_t bload<uint32_t>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0);
}
template <>
__device__ __forceinline__ u32x2 bload<u32x2>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
}
template <>
__device__ __forceinline__ u32x4 bload<u32x4>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
}

template <uint32_t SRC, uint32_t F, bool MAT>
__device__ __forceinline__ LayerArgs layer_args(CDesc d, const Layer& L) {
  LayerArgs a;
  const uint32_t pb = L.prog_begin;
  const uint32_t s0 = instr_slot(d, pb);
  a.stride = (uint32_t)d->slot[s0].row_stride;
  const int bytes = (int)(a.stride * L.src_h);  // < 4 GiB: the 32-bit range field
  a.src0 = __builtin_amdgcn_make_buffer_rsrc((void*)d->slot[s0].ptr, (short)0, bytes, kBufferFlags);
  a.src1 = a.src0;
  if constexpr (SRC == WSRC_OVER2)
    a.src1 = __builtin_amdgcn_make_buffer_rsrc((void*)d->slot[instr_slot(d, pb + 2)].ptr, (short)0, bytes,
                                               kBufferFlags);
  if constexpr (MAT) {
    const uint32_t q = d->code[pb + 1].arg;
#pragma unroll
    for (int k = 0; k < 9; ++k) a.m[k] = d->params[q + k];
  }
  a.sw = (int)L.src_w;
  a.sh = (int)L.src_h;
  return a;
}

// Does any bilinear tap of the tile's pixel centres land in the source?
// (Affine maps: the corners bound the footprint.)
__device__ __forceinline__ bool layer_hits(const Layer& L, double ex0, double ey0, double ex1, double ey1) {
  const double* m = L.inv;
  const double u00 = m[0] * ex0 + m[1] * ey0 + m[2], u10 = m[0] * ex1 + m[1] * ey0 + m[2];
  const double u01 = m[0] * ex0 + m[1] * ey1 + m[2], u11 = m[0] * ex1 + m[1] * ey1 + m[2];
  const double v00 = m[3] * ex0 + m[4] * ey0 + m[5], v10 = m[3] * ex1 + m[4] * ey0 + m[5];
  const double v01 = m[3] * ex0 + m[4] * ey1 + m[5], v11 = m[3] * ex1 + m[4] * ey1 + m[5];
  const double umin = fmin(fmin(u00, u10), fmin(u01, u11)) - 0.5, umax = fmax(fmax(u00, u10), fmax(u01, u11)) - 0.5;
  const double vmin = fmin(fmin(v00, v10), fmin(v01, v11)) - 0.5, vmax = fmax(fmax(v00, v10), fmax(v01, v11)) - 0.5;
  return umax > -1.0 && vmax > -1.0 && umin < (double)L.src_w && vmin < (double)L.src_h;
}

// layer_hits over the pixel-centre extent of tile (tx, ty) (clipped to the
// image).  The fallback for stages without the host's per-tile masks; kept
// out of line so that its FP64 set-up is not hoisted onto every tile's path.
__device__ __attribute__((noinline)) bool tile_hits(CDesc d, uint32_t l, uint32_t tx, uint32_t ty) {
  const int TW = (int)d->tile_w, TH = (int)d->tile_h, W = (int)d->width, H = (int)d->height;
  const double ex0 = (double)tx * TW + 0.5, ey0 = (double)ty * TH + 0.5;
  const double ex1 = min((int)tx * TW + TW, W) - 0.5, ey1 = min((int)ty * TH + TH, H) - 0.5;
  return layer_hits(ldc(&d->layer[l]), ex0, ey0, ex1, ey1);
}

// Warp variant: source form, source format, mat3 flag, post program, and
// whether every layer's map is axis-aligned (separable taps).
template <uint32_t SRC_, uint32_t F_, bool MAT_, class Post_, bool SEP_>
struct Warp {
  static constexpr uint32_t SRC = SRC_, F = F_;
  static constexpr bool MAT = MAT_, SEP = SEP_;
  typedef Post_ Post;
};

// Map of a lane's pixels into one layer's source, in 32.32 fixed point:
// (u0, v0) = the lane's first pixel centre mapped, minus the half texel;
// (su, sv) = the step between the lane's pixels (lanes_x output pixels).
struct LaneMap {
  int64_t u0, v0, su, sv;
};

// Integer part and fraction of a 32.32 value.  The fraction rounds to the
// nearest float and can reach 1.0 from just below it: the weights are
// continuous there, so the taps' blend is the same.
__device__ __forceinline__ int fx_int(int64_t v) { return (int)(v >> 32); }
// A value equal in every lane, moved to scalar registers.
__device__ __forceinline__ int64_t uniform64(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float fx_frac(int64_t v) { return (float)(uint32_t)v * 0x1p-32f; }

// x a for a step a in 8.56 fixed point (|a| < 2^7) and 0 <= x < 2^24, in
// 32.32 fixed point (the 24 lowest fraction bits dropped, rounding toward
// -inf): the 88-bit product assembled from two 32 x 32 products.
__device__ __forceinline__ int64_t mul56(uint32_t x, int64_t a) {
  const uint64_t lo = (uint64_t)x * (uint32_t)a;         // < 2^56
  const int64_t hi = (int64_t)x * (int32_t)(a >> 32);    // |.| < 2^55
  return hi * 256 + (int64_t)(lo >> 24);
}

// The lane map of one layer for the tile whose first pixel is (X0, Y0) and
// the lane's first pixel (X0 + lx, Y0 + ly): integer arithmetic on the
// layer's fixed-point map (Layer::fx_step, fx_org), the tile's part equal in
// every lane.  Each term errs by under 2^-32 texel, the map's steps by
// under 2^-57 per pixel: about 1e-9 texel at any image size.
template <bool SEP>
__device__ __forceinline__ LaneMap lane_map(const Layer& Ly, uint32_t X0, uint32_t Y0, int lx, int ly,
                                            int lanes_x) {
  const int64_t ut = Ly.fx_org[0] + mul56(X0, Ly.fx_step[0]) + mul56(Y0, Ly.fx_step[1]);
  const int64_t vt = Ly.fx_org[1] + mul56(X0, Ly.fx_step[2]) + mul56(Y0, Ly.fx_step[3]);
  LaneMap mp;
  // Separable maps have no cross terms (m1 = m3 = 0).
  mp.u0 = ut + mul56((uint32_t)lx, Ly.fx_step[0]) + (SEP ? 0 : mul56((uint32_t)ly, Ly.fx_step[1]));
  mp.v0 = vt + (SEP ? 0 : mul56((uint32_t)lx, Ly.fx_step[2])) + mul56((uint32_t)ly, Ly.fx_step[3]);
  mp.su = uniform64(mul56((uint32_t)lanes_x, Ly.fx_step[0]));
  mp.sv = uniform64(mul56((uint32_t)lanes_x, Ly.fx_step[2]));
  return mp;
}

// Composite one layer's bilinear samples over the lane's kPix pixels.
// Taps of all pixels are addressed first (weight 0 for taps outside the
// source or missing pixels, whose loads read 0 or any in-range texel), then
// loaded, then blended:
// 4 * kPix loads in flight per lane.  Straight-alpha sources are weighted by
// w * alpha (premultiplying inside the weighted sum).
// Two horizontally adjacent texels of 4 or 8 bytes, fetched by one load.
template <uint32_t B> struct PairT { typedef u32x4 type; };
template <> struct PairT<4> { typedef u32x2 type; };
template <uint32_t F>
__device__ __forceinline__ typename RawT<F>::type pair_half(const typename PairT<fmt_bpp(F)>::type& p, int h) {
  if constexpr (fmt_bpp(F) == 8) {
    u32x2 r;
    r.x = h ? p.z : p.x;
    r.y = h ? p.w : p.y;
    return r;
  } else {
    return h ? p.y : p.x;
  }
}

template <class V, bool SEP>
__device__ __forceinline__ void layer_taps(const LayerArgs& a, const LaneMap& mp, int n, Px& X, const Luts* L) {
  typedef typename RawT<V::F>::type R;
  constexpr uint32_t B = fmt_bpp(V::F);
  // 4- and 8-byte texels: the two taps of a row are one load.  The left tap
  // is clamped to [0, sw - 2] (the planner admits sources at least 2 texels
  // wide), so the pair lies inside its row, and each texel of the pair gets
  // the tent weight max(0, 1 - |sample - texel|): the bilinear weight for
  // the sample's own taps, 0 for a texel the clamp substituted at a border.
  constexpr bool PAIR = B <= 8;
  typedef typename PairT<B <= 8 ? B : 8>::type PR;
  // One-source 8-bit formats weigh each tap by its alpha byte: the 1/255 of
  // the byte's value goes into the row weights instead of every tap.
  constexpr bool A8 = B == 4 && V::SRC == WSRC_ONE;
  constexpr float kA = A8 ? 1.0f / 255.0f : 1.0f;
  uint32_t off[kPix][4];
  float w[kPix][4];
  bool any = false;
  // Rows: shared by the lane's pixels when the map is separable.
  float wy0s = 0.0f, wy1s = 0.0f;
  uint32_t r0s = 0, r1s = 0;
  bool yas = false, ybs = false;
  if constexpr (SEP) {
    const int j = fx_int(mp.v0);
    const float ay = fx_frac(mp.v0);
    yas = (unsigned)j < (unsigned)a.sh;
    ybs = (unsigned)(j + 1) < (unsigned)a.sh;
    wy0s = yas ? (1.0f - ay) * kA : 0.0f;
    wy1s = ybs ? ay * kA : 0.0f;
    r0s = (uint32_t)__mul24(j, (int)a.stride);
    r1s = r0s + a.stride;
  }
#pragma unroll
  for (int k = 0; k < kPix; ++k) {
    const int64_t uk = mp.u0 + mp.su * k;
    const int i = fx_int(uk);
    const float ax = fx_frac(uk);
    const bool in = k < n;
    const bool xa = in && (unsigned)i < (unsigned)a.sw, xb = in && (unsigned)(i + 1) < (unsigned)a.sw;
    bool ya, yb;
    float wy0, wy1;
    uint32_t r0, r1;
    if constexpr (SEP) {
      ya = yas;
      yb = ybs;
      wy0 = wy0s;
      wy1 = wy1s;
      r0 = r0s;
      r1 = r1s;
    } else {
      const int64_t vk = mp.v0 + mp.sv * k;
      const int j = fx_int(vk);
      const float ay = fx_frac(vk);
      ya = (unsigned)j < (unsigned)a.sh;
      yb = (unsigned)(j + 1) < (unsigned)a.sh;
      wy0 = ya ? (1.0f - ay) * kA : 0.0f;
      wy1 = yb ? ay * kA : 0.0f;
      r0 = (uint32_t)__mul24(j, (int)a.stride);
      r1 = r0 + a.stride;
    }
    any = any || ((xa || xb) && (ya || yb));
    if constexpr (PAIR) {
      const int ic = min(max(i, 0), a.sw - 2);
      const float sx = (float)(i - ic) + ax;
      const float wl = __builtin_amdgcn_fmed3f(1.0f - fabsf(sx), 0.0f, 1.0f);
      const float wh = __builtin_amdgcn_fmed3f(1.0f - fabsf(sx - 1.0f), 0.0f, 1.0f);
      w[k][0] = wl * wy0;
      w[k][1] = wh * wy0;
      w[k][2] = wl * wy1;
      w[k][3] = wh * wy1;
      const uint32_t c0 = (uint32_t)ic * B;
      off[k][0] = opaque(r0 + c0);
      off[k][1] = opaque(r1 + c0);
    } else {
      const float wx0 = xa ? 1.0f - ax : 0.0f, wx1 = xb ? ax : 0.0f;
      w[k][0] = wx0 * wy0;
      w[k][1] = wx1 * wy0;
      w[k][2] = wx0 * wy1;
      w[k][3] = wx1 * wy1;
      const uint32_t c0 = (uint32_t)i * B;
      off[k][0] = opaque(r0 + c0);
      off[k][1] = opaque(off[k][0] + B);
      off[k][2] = opaque(r1 + c0);
      off[k][3] = opaque(off[k][2] + B);
    }
  }
  if (!any) return;
  R t0[kPix][4], t1[kPix][4];
  PR p0[kPix][2], p1[kPix][2];
  if constexpr (PAIR) {
#pragma unroll
    for (int k = 0; k < kPix; ++k)
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        p0[k][r] = bload<PR>(a.src0, off[k][r]);
        if constexpr (V::SRC == WSRC_OVER2) p1[k][r] = bload<PR>(a.src1, off[k][r]);
      }
  } else {
#pragma unroll
    for (int k = 0; k < kPix; ++k)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        t0[k][t] = bload<R>(a.src0, off[k][t]);
        if constexpr (V::SRC == WSRC_OVER2) t1[k][t] = bload<R>(a.src1, off[k][t]);
      }
  }
  // Keep every load above this point: left free, the scheduler may trade
  // the loads in flight for registers and wait on each tap in turn.
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int k = 0; k < kPix; ++k) {
    float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float4 c = raw_dec<V::F>(PAIR ? pair_half<V::F>(p0[k][t >> 1], t & 1) : t0[k][t], L);
      if constexpr (V::SRC == WSRC_OVER2) {
        // Premultiplied 'over' of the two sources, then weighted.
        const float4 lo = premul4(c);
        const float4 h = premul4(raw_dec<V::F>(PAIR ? pair_half<V::F>(p1[k][t >> 1], t & 1) : t1[k][t], L));
        const float q = 1.0f - h.w, wt = w[k][t];
        const float4 m = make_float4(fmaf(lo.x, q, h.x), fmaf(lo.y, q, h.y), fmaf(lo.z, q, h.z), fmaf(lo.w, q, h.w));
        s = make_float4(fmaf(m.x, wt, s.x), fmaf(m.y, wt, s.y), fmaf(m.z, wt, s.z), fmaf(m.w, wt, s.w));
      } else {
        float al = c.w;
        if constexpr (A8) al = (float)((PAIR ? pair_half<V::F>(p0[k][t >> 1], t & 1) : t0[k][t]) >> 24);
        // Half alphas: w * a as fma(a, w, +0), which compiles to one mixed-
        // precision fma reading the f16 half directly (the same value but
        // for the sign of a zero product).
        const float wa = V::F == FMT_F16 ? __builtin_fmaf(al, w[k][t], 0.0f) : w[k][t] * al;
        s = make_float4(fmaf(c.x, wa, s.x), fmaf(c.y, wa, s.y), fmaf(c.z, wa, s.z), s.w + wa);
      }
    }
    if constexpr (V::MAT) {
      const float r = a.m[0] * s.x + a.m[1] * s.y + a.m[2] * s.z;
      const float g = a.m[3] * s.x + a.m[4] * s.y + a.m[5] * s.z;
      const float b = a.m[6] * s.x + a.m[7] * s.y + a.m[8] * s.z;
      s = make_float4(r, g, b, s.w);
    }
    const float kk = 1.0f - s.w;
    X[k] = make_float4(fmaf(X[k].x, kk, s.x), fmaf(X[k].y, kk, s.y), fmaf(X[k].z, kk, s.z), fmaf(X[k].w, kk, s.w));
  }
}

// One output tile (tx, ty) of stage d.  The tile is tile_w x tile_h pixels;
// lane (lx, ly) owns pixels x = tx * tile_w + lx + k * lanes_x (k < kPix) of
// row ty * tile_h + ly, so every load and store instruction covers runs of
// lanes_x adjacent pixels.
template <class V>
__device__ __forceinline__ void warp_tile(CDesc d, uint32_t tx, uint32_t ty, const Luts* L) {
  typedef typename V::Post Post;
  typedef typename RawT<V::F>::type R;
  constexpr uint32_t B = fmt_bpp(V::F);
  const int TW = (int)d->tile_w, TH = (int)d->tile_h;
  // Tile widths are powers of two of at least kPix pixels (the planner's
  // and the autotuner's shapes): the lane split is a shift.
  const int lanes_x = TW / kPix;
  const int lsh = __builtin_ctz((unsigned)lanes_x);
  const int lx = (int)threadIdx.x & (lanes_x - 1), ly = (int)threadIdx.x >> lsh;
  const int W = (int)d->width, H = (int)d->height;
  const int x0 = (int)tx * TW + lx, y = (int)ty * TH + ly;
  const int n = (y < H && x0 < W) ? min(kPix, (W - x0 + lanes_x - 1) >> lsh) : 0;

  // Base: a premultiplied solid colour (transparent when the program is empty).
  float4 base = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (d->base_end > d->base_begin) {
    const uint32_t q = d->code[d->base_begin].arg;
    base = premul4(make_float4(d->params[q], d->params[q + 1], d->params[q + 2], d->params[q + 3]));
  }
  Px X;
#pragma unroll
  for (int k = 0; k < kPix; ++k) X[k] = base;

  const uint32_t nl = d->n_layers;
  // Layers meeting this tile: the host's per-tile mask when there is one.
  const uint32_t* tl = d->tile_layers;
  const uint32_t t = ty * d->tiles_x + tx;
  const uint32_t lmask = tl ? (cword(tl, t >> 2) >> ((t & 3u) * 8u)) & 255u : 255u;
  for (uint32_t l = 0; l < nl; ++l) {
    if (!((lmask >> l) & 1u)) continue;  // uniform
    const Layer Ly = ldc(&d->layer[l]);
    if (!tl && !Ly.covers_all && !tile_hits(d, l, tx, ty)) continue;
    const LayerArgs a = layer_args<V::SRC, V::F, V::MAT>(d, Ly);
    // Each lane's first pixel and the step between a lane's pixels in fixed
    // point.  Axis-aligned maps (pure scale and shift; the planner picks the
    // SEP variant only when every layer has one) are separable: the lane's
    // pixels share one source row pair, computed once.
    const LaneMap mp = lane_map<V::SEP>(Ly, tx * (uint32_t)TW, ty * (uint32_t)TH, lx, ly, lanes_x);
    layer_taps<V, V::SEP>(a, mp, n, X, L);
  }

  if (n == 0) return;
  // Post program on the composite; its last element stores the lane's pixels.
  Uni<Post> u;
  resolve<Post>(d, u, d->post_begin);
  const uint32_t so = instr_slot(d, d->post_end - 1);
  const uint64_t stride_tx = d->slot[so].row_stride / fmt_bpp(el_fmt(Post::el[Post::n - 1]));
  const uint64_t xi = (uint64_t)y * stride_tx + (uint64_t)x0;
  Px Y, Z;
  exec<Post, false>(u, xi, n, nullptr, X, Y, Z, L, lanes_x);
}

// Tile t of a stage with tiles_x tiles per row; tiles are spread over the
// grid so that each XCD (blocks are dealt to the 8 XCDs round-robin) works on
// a contiguous run of tiles per round and shares their halo texels in its L2.
__device__ __forceinline__ uint32_t virtual_block() {
  const uint32_t g = gridDim.x, b = blockIdx.x;
  if (g % 8u != 0u) return b;
  return (b % 8u) * (g / 8u) + b / 8u;
}

// Tile t of a stage in dealing order -> (tx, ty).  With strip_w > 0 the tiles
// are taken in vertical strips strip_w tiles wide, by rows within a strip, so
// a run of consecutive tiles (what one XCD works on at once) is a compact
// block: a rotated or scaled source footprint with a shorter perimeter, fewer
// texels fetched twice by neighbouring XCDs, and fewer partly used lines.
__device__ __forceinline__ void tile_xy(CDesc d, uint32_t t, uint32_t& tx, uint32_t& ty) {
  const uint32_t nx = d->tiles_x, sw = d->strip_w;
  if (sw == 0 || sw >= nx) {
    tx = t % nx;
    ty = t / nx;
    return;
  }
  const uint32_t per = sw * d->tiles_y;
  const uint32_t s = t / per, r = t - s * per;
  const uint32_t w = min(sw, nx - s * sw);
  ty = r / w;
  tx = s * sw + (r - ty * w);
}

// The 8-bit sRGB tables are staged in LDS only where a source or the post
// program codes 8-bit sRGB texels.  (Decoding straight from the 256-entry
// table beats a bank-replicated copy here: a tile is too short to repay the
// copy's fill.)
template <class V>
constexpr bool needs_lut() {
  if (V::F == FMT_U8_SRGB_RGBA || V::F == FMT_U8_SRGB_BGRA) return true;
  for (int k = 0; k < V::Post::n; ++k) {
    const uint32_t f = el_fmt(V::Post::el[k]);
    if (f == FMT_U8_SRGB_RGBA || f == FMT_U8_SRGB_BGRA) return true;
  }
  return false;
}

// ---- staged tiles ------------------------------------------------------------
// A one-layer stage with a separable map (WSRC_ONE) whose per-tile source box
// fits kBoxTexels -- magnifying maps and mild minification, the planner's
// test -- reads each source texel for up to 4 / scale^2 taps.  Staged, a
// tile's box (the taps of its first and last pixel centres along each axis,
// the map being monotonic, widened by one texel per side) is decoded and
// premultiplied once into LDS, texels outside the source stored as 0 (the
// transparent border), and every pixel interpolates its four taps from
// there: the decode, the tap addressing and the border tests leave the
// per-pixel path.  The lanes' maps are the direct kernel's (32.32 fixed
// point), so both read the same taps.
//
// Each workgroup pipelines its tiles: the raw texels of its next tile are
// fetched by buffer loads that write LDS directly (no registers held), one
// dword per texel word, while the current tile interpolates; then, behind a
// wait and a barrier, they are decoded into the box.
template <uint32_t F> constexpr int raw_words() { return (int)(fmt_bpp(F) / 4); }
constexpr int kBoxLoads = (kBoxTexels + kBlock - 1) / kBlock;
static_assert(kBoxTexels % 64 == 0, "box loads write whole 64-lane chunks");

// Uniform state of one staged tile (kept in scalar registers while the next
// tile's texels are in flight).
struct BoxTile {
  CDesc d;
  uint32_t tx, ty;
  int hit;  // the layer meets the tile (else the tile is the base colour)
  int i0, j0, bw, bh, nb;
  int pitch;  // texels per box row in LDS: bw rounded up to odd, so rows start on different banks
};
// Words of the raw texels are planes kRawPlane dwords apart (one dword past
// the box size, so a texel's words sit in different banks).
constexpr int kRawPlane = kBoxTexels + 1;

// Every field of a tile's state is equal across the workgroup; moved to
// scalar registers explicitly, it stays there through the loop that carries
// it from one tile to the next (left to the compiler, the descriptor pointer
// lands in vector registers and every descriptor read becomes a vector load).
__device__ __forceinline__ BoxTile uniform_tile(BoxTile s) {
  const uint64_t p = (uint64_t)(uintptr_t)s.d;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)p), hi = __builtin_amdgcn_readfirstlane((uint32_t)(p >> 32));
  s.d = (CDesc)(uintptr_t)(((uint64_t)hi << 32) | lo);
  s.tx = __builtin_amdgcn_readfirstlane(s.tx);
  s.ty = __builtin_amdgcn_readfirstlane(s.ty);
  s.hit = __builtin_amdgcn_readfirstlane(s.hit);
  s.i0 = __builtin_amdgcn_readfirstlane(s.i0);
  s.j0 = __builtin_amdgcn_readfirstlane(s.j0);
  s.bw = __builtin_amdgcn_readfirstlane(s.bw);
  s.bh = __builtin_amdgcn_readfirstlane(s.bh);
  s.nb = __builtin_amdgcn_readfirstlane(s.nb);
  s.pitch = __builtin_amdgcn_readfirstlane(s.pitch);
  return s;
}

// Box texel e = r bw + c: r = floor((e + 0.5) / bw) through a 1-ulp
// reciprocal.  For e < 2^11 and 4 <= bw <= 400 the quotient lies at least
// 0.5 / bw = 1.25e-3 from an integer and errs by under 2^-21 e / bw.
__device__ __forceinline__ int box_row(int e, float rbw) { return (int)(((float)e + 0.5f) * rbw); }

__device__ __forceinline__ BoxTile box_tile(CDesc d, uint32_t tx, uint32_t ty) {
  BoxTile s;
  s.d = d;
  s.tx = tx;
  s.ty = ty;
  s.i0 = s.j0 = 0;
  s.bw = s.bh = 4;
  s.nb = 0;
  s.pitch = 5;
  const auto& Ly = d->layer[0];  // constant address space: fields through ldc
  const int TW = (int)d->tile_w, TH = (int)d->tile_h;
  const double ex0 = (double)tx * TW + 0.5, ey0 = (double)ty * TH + 0.5;
  const uint32_t* tl = d->tile_layers;
  const uint32_t t = ty * d->tiles_x + tx;
  if (tl) {
    s.hit = (cword(tl, t >> 2) >> ((t & 3u) * 8u)) & 1u;
  } else {
    const int W = (int)d->width, H = (int)d->height;
    const double ex1 = min((int)tx * TW + TW, W) - 0.5, ey1 = min((int)ty * TH + TH, H) - 0.5;
    s.hit = ldc(&Ly.covers_all) || layer_hits(ldc(&Ly), ex0, ey0, ex1, ey1);
  }
  if (!s.hit) return s;
  const double m0 = ldc(&Ly.inv[0]), m1 = ldc(&Ly.inv[1]), m2 = ldc(&Ly.inv[2]);
  const double m3 = ldc(&Ly.inv[3]), m4 = ldc(&Ly.inv[4]), m5 = ldc(&Ly.inv[5]);
  const double uo = m0 * ex0 + m1 * ey0 + m2, vo = m3 * ex0 + m4 * ey0 + m5;
  // The box over the whole tile (pixels past the image edge included, so
  // every lane's taps lie inside it).
  const double ua = uo - 0.5, ub = fma(m0, (double)(TW - 1), uo) - 0.5;
  const double va = vo - 0.5, vb = fma(m4, (double)(TH - 1), vo) - 0.5;
  s.i0 = (int)floor(fmin(ua, ub)) - 1;
  s.j0 = (int)floor(fmin(va, vb)) - 1;
  s.bw = (int)floor(fmax(ua, ub)) + 3 - s.i0;
  s.bh = (int)floor(fmax(va, vb)) + 3 - s.j0;
  s.pitch = s.bw | 1;
  if (s.bw > kBoxTexels / 4 || s.bh > kBoxTexels / 4 || s.pitch * s.bh > kBoxTexels) {
    // The planner's bound excludes it; the tile is left as the base colour.
    if (threadIdx.x == 0) *d->error_flag = 1u;
    s.hit = 0;
    s.bw = s.bh = 4;
    s.pitch = 5;
    return s;
  }
  s.nb = s.bw * s.bh;
  return s;
}

// Issue the loads of tile s's box texels into `raw` (word w of texel e at
// raw[w * kRawPlane + e]).  Lane l of a wave writes the dword at the wave's
// LDS base + 4 l, so every 64-lane chunk of texels is one load per word.
// Texels outside the source load a clamped, valid texel (decoded as 0).
template <class V>
__device__ __forceinline__ void box_issue(const BoxTile& s, uint32_t* raw) {
  constexpr uint32_t B = fmt_bpp(V::F);
  constexpr int NW = raw_words<V::F>();
  if (!s.hit) return;
  const LayerArgs a = layer_args<V::SRC, V::F, false>(s.d, ldc(&s.d->layer[0]));
  const int tid = (int)threadIdx.x, wbase = tid & ~63;
  const float rbw = __builtin_amdgcn_rcpf((float)s.bw);
#pragma unroll
  for (int q = 0; q < kBoxLoads; ++q) {
    const int cb = q * kBlock + wbase;
    if (cb < s.nb) {  // uniform in the wave
      const int e = q * kBlock + tid;
      const int r = box_row(e, rbw), c = e - r * s.bw;
      const int si = min(max(s.i0 + c, 0), a.sw - 1), sj = min(max(s.j0 + r, 0), a.sh - 1);
      const uint32_t off = __umul24((uint32_t)sj, a.stride) + (uint32_t)si * B;
#pragma unroll
      for (int w = 0; w < NW; ++w)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            a.src0, (__attribute__((address_space(3))) void*)(raw + w * kRawPlane + cb), 4, (int)(off + 4 * w), 0, 0,
            0);
    }
  }
}

// Decode tile s's raw texels into the box: premultiplied float4, 0 outside
// the source.
template <class V>
__device__ __forceinline__ void box_decode(const BoxTile& s, const uint32_t* raw, float4* box, const Luts* L) {
  typedef typename RawT<V::F>::type R;
  constexpr int NW = raw_words<V::F>();
  if (!s.hit) return;
  const int sw = (int)ldc(&s.d->layer[0].src_w), sh = (int)ldc(&s.d->layer[0].src_h);
  const int tid = (int)threadIdx.x;
  const float rbw = __builtin_amdgcn_rcpf((float)s.bw);
#pragma unroll
  for (int q = 0; q < kBoxLoads; ++q) {
    const int e = q * kBlock + tid;
    if (q * kBlock < s.nb && e < s.nb) {
      const int r = box_row(e, rbw), c = e - r * s.bw;
      const bool in = (unsigned)(s.i0 + c) < (unsigned)sw && (unsigned)(s.j0 + r) < (unsigned)sh;
      uint32_t wv[NW];
#pragma unroll
      for (int w = 0; w < NW; ++w) wv[w] = raw[w * kRawPlane + e];
      R rv;
      if constexpr (NW == 1) rv = wv[0];
      else if constexpr (NW == 2) { rv.x = wv[0]; rv.y = wv[1]; }
      else { rv.x = wv[0]; rv.y = wv[1]; rv.z = wv[2]; rv.w = wv[3]; }
      const float4 t = raw_dec<V::F>(rv, L);
      const float al = in ? t.w : 0.0f;
      box[r * s.pitch + c] = make_float4(t.x * al, t.y * al, t.z * al, al);
    }
  }
}

// The pixels of tile s: four taps from the box, the source matrix, 'over'
// the base, then the post program (which stores them).
template <class V>
__device__ __forceinline__ void box_pixels(const BoxTile& s, const float4* box, const Luts* L) {
  typedef typename V::Post Post;
  const CDesc d = s.d;
  const int TW = (int)d->tile_w, TH = (int)d->tile_h;
  const int lanes_x = TW / kPix;
  const int lsh = __builtin_ctz((unsigned)lanes_x);
  const int tid = (int)threadIdx.x;
  const int lx = tid & (lanes_x - 1), ly = tid >> lsh;
  const int W = (int)d->width, H = (int)d->height;
  const int x0 = (int)s.tx * TW + lx, y = (int)s.ty * TH + ly;
  const int n = (y < H && x0 < W) ? min(kPix, (W - x0 + lanes_x - 1) >> lsh) : 0;

  float4 base = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (d->base_end > d->base_begin) {
    const uint32_t q = d->code[d->base_begin].arg;
    base = premul4(make_float4(d->params[q], d->params[q + 1], d->params[q + 2], d->params[q + 3]));
  }
  Px X;
#pragma unroll
  for (int k = 0; k < kPix; ++k) X[k] = base;
  if (s.hit) {
    const Layer Ly = ldc(&d->layer[0]);
    float m[9];
    if constexpr (V::MAT) {
      const uint32_t q = d->code[Ly.prog_begin + 1].arg;
#pragma unroll
      for (int k = 0; k < 9; ++k) m[k] = d->params[q + k];
    }
    const LaneMap mp = lane_map<true>(Ly, s.tx * (uint32_t)TW, s.ty * (uint32_t)TH, lx, ly, lanes_x);
    const int64_t u0 = mp.u0, v0 = mp.v0, su = mp.su;
    const float fy = fx_frac(v0), gy = 1.0f - fy;
    const int bw = s.bw, pitch = s.pitch;
    const float4* row = box + min(max(fx_int(v0) - s.j0, 0), s.bh - 2) * pitch;
#pragma unroll
    for (int k = 0; k < kPix; ++k) {
      const int64_t uk = u0 + su * k;
      const float fx = fx_frac(uk), gx = 1.0f - fx;
      const float4* p = row + min(max(fx_int(uk) - s.i0, 0), bw - 2);
      const float4 t00 = p[0], t01 = p[1], t10 = p[pitch], t11 = p[pitch + 1];
      const float w00 = gx * gy, w01 = fx * gy, w10 = gx * fy, w11 = fx * fy;
      float4 c;
      c.x = fmaf(t11.x, w11, fmaf(t10.x, w10, fmaf(t01.x, w01, t00.x * w00)));
      c.y = fmaf(t11.y, w11, fmaf(t10.y, w10, fmaf(t01.y, w01, t00.y * w00)));
      c.z = fmaf(t11.z, w11, fmaf(t10.z, w10, fmaf(t01.z, w01, t00.z * w00)));
      c.w = fmaf(t11.w, w11, fmaf(t10.w, w10, fmaf(t01.w, w01, t00.w * w00)));
      if constexpr (V::MAT) {
        const float r = m[0] * c.x + m[1] * c.y + m[2] * c.z;
        const float g = m[3] * c.x + m[4] * c.y + m[5] * c.z;
        const float b = m[6] * c.x + m[7] * c.y + m[8] * c.z;
        c = make_float4(r, g, b, c.w);
      }
      const float kk = 1.0f - c.w;
      X[k] = make_float4(fmaf(X[k].x, kk, c.x), fmaf(X[k].y, kk, c.y), fmaf(X[k].z, kk, c.z), fmaf(X[k].w, kk, c.w));
    }
  }
  if (n == 0) return;
  Uni<Post> u;
  resolve<Post>(d, u, d->post_begin);
  const uint32_t so = instr_slot(d, d->post_end - 1);
  const uint64_t stride_tx = d->slot[so].row_stride / fmt_bpp(el_fmt(Post::el[Post::n - 1]));
  const uint64_t xi = (uint64_t)y * stride_tx + (uint64_t)x0;
  Px Y, Z;
  exec<Post, false>(u, xi, n, nullptr, X, Y, Z, L, lanes_x);
}

// The pipelined loop over a workgroup's tiles; next(j, &s) gives the tile of
// queue position j.  Per tile: wait for its texels and for every lane to
// leave the previous box, decode, barrier, issue the next tile's loads,
// interpolate.
template <class V, class Next>
__device__ __forceinline__ void box_loop(uint32_t j, uint32_t total, Next next, uint32_t* raw, float4* box,
                                         const Luts* L) {
  BoxTile cur = next(j);
  box_issue<V>(cur, raw);
  while (true) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    box_decode<V>(cur, raw, box, L);
    __syncthreads();
    const uint32_t jn = j + gridDim.x;
    const bool more = jn < total;
    BoxTile nxt = cur;
    if (more) {
      nxt = next(jn);
      box_issue<V>(nxt, raw);
    }
    box_pixels<V>(cur, box, L);
    if (!more) break;
    cur = nxt;
    j = jn;
  }
}

template <class V>
__device__ __forceinline__ void warp_staged_stage(const StageDesc* dd) {
  static_assert(V::SEP && V::SRC == WSRC_ONE, "staged tiles serve one-source separable forms");
  const CDesc d = as_const(dd);
  __shared__ Luts luts;
  __shared__ float4 box[kBoxTexels];
  __shared__ uint32_t raw[raw_words<V::F>() * kRawPlane];
  if constexpr (needs_lut<V>()) {
    load_lut(d, &luts);
    __syncthreads();
  }
  const uint32_t tiles = d->tiles_x * d->tiles_y;
  const uint32_t j0 = virtual_block();
  if (j0 >= tiles) return;
  box_loop<V>(j0, tiles,
              [&](uint32_t t) {
                uint32_t tx, ty;
                tile_xy(d, t, tx, ty);
                return uniform_tile(box_tile(d, tx, ty));
              },
              raw, box, &luts);
}

template <class V>
__device__ __forceinline__ void warp_staged_batch(const BatchDesc* bb) {
  static_assert(V::SEP && V::SRC == WSRC_ONE, "staged tiles serve one-source separable forms");
  const CBatch b = (CBatch)(uintptr_t)bb;
  __shared__ Luts luts;
  __shared__ float4 box[kBoxTexels];
  __shared__ uint32_t raw[raw_words<V::F>() * kRawPlane];
  if constexpr (needs_lut<V>()) {
    load_lut(as_const(b->stages), &luts);
    __syncthreads();
  }
  const uint32_t total = b->total_tiles;
  const uint32_t j0 = virtual_block();
  if (j0 >= total) return;
  uint32_t lo = 0, hi = b->n_stages;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (cword(b->tile_prefix, mid) <= j0) lo = mid; else hi = mid;
  }
  box_loop<V>(j0, total,
              [&](uint32_t j) {
                while (cword(b->tile_prefix, lo + 1) <= j) ++lo;  // uniform; tiles only move forward
                const CDesc d = as_const(b->stages + cword(b->stage_ids, lo));
                uint32_t tx, ty;
                tile_xy(d, j - cword(b->tile_prefix, lo), tx, ty);
                return uniform_tile(box_tile(d, tx, ty));
              },
              raw, box, &luts);
}

template <class V>
__device__ __forceinline__ void warp_stage(const StageDesc* dd) {
  const CDesc d = as_const(dd);
  __shared__ Luts luts;
  if constexpr (needs_lut<V>()) {
    load_lut(d, &luts);
    __syncthreads();
  }
  const uint32_t tiles = d->tiles_x * d->tiles_y;
  for (uint32_t t = virtual_block(); t < tiles; t += gridDim.x) {
    uint32_t tx, ty;
    tile_xy(d, t, tx, ty);
    warp_tile<V>(d, tx, ty, &luts);
  }
}

// Persistent grid over the tiles of a batch of stages of one form (images of
// uneven sizes): tile j belongs to the last stage s with prefix[s] <= j.  The
// tiles are dealt exactly as one stage's tiles are (grid-stride, XCD-aware),
// so at any moment the grid works on one contiguous band of tiles and their
// source rows stay in the L2 and the TLBs.  A workgroup's tiles only move
// forward, so after one binary search its stage is found by walking the
// prefix, less than one scalar load per tile on average.
template <class V>
__device__ __forceinline__ void warp_batch(const BatchDesc* bb) {
  const CBatch b = (CBatch)(uintptr_t)bb;
  __shared__ Luts luts;
  if constexpr (needs_lut<V>()) {
    load_lut(as_const(b->stages), &luts);
    __syncthreads();
  }
  const uint32_t total = b->total_tiles;
  const uint32_t j0 = virtual_block();
  if (j0 >= total) return;
  uint32_t lo = 0, hi = b->n_stages;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (cword(b->tile_prefix, mid) <= j0) lo = mid; else hi = mid;
  }
  for (uint32_t j = j0; j < total; j += gridDim.x) {
    while (cword(b->tile_prefix, lo + 1) <= j) ++lo;  // uniform
    const CDesc d = as_const(b->stages + cword(b->stage_ids, lo));
    uint32_t tx, ty;
    tile_xy(d, j - cword(b->tile_prefix, lo), tx, ty);
    warp_tile<V>(d, tx, ty, &luts);
  }
}

// Two warp forms in one persistent grid: the tiles of batch a (form VA) and
// of batch b (form VB) are dealt alternately -- queue position 2 i is a's
// tile i, 2 i + 1 is b's tile i -- while both have tiles, then the rest of
// the larger batch in order.  Every compute unit then holds tiles of both
// forms at once, so the vector work of one fills the other's memory waits.
// Within each batch, tiles are dealt as warp_batch deals them (each
// workgroup's tiles only move forward, so each batch's stage cursor only
// walks forward).
__device__ __forceinline__ uint32_t first_stage(CBatch b, uint32_t t) {
  uint32_t lo = 0, hi = b->n_stages;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (cword(b->tile_prefix, mid) <= t) lo = mid; else hi = mid;
  }
  return lo;
}

template <class VA, class VB>
__device__ __forceinline__ void warp_pair(const BatchDesc* ba, const BatchDesc* bb) {
  const CBatch A = (CBatch)(uintptr_t)ba, B = (CBatch)(uintptr_t)bb;
  __shared__ Luts luts;
  if constexpr (needs_lut<VA>() || needs_lut<VB>()) {
    load_lut(as_const(A->stages), &luts);
    __syncthreads();
  }
  const uint32_t ta = A->total_tiles, tb = B->total_tiles;
  const uint32_t m = min(ta, tb), total = ta + tb;
  const uint32_t j0 = virtual_block();
  if (j0 >= total) return;
  // Queue position j -> (batch, tile of that batch).
  auto which = [&](uint32_t j, uint32_t& t) {
    if (j < 2 * m) {
      t = j >> 1;
      return (j & 1u) == 0u;
    }
    t = j - m;
    return ta > tb;
  };
  // Both batches' tiles at queue positions j >= j0 are at least t_lb (j >> 1
  // while both are dealt, j - m after): each cursor starts at that tile's
  // stage.
  const uint32_t t_lb = j0 < 2 * m ? j0 >> 1 : j0 - m;
  uint32_t lo_a = first_stage(A, min(t_lb, ta ? ta - 1 : 0u));
  uint32_t lo_b = first_stage(B, min(t_lb, tb ? tb - 1 : 0u));
  for (uint32_t j = j0; j < total; j += gridDim.x) {
    uint32_t t;
    if (which(j, t)) {
      while (cword(A->tile_prefix, lo_a + 1) <= t) ++lo_a;  // uniform
      const CDesc d = as_const(A->stages + cword(A->stage_ids, lo_a));
      uint32_t tx, ty;
      tile_xy(d, t - cword(A->tile_prefix, lo_a), tx, ty);
      warp_tile<VA>(d, tx, ty, &luts);
    } else {
      while (cword(B->tile_prefix, lo_b + 1) <= t) ++lo_b;  // uniform
      const CDesc d = as_const(B->stages + cword(B->stage_ids, lo_b));
      uint32_t tx, ty;
      tile_xy(d, t - cword(B->tile_prefix, lo_b), tx, ty);
      warp_tile<VB>(d, tx, ty, &luts);
    }
  }
}

}  // namespace spec
}  // namespace zh
