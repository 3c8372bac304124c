// zh_prog.h -- compile-time stage programs: the building blocks of the
// specialised pointwise kernels (zh_spec.hip) and of the specialised affine
// layer kernels (zh_warp.hip).
//
// A program is a parameter pack of signature elements (zh_spec.h); its
// uniform arguments (slot addresses, matrices) are resolved once per kernel
// into scalar registers, and exec() runs the elements as straight-line code
// on a lane's group of kPix pixels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zh_device.h"
#include "zh_spec.h"

namespace zh {
namespace spec {

constexpr uint32_t el_op(uint16_t e) { return e & 0xffu; }
constexpr uint32_t el_fmt(uint16_t e) { return (e >> 8) & 0xfu; }
constexpr bool el_push(uint16_t e) { return ((e >> 12) & 1u) != 0; }

constexpr uint32_t fmt_bpp(uint32_t f) { return f == FMT_F16 ? 8u : f == FMT_F32 ? 16u : 4u; }
// 16-byte words of one lane's group of kPix texels.
constexpr int fmt_words(uint32_t f) { return (int)(fmt_bpp(f) * kPix / 16); }

template <uint16_t... S>
struct Prog {
  static constexpr uint16_t el[sizeof...(S)] = {S...};
  static constexpr int n = (int)sizeof...(S);
  // Raw-word offset of element i (the words of every load before it).
  static constexpr int woff(int i) {
    int o = 0;
    for (int k = 0; k < i; ++k)
      if (el_op(el[k]) == OP_LOAD) o += fmt_words(el_fmt(el[k]));
    return o;
  }
  static constexpr int nwords = woff(n);
  // Parameter-register offset of element i (9 floats per matrix).
  static constexpr int poff(int i) {
    int o = 0;
    for (int k = 0; k < i; ++k)
      if (el_op(el[k]) == OP_MAT3) o += 9;
    return o;
  }
  static constexpr int nparams = poff(n) > 0 ? poff(n) : 1;
  // The program's only load (element 0) and its store (the last element) are
  // 8-bit, and nothing between them writes alpha (the colour ops touch rgb
  // only): the stored alpha byte is the loaded one, since
  // rint((a / 255) * 255) == a for every byte a in FP32.
  static constexpr bool alpha_byte_passes() {
    if (n < 2 || el_op(el[0]) != OP_LOAD || el_op(el[n - 1]) != OP_STORE) return false;
    if (fmt_bpp(el_fmt(el[0])) != 4 || fmt_bpp(el_fmt(el[n - 1])) != 4) return false;
    for (int k = 1; k < n - 1; ++k) {
      const uint32_t o = el_op(el[k]);
      if (o != OP_MAT3 && o != OP_CBRT && o != OP_CUBE && o != OP_CLAMP01) return false;
    }
    return true;
  }
  // Elements i .. i + 4 composite a loaded layer over the value with straight
  // alpha in and out: premul, load (push), premul, over, unpremul.
  static constexpr bool straight_over(int i) {
    return i + 4 < n && el_op(el[i]) == OP_PREMUL && el_op(el[i + 1]) == OP_LOAD && el_push(el[i + 1]) &&
           el_op(el[i + 2]) == OP_PREMUL && el_op(el[i + 3]) == OP_OVER && el_op(el[i + 4]) == OP_UNPREMUL;
  }
};

// Straight-alpha 'up over dn' as one step: with a = a_u + a_d (1 - a_u) (the
// alpha of the premultiplied composite) the un-premultiplied colour is the
// convex combination up * (a_u / a) + dn * (a_d (1 - a_u) / a), 0 where a = 0.
// Equal to premul -> over -> unpremul up to FP32 rounding, in fewer
// instructions, and a weighted mean cannot amplify the inputs' rounding.
// Where a = 0 both weights are 0 whatever finite r is, so the reciprocal is
// taken of max(a, FLT_MIN) (a is never positive and below FLT_MIN for 8- and
// 16-bit alphas; an FP32 alpha that small is treated as FLT_MIN).
__device__ __forceinline__ float4 over_straight(float4 up, float4 dn) {
  const float k = 1.0f - up.w;
  const float a = fmaf(dn.w, k, up.w);
  const float r = __builtin_amdgcn_rcpf(fmaxf(a, 1.17549435e-38f));
  const float wu = up.w * r, wd = dn.w * k * r;
  return make_float4(fmaf(up.x, wu, dn.x * wd), fmaf(up.y, wu, dn.y * wd), fmaf(up.z, wu, dn.z * wd), a);
}

// Uniform arguments of one stage, resolved before the pixel loop.
template <class P>
struct Uni {
  const gbyte* base[P::n];  // loads and stores: slot base address
  float m[P::nparams];      // matrices, row-major
};

// Element I of the program is instruction begin + I of the stage.
template <class P, int I = 0>
__device__ __forceinline__ void resolve(CDesc d, Uni<P>& u, uint32_t begin = 0) {
  if constexpr (I < P::n) {
    constexpr uint16_t e = P::el[I];
    if constexpr (el_op(e) == OP_LOAD || el_op(e) == OP_STORE) {
      u.base[I] = (const gbyte*)d->slot[instr_slot(d, begin + I)].ptr;
    } else if constexpr (el_op(e) == OP_MAT3) {
      const uint32_t a = d->code[begin + I].arg;
#pragma unroll
      for (int k = 0; k < 9; ++k) u.m[P::poff(I) + k] = d->params[a + k];
    }
    resolve<P, I + 1>(d, u, begin);
  }
}

// Raw words of every load of the group at pixel x0.  FULL: the group's kPix
// texels exist (16-byte loads); otherwise texels [0, n) are read one by one
// and the rest are zero.
template <class P, bool FULL, int I = 0>
__device__ __forceinline__ void fetch(const Uni<P>& u, uint64_t x0, int n, uint4* raw) {
  if constexpr (I < P::n) {
    constexpr uint16_t e = P::el[I];
    if constexpr (el_op(e) == OP_LOAD) {
      constexpr uint32_t F = el_fmt(e), B = fmt_bpp(F);
      constexpr int W = fmt_words(F), O = P::woff(I);
      const gbyte* q = u.base[I] + x0 * B;
      if constexpr (FULL) {
#pragma unroll
        for (int w = 0; w < W; ++w) raw[O + w] = nt_load(q + 16 * w);
      } else {
        uint32_t t[4 * W];
#pragma unroll
        for (int k = 0; k < 4 * W; ++k) t[k] = 0u;
#pragma unroll
        for (int i = 0; i < kPix; ++i)
          if (i < n) {
#pragma unroll
            for (int k = 0; k < (int)(B / 4); ++k) t[i * (B / 4) + k] = gld<uint32_t>(q + i * B + 4 * k);
          }
#pragma unroll
        for (int w = 0; w < W; ++w) raw[O + w] = make_uint4(t[4 * w], t[4 * w + 1], t[4 * w + 2], t[4 * w + 3]);
      }
    }
    fetch<P, FULL, I + 1>(u, x0, n, raw);
  }
}

template <uint32_t F>
__device__ __forceinline__ void decode(const uint4* r, Px& X, const Luts* L) {
  if constexpr (F == FMT_F16) {
    X[0] = dec_f16(r[0].x, r[0].y);
    X[1] = dec_f16(r[0].z, r[0].w);
    X[2] = dec_f16(r[1].x, r[1].y);
    X[3] = dec_f16(r[1].z, r[1].w);
  } else if constexpr (F == FMT_F32) {
#pragma unroll
    for (int i = 0; i < kPix; ++i)
      X[i] = make_float4(__uint_as_float(r[i].x), __uint_as_float(r[i].y), __uint_as_float(r[i].z),
                         __uint_as_float(r[i].w));
  } else {
    X[0] = dec_u8<F>(r[0].x, L);
    X[1] = dec_u8<F>(r[0].y, L);
    X[2] = dec_u8<F>(r[0].z, L);
    X[3] = dec_u8<F>(r[0].w, L);
  }
}

// Store the lane's pixels: FULL writes kPix consecutive texels with 16-byte
// stores; otherwise texels [0, n) are written one by one, pstride texels
// apart.
// alpha_src: when given (8-bit formats), the 4 words whose top bytes are the
// alpha bytes to store, instead of encoding X[i].w.
template <uint32_t F, bool FULL>
__device__ __forceinline__ void encode(gbyte* q, int n, const Px& X, const Luts* L, int pstride = 1,
                                       const uint4* alpha_src = nullptr) {
  constexpr int W = fmt_words(F);
  uint4 o[W];
  if constexpr (F == FMT_F16) {
    o[0] = make_uint4(pack_f16(X[0].x, X[0].y), pack_f16(X[0].z, X[0].w), pack_f16(X[1].x, X[1].y),
                      pack_f16(X[1].z, X[1].w));
    o[1] = make_uint4(pack_f16(X[2].x, X[2].y), pack_f16(X[2].z, X[2].w), pack_f16(X[3].x, X[3].y),
                      pack_f16(X[3].z, X[3].w));
  } else if constexpr (F == FMT_F32) {
#pragma unroll
    for (int i = 0; i < kPix; ++i)
      o[i] = make_uint4(__float_as_uint(X[i].x), __float_as_uint(X[i].y), __float_as_uint(X[i].z),
                        __float_as_uint(X[i].w));
  } else if (alpha_src) {
    const uint4 a = alpha_src[0];
    o[0] = make_uint4(enc_u8_rgb<F>(X[0], L) | (a.x & 0xff000000u), enc_u8_rgb<F>(X[1], L) | (a.y & 0xff000000u),
                      enc_u8_rgb<F>(X[2], L) | (a.z & 0xff000000u), enc_u8_rgb<F>(X[3], L) | (a.w & 0xff000000u));
  } else {
    o[0] = make_uint4(enc_u8<F>(X[0], L), enc_u8<F>(X[1], L), enc_u8<F>(X[2], L), enc_u8<F>(X[3], L));
  }
  if constexpr (FULL) {
#pragma unroll
    for (int w = 0; w < W; ++w) nt_store(o[w], q + 16 * w);
  } else {
    constexpr uint32_t B = fmt_bpp(F);
    uint32_t t[4 * W];
#pragma unroll
    for (int w = 0; w < W; ++w) { t[4 * w] = o[w].x; t[4 * w + 1] = o[w].y; t[4 * w + 2] = o[w].z; t[4 * w + 3] = o[w].w; }
#pragma unroll
    for (int i = 0; i < kPix; ++i)
      if (i < n) {
#pragma unroll
        for (int k = 0; k < (int)(B / 4); ++k) gst<uint32_t>(q + i * pstride * B + 4 * k, t[i * (B / 4) + k]);
      }
  }
}

// The stage program, element by element, on the lane's group at x0.  The
// register stack follows the interpreter: a pushing load moves X to Y and Y
// to Z; 'over' composites X onto Y and pops.
template <class P, bool FULL, int I = 0>
__device__ __forceinline__ void exec(const Uni<P>& u, uint64_t x0, int n, const uint4* raw, Px& X, Px& Y, Px& Z,
                                     const Luts* L, int pstride = 1) {
  if constexpr (I < P::n && P::straight_over(I)) {
    Px U;
    decode<el_fmt(P::el[I + 1])>(raw + P::woff(I + 1), U, L);
#pragma unroll
    for (int i = 0; i < kPix; ++i) X[i] = over_straight(U[i], X[i]);
    exec<P, FULL, I + 5>(u, x0, n, raw, X, Y, Z, L, pstride);
  } else if constexpr (I < P::n) {
    constexpr uint16_t e = P::el[I];
    constexpr uint32_t op = el_op(e);
    if constexpr (op == OP_LOAD) {
      if constexpr (el_push(e)) {
#pragma unroll
        for (int i = 0; i < kPix; ++i) { Z[i] = Y[i]; Y[i] = X[i]; }
      }
      decode<el_fmt(e)>(raw + P::woff(I), X, L);
    } else if constexpr (op == OP_MAT3) {
      const float* m = u.m + P::poff(I);
#pragma unroll
      for (int i = 0; i < kPix; ++i) {
        const float r = m[0] * X[i].x + m[1] * X[i].y + m[2] * X[i].z;
        const float g = m[3] * X[i].x + m[4] * X[i].y + m[5] * X[i].z;
        const float b = m[6] * X[i].x + m[7] * X[i].y + m[8] * X[i].z;
        X[i] = make_float4(r, g, b, X[i].w);
      }
    } else if constexpr (op == OP_CBRT && el_fmt(e) == 1) {  // input proved non-negative
      constexpr float t = 1.0f / 3.0f;
#pragma unroll
      for (int i = 0; i < kPix; ++i)
        X[i] = make_float4(pow_pos(X[i].x, t), pow_pos(X[i].y, t), pow_pos(X[i].z, t), X[i].w);
    } else if constexpr (op == OP_CBRT) {
#pragma unroll
      for (int i = 0; i < kPix; ++i) X[i] = make_float4(cbrt_s(X[i].x), cbrt_s(X[i].y), cbrt_s(X[i].z), X[i].w);
    } else if constexpr (op == OP_CUBE) {
#pragma unroll
      for (int i = 0; i < kPix; ++i)
        X[i] = make_float4(X[i].x * X[i].x * X[i].x, X[i].y * X[i].y * X[i].y, X[i].z * X[i].z * X[i].z, X[i].w);
    } else if constexpr (op == OP_CLAMP01) {
#pragma unroll
      for (int i = 0; i < kPix; ++i) X[i] = make_float4(clamp01(X[i].x), clamp01(X[i].y), clamp01(X[i].z), X[i].w);
    } else if constexpr (op == OP_PREMUL) {
#pragma unroll
      for (int i = 0; i < kPix; ++i) X[i] = make_float4(X[i].x * X[i].w, X[i].y * X[i].w, X[i].z * X[i].w, X[i].w);
    } else if constexpr (op == OP_UNPREMUL) {
#pragma unroll
      for (int i = 0; i < kPix; ++i) {
        const float a = X[i].w;
        const float r = a > 0.0f ? __builtin_amdgcn_rcpf(a) : 0.0f;
        X[i] = make_float4(X[i].x * r, X[i].y * r, X[i].z * r, a);
      }
    } else if constexpr (op == OP_OVER) {
#pragma unroll
      for (int i = 0; i < kPix; ++i) {
        const float k = 1.0f - X[i].w;
        X[i] = make_float4(fmaf(Y[i].x, k, X[i].x), fmaf(Y[i].y, k, X[i].y), fmaf(Y[i].z, k, X[i].z),
                           fmaf(Y[i].w, k, X[i].w));
        Y[i] = Z[i];
      }
    } else if constexpr (op == OP_STORE) {
      constexpr uint32_t F = el_fmt(e);
      if constexpr (P::alpha_byte_passes()) {
        // Alpha bytes straight from the loaded texels (raw[0]: the 4 words).
        encode<F, FULL>((gbyte*)u.base[I] + x0 * fmt_bpp(F), n, X, L, pstride, raw);
      } else {
        encode<F, FULL>((gbyte*)u.base[I] + x0 * fmt_bpp(F), n, X, L, pstride);
      }
    } else {
      static_assert(op == OP_NOP, "opcode without a specialised body");
    }
    exec<P, FULL, I + 1>(u, x0, n, raw, X, Y, Z, L, pstride);
  }
}

}  // namespace spec
}  // namespace zh
