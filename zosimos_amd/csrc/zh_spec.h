// zh_spec.h -- registered signatures of pointwise programs that run as
// compile-time specialised kernels instead of through the interpreter.
//
// A signature is the stage's micro-op sequence with the format code of every
// load and store (zh_stage.h FmtCode).  The planner emits its usual program;
// when the sequence matches a registered signature (and the stage is flat),
// the runtime launches the specialised instantiation, which executes the
// same instructions as straight-line code: instruction arguments are uniform
// loads hoisted out of the pixel loop, so no dispatch remains per pixel.
#pragma once
#include <stdint.h>

#include "zh_stage.h"

namespace zh {

// One signature element: opcode | (format << 8) | (push << 12).
constexpr uint16_t sg(uint32_t op, uint32_t fmt = 0, uint32_t push = 0) {
  return (uint16_t)(op | (fmt << 8) | (push << 12));
}

#define ZH_LD(f) sg(OP_LOAD, f)
#define ZH_LDP(f) sg(OP_LOAD, f, 1)
#define ZH_ST(f) sg(OP_STORE, f)
#define ZH_M sg(OP_MAT3)
#define ZH_CB sg(OP_CBRT)
#define ZH_CBP sg(OP_CBRT, 1)  // cube root of a value the planner proved non-negative
#define ZH_CU sg(OP_CUBE)
#define ZH_CL sg(OP_CLAMP01)
#define ZH_PM sg(OP_PREMUL)
#define ZH_UP sg(OP_UNPREMUL)
#define ZH_OV sg(OP_OVER)

#define ZH_S8 FMT_U8_SRGB_RGBA
#define ZH_L8 FMT_U8_LIN_RGBA
#define ZH_H FMT_F16
#define ZH_F FMT_F32

// X(name, elements...) for every registered signature.
#define ZH_SPEC_LIST(X)                                                                                  \
  /* Rec.709 -> P3 -> Oklab -> chroma -> back -> 'over' an f16 layer -> RGBA8 sRGB */                   \
  X(c2_over, ZH_LD(ZH_S8), ZH_M, ZH_CB, ZH_M, ZH_CU, ZH_M, ZH_CL, ZH_PM, ZH_LDP(ZH_H), ZH_PM, ZH_OV, ZH_UP, \
    ZH_ST(ZH_S8))                                                                                         \
  X(c2_over_p, ZH_LD(ZH_S8), ZH_M, ZH_CBP, ZH_M, ZH_CU, ZH_M, ZH_CL, ZH_PM, ZH_LDP(ZH_H), ZH_PM, ZH_OV,      \
    ZH_UP, ZH_ST(ZH_S8))                                                                                  \
  /* Oklab round trip with chroma / hue (the learned-chain workload's plain chain) */                    \
  X(oklab_rt_s8, ZH_LD(ZH_S8), ZH_M, ZH_CB, ZH_M, ZH_CU, ZH_M, ZH_CL, ZH_ST(ZH_S8))                       \
  X(oklab_rt_s8p, ZH_LD(ZH_S8), ZH_M, ZH_CBP, ZH_M, ZH_CU, ZH_M, ZH_CL, ZH_ST(ZH_S8))                     \
  X(oklab_rt_h, ZH_LD(ZH_H), ZH_M, ZH_CB, ZH_M, ZH_CU, ZH_M, ZH_CL, ZH_ST(ZH_H))                          \
  X(oklab_rt_f, ZH_LD(ZH_F), ZH_M, ZH_CB, ZH_M, ZH_CU, ZH_M, ZH_CL, ZH_ST(ZH_F))                          \
  /* to / from Oklab */                                                                                  \
  X(oklab_enc_s8h, ZH_LD(ZH_S8), ZH_M, ZH_CB, ZH_M, ZH_ST(ZH_H))                                          \
  X(oklab_enc_s8hp, ZH_LD(ZH_S8), ZH_M, ZH_CBP, ZH_M, ZH_ST(ZH_H))                                        \
  X(oklab_dec_hs8, ZH_LD(ZH_H), ZH_M, ZH_CU, ZH_M, ZH_CL, ZH_ST(ZH_S8))                                   \
  /* colour matrices */                                                                                  \
  X(mat_s8s8, ZH_LD(ZH_S8), ZH_M, ZH_ST(ZH_S8))                                                           \
  X(mat_hh, ZH_LD(ZH_H), ZH_M, ZH_ST(ZH_H))                                                               \
  X(mat_s8h, ZH_LD(ZH_S8), ZH_M, ZH_ST(ZH_H))                                                             \
  X(mat_hs8, ZH_LD(ZH_H), ZH_M, ZH_ST(ZH_S8))                                                             \
  X(mat_ff, ZH_LD(ZH_F), ZH_M, ZH_ST(ZH_F))                                                               \
  /* format conversions */                                                                               \
  X(cv_s8s8, ZH_LD(ZH_S8), ZH_ST(ZH_S8))                                                                  \
  X(cv_l8l8, ZH_LD(ZH_L8), ZH_ST(ZH_L8))                                                                  \
  X(cv_s8l8, ZH_LD(ZH_S8), ZH_ST(ZH_L8))                                                                  \
  X(cv_l8s8, ZH_LD(ZH_L8), ZH_ST(ZH_S8))                                                                  \
  X(cv_s8h, ZH_LD(ZH_S8), ZH_ST(ZH_H))                                                                    \
  X(cv_hs8, ZH_LD(ZH_H), ZH_ST(ZH_S8))                                                                    \
  X(cv_hh, ZH_LD(ZH_H), ZH_ST(ZH_H))                                                                      \
  X(cv_hf, ZH_LD(ZH_H), ZH_ST(ZH_F))                                                                      \
  X(cv_fh, ZH_LD(ZH_F), ZH_ST(ZH_H))                                                                      \
  X(cv_s8f, ZH_LD(ZH_S8), ZH_ST(ZH_F))                                                                    \
  X(cv_fs8, ZH_LD(ZH_F), ZH_ST(ZH_S8))                                                                    \
  /* full-frame 'over' */                                                                                \
  X(over_s8, ZH_LD(ZH_S8), ZH_PM, ZH_LDP(ZH_S8), ZH_PM, ZH_OV, ZH_UP, ZH_ST(ZH_S8))                      \
  X(over_h, ZH_LD(ZH_H), ZH_PM, ZH_LDP(ZH_H), ZH_PM, ZH_OV, ZH_UP, ZH_ST(ZH_H))

// ---- affine layer stacks (zh_warp.hip) ----------------------------------------
// A resample stage whose layers are all bilinear 'over' through affine maps,
// whose base is a solid colour (or transparent) and whose layer source
// programs share one form runs as a specialised warp kernel:
//   WSRC_ONE    layer program [load(F) (mat3) premul]
//   WSRC_OVER2  layer program [load(F) premul load(F) premul over]
// X(name, source form, F, mat3 in the source program, post elements...)
enum WarpSrc : uint32_t { WSRC_ONE = 0, WSRC_OVER2 = 1 };

#define ZH_WARP_LIST(X)                                                                        \
  X(w_over2_s8, WSRC_OVER2, ZH_S8, 0, ZH_UP, ZH_ST(ZH_S8))                                      \
  X(w_s8m_s8, WSRC_ONE, ZH_S8, 1, ZH_UP, ZH_ST(ZH_S8))                                          \
  X(w_hm_h, WSRC_ONE, ZH_H, 1, ZH_UP, ZH_ST(ZH_H))                                              \
  X(w_s8_s8, WSRC_ONE, ZH_S8, 0, ZH_UP, ZH_ST(ZH_S8))                                           \
  X(w_h_h, WSRC_ONE, ZH_H, 0, ZH_UP, ZH_ST(ZH_H))                                               \
  X(w_h_s8, WSRC_ONE, ZH_H, 0, ZH_UP, ZH_ST(ZH_S8))                                             \
  X(w_h_chain_s8, WSRC_ONE, ZH_H, 0, ZH_UP, ZH_M, ZH_CB, ZH_M, ZH_CU, ZH_M, ZH_CL, ZH_ST(ZH_S8))

}  // namespace zh

extern "C" {
// Index of the specialised kernel for a signature, or -1 (host function,
// defined next to the kernels).
int zh_spec_lookup(const uint16_t* sig, int n);
const char* zh_spec_name(int index);
// Index of the warp kernel for (source form, format, mat3, post signature)
// with general affine maps, or -1; the next index is the same form's kernel
// for axis-aligned maps.
int zh_warp_lookup(uint32_t src, uint32_t fmt, uint32_t mat, const uint16_t* post, int n);
const char* zh_warp_name(int index);
// Does warp kernel `index` have an LDS-staged form (separable, one source)?
int zh_warp_has_staged(int index);
// The form of warp kernel `index` (source form, format, mat3 flag and post
// signature; odd indices are the separable kernels); 0 for no such kernel.
int zh_warp_form_of(int index, uint32_t* src, uint32_t* fmt, uint32_t* mat, const uint16_t** post, int* n);
}
