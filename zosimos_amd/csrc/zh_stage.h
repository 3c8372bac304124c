// zh_stage.h -- the compiled form of one fused stage, shared by the host
// planner and the gfx950 kernels.
//
// A stage writes one output image.  Its per-pixel work is a short program of
// micro-ops over a three-entry register stack (X, Y, Z of float4 per pixel),
// evaluated for P = 4 horizontally adjacent pixels per lane.  Every
// instruction, slot and parameter is uniform across the grid, so the
// interpreter's dispatch is scalar; the vector work is the colour math.
//
// Stage kinds
//   STAGE_POINTWISE  out(p) = prog(p): loads of same-sized images at p.
//   STAGE_RESAMPLE   out(p) = post(compose_l layer_l(p) onto base(p)); a
//                    layer is a source program evaluated on a source grid
//                    and sampled through an affine map (bilinear, staged in
//                    LDS per output tile, or nearest with quad coverage).
//   STAGE_COPY       byte copy of rows (transmute / same-format inscribe).
//   STAGE_DYNAMIC    a user kernel loaded from a code object; it reads and
//                    writes RGBA f32 intermediates, not a StageDesc.
#pragma once
#include <stdint.h>

namespace zh {

constexpr int kPix = 4;          // pixels per lane
constexpr int kMaxSlots = 12;    // images referenced by one stage
constexpr int kMaxInstr = 64;
constexpr int kMaxParams = 256;  // floats
constexpr int kMaxLayers = 8;
constexpr int kBlock = 256;      // threads per workgroup
// A batch of flat pointwise stages is dealt in chunks of kSpecChunkGroups
// groups of kPix pixels per lane of one workgroup (16,384 pixels).
constexpr uint32_t kSpecChunkGroups = 16;
constexpr uint64_t kSpecChunkPixels = (uint64_t)kBlock * kPix * kSpecChunkGroups;
// LDS bytes available to one resample tile (float4 texels).
constexpr int kTileLdsBytes = 64 * 1024;
// Staged warp tiles: float4 texels of a tile's source box in LDS (25.6 KB;
// a 64 x 16 tile under a magnifying map needs at most 69 x 21).
constexpr int kBoxTexels = 1600;

enum Opcode : uint8_t {
  OP_NOP = 0,
  OP_LOAD,      // a = slot; b = push.  X <- decode(slot at the pixel)
  OP_SOLID,     // arg = param index of 4 floats; b = push
  OP_MAT3,      // arg = param index of 9 floats (row-major); X.rgb <- M X.rgb
  OP_CBRT,      // X.rgb <- sign(x) |x|^(1/3); a = 1: the planner proved X.rgb >= 0 (no sign handling)
  OP_CUBE,      // X.rgb <- x^3
  OP_CLAMP01,   // X.rgb <- clamp(x, 0, 1)
  OP_PREMUL,    // X.rgb <- X.rgb * X.a
  OP_UNPREMUL,  // X.rgb <- X.a > 0 ? X.rgb / X.a : 0
  OP_OVER,      // X <- X + Y (1 - X.a); pop (Y <- Z)
  OP_SWAP,      // X <-> Y
  OP_EOTF,      // arg = transfer id; X.rgb electrical -> linear
  OP_OETF,      // arg = transfer id; X.rgb linear -> electrical
  OP_QUANT,     // a = slot whose format X is rounded through (encode+decode)
  OP_STORE,     // a = slot; encode X at the pixel
  OP_OFFSET,    // a = 1: following loads read at (x + dx, y + dy); arg = params (dx, dy); a = 0 resets
  OP_GRID,      // arg = params of 6 vec4 {u_min, u_max, v_min, v_max, uv_min, uv_max}; b = push.
                //   X <- mix(u_min, u_max, u) + mix(v_min, v_max, v) + mix(uv_min, uv_max, u v) at the
                //   pixel centre, (u, v) = ((x + 0.5) / width, (y + 0.5) / height) (bilinear.frag)
  OP_BOX3,      // a = slot; arg = 9 weights, row-major by y offset; b = push.  X.rgb <- sum of
                //   w[dy + 1][dx + 1] * texel (x + dx, y + dy), coordinates clamped to the slot; X.a <- 1
                //   (box3.frag)
  OP_NORMAL2D,  // arg = params {mu.x, mu.y, C00, C01, C10, C11, det}; b = push.  At the pixel centre,
                //   p = 2 (u, v) - 1 - mu, X <- (g, g, g, 1) with g = exp(-p.C p / 2) / sqrt(det)
                //   (distribution_normal2d.frag)
  OP_SRLAB,     // X.rgb <- the SrLab2 non-linearity per channel: 24389/2700 x below 216/24389 in
                //   magnitude, 1.16 x^(1/3) - 0.16 above (srlab2.frag)
  OP_SRLAB_INV, // X.rgb <- its inverse: 2700/24389 x below 0.08 in magnitude, ((x + 0.16) / 1.16)^3 above
  OP_PALETTE,   // a = slot; arg = params {mu (4), mv (4), bias u, bias v}.  X <- decode(slot texel nearest to
                //   (mu.X + bias u, mv.X + bias v) in [0, 1]^2, clamped to the edge) (palette.frag)
  OP_CHANNEL,   // arg = params {mul (4), add (4)}.  X <- X * mul + add per component (alpha included)
  OP_INJECT,    // arg = params {select (4), color (4)}.  X <- mix(Y, dot(X, color), select); pop (Y <- Z)
                //   (inject.frag: Y the image injected into, X the injected one)
  OP_NOISE,     // arg = params {scale.x, scale.y, amplitude, damping, octaves}; b = push.  X <- fractal
                //   (octave-summed) value noise of pcg4d-hashed cell corners at the pixel centre
                //   (fractal_noise.frag)
};

struct Instr {
  uint8_t op, a, b, c;
  uint32_t arg;
};

// 8-bit sRGB tables (built once per context on the host in FP64):
//   dec[k]  linear value of code k, idx[i] code of i / 4096 (i <= 4096), thr[k] smallest
//   f32 whose code is >= k (thr[256..259] = +inf).
struct Luts {
  float dec[256];
  float thr[260];
  uint8_t idx[4100];  // 4097 buckets (i / 4096 for i <= 4096), padded to a word
};

struct Slot {
  uint64_t ptr;         // device address of texel (0, 0)
  uint64_t row_stride;  // bytes
  uint32_t width, height;
  uint32_t bits, parts, transfer;
  uint32_t bpp;         // bytes per texel
  uint32_t native;      // 1: native unorm (round to nearest, proper BGRA swizzle)
  uint32_t vec_ok;      // 1: ptr and row_stride are 16-byte aligned
  uint32_t fmt;         // FMT_* code of the in-line codecs, FMT_GENERIC otherwise
  uint32_t pad;
};

// Formats with a specialised in-line codec; one uniform switch per micro-op
// selects the body, so per-pixel code carries no format tests.
enum FmtCode : uint32_t {
  FMT_GENERIC = 0,
  FMT_U8_SRGB_RGBA = 1,
  FMT_U8_SRGB_BGRA = 2,
  FMT_U8_LIN_RGBA = 3,
  FMT_U8_LIN_BGRA = 4,
  FMT_F16 = 5,  // linear, rgba-identity parts
  FMT_F32 = 6,
};

enum LayerMode : uint32_t {
  LAYER_BILINEAR_OVER = 0,  // premultiplied bilinear sample, transparent border, 'over'
  LAYER_NEAREST_REPLACE = 1,  // nearest texel where the pixel centre is inside the quad
  LAYER_NEAREST_OVER = 2,     // nearest texel, composited 'over'
};

struct Layer {
  uint32_t mode;
  uint32_t prog_begin, prog_end;  // source program (evaluated on the source grid)
  uint32_t src_w, src_h;          // source grid
  uint32_t covers_all;            // affine map: every output pixel's taps meet the source (no tile test)
  // Output pixel centre (x + 0.5, y + 0.5) -> source texel space, in FP64
  // (at 8K an FP32 ulp is 5e-4 px):
  //   u = (m0 px + m1 py + m2) / (m6 px + m7 py + m8), v likewise with m3..m5.
  double inv[9];
  // Nearest modes: coverage rectangle in source texel space [x0, x1) x [y0, y1).
  double cover[4];
  // Affine maps of the warp kernels, for their integer lane maps: the steps
  // m0, m1, m3, m4 in 8.56 fixed point (|m| < 127), and (u, v) of the pixel
  // centre (0.5, 0.5) minus the half texel in 32.32 fixed point.
  int64_t fx_step[4];
  int64_t fx_org[2];
};

enum StageKind : uint32_t { STAGE_POINTWISE = 0, STAGE_RESAMPLE = 1, STAGE_COPY = 2, STAGE_DYNAMIC = 3 };

struct StageDesc {
  uint32_t kind;
  uint32_t width, height;     // output grid
  uint32_t flat;              // pointwise: all slots tight and same size -> 1-D
  uint32_t n_instr;
  uint32_t base_begin, base_end;  // resample: base program (empty = transparent)
  uint32_t post_begin, post_end;  // resample: post program (ends with a STORE)
  uint32_t n_layers;
  uint32_t tile_w, tile_h;        // resample tile (tile_w * tile_h == kBlock * kPix)
  uint32_t tiles_x, tiles_y;
  uint32_t lds_texels;        // resample: float4 texels of LDS available per tile
  uint32_t band_h;            // resample: rows per LDS fill (divides tile_h)
  int32_t pre_slot[2];        // pointwise (flat): slots whose raw texels are prefetched, -1 = none
  int32_t variant;            // resample: warp kernel index (zh_spec.h), -1 = none
  uint32_t strip_w;           // warp kernels: tiles dealt in vertical strips strip_w tiles wide (0: by rows)
  // copy stage
  uint64_t copy_src, copy_dst, copy_src_stride, copy_dst_stride, copy_row_bytes;
  uint32_t copy_rows, copy_pad;
  const void* srgb_lut;       // the context's 8-bit sRGB tables (Luts of zh_codec.h)
  // Warp stages: per tile (row-major) a byte of the layers whose footprint
  // meets the tile, four tiles per word (bit l: layer l); null: the kernel
  // tests each layer itself.
  const uint32_t* tile_layers;
  uint32_t* error_flag;       // set to 1 by a kernel that met an impossible case
  Layer layer[kMaxLayers];
  Slot slot[kMaxSlots];
  Instr code[kMaxInstr];
  float params[kMaxParams];
};

// A batch of resample stages executed by one persistent tile-queue launch.
struct BatchDesc {
  uint32_t n_stages;
  uint32_t total_tiles;
  uint32_t* counter;            // reset by a one-lane kernel before each launch
  const StageDesc* stages;      // the program's descriptors
  const uint32_t* stage_ids;    // n_stages indices into `stages`
  const uint32_t* tile_prefix;  // n_stages + 1 exclusive prefix of tile counts
};

}  // namespace zh
