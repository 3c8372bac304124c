// zh_planner.cpp -- lowers a High op list (program.rs:88-240) into fused
// stages (zh_stage.h).
//
// The reference lowers every High::DrawInto into its own render pass with a
// staging decode before and an encode after (encoder.rs:871-956,
// program.rs:1486-1533).  Here the draws become a value graph:
//   * full-screen shaders are per-pixel nodes;
//   * PaintToSelection draws are layers (nearest with quad coverage, or
//     bilinear with a transparent border) composited onto the target's
//     previous value (Target::Load) or the clear colour (Target::Discard);
//   * High::Copy is a byte reinterpretation.
// Values are materialised only where something needs bytes: outputs, copies,
// values read by several consumers, and layers nested inside per-pixel chains.
// Everything else fuses into one kernel per materialised value; runs of 3x3
// matrices (colour conversion, Oklab M1/M2, chroma/hue) fold in FP64 first.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <sstream>

#include "zh_internal.h"
#include "zh_spec.h"

namespace zh {

// ---------------------------------------------------------------------------
// format helpers
// ---------------------------------------------------------------------------
uint32_t bits_bytes(uint32_t bits) {
  static const uint32_t b[20] = {1, 1, 1, 2, 2, 2, 2, 2, 2, 3, 4, 4, 6, 8, 4, 4, 4, 4, 8, 16};
  return bits < 20 ? b[bits] : 0;
}

bool is_native(const zh_descriptor& d) {
  return d.texel.bits == ZH_BITS_INT8X4 && (d.texel.parts == ZH_PARTS_RGBA || d.texel.parts == ZH_PARTS_BGRA) &&
         d.color.model == ZH_COLOR_RGB &&
         (d.color.transfer == ZH_TRANSFER_SRGB || d.color.transfer == ZH_TRANSFER_LINEAR);
}

uint32_t stage_transfer(const zh_descriptor& d) {
  if (d.color.model == ZH_COLOR_OKLAB || d.color.model == ZH_COLOR_SRLAB2)
    return d.texel.parts == ZH_PARTS_LCHA ? ZH_TRANSFER_LABLCH : ZH_TRANSFER_LINEAR;
  return d.color.transfer;
}

uint32_t stage_parts(const zh_descriptor& d) {
  // Oklab and SrLab2 texels stage as LchA (program.rs:873-930), which passes
  // components.
  if (d.color.model == ZH_COLOR_OKLAB || d.color.model == ZH_COLOR_SRLAB2) return ZH_PARTS_LCHA;
  return d.texel.parts;
}

uint64_t tight_stride(const zh_descriptor& d) { return (uint64_t)d.width * bits_bytes(d.texel.bits); }
uint64_t row_stride_of(const zh_descriptor& d) { return d.row_stride ? d.row_stride : tight_stride(d); }

bool descriptor_valid(const zh_descriptor& d, std::string* why) {
  const uint32_t bpp = bits_bytes(d.texel.bits);
  if (bpp == 0) { if (why) *why = "unknown sample bits id"; return false; }
  if (d.texel_stride && d.texel_stride != bpp) { if (why) *why = "texel_stride does not match sample bits"; return false; }
  if (d.width == 0 || d.height == 0) { if (why) *why = "empty image"; return false; }
  if (d.row_stride && d.row_stride < tight_stride(d)) { if (why) *why = "row_stride shorter than a row"; return false; }
  if (d.texel.parts > 20) { if (why) *why = "unknown sample parts id"; return false; }
  if (d.color.model > ZH_COLOR_SRLAB2) { if (why) *why = "unknown color model"; return false; }
  if ((d.color.model == ZH_COLOR_OKLAB || d.color.model == ZH_COLOR_SRLAB2) && d.texel.parts != ZH_PARTS_LCHA &&
      d.texel.parts != 18) {
    if (why) *why = "Oklab and SrLab2 need LabA or LchA parts";
    return false;
  }
  return true;
}

bool same_layout(const zh_descriptor& a, const zh_descriptor& b) {
  return a.width == b.width && a.height == b.height && a.texel.bits == b.texel.bits && a.texel.parts == b.texel.parts &&
         stage_transfer(a) == stage_transfer(b) && a.color.model == b.color.model &&
         row_stride_of(a) == row_stride_of(b);
}

Mat3 mat_identity() { return {1, 0, 0, 0, 1, 0, 0, 0, 1}; }

Mat3 mat_mul(const Mat3& a, const Mat3& b) {
  Mat3 r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
  return r;
}

bool mat_inv(const Mat3& m, Mat3* out) {
  const double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
  const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
  const double det = a * A + b * B + c * C;
  if (!(std::fabs(det) > 1e-300)) return false;
  const double k = 1.0 / det;
  *out = {A * k, -(b * i - c * h) * k, (b * f - c * e) * k, B * k, (a * i - c * g) * k, -(a * f - c * d) * k,
          C * k, -(a * h - b * g) * k, (a * e - b * d) * k};
  return true;
}

Mat3 mat_from_std140(const void* blob) {
  float f[12];
  std::memcpy(f, blob, sizeof(f));
  // std140 mat3: three columns of vec4 (color_matrix.rs:151-161).
  return {f[0], f[4], f[8], f[1], f[5], f[9], f[2], f[6], f[10]};
}

// Oklab constants (oklab.frag:14-24, row-major).
static const Mat3 kM1 = {0.8189330101, 0.3618667424, -0.1288597137, 0.0329845436, 0.9293118715,
                         0.0361456387, 0.0482003018, 0.2643662691, 0.6338517070};
// CIECAM02's CAT02 and Hunt-Pointer-Estevez matrices, row-major
// (srlab2.frag:13-23 writes them column by column).
static const Mat3 kCAT02 = {0.7328, 0.4296, -0.1624, -0.7036, 1.6975, 0.0061, 0.0030, 0.0136, 0.9834};
static const Mat3 kHPE = {0.38971, 0.68898, -0.07868, -0.22981, 1.18340, 0.04641, 0.0, 0.0, 1.0};
static const Mat3 kM2 = {0.2104542553, 0.7936177850, -0.0040720468, 1.9779984951, -2.4285922050,
                         0.4505937099, 0.0259040371, 0.7827717662, -0.8086757660};

// ---------------------------------------------------------------------------
// value graph
// ---------------------------------------------------------------------------
// N_GRID: a generator of pixel coordinates (bilinear.frag,
// distribution_normal2d.frag: micro-op gen_op); N_STENCIL: box3.frag over a physical
// texture (its operand is materialised).  Both are always the root of a stage
// of their own, which runs on the 2-D interpreter: they need pixel
// coordinates.
// N_DYNAMIC: a user stage kernel over RGBA f32 copies of its operands (a, b:
// identity nodes stored as RGBA f32), itself stored as RGBA f32; the
// register's value is an identity node over it that rounds through the
// register's declared texel, as the reference's target texture holds it.
enum NodeKind { N_INPUT, N_SOLID, N_UNARY, N_BLEND, N_PAINT, N_BYTES, N_GRID, N_STENCIL, N_PALETTE, N_INJECT, N_DYNAMIC };
enum MicroKind { M_MAT3, M_CBRT, M_CUBE, M_CLAMP, M_PREMUL, M_UNPREMUL, M_SRLAB, M_SRLAB_INV, M_CHANNEL };

struct Micro {
  MicroKind k;
  Mat3 m;  // M_CHANNEL: m[0..3] the per-component factor, m[4..7] the addend
};

struct Node {
  NodeKind kind = N_SOLID;
  zh_descriptor desc{};  // declared descriptor of the register holding this value
  uint32_t w = 0, h = 0;
  int input_slot = -1;
  float color[4] = {0, 0, 0, 0};
  float grid[24] = {};  // N_GRID: the generator's parameters; N_STENCIL: 9 weights
  uint8_t gen_op = OP_GRID;
  uint32_t gen_params = 24;
  std::vector<Micro> micro;
  int a = -1, b = -1;  // UNARY: a; BLEND: a below, b above; PAINT: a base, b source; BYTES: a;
                       // PALETTE: a palette (sampled), b indices; INJECT: a below, b above
  uint32_t mode = 0;
  Mat3 inv{};
  double cover[4] = {0, 0, 0, 0};
  uint32_t src_w = 0, src_h = 0;
  std::string label;
  // analysis
  int uses = 0;
  std::vector<int> out_slots;
  bool mat = false;     // materialised by its own stage
  bool sampled = false; // read as a texture by a stencil: stored even when constant
  int buffer = -1;      // intermediate buffer id
  int dyn = -1;         // N_DYNAMIC: index into Plan::dynamics
  bool f32_buffer = false;  // stored as RGBA f32 linear (a user kernel's operand or result)
  bool declared = false;    // rounded through `desc` wherever it is read, in every mode
};

struct Builder {
  std::vector<Node> nodes;
  std::vector<int> cur;  // texture -> node
  int add(Node n) {
    nodes.push_back(std::move(n));
    return (int)nodes.size() - 1;
  }
};

static bool is_layer_class_bilinear(uint32_t mode) { return mode == LAYER_BILINEAR_OVER; }

// Pixel -> texel affine map of a PaintToSelection draw: the quad corners are
// placed in the target by box.vert (position / viewport size over the whole
// target, box.vert:141-148) and carry the selection's texel corners as uv
// (encoder.rs:1831-1849).
static bool selection_map(const zh_op& op, uint32_t W, uint32_t H, Mat3* inv) {
  const double vw = op.viewport.max_x > op.viewport.x ? (double)(op.viewport.max_x - op.viewport.x) : 0.0;
  const double vh = op.viewport.max_y > op.viewport.y ? (double)(op.viewport.max_y - op.viewport.y) : 0.0;
  if (vw <= 0 || vh <= 0) return false;
  double P[4][2], S[4][2];
  for (int k = 0; k < 4; ++k) {
    P[k][0] = ((double)op.quad[2 * k] - op.viewport.x) * (double)W / vw;
    P[k][1] = ((double)op.quad[2 * k + 1] - op.viewport.y) * (double)H / vh;
  }
  const double sx0 = op.selection.x, sy0 = op.selection.y, sx1 = op.selection.max_x, sy1 = op.selection.max_y;
  S[0][0] = sx0; S[0][1] = sy0;
  S[1][0] = sx1; S[1][1] = sy0;
  S[2][0] = sx1; S[2][1] = sy1;
  S[3][0] = sx0; S[3][1] = sy1;
  // Affine A with A P_k = S_k for k = 0, 1, 3 (the quad is a parallelogram).
  const double e1x = P[1][0] - P[0][0], e1y = P[1][1] - P[0][1];
  const double e2x = P[3][0] - P[0][0], e2y = P[3][1] - P[0][1];
  const double f1x = S[1][0] - S[0][0], f1y = S[1][1] - S[0][1];
  const double f2x = S[3][0] - S[0][0], f2y = S[3][1] - S[0][1];
  const double det = e1x * e2y - e2x * e1y;
  if (!(std::fabs(det) > 1e-12)) return false;
  // [e1 e2]^-1
  const double i00 = e2y / det, i01 = -e2x / det, i10 = -e1y / det, i11 = e1x / det;
  const double a00 = f1x * i00 + f2x * i10, a01 = f1x * i01 + f2x * i11;
  const double a10 = f1y * i00 + f2y * i10, a11 = f1y * i01 + f2y * i11;
  const double b0 = S[0][0] - (a00 * P[0][0] + a01 * P[0][1]);
  const double b1 = S[0][1] - (a10 * P[0][0] + a11 * P[0][1]);
  *inv = {a00, a01, b0, a10, a11, b1, 0, 0, 1};
  return true;
}

static bool is_identity_map(const Mat3& m) {
  const Mat3 I = mat_identity();
  for (int i = 0; i < 9; ++i)
    if (std::fabs(m[i] - I[i]) > 1e-9) return false;
  return true;
}

static int solid(Builder& b, const zh_descriptor& d, float r, float g, float bl, float a, const char* label) {
  Node n;
  n.kind = N_SOLID;
  n.desc = d;
  n.w = d.width;
  n.h = d.height;
  n.color[0] = r; n.color[1] = g; n.color[2] = bl; n.color[3] = a;
  n.label = label;
  return b.add(n);
}

static zh_status parse(const std::vector<zh_descriptor>& tex, const std::vector<zh_op>& ops,
                       const std::vector<std::vector<uint8_t>>& params, Builder& B, Plan& plan,
                       std::vector<std::pair<int, int>>& outputs) {
  B.cur.assign(tex.size(), -1);
  auto tex_ok = [&](uint32_t t) { return t < tex.size(); };
  std::vector<std::pair<int, uint32_t>> buffers;  // (node, op) of each ZH_HIGH_BUFFER
  for (size_t oi = 0; oi < ops.size(); ++oi) {
    const zh_op& op = ops[oi];
    const std::vector<uint8_t>& blob = params[oi];
    const std::string at = " (op " + std::to_string(oi) + ")";
    switch (op.high) {
      case ZH_HIGH_INPUT: {
        if (!tex_ok(op.dst)) return fail(ZH_ERR_INVALID_ARGUMENT, "input texture id out of range" + at);
        Node n;
        n.kind = N_INPUT;
        n.desc = tex[op.dst];
        n.w = n.desc.width;
        n.h = n.desc.height;
        n.input_slot = (int)op.src;
        n.label = "input" + std::to_string(op.src);
        if (plan.input_desc.size() <= op.src) {
          plan.input_desc.resize(op.src + 1);
          plan.input_used.resize(op.src + 1, false);
        }
        plan.input_desc[op.src] = n.desc;
        plan.input_used[op.src] = true;
        B.cur[op.dst] = B.add(n);
        break;
      }
      case ZH_HIGH_BUFFER: {
        // Read in place like an input; its slot is numbered after the
        // caller's inputs once they are all known.
        if (!tex_ok(op.dst)) return fail(ZH_ERR_INVALID_ARGUMENT, "buffer texture id out of range" + at);
        const zh_descriptor& d = tex[op.dst];
        if ((uint64_t)blob.size() < row_stride_of(d) * d.height)
          return fail(ZH_ERR_INVALID_ARGUMENT, "buffer of " + std::to_string(blob.size()) +
                                                   " bytes is shorter than the image it holds" + at);
        Node n;
        n.kind = N_INPUT;
        n.desc = d;
        n.w = d.width;
        n.h = d.height;
        n.label = "buffer";
        const int id = B.add(n);
        buffers.push_back({id, (uint32_t)oi});
        B.cur[op.dst] = id;
        break;
      }
      case ZH_HIGH_OUTPUT: {
        if (!tex_ok(op.src) || B.cur[op.src] < 0)
          return fail(ZH_ERR_INVALID_ARGUMENT, "output of an unwritten texture" + at);
        if (plan.output_desc.size() <= op.dst) {
          plan.output_desc.resize(op.dst + 1);
          plan.output_used.resize(op.dst + 1, false);
        }
        plan.output_desc[op.dst] = tex[op.src];
        plan.output_used[op.dst] = true;
        outputs.push_back({(int)op.dst, B.cur[op.src]});
        break;
      }
      case ZH_HIGH_UNINIT: {
        if (!tex_ok(op.dst)) return fail(ZH_ERR_INVALID_ARGUMENT, "texture id out of range" + at);
        B.cur[op.dst] = solid(B, tex[op.dst], 0, 0, 0, 0, "uninit");
        break;
      }
      case ZH_HIGH_COPY: {
        if (!tex_ok(op.dst) || !tex_ok(op.src) || B.cur[op.src] < 0)
          return fail(ZH_ERR_INVALID_ARGUMENT, "copy between unknown textures" + at);
        const zh_descriptor& s = tex[op.src];
        const zh_descriptor& d = tex[op.dst];
        if ((uint64_t)s.height * tight_stride(s) != (uint64_t)d.height * tight_stride(d))
          return fail(ZH_ERR_TYPE, "High::Copy between textures of different byte size" + at);
        Node n;
        n.kind = N_BYTES;
        n.desc = d;
        n.w = d.width;
        n.h = d.height;
        n.a = B.cur[op.src];
        n.label = "transmute";
        B.cur[op.dst] = B.add(n);
        break;
      }
      case ZH_HIGH_DRAW: {
        if (!tex_ok(op.dst)) return fail(ZH_ERR_INVALID_ARGUMENT, "draw target out of range" + at);
        const zh_descriptor& td = tex[op.dst];
        for (uint32_t k = 0; k < op.num_operands && k < 2; ++k)
          if (!tex_ok(op.operands[k]) || B.cur[op.operands[k]] < 0)
            return fail(ZH_ERR_INVALID_ARGUMENT, "operand texture not written" + at);
        auto need_blob = [&](size_t n) { return blob.size() >= n; };
        Node n;
        n.desc = td;
        n.w = td.width;
        n.h = td.height;
        if (op.initializer == ZH_INIT_PAINT_FULLSCREEN) {
          const int src = op.num_operands > 0 ? B.cur[op.operands[0]] : -1;
          if (op.shader == ZH_SHADER_DYNAMIC) {
            // FragmentShaderInvocation::Runtime (command.rs:2767-2830).
            zh_dynamic_call call;
            if (!need_blob(sizeof(call))) return fail(ZH_ERR_INVALID_ARGUMENT, "dynamic blob shorter than its header" + at);
            std::memcpy(&call, blob.data(), sizeof(call));
            if (!call.shader) return fail(ZH_ERR_INVALID_ARGUMENT, "dynamic call without a shader" + at);
            if (op.knob >= 0) return fail(ZH_ERR_UNSUPPORTED, "a dynamic shader's data cannot be a knob" + at);
            if (call.num_args > 2 || call.num_args != op.num_operands)
              return fail(ZH_ERR_INVALID_ARGUMENT, "dynamic shader takes " + std::to_string(call.num_args) +
                                                       " operands, the draw pushes " + std::to_string(op.num_operands) + at);
            if (blob.size() < sizeof(call) + (size_t)call.data_len)
              return fail(ZH_ERR_INVALID_ARGUMENT, "dynamic blob shorter than its shader data" + at);
            DynamicCall dc;
            dc.shader = ShaderRef(call.shader);
            dc.data.assign(blob.begin() + sizeof(call), blob.begin() + sizeof(call) + call.data_len);
            dc.num_args = call.num_args;
            plan.dynamics.push_back(std::move(dc));
            auto f32 = [](uint32_t w, uint32_t h) {
              zh_descriptor d;
              std::memset(&d, 0, sizeof(d));
              d.width = w;
              d.height = h;
              d.texel = zh_texel{ZH_BITS_FLOAT32X4, ZH_PARTS_RGBA};
              d.texel_stride = 16;
              d.color.model = ZH_COLOR_SCALARS;
              d.color.transfer = ZH_TRANSFER_LINEAR;
              return d;
            };
            Node r;
            r.kind = N_DYNAMIC;
            r.dyn = (int)plan.dynamics.size() - 1;
            r.desc = f32(td.width, td.height);
            r.w = td.width;
            r.h = td.height;
            r.f32_buffer = true;
            r.label = std::string("dynamic:") + shader_kernel_name(call.shader);
            for (uint32_t k = 0; k < call.num_args; ++k) {
              const int x = B.cur[op.operands[k]];
              Node c;
              c.kind = N_UNARY;
              c.a = x;
              c.desc = f32(B.nodes[x].w, B.nodes[x].h);
              c.w = B.nodes[x].w;
              c.h = B.nodes[x].h;
              c.f32_buffer = true;
              c.label = "dynamic-operand";
              (k == 0 ? r.a : r.b) = B.add(c);
            }
            const int rid = B.add(r);
            Node u;
            u.kind = N_UNARY;
            u.a = rid;
            u.desc = td;
            u.w = td.width;
            u.h = td.height;
            u.declared = true;
            u.label = "dynamic-result";
            B.cur[op.dst] = B.add(u);
            continue;
          }
          if (op.shader != ZH_SHADER_SOLID_RGB && op.shader != ZH_SHADER_BILINEAR &&
              op.shader != ZH_SHADER_DISTRIBUTION_NORMAL2D && op.shader != ZH_SHADER_FRACTAL_NOISE) {
            if (src < 0) return fail(ZH_ERR_INVALID_ARGUMENT, "full-screen shader without operand" + at);
            // A palette's first operand is sampled anywhere; its indices
            // (second operand) are read at the pixel.
            const Node& sn = B.nodes[op.shader == ZH_SHADER_PALETTE && op.num_operands > 1 ? B.cur[op.operands[1]] : src];
            if (sn.w != td.width || sn.h != td.height)
              return fail(ZH_ERR_UNSUPPORTED, "full-screen paint of a differently sized operand (resize)" + at);
          }
          switch (op.shader) {
            case ZH_SHADER_PAINT_COPY:
              B.cur[op.dst] = src;
              continue;
            case ZH_SHADER_SOLID_RGB: {
              if (!need_blob(16)) return fail(ZH_ERR_INVALID_ARGUMENT, "solid colour needs 16 bytes" + at);
              float c[4];
              std::memcpy(c, blob.data(), 16);
              B.cur[op.dst] = solid(B, td, c[0], c[1], c[2], c[3], "solid");
              continue;
            }
            case ZH_SHADER_BILINEAR: {
              if (!need_blob(96)) return fail(ZH_ERR_INVALID_ARGUMENT, "bilinear needs 96 bytes" + at);
              Node g;
              g.kind = N_GRID;
              g.desc = td;
              g.w = td.width;
              g.h = td.height;
              std::memcpy(g.grid, blob.data(), 96);
              g.label = "bilinear";
              B.cur[op.dst] = B.add(g);
              continue;
            }
            case ZH_SHADER_DISTRIBUTION_NORMAL2D: {
              if (!need_blob(28)) return fail(ZH_ERR_INVALID_ARGUMENT, "normal2d needs 28 bytes" + at);
              Node g;
              g.kind = N_GRID;
              g.desc = td;
              g.w = td.width;
              g.h = td.height;
              std::memcpy(g.grid, blob.data(), 28);
              g.gen_op = OP_NORMAL2D;
              g.gen_params = 7;
              g.label = "normal2d";
              B.cur[op.dst] = B.add(g);
              continue;
            }
            case ZH_SHADER_FRACTAL_NOISE: {
              if (!need_blob(20)) return fail(ZH_ERR_INVALID_ARGUMENT, "fractal noise needs 20 bytes" + at);
              uint32_t octaves;
              std::memcpy(&octaves, blob.data() + 16, 4);
              // Each octave is a uniform loop trip of the kernel; bound the launch.
              if (octaves > 4096) return fail(ZH_ERR_UNSUPPORTED, "more than 4096 noise octaves" + at);
              Node g;
              g.kind = N_GRID;
              g.desc = td;
              g.w = td.width;
              g.h = td.height;
              std::memcpy(g.grid, blob.data(), 16);
              g.grid[4] = (float)octaves;
              g.gen_op = OP_NOISE;
              g.gen_params = 5;
              g.label = "noise";
              B.cur[op.dst] = B.add(g);
              continue;
            }
            case ZH_SHADER_LINEAR_COLOR_MATRIX:
              if (!need_blob(48)) return fail(ZH_ERR_INVALID_ARGUMENT, "matrix blob needs 48 bytes" + at);
              n.kind = N_UNARY;
              n.micro = {{M_MAT3, mat_from_std140(blob.data())}};
              n.label = "linear-matrix";
              break;
            case ZH_SHADER_OKLAB_ENCODE:
              if (!need_blob(48)) return fail(ZH_ERR_INVALID_ARGUMENT, "matrix blob needs 48 bytes" + at);
              n.kind = N_UNARY;
              n.micro = {{M_MAT3, mat_mul(kM1, mat_from_std140(blob.data()))}, {M_CBRT, {}}, {M_MAT3, kM2}};
              n.label = "oklab-encode";
              break;
            case ZH_SHADER_OKLAB_DECODE: {
              if (!need_blob(48)) return fail(ZH_ERR_INVALID_ARGUMENT, "matrix blob needs 48 bytes" + at);
              Mat3 m1i, m2i;
              mat_inv(kM1, &m1i);
              mat_inv(kM2, &m2i);
              n.kind = N_UNARY;
              n.micro = {{M_MAT3, m2i}, {M_CUBE, {}}, {M_MAT3, mat_mul(mat_from_std140(blob.data()), m1i)}, {M_CLAMP, {}}};
              n.label = "oklab-decode";
              break;
            }
            case ZH_SHADER_SRLAB2_ENCODE:
            case ZH_SHADER_SRLAB2_DECODE: {
              // srlab2.frag: XYZ into Hunt-Pointer-Estevez cone space, the
              // non-linearity, and back through the difference scheme
              // (L, a, b) = (y, (x - y) 5 / 1.16, (z - y) 2 / 1.16).  The
              // encode's CAT02 round trip cancels; the decode scales the CAT02
              // cone responses by those of the white point, which the encode
              // does not divide out (srlab2.frag:35-81).
              if (!need_blob(64)) return fail(ZH_ERR_INVALID_ARGUMENT, "srlab2 blob needs 64 bytes" + at);
              float wpf[4];
              std::memcpy(wpf, blob.data(), 16);
              const Mat3 m = mat_from_std140(blob.data() + 16);
              const Mat3 E = {0, 1, 0, 5 / 1.16, -5 / 1.16, 0, 0, -2 / 1.16, 2 / 1.16};
              Mat3 Ei, Hi, Ci;
              mat_inv(E, &Ei);
              mat_inv(kHPE, &Hi);
              mat_inv(kCAT02, &Ci);
              n.kind = N_UNARY;
              if (op.shader == ZH_SHADER_SRLAB2_ENCODE) {
                n.micro = {{M_MAT3, mat_mul(kHPE, m)}, {M_SRLAB, {}}, {M_MAT3, mat_mul(E, Hi)}};
                n.label = "srlab2-encode";
              } else {
                double w[3];
                for (int r = 0; r < 3; ++r)
                  w[r] = kCAT02[r * 3] * wpf[0] + kCAT02[r * 3 + 1] * wpf[1] + kCAT02[r * 3 + 2] * wpf[2];
                const Mat3 D = {w[0], 0, 0, 0, w[1], 0, 0, 0, w[2]};
                const Mat3 back = mat_mul(m, mat_mul(Ci, mat_mul(D, mat_mul(kCAT02, Hi))));
                n.micro = {{M_MAT3, mat_mul(kHPE, Ei)}, {M_SRLAB_INV, {}}, {M_MAT3, back}, {M_CLAMP, {}}};
                n.label = "srlab2-decode";
              }
              break;
            }
            case ZH_SHADER_OKLAB_ADJUST: {
              if (!need_blob(8)) return fail(ZH_ERR_INVALID_ARGUMENT, "adjust blob needs 8 bytes" + at);
              float p[2];
              std::memcpy(p, blob.data(), 8);
              const double k = p[0], c = std::cos((double)p[1]), s = std::sin((double)p[1]);
              n.kind = N_UNARY;
              n.micro = {{M_MAT3, {1, 0, 0, 0, k * c, -k * s, 0, k * s, k * c}}};
              n.label = "oklab-adjust";
              break;
            }
            case ZH_SHADER_BOX3: {
              if (!need_blob(48)) return fail(ZH_ERR_INVALID_ARGUMENT, "box3 blob needs 48 bytes" + at);
              const Mat3 m = mat_from_std140(blob.data());
              n.kind = N_STENCIL;
              n.a = src;
              for (int k = 0; k < 9; ++k) n.grid[k] = (float)m[k];
              n.label = "box3";
              B.cur[op.dst] = B.add(n);
              continue;
            }
            case ZH_SHADER_PALETTE: {
              if (op.num_operands < 2) return fail(ZH_ERR_INVALID_ARGUMENT, "palette needs two operands" + at);
              if (!need_blob(32)) return fail(ZH_ERR_INVALID_ARGUMENT, "palette blob needs 32 bytes" + at);
              float m[8];
              std::memcpy(m, blob.data(), 32);
              n.kind = N_PALETTE;
              n.a = src;
              n.b = B.cur[op.operands[1]];
              for (int k = 0; k < 4; ++k) {
                n.grid[k] = m[2 * k];          // u row: x_k
                n.grid[4 + k] = m[2 * k + 1];  // v row: y_k
              }
              // palette.frag's bias is half a texel of the indices' image.
              n.grid[8] = 0.5f / (float)B.nodes[n.b].w;
              n.grid[9] = 0.5f / (float)B.nodes[n.b].h;
              n.label = "palette";
              B.cur[op.dst] = B.add(n);
              continue;
            }
            case ZH_SHADER_EXTRACT: {
              if (!need_blob(4)) return fail(ZH_ERR_INVALID_ARGUMENT, "extract blob needs 4 bytes" + at);
              uint32_t pos;
              std::memcpy(&pos, blob.data(), 4);
              if (pos > 2) return fail(ZH_ERR_INVALID_ARGUMENT, "extract of a channel other than R, G, B" + at);
              // The single-channel texel decodes to its channel at its
              // position, 0 elsewhere and alpha 1 (stage.frag parts_normalize).
              Mat3 m{};
              m[pos] = 1.0;
              m[7] = 1.0;
              n.kind = N_UNARY;
              n.micro = {{M_CHANNEL, m}};
              n.label = "extract";
              break;
            }
            case ZH_SHADER_INJECT: {
              if (op.num_operands < 2) return fail(ZH_ERR_INVALID_ARGUMENT, "inject needs two operands" + at);
              if (!need_blob(32)) return fail(ZH_ERR_INVALID_ARGUMENT, "inject blob needs 32 bytes" + at);
              const Node& above = B.nodes[B.cur[op.operands[1]]];
              if (above.w != td.width || above.h != td.height)
                return fail(ZH_ERR_UNSUPPORTED, "inject of a differently sized image" + at);
              n.kind = N_INJECT;
              n.a = src;
              n.b = B.cur[op.operands[1]];
              std::memcpy(n.grid, blob.data(), 32);
              n.label = "inject";
              B.cur[op.dst] = B.add(n);
              continue;
            }
            case ZH_SHADER_PREMULTIPLY:
              n.kind = N_UNARY;
              n.micro = {{M_PREMUL, {}}};
              n.label = "premultiply";
              break;
            case ZH_SHADER_UNPREMULTIPLY:
              n.kind = N_UNARY;
              n.micro = {{M_UNPREMUL, {}}};
              n.label = "unpremultiply";
              break;
            case ZH_SHADER_BLEND_OVER: {
              if (op.num_operands < 2) return fail(ZH_ERR_INVALID_ARGUMENT, "blend needs two operands" + at);
              const Node& above = B.nodes[B.cur[op.operands[1]]];
              if (above.w != td.width || above.h != td.height)
                return fail(ZH_ERR_UNSUPPORTED, "full-screen blend of a differently sized layer" + at);
              n.kind = N_BLEND;
              n.a = src;
              n.b = B.cur[op.operands[1]];
              n.label = "over";
              B.cur[op.dst] = B.add(n);
              continue;
            }
            default:
              return fail(ZH_ERR_UNSUPPORTED, "shader key " + std::to_string(op.shader) + " not supported" + at);
          }
          n.a = src;
          B.cur[op.dst] = B.add(n);
          continue;
        }
        if (op.initializer != ZH_INIT_PAINT_TO_SELECTION)
          return fail(ZH_ERR_INVALID_ARGUMENT, "unknown initializer" + at);
        if (op.num_operands < 1) return fail(ZH_ERR_INVALID_ARGUMENT, "paint without operand" + at);
        const int src = B.cur[op.operands[0]];
        const zh_descriptor& sd = tex[op.operands[0]];
        // The clear colour of Target::Discard is opaque blue (program.rs:1495-1499).
        const int base = op.target == ZH_TARGET_LOAD ? B.cur[op.dst] : solid(B, td, 0, 0, 1, 1, "clear");
        if (base < 0) return fail(ZH_ERR_INVALID_ARGUMENT, "Load of an unwritten target" + at);
        Mat3 inv;
        uint32_t mode;
        if (op.shader == ZH_SHADER_PAINT_COPY) {
          mode = LAYER_NEAREST_REPLACE;
        } else if (op.shader == ZH_SHADER_BLEND_OVER) {
          mode = LAYER_NEAREST_OVER;
        } else if (op.shader == ZH_SHADER_AFFINE_BILINEAR) {
          mode = LAYER_BILINEAR_OVER;
        } else {
          return fail(ZH_ERR_UNSUPPORTED, "paint-to-selection with shader " + std::to_string(op.shader) + at);
        }
        if (mode == LAYER_BILINEAR_OVER) {
          if (!need_blob(48)) return fail(ZH_ERR_INVALID_ARGUMENT, "affine blob needs 48 bytes" + at);
          // The blob is the row-major Affine matrix (source px -> target px).
          Mat3 fwd = mat_from_std140(blob.data());
          if (!mat_inv(fwd, &inv)) { B.cur[op.dst] = base; continue; }
        } else if (!selection_map(op, td.width, td.height, &inv)) {
          B.cur[op.dst] = base;  // degenerate quad: nothing is rasterised
          continue;
        }
        // A nearest paint of a same-sized image at identity that covers the
        // whole target is the image itself (or an 'over' of it).
        const bool full_sel = op.selection.x == 0 && op.selection.y == 0 && op.selection.max_x == sd.width &&
                              op.selection.max_y == sd.height && sd.width == td.width && sd.height == td.height;
        if (mode != LAYER_BILINEAR_OVER && full_sel && is_identity_map(inv)) {
          if (mode == LAYER_NEAREST_REPLACE) {
            B.cur[op.dst] = src;
          } else {
            n.kind = N_BLEND;
            n.a = base;
            n.b = src;
            n.label = "over";
            B.cur[op.dst] = B.add(n);
          }
          continue;
        }
        n.kind = N_PAINT;
        n.a = base;
        n.b = src;
        n.mode = mode;
        n.inv = inv;
        n.cover[0] = op.selection.x;
        n.cover[1] = op.selection.y;
        n.cover[2] = op.selection.max_x;
        n.cover[3] = op.selection.max_y;
        n.src_w = sd.width;
        n.src_h = sd.height;
        n.label = mode == LAYER_BILINEAR_OVER ? "affine-bilinear" : (mode == LAYER_NEAREST_OVER ? "paint-over" : "paint");
        B.cur[op.dst] = B.add(n);
        break;
      }
      default:
        return fail(ZH_ERR_INVALID_ARGUMENT, "unknown High kind" + at);
    }
  }
  plan.user_inputs = (uint32_t)plan.input_desc.size();
  for (const auto& b : buffers) {
    B.nodes[b.first].input_slot = (int)plan.input_desc.size();
    plan.input_desc.push_back(B.nodes[b.first].desc);
    plan.input_used.push_back(true);
    plan.buffer_ops.push_back(b.second);
  }
  return ZH_OK;
}

// ---------------------------------------------------------------------------
// stage emission
// ---------------------------------------------------------------------------
struct Emitter {
  Builder& B;
  Plan& plan;
  StagePlan& sp;
  bool strict;
  std::vector<Instr> code;
  std::vector<float> params;
  std::map<std::pair<int64_t, std::string>, int> slot_index;
  int depth = 0;
  std::string err;

  Emitter(Builder& b, Plan& p, StagePlan& s, bool st) : B(b), plan(p), sp(s), strict(st) {}

  static std::string fmt_key(const zh_descriptor& d) {
    std::ostringstream o;
    o << d.width << 'x' << d.height << ':' << d.texel.bits << ':' << d.texel.parts << ':' << stage_transfer(d) << ':'
      << d.color.model << ':' << row_stride_of(d);
    return o.str();
  }

  int slot(int64_t ref, const zh_descriptor& d) {
    auto key = std::make_pair(ref, fmt_key(d));
    auto it = slot_index.find(key);
    if (it != slot_index.end()) return it->second;
    const int k = (int)sp.slot_ref.size();
    if (k >= kMaxSlots) {
      err = "stage references too many images";
      return 0;
    }
    sp.slot_ref.push_back(ref);
    Slot& s = sp.desc.slot[k];
    std::memset(&s, 0, sizeof(s));
    s.ptr = 0;
    s.row_stride = row_stride_of(d);
    s.width = d.width;
    s.height = d.height;
    s.bits = d.texel.bits;
    s.parts = stage_parts(d);
    s.transfer = stage_transfer(d);
    s.bpp = bits_bytes(d.texel.bits);
    s.native = is_native(d) ? 1u : 0u;
    s.vec_ok = 0;  // set at bind time from the actual pointer
    s.fmt = FMT_GENERIC;
    if (s.native) {
      const bool bgra = d.texel.parts == ZH_PARTS_BGRA, srgb = s.transfer == ZH_TRANSFER_SRGB;
      s.fmt = srgb ? (bgra ? FMT_U8_SRGB_BGRA : FMT_U8_SRGB_RGBA) : (bgra ? FMT_U8_LIN_BGRA : FMT_U8_LIN_RGBA);
    } else if ((s.bits == ZH_BITS_FLOAT16X4 || s.bits == ZH_BITS_FLOAT32X4) && s.transfer == ZH_TRANSFER_LINEAR &&
               (s.parts == ZH_PARTS_RGBA || s.parts == 18 || s.parts == ZH_PARTS_LCHA)) {
      s.fmt = s.bits == ZH_BITS_FLOAT16X4 ? FMT_F16 : FMT_F32;
    }
    slot_index[key] = k;
    return k;
  }

  void ins(uint8_t op, uint8_t a = 0, uint8_t b = 0, uint32_t arg = 0) {
    if (code.size() >= (size_t)kMaxInstr) {
      err = "stage program longer than the instruction budget";
      return;
    }
    code.push_back(Instr{op, a, b, 0, arg});
  }
  uint32_t param(const float* v, int n) {
    const uint32_t at = (uint32_t)params.size();
    if (params.size() + n > (size_t)kMaxParams) {
      err = "stage parameters exceed the budget";
      return 0;
    }
    params.insert(params.end(), v, v + n);
    return at;
  }
  void mat(const Mat3& m) {
    if (is_identity_map(m) && !strict) return;
    // Fold into a directly preceding matrix.
    if (!code.empty() && code.back().op == OP_MAT3 && pending_mat_valid) {
      pending_mat = mat_mul(m, pending_mat);
      write_pending();
      return;
    }
    pending_mat = m;
    pending_mat_valid = true;
    float f[9];
    for (int i = 0; i < 9; ++i) f[i] = (float)m[i];
    ins(OP_MAT3, 0, 0, param(f, 9));
  }
  Mat3 pending_mat{};
  bool pending_mat_valid = false;
  void write_pending() {
    const uint32_t at = code.back().arg;
    for (int i = 0; i < 9; ++i) params[at + i] = (float)pending_mat[i];
  }
  void op(uint8_t o, uint8_t a = 0, uint8_t b = 0, uint32_t arg = 0) {
    // UNPREMUL directly followed by PREMUL is the identity on the representation.
    if (o == OP_PREMUL && !code.empty() && code.back().op == OP_UNPREMUL) {
      code.pop_back();
      return;
    }
    pending_mat_valid = false;
    ins(o, a, b, arg);
  }

  // Whether X's colour is known non-negative: the program so far ends with
  // a matrix of non-negative entries applied to a load of an 8-bit unorm
  // texel (whose decoded colour lies in [0, 1]).
  bool rgb_nonneg() const {
    const size_t n = code.size();
    if (n < 2 || code[n - 1].op != OP_MAT3 || code[n - 2].op != OP_LOAD) return false;
    for (int k = 0; k < 9; ++k)
      if (!(params[code[n - 1].arg + k] >= 0.0f)) return false;
    const uint32_t f = sp.desc.slot[code[n - 2].a].fmt;
    return f == FMT_U8_SRGB_RGBA || f == FMT_U8_SRGB_BGRA || f == FMT_U8_LIN_RGBA || f == FMT_U8_LIN_BGRA;
  }
  void cbrt() { op(OP_CBRT, rgb_nonneg() ? 1 : 0); }

  // X <- X * mul + add per component (M_CHANNEL).
  void channel(const Mat3& m) {
    float f[8];
    for (int k = 0; k < 8; ++k) f[k] = (float)m[k];
    pending_mat_valid = false;
    ins(OP_CHANNEL, 0, 0, param(f, 8));
  }

  bool leaf_physical(int id) const {
    const Node& n = B.nodes[id];
    return n.kind == N_INPUT || n.mat;
  }

  // Where the bytes of node `id` live (input slot or intermediate buffer).
  int64_t phys_ref(int id) const {
    const Node& n = B.nodes[id];
    if (n.kind == N_INPUT) return n.input_slot;
    return kIntermediateBase + n.buffer;
  }
  zh_descriptor phys_desc(int id) const {
    const Node& n = B.nodes[id];
    if (n.kind == N_INPUT) return n.desc;
    return plan.intermediates[n.buffer];
  }

  void load(int64_t ref, const zh_descriptor& d) {
    const int s = slot(ref, d);
    pending_mat_valid = false;
    ins(OP_LOAD, (uint8_t)s, depth > 0 ? 1 : 0);
    depth++;
  }

  // Emit the value of node id into X.  Returns true when X holds a
  // premultiplied representation.
  bool value(int id, int root) {
    const Node& n = B.nodes[id];
    if (id != root && leaf_physical(id)) {
      load(phys_ref(id), phys_desc(id));
      return false;
    }
    switch (n.kind) {
      case N_INPUT:
        load(phys_ref(id), phys_desc(id));
        return false;
      case N_BYTES: {
        // Reinterpretation of materialised bytes: read the child's storage
        // with this node's descriptor.
        const Node& c = B.nodes[n.a];
        zh_descriptor d = n.desc;
        if (c.kind == N_INPUT) {
          d.row_stride = row_stride_of(c.desc) * c.desc.height / std::max(1u, d.height);
          load(c.input_slot, d);
        } else {
          load(kIntermediateBase + c.buffer, d);
        }
        return false;
      }
      case N_SOLID: {
        pending_mat_valid = false;
        ins(OP_SOLID, 0, depth > 0 ? 1 : 0, param(n.color, 4));
        depth++;
        return quantise(n, false, root);
      }
      case N_GRID: {
        pending_mat_valid = false;
        ins(n.gen_op, 0, depth > 0 ? 1 : 0, param(n.grid, n.gen_params));
        depth++;
        return quantise(n, false, root);
      }
      case N_STENCIL: {
        const int s = slot(phys_ref(n.a), phys_desc(n.a));
        pending_mat_valid = false;
        ins(OP_BOX3, (uint8_t)s, depth > 0 ? 1 : 0, param(n.grid, 9));
        depth++;
        return quantise(n, false, root);
      }
      case N_PALETTE: {
        // The indices' straight colour replaces itself with the palette's.
        if (value(n.b, root)) op(OP_UNPREMUL);
        // The lookup is discontinuous in its index: an index computed in
        // this stage (a fused generator) is rounded through its register's
        // declared format, as the reference's register texture holds it
        // (resize's RGBA8 ramp, command.rs:1677).  Strict mode rounded it
        // already.
        if (!strict && !leaf_physical(n.b)) {
          const int qs = slot(-1000000 - (int64_t)n.b, B.nodes[n.b].desc);
          op(OP_QUANT, (uint8_t)qs);
        }
        const int s = slot(phys_ref(n.a), phys_desc(n.a));
        pending_mat_valid = false;
        ins(OP_PALETTE, (uint8_t)s, 0, param(n.grid, 10));
        return quantise(n, false, root);
      }
      case N_UNARY: {
        bool pm = value(n.a, root);
        for (const Micro& m : n.micro) {
          if (pm) { op(OP_UNPREMUL); pm = false; }
          switch (m.k) {
            case M_MAT3: mat(m.m); break;
            case M_CBRT: cbrt(); break;
            case M_CUBE: op(OP_CUBE); break;
            case M_CLAMP: op(OP_CLAMP01); break;
            case M_PREMUL: op(OP_PREMUL); break;
            case M_UNPREMUL: op(OP_UNPREMUL); break;
            case M_SRLAB: op(OP_SRLAB); break;
            case M_SRLAB_INV: op(OP_SRLAB_INV); break;
            case M_CHANNEL: channel(m.m); break;
          }
        }
        return quantise(n, pm, root);
      }
      case N_INJECT: {
        if (value(n.a, root)) op(OP_UNPREMUL);
        if (value(n.b, root)) op(OP_UNPREMUL);
        if (depth > 3) err = "expression deeper than the register stack";
        pending_mat_valid = false;
        ins(OP_INJECT, 0, 0, param(n.grid, 8));
        depth--;
        return quantise(n, false, root);
      }
      case N_BLEND: {
        if (!value(n.a, root)) op(OP_PREMUL);
        if (!value(n.b, root)) op(OP_PREMUL);
        if (depth > 3) err = "expression deeper than the register stack";
        op(OP_OVER);
        depth--;
        return quantise(n, true, root);
      }
      case N_PAINT:
        err = "internal: paint node inside a per-pixel expression";
        return false;
      case N_DYNAMIC:
        err = "internal: user kernel inside a per-pixel expression";
        return false;
    }
    return false;
  }

  // Round the value through its register's declared texel format (strict mode).
  // Skipped at the stage's root, whose stores encode into the declared
  // formats themselves.  Returns whether X is still premultiplied.
  bool quantise(const Node& n, bool pm, int root) {
    const int64_t id = (int64_t)(&n - &B.nodes[0]);
    if ((!strict && !n.declared) || id == root) return pm;
    if (pm) op(OP_UNPREMUL);
    const int s = slot(-1000000 - id, n.desc);
    op(OP_QUANT, (uint8_t)s);
    return false;
  }

  void finish() {
    if (!err.empty()) return;
    std::memcpy(sp.desc.code, code.data(), code.size() * sizeof(Instr));
    std::memcpy(sp.desc.params, params.data(), params.size() * sizeof(float));
    sp.desc.n_instr = (uint32_t)code.size();
  }
};

static const char* op_name(uint8_t o) {
  static const char* n[] = {"nop",     "load", "solid", "mat3", "cbrt",  "cube",  "clamp01", "premul", "unpremul",
                            "over",    "swap", "eotf",  "oetf", "quant", "store", "offset",  "grid",    "box3", "normal2d", "srlab", "srlab_inv", "palette", "channel", "inject", "noise"};
  return o < sizeof(n) / sizeof(n[0]) ? n[o] : "?";
}

static std::string describe_code(const StageDesc& d, uint32_t b, uint32_t e) {
  std::ostringstream o;
  for (uint32_t i = b; i < e; ++i) {
    if (i > b) o << ' ';
    o << op_name(d.code[i].op);
    if (d.code[i].op == OP_LOAD || d.code[i].op == OP_STORE || d.code[i].op == OP_QUANT || d.code[i].op == OP_BOX3)
      o << '#' << (int)d.code[i].a;
  }
  return o.str();
}

// Bytes of a slot's image that a stage reads or writes once.
static uint64_t slot_bytes(const Slot& s) { return (uint64_t)s.width * s.height * s.bpp; }

// Source box (texels) of one interior tile for a layer, with the device's
// one-texel margin plus one more for rounding.
static uint64_t tile_box_texels(const Mat3& inv, int TW, int TH) {
  double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
  const double xs[2] = {0.5, TW - 0.5}, ys[2] = {0.5, TH - 0.5};
  for (double x : xs)
    for (double y : ys) {
      const double q = inv[6] * x + inv[7] * y + inv[8];
      const double u = (inv[0] * x + inv[1] * y + inv[2]) / q, v = (inv[3] * x + inv[4] * y + inv[5]) / q;
      umin = std::min(umin, u); umax = std::max(umax, u);
      vmin = std::min(vmin, v); vmax = std::max(vmax, v);
    }
  const double bw = std::floor(umax) - std::floor(umin) + 5, bh = std::floor(vmax) - std::floor(vmin) + 5;
  return (uint64_t)(bw * bh);
}

// Fraction of a layer's source texels whose centre lands on the output
// canvas (for algorithmic bytes).
// Registered warp form of a bilinear resample stage (zh_spec.h), or -1:
// affine layers, a solid (or empty) base, one source form shared by every
// layer and a registered post program.
static int warp_variant(const StageDesc& d, uint32_t flags) {
  const uint32_t nb = d.base_end - d.base_begin;
  if (!(nb == 0 || (nb == 2 && d.code[d.base_begin].op == OP_SOLID && d.code[d.base_begin + 1].op == OP_PREMUL)))
    return -1;
  uint32_t src = ~0u, fmt = 0, mat = 0;
  for (uint32_t l = 0; l < d.n_layers; ++l) {
    const Layer& L = d.layer[l];
    if (L.mode != LAYER_BILINEAR_OVER) return -1;
    if (L.inv[6] != 0.0 || L.inv[7] != 0.0 || std::fabs(L.inv[8] - 1.0) > 1e-12) return -1;
    // The lane maps hold the steps in 8.56 fixed point and coordinates in
    // 32.32 (Layer::fx_step, fx_org).
    if (std::fabs(L.inv[0]) >= 127.0 || std::fabs(L.inv[1]) >= 127.0 || std::fabs(L.inv[3]) >= 127.0 ||
        std::fabs(L.inv[4]) >= 127.0 || std::fabs(L.inv[2]) >= 1e9 || std::fabs(L.inv[5]) >= 1e9)
      return -1;
    const Instr* c = d.code + L.prog_begin;
    const uint32_t n = L.prog_end - L.prog_begin;
    uint32_t s_, f_, m_ = 0;
    if (n == 2 && c[0].op == OP_LOAD && c[1].op == OP_PREMUL) {
      s_ = WSRC_ONE;
    } else if (n == 3 && c[0].op == OP_LOAD && c[1].op == OP_MAT3 && c[2].op == OP_PREMUL) {
      s_ = WSRC_ONE;
      m_ = 1;
    } else if (n == 5 && c[0].op == OP_LOAD && c[1].op == OP_PREMUL && c[2].op == OP_LOAD && c[2].b &&
               c[3].op == OP_PREMUL && c[4].op == OP_OVER && d.slot[c[0].a].fmt == d.slot[c[2].a].fmt) {
      s_ = WSRC_OVER2;
    } else {
      return -1;
    }
    f_ = d.slot[c[0].a].fmt;
    if (f_ == FMT_GENERIC) return -1;
    // The kernel addresses every source over the whole layer grid with
    // 32-bit byte offsets, one row stride for both sources of WSRC_OVER2, and
    // fetches horizontal tap pairs (sources at least 2 texels wide).
    for (uint32_t i = 0; i < n; ++i)
      if (c[i].op == OP_LOAD) {
        const Slot& sl = d.slot[c[i].a];
        if (sl.width != L.src_w || sl.height != L.src_h || sl.width < 2 || sl.row_stride != d.slot[c[0].a].row_stride ||
            sl.row_stride * sl.height >= (1ull << 32) || sl.row_stride >= (1u << 23) || sl.height >= (1u << 23))
          return -1;
      }
    if (src == ~0u) {
      src = s_;
      fmt = f_;
      mat = m_;
    } else if (src != s_ || fmt != f_ || mat != m_) {
      return -1;
    }
  }
  if (src == ~0u) return -1;
  uint16_t sig[kMaxInstr];
  const uint32_t np = d.post_end - d.post_begin;
  for (uint32_t i = 0; i < np; ++i) {
    const Instr& in = d.code[d.post_begin + i];
    if (in.op == OP_LOAD) return -1;
    if (in.op == OP_STORE && (d.slot[in.a].width != d.width || d.slot[in.a].height != d.height)) return -1;
    sig[i] = sg(in.op, in.op == OP_STORE ? d.slot[in.a].fmt : 0u);
  }
  int v = zh_warp_lookup(src, fmt, mat, sig, (int)np);
  // A form without a registered kernel gets one compiled at run time.
  if (v < 0 && !(flags & ZH_PLAN_NO_JIT) && jit_warp_supported(sig, (int)np))
    v = kJitWarpBase + 2 * jit_warp_form(src, fmt, mat, sig, (int)np);
  if (v < 0) return -1;
  bool sep = true;
  for (uint32_t l = 0; l < d.n_layers; ++l) sep = sep && d.layer[l].inv[1] == 0.0 && d.layer[l].inv[3] == 0.0;
  return sep ? v + 1 : v;
}

bool warp_staged_fits(const StageDesc& d, uint32_t tw, uint32_t th) {
  if (d.kind != STAGE_RESAMPLE || d.variant < 0 || !(d.variant & 1) || d.n_layers != 1) return false;
  const bool has = d.variant >= kJitWarpBase ? jit_warp_has_staged((d.variant - kJitWarpBase) / 2)
                                             : zh_warp_has_staged(d.variant) != 0;
  if (!has || tw == 0 || th == 0) return false;
  // The kernel's box spans floor(u_max) - floor(u_min) + 4 columns with
  // u_max - u_min = |m0| (tw - 1) up to rounding (rows likewise): at most
  // floor(|m0| (tw - 1)) + 6.
  const Layer& L = d.layer[0];
  const double bw = std::floor(std::fabs(L.inv[0]) * (tw - 1)) + 6.0;
  const double bh = std::floor(std::fabs(L.inv[4]) * (th - 1)) + 6.0;
  // Rows are laid out an odd number of texels apart: one more at most.
  return bw <= kBoxTexels / 4 && bh <= kBoxTexels / 4 && (bw + 1) * bh <= kBoxTexels;
}

// x in 32.32 fixed point (|x| < 2^31; the fraction truncated to 2^-32), as
// the warp kernels hold mapped coordinates.
static int64_t fixed32(double x) {
  const double f = std::floor(x);
  const uint32_t lo = (uint32_t)((x - f) * 4294967296.0);
  return (int64_t)(((uint64_t)(uint32_t)(int64_t)f << 32) | lo);
}

// The fixed-point form of a layer's affine map (Layer::fx_step, fx_org);
// steps of 127 or more in magnitude are left 0 (warp_variant admits none).
static void fixed_map(Layer& ly) {
  const int idx[4] = {0, 1, 3, 4};
  for (int k = 0; k < 4; ++k) {
    const double m = ly.inv[idx[k]];
    ly.fx_step[k] = std::fabs(m) < 127.0 ? (int64_t)std::llround(std::ldexp(m, 56)) : 0;
  }
  const double u = ly.inv[0] * 0.5 + ly.inv[1] * 0.5 + ly.inv[2] - 0.5;
  const double v = ly.inv[3] * 0.5 + ly.inv[4] * 0.5 + ly.inv[5] - 0.5;
  ly.fx_org[0] = std::fabs(u) < 2147483648.0 ? fixed32(u) : 0;
  ly.fx_org[1] = std::fabs(v) < 2147483648.0 ? fixed32(v) : 0;
}

static double visible_fraction(const Mat3& inv, uint32_t sw, uint32_t sh, uint32_t W, uint32_t H) {
  Mat3 fwd;
  if (!mat_inv(inv, &fwd)) return 1.0;
  const int N = 128;
  int hit = 0;
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < N; ++i) {
      const double u = (i + 0.5) * sw / N, v = (j + 0.5) * sh / N;
      const double q = fwd[6] * u + fwd[7] * v + fwd[8];
      const double x = (fwd[0] * u + fwd[1] * v + fwd[2]) / q, y = (fwd[3] * u + fwd[4] * v + fwd[5]) / q;
      if (x >= 0 && y >= 0 && x < W && y < H) hit++;
    }
  return (double)hit / (N * N);
}

zh_status plan_program(const std::vector<zh_descriptor>& textures, const std::vector<zh_op>& ops,
                       const std::vector<std::vector<uint8_t>>& params, uint32_t flags, Plan* out) {
  Plan plan;
  Builder B;
  std::vector<std::pair<int, int>> outputs;  // (output slot, node)
  for (size_t i = 0; i < textures.size(); ++i) {
    std::string why;
    if (!descriptor_valid(textures[i], &why))
      return fail(ZH_ERR_INVALID_ARGUMENT, "texture " + std::to_string(i) + ": " + why);
  }
  zh_status st = parse(textures, ops, params, B, plan, outputs);
  if (st != ZH_OK) return st;
  if (outputs.empty()) return fail(ZH_ERR_INVALID_ARGUMENT, "program has no output");
  const bool strict = (flags & ZH_PLAN_STRICT_REGISTERS) != 0;

  // Uses reachable from the outputs.
  std::vector<bool> live(B.nodes.size(), false);
  std::vector<int> stack;
  for (auto& o : outputs) {
    B.nodes[o.second].out_slots.push_back(o.first);
    stack.push_back(o.second);
  }
  while (!stack.empty()) {
    const int id = stack.back();
    stack.pop_back();
    if (live[id]) continue;
    live[id] = true;
    const Node& n = B.nodes[id];
    if (n.a >= 0) stack.push_back(n.a);
    if (n.b >= 0) stack.push_back(n.b);
  }
  for (size_t id = 0; id < B.nodes.size(); ++id) {
    if (!live[id]) continue;
    const Node& n = B.nodes[id];
    if (n.a >= 0) B.nodes[n.a].uses++;
    if (n.b >= 0) B.nodes[n.b].uses++;
  }

  // Materialisation: outputs, byte reinterpretations, shared values, and
  // paints that cannot be the core of a resample stage.
  auto mark = [&](int id) {
    Node& n = B.nodes[id];
    // Inputs and constants are read in place unless an output or a stencil
    // needs their texels.
    if ((n.kind == N_INPUT || (n.kind == N_SOLID && !n.sampled)) && n.out_slots.empty()) return false;
    if (n.mat) return false;
    n.mat = true;
    return true;
  };
  for (size_t id = 0; id < B.nodes.size(); ++id) {
    if (!live[id]) continue;
    Node& n = B.nodes[id];
    if (!n.out_slots.empty()) mark((int)id);
    // Generators run on the 2-D interpreter, so they get a stage of their
    // own; a coordinate grid read only by a palette lookup (itself 2-D) is
    // evaluated inside the lookup's stage instead, at full precision.
    bool palette_only = n.kind == N_GRID && n.uses == 1 && n.out_slots.empty();
    if (palette_only) {
      palette_only = false;
      for (size_t j = 0; j < B.nodes.size(); ++j)
        if (live[j] && B.nodes[j].kind == N_PALETTE && B.nodes[j].b == (int)id) palette_only = true;
    }
    if ((n.kind == N_GRID && !palette_only) || n.kind == N_STENCIL) mark((int)id);
    if ((n.kind == N_STENCIL || n.kind == N_PALETTE) && B.nodes[n.a].kind != N_INPUT) {
      B.nodes[n.a].sampled = true;
      mark(n.a);
    }
    if (n.kind == N_PALETTE) mark((int)id);
    if (n.kind == N_DYNAMIC) {
      mark((int)id);
      if (n.a >= 0) mark(n.a);
      if (n.b >= 0) mark(n.b);
    }
    if (n.kind == N_BYTES && B.nodes[n.a].kind != N_INPUT) mark(n.a);
    if (n.uses > 1 && n.kind != N_BYTES) mark((int)id);
  }
  // Fixed point: walk each stage's structure; paints met inside per-pixel
  // expressions or layer sources are materialised on their own.
  bool changed = true;
  auto walk_expr = [&](int id, int root, auto&& self) -> void {
    const Node& n = B.nodes[id];
    if (id != root && (n.kind == N_INPUT || n.mat)) return;
    if (n.kind == N_PAINT) {
      if (id != root) { changed |= mark(id); return; }
    }
    if (n.kind == N_BYTES) return;
    if (n.kind == N_UNARY) self(n.a, root, self);
    if (n.kind == N_BLEND) { self(n.a, root, self); self(n.b, root, self); }
  };
  auto stage_core = [&](int root, int& paint) {
    // Peel single-use per-pixel ops above a paint.
    int id = root;
    while (true) {
      const Node& n = B.nodes[id];
      if (n.kind == N_PAINT) { paint = id; return; }
      if (n.kind == N_UNARY && (n.a >= 0) && !B.nodes[n.a].mat && B.nodes[n.a].kind != N_INPUT && B.nodes[n.a].uses == 1) {
        id = n.a;
        continue;
      }
      paint = -1;
      return;
    }
  };
  while (changed) {
    changed = false;
    for (size_t id = 0; id < B.nodes.size(); ++id) {
      if (!live[id] || !B.nodes[id].mat) continue;
      int paint;
      stage_core((int)id, paint);
      if (paint < 0) {
        walk_expr((int)id, (int)id, walk_expr);
        continue;
      }
      // unary chain between root and paint is part of this stage
      // layers: paint, then its base while it is a same-class single-use paint
      const bool bil = is_layer_class_bilinear(B.nodes[paint].mode);
      int layer = paint, count = 0;
      while (true) {
        const Node& L = B.nodes[layer];
        walk_expr(L.b, -1, walk_expr);
        if (L.b >= 0 && B.nodes[L.b].kind == N_PAINT && !B.nodes[L.b].mat) changed |= mark(L.b);
        count++;
        const int base = L.a;
        const Node& bn = B.nodes[base];
        if (bn.kind == N_PAINT && !bn.mat && bn.uses == 1 && bn.mode == L.mode && bn.w == L.w && bn.h == L.h &&
            count < kMaxLayers) {
          layer = base;
          continue;
        }
        walk_expr(base, -1, walk_expr);
        if (bn.kind == N_PAINT && !bn.mat) changed |= mark(base);
        break;
      }
    }
  }

  // Intermediate buffers for materialised values read by other stages.
  for (size_t id = 0; id < B.nodes.size(); ++id) {
    Node& n = B.nodes[id];
    if (!live[id] || !n.mat || n.kind == N_INPUT || (n.kind == N_SOLID && !n.sampled)) continue;
    const bool internal = n.uses > 0;
    if (!internal) continue;
    zh_descriptor d = n.desc;
    bool bytes_child = false;
    for (size_t j = 0; j < B.nodes.size(); ++j)
      if (live[j] && B.nodes[j].kind == N_BYTES && B.nodes[j].a == (int)id) bytes_child = true;
    // Strict registers keep their declared texels, so a later stage reads
    // exactly what the reference's register would hold.
    if (n.f32_buffer) {
      d = n.desc;  // RGBA f32 linear, the user kernel's layout
    } else if (!bytes_child && !strict && !n.declared) {
      // FP16 where the declared register allows it, FP32 otherwise.
      const bool wide = d.texel.bits == ZH_BITS_FLOAT32X4 || d.texel.bits == ZH_BITS_INT16X4 ||
                        d.texel.bits == ZH_BITS_INT16 || d.texel.bits == 11 || d.texel.bits == 12;
      d.texel.bits = wide ? ZH_BITS_FLOAT32X4 : ZH_BITS_FLOAT16X4;
      d.texel.parts = ZH_PARTS_RGBA;
      d.texel_stride = bits_bytes(d.texel.bits);
      d.color.model = ZH_COLOR_SCALARS;
      d.color.transfer = ZH_TRANSFER_LINEAR;
    }
    d.row_stride = 0;
    n.buffer = (int)plan.intermediates.size();
    plan.intermediates.push_back(d);
  }

  // Emit stages in program order.
  std::ostringstream text;
  for (size_t id = 0; id < B.nodes.size(); ++id) {
    Node& n = B.nodes[id];
    if (!live[id] || !n.mat) continue;
    StagePlan sp;
    std::memset(&sp.desc, 0, sizeof(sp.desc));
    sp.desc.pre_slot[0] = sp.desc.pre_slot[1] = -1;
    sp.desc.variant = -1;
    sp.desc.width = n.w;
    sp.desc.height = n.h;
    Emitter E(B, plan, sp, strict);
    int paint;
    stage_core((int)id, paint);

    auto emit_stores = [&](bool pm) {
      if (pm) E.op(OP_UNPREMUL);
      for (int slot_out : n.out_slots) {
        const int s = E.slot(~(int64_t)slot_out, plan.output_desc[slot_out]);
        E.op(OP_STORE, (uint8_t)s);
      }
      if (n.buffer >= 0) {
        const int s = E.slot(kIntermediateBase + n.buffer, plan.intermediates[n.buffer]);
        E.op(OP_STORE, (uint8_t)s);
      }
    };

    if (n.kind == N_BYTES && n.out_slots.size() > 0 && n.buffer < 0 && !strict) {
      // Pure reinterpretation to an output: a byte copy (bit-exact).
      const Node& c = B.nodes[n.a];
      sp.desc.kind = STAGE_COPY;
      zh_descriptor src_d = c.kind == N_INPUT ? c.desc : plan.intermediates[c.buffer];
      const int64_t src_ref = c.kind == N_INPUT ? c.input_slot : kIntermediateBase + c.buffer;
      sp.slot_ref.push_back(src_ref);
      sp.slot_ref.push_back(~(int64_t)n.out_slots[0]);
      sp.desc.copy_rows = src_d.height;
      sp.desc.copy_row_bytes = tight_stride(src_d);
      sp.desc.copy_src_stride = row_stride_of(src_d);
      sp.desc.copy_dst_stride = row_stride_of(plan.output_desc[n.out_slots[0]]);
      if (tight_stride(n.desc) * n.desc.height != sp.desc.copy_row_bytes * sp.desc.copy_rows)
        return fail(ZH_ERR_TYPE, "transmute changes the byte size");
      if (tight_stride(n.desc) != sp.desc.copy_row_bytes) {
        // Different row shapes: both sides must be tight to copy as one run.
        if (row_stride_of(src_d) != tight_stride(src_d) ||
            row_stride_of(plan.output_desc[n.out_slots[0]]) != tight_stride(n.desc))
          return fail(ZH_ERR_UNSUPPORTED, "transmute between padded layouts of different row size");
        sp.desc.copy_row_bytes = sp.desc.copy_row_bytes * sp.desc.copy_rows;
        sp.desc.copy_rows = 1;
      }
      sp.alg_bytes = 2 * sp.desc.copy_row_bytes * sp.desc.copy_rows;
      sp.text = "copy " + std::to_string(sp.desc.copy_row_bytes * sp.desc.copy_rows) + " bytes";
      for (size_t k = 1; k < n.out_slots.size(); ++k) {
        StagePlan extra = sp;
        extra.slot_ref[1] = ~(int64_t)n.out_slots[k];
        extra.desc.copy_dst_stride = row_stride_of(plan.output_desc[n.out_slots[k]]);
        plan.stages.push_back(extra);
      }
      plan.stages.push_back(sp);
      continue;
    }

    if (n.kind == N_DYNAMIC) {
      // The user kernel: operands and result in RGBA f32 intermediates.
      sp.desc.kind = STAGE_DYNAMIC;
      sp.dynamic = n.dyn;
      std::ostringstream t;
      t << "dynamic " << n.w << "x" << n.h << ": ";
      for (int operand : {n.a, n.b}) {
        if (operand < 0) continue;
        const Node& c = B.nodes[operand];
        sp.slot_ref.push_back(kIntermediateBase + c.buffer);
        sp.alg_bytes += (uint64_t)c.w * c.h * 16;
      }
      if (n.buffer < 0) return fail(ZH_ERR_UNSUPPORTED, "internal: a user kernel's result has no buffer");
      sp.slot_ref.push_back(kIntermediateBase + n.buffer);
      sp.alg_bytes += (uint64_t)n.w * n.h * 16;
      t << plan.dynamics[n.dyn].num_args << " operand(s) [kernel " << n.label << "]";
      sp.text = t.str();
      plan.stages.push_back(sp);
      continue;
    }
    if (n.kind == N_INPUT) {
      // Passthrough of an input (tests/color_support.rs): row copies when the
      // texels agree (the row pitches may differ), a decode/encode stage
      // otherwise.
      for (int slot_out : n.out_slots) {
        const zh_descriptor& od = plan.output_desc[slot_out];
        StagePlan cp;
        std::memset(&cp.desc, 0, sizeof(cp.desc));
        cp.desc.pre_slot[0] = cp.desc.pre_slot[1] = -1;
        cp.desc.variant = -1;
        zh_descriptor pitched = n.desc;
        pitched.row_stride = od.row_stride;
        if (same_layout(pitched, od)) {
          cp.desc.kind = STAGE_COPY;
          cp.slot_ref = {n.input_slot, ~(int64_t)slot_out};
          cp.desc.copy_rows = n.desc.height;
          cp.desc.copy_row_bytes = tight_stride(n.desc);
          cp.desc.copy_src_stride = row_stride_of(n.desc);
          cp.desc.copy_dst_stride = row_stride_of(od);
          cp.alg_bytes = 2 * cp.desc.copy_row_bytes * cp.desc.copy_rows;
          cp.text = "copy " + std::to_string(cp.alg_bytes / 2) + " bytes";
        } else {
          cp.desc.kind = STAGE_POINTWISE;
          cp.desc.width = n.w;
          cp.desc.height = n.h;
          Emitter E2(B, plan, cp, false);
          E2.load(n.input_slot, n.desc);
          const int so = E2.slot(~(int64_t)slot_out, od);
          E2.op(OP_STORE, (uint8_t)so);
          E2.finish();
          cp.alg_bytes = slot_bytes(cp.desc.slot[0]) + slot_bytes(cp.desc.slot[so]);
          cp.text = "pointwise " + std::to_string(n.w) + "x" + std::to_string(n.h) + ": load store";
        }
        plan.stages.push_back(cp);
      }
      continue;
    }
    if (paint < 0) {
      sp.desc.kind = STAGE_POINTWISE;
      const bool pm = E.value((int)id, (int)id);
      emit_stores(pm);
      if (!E.err.empty()) return fail(ZH_ERR_UNSUPPORTED, "stage " + std::to_string(plan.stages.size()) + ": " + E.err);
      E.finish();
      // Flat (1-D) indexing when every slot is tight and the same size, and
      // no generator needs the pixel coordinates.
      bool flat = true;
      for (uint32_t i = 0; i < sp.desc.n_instr; ++i)
        if (sp.desc.code[i].op == OP_GRID || sp.desc.code[i].op == OP_BOX3 || sp.desc.code[i].op == OP_NORMAL2D ||
            sp.desc.code[i].op == OP_NOISE || sp.desc.code[i].op == OP_PALETTE)
          flat = false;
      for (size_t k = 0; k < sp.slot_ref.size(); ++k) {
        const Slot& s = sp.desc.slot[k];
        if (s.width != n.w || s.height != n.h || s.row_stride != (uint64_t)s.width * s.bpp) flat = false;
      }
      sp.desc.flat = flat ? 1u : 0u;
      // In-line codecs only (k_pointwise<true>): every slot has a FmtCode
      // and no out-of-line transfer ops.
      bool fast = flat;
      for (size_t k = 0; k < sp.slot_ref.size(); ++k) fast = fast && sp.desc.slot[k].fmt != FMT_GENERIC;
      for (uint32_t i = 0; i < sp.desc.n_instr; ++i) {
        const uint8_t o = sp.desc.code[i].op;
        if (o == OP_EOTF || o == OP_OETF || o == OP_QUANT || o == OP_OFFSET) fast = false;
      }
      sp.fast = fast;
      // A registered signature runs as a specialised kernel (zh_spec.h).
      sp.spec = -1;
      if (fast) {
        uint16_t sig[kMaxInstr];
        for (uint32_t i = 0; i < sp.desc.n_instr; ++i) {
          const Instr& in = sp.desc.code[i];
          const bool io = in.op == OP_LOAD || in.op == OP_STORE;
          sig[i] = sg(in.op, io ? sp.desc.slot[in.a].fmt : in.op == OP_CBRT ? in.a : 0u,
                      in.op == OP_LOAD && in.b ? 1u : 0u);
        }
        sp.spec = zh_spec_lookup(sig, (int)sp.desc.n_instr);
        // Any other chain of the operations the specialised bodies run gets
        // its kernel compiled at run time (zh_jit.cpp).
        if (sp.spec < 0 && !(flags & ZH_PLAN_NO_JIT) && jit_supported(sig, (int)sp.desc.n_instr))
          sp.jit = jit_signature(sig, (int)sp.desc.n_instr);
      }
      if (flat) {
        // The first two distinct 4- or 8-byte load slots are prefetched one
        // grid-stride step ahead.
        int np = 0;
        for (uint32_t i = 0; i < sp.desc.n_instr && np < 2; ++i) {
          const Instr& in = sp.desc.code[i];
          if (in.op != OP_LOAD) continue;
          const uint32_t bpp = sp.desc.slot[in.a].bpp;
          if ((bpp == 4 || bpp == 8) && (np == 0 || sp.desc.pre_slot[0] != in.a)) sp.desc.pre_slot[np++] = in.a;
        }
      }
      if (flat) {
        for (size_t k = 0; k < sp.slot_ref.size(); ++k) {
          Slot& s = sp.desc.slot[k];
          s.width = n.w * n.h;
          s.height = 1;
          s.row_stride = (uint64_t)s.width * s.bpp;
        }
      }
      for (size_t k = 0; k < sp.slot_ref.size(); ++k)
        if (sp.slot_ref[k] > -1000000 || sp.slot_ref[k] < -2000000000ll) {
          // quantisation-only slots (strict mode) carry no traffic
        }
      for (uint32_t i = 0; i < sp.desc.n_instr; ++i) {
        const Instr& in = sp.desc.code[i];
        if (in.op == OP_LOAD || in.op == OP_STORE || in.op == OP_BOX3) sp.alg_bytes += slot_bytes(sp.desc.slot[in.a]);
      }
      std::ostringstream t;
      t << "pointwise " << n.w << "x" << n.h << (flat ? " flat" : "") << ": "
        << describe_code(sp.desc, 0, sp.desc.n_instr);
      if (sp.spec >= 0)
        t << " [kernel " << zh_spec_name(sp.spec) << "]";
      else if (sp.jit >= 0)
        t << " [kernel jit#" << sp.jit << "]";
      else if (sp.fast)
        t << " [kernel interp-inline]";
      sp.text = t.str();
      plan.stages.push_back(sp);
      continue;
    }

    // Resample stage: layers bottom-up, then the base, then the post chain.
    sp.desc.kind = STAGE_RESAMPLE;
    std::vector<int> chain;  // unary nodes from root down to the paint
    for (int u = (int)id; u != paint; u = B.nodes[u].a) chain.push_back(u);
    std::vector<int> layers;
    const bool bil = is_layer_class_bilinear(B.nodes[paint].mode);
    int base = paint;
    while (true) {
      layers.push_back(base);
      const Node& L = B.nodes[base];
      const Node& bn = B.nodes[L.a];
      if ((int)layers.size() < kMaxLayers && bn.kind == N_PAINT && !bn.mat && bn.uses == 1 && bn.mode == L.mode &&
          bn.w == L.w && bn.h == L.h) {
        base = L.a;
        continue;
      }
      base = L.a;
      break;
    }
    std::reverse(layers.begin(), layers.end());
    // Replacing paints move texels unchanged, so they stay in straight
    // alpha (a premultiply round trip would lose the colour of alpha = 0
    // texels); compositing layers work premultiplied.
    const bool straight = B.nodes[paint].mode == LAYER_NEAREST_REPLACE;
    auto represent = [&](bool pm_now) {
      if (straight && pm_now) E.op(OP_UNPREMUL);
      if (!straight && !pm_now) E.op(OP_PREMUL);
    };
    // base program
    sp.desc.base_begin = 0;
    E.depth = 0;
    represent(E.value(base, -1));
    sp.desc.base_end = (uint32_t)E.code.size();
    // layer source programs
    sp.desc.n_layers = (uint32_t)layers.size();
    uint64_t max_box = 0;
    int TW = 32, TH = 32, BH = 32;
    // Tile shape and band height of the LDS-staged kernel: the candidate with
    // the least LDS fill per output pixel whose worst band box fits the LDS
    // budget; no fit -> four gathered taps per pixel (TILE_DIRECT).
    auto choose_tiles = [&](const std::vector<std::array<int, 2>>& cands) {
      double best = 1e300;
      uint64_t best_box = 0;
      for (auto& cd : cands) {
        for (int bh = cd[1]; bh >= 1; bh /= 2) {
          uint64_t worst = 0, sum = 0;
          for (int L : layers) {
            const uint64_t b = tile_box_texels(B.nodes[L].inv, cd[0], bh);
            worst = std::max(worst, b);
            sum += b * (uint64_t)(cd[1] / bh);
          }
          if (worst * sizeof(float) * 4 > (uint64_t)kTileLdsBytes) continue;
          if ((double)sum < best) {
            best = (double)sum;
            TW = cd[0];
            TH = cd[1];
            BH = bh;
            best_box = worst;
          }
          break;  // the tallest band that fits has the least refetch
        }
      }
      if (best_box == 0) {
        sp.tile_mode = 1;  // TILE_DIRECT
        TW = cands[0][0];
        TH = cands[0][1];
        BH = TH;
      } else {
        sp.tile_mode = 0;  // TILE_STAGED
        max_box = best_box;
      }
    };
    if (bil) {
      choose_tiles({{32, 32}, {64, 16}, {16, 64}, {128, 8}, {8, 128}});
    } else {
      sp.tile_mode = 2;  // TILE_NEAREST
    }
    sp.desc.tile_w = TW;
    sp.desc.tile_h = TH;
    sp.desc.band_h = BH;
    sp.desc.tiles_x = (n.w + TW - 1) / TW;
    sp.desc.tiles_y = (n.h + TH - 1) / TH;
    sp.desc.lds_texels = (uint32_t)std::min<uint64_t>(max_box, kTileLdsBytes / 16);
    sp.lds_bytes = sp.tile_mode == 0 ? (uint32_t)(sp.desc.lds_texels * 16) : 0;

    std::ostringstream t;
    t << "resample " << n.w << "x" << n.h << " tile " << TW << "x" << TH
      << (sp.tile_mode == 0 ? " lds" : (sp.tile_mode == 1 ? " direct" : " nearest"));
    t << " base[" << describe_code(sp.desc, 0, 0) << "]";
    for (size_t li = 0; li < layers.size(); ++li) {
      const Node& L = B.nodes[layers[li]];
      Layer& ly = sp.desc.layer[li];
      ly.mode = L.mode;
      ly.src_w = L.src_w;
      ly.src_h = L.src_h;
      for (int k = 0; k < 9; ++k) ly.inv[k] = L.inv[k];
      for (int k = 0; k < 4; ++k) ly.cover[k] = L.cover[k];
      fixed_map(ly);
      // An affine map takes the output's pixel-centre hull to a
      // parallelogram; when its corners (minus the half texel) lie inside
      // (-1, w) x (-1, h), so does every tile's footprint, and the kernels'
      // per-tile test would always pass.
      ly.covers_all = 0;
      if (L.inv[6] == 0.0 && L.inv[7] == 0.0 && L.inv[8] == 1.0) {
        bool inside = true;
        const double xs[2] = {0.5, n.w - 0.5}, ys[2] = {0.5, n.h - 0.5};
        for (double x : xs)
          for (double y : ys) {
            const double u = L.inv[0] * x + L.inv[1] * y + L.inv[2] - 0.5;
            const double v = L.inv[3] * x + L.inv[4] * y + L.inv[5] - 0.5;
            inside = inside && u > -1.0 && v > -1.0 && u < (double)L.src_w && v < (double)L.src_h;
          }
        ly.covers_all = inside ? 1u : 0u;
      }
      ly.prog_begin = (uint32_t)E.code.size();
      E.depth = 0;
      E.pending_mat_valid = false;
      represent(E.value(L.b, -1));
      ly.prog_end = (uint32_t)E.code.size();
      // algorithmic bytes: the visible part of each source read once
      const double frac = L.mode == LAYER_BILINEAR_OVER ? visible_fraction(L.inv, L.src_w, L.src_h, n.w, n.h) : 1.0;
      for (uint32_t i = ly.prog_begin; i < ly.prog_end; ++i) {
        const Instr& in = E.code[i];
        if (in.op == OP_LOAD) sp.alg_bytes += (uint64_t)(frac * slot_bytes(sp.desc.slot[in.a]));
      }
    }
    for (uint32_t i = sp.desc.base_begin; i < sp.desc.base_end; ++i)
      if (E.code[i].op == OP_LOAD) sp.alg_bytes += slot_bytes(sp.desc.slot[E.code[i].a]);
    // post program: per-pixel chain above the paint, then stores
    sp.desc.post_begin = (uint32_t)E.code.size();
    E.depth = 1;
    E.pending_mat_valid = false;
    bool pm = !straight;
    for (auto it = chain.rbegin(); it != chain.rend(); ++it) {
      const Node& u = B.nodes[*it];
      for (const Micro& m : u.micro) {
        if (pm) { E.op(OP_UNPREMUL); pm = false; }
        switch (m.k) {
          case M_MAT3: E.mat(m.m); break;
          case M_CBRT: E.cbrt(); break;
          case M_CUBE: E.op(OP_CUBE); break;
          case M_CLAMP: E.op(OP_CLAMP01); break;
          case M_PREMUL: E.op(OP_PREMUL); break;
          case M_UNPREMUL: E.op(OP_UNPREMUL); break;
          case M_SRLAB: E.op(OP_SRLAB); break;
          case M_SRLAB_INV: E.op(OP_SRLAB_INV); break;
          case M_CHANNEL: E.channel(m.m); break;
        }
      }
      pm = E.quantise(u, pm, (int)id);
    }
    emit_stores(pm);
    sp.desc.post_end = (uint32_t)E.code.size();
    if (!E.err.empty()) return fail(ZH_ERR_UNSUPPORTED, "stage " + std::to_string(plan.stages.size()) + ": " + E.err);
    E.finish();
    for (uint32_t i = sp.desc.post_begin; i < sp.desc.post_end; ++i)
      if (sp.desc.code[i].op == OP_STORE) sp.alg_bytes += slot_bytes(sp.desc.slot[sp.desc.code[i].a]);
    // A registered affine layer form runs as a warp kernel (zh_warp.hip) on
    // 64 x 16 tiles; the interpreter kernels stay valid on the same tiles as
    // the fallback for images the warp kernel cannot address.
    sp.desc.variant = bil ? warp_variant(sp.desc, flags) : -1;
    if (sp.desc.variant >= 0) {
      max_box = 0;
      choose_tiles({{64, 16}});
      sp.desc.tile_w = TW;
      sp.desc.tile_h = TH;
      sp.desc.band_h = BH;
      sp.desc.tiles_x = (n.w + TW - 1) / TW;
      sp.desc.tiles_y = (n.h + TH - 1) / TH;
      sp.desc.lds_texels = (uint32_t)std::min<uint64_t>(max_box, kTileLdsBytes / 16);
      sp.lds_bytes = sp.tile_mode == 0 ? (uint32_t)(sp.desc.lds_texels * 16) : 0;
      sp.staged = (flags & ZH_PLAN_STAGED) && warp_staged_fits(sp.desc, TW, TH);
    }
    {
      std::ostringstream t2;
      t2 << "resample " << n.w << "x" << n.h << " tile " << TW << "x" << TH << " band " << BH
         << (sp.tile_mode == 0 ? " lds" : (sp.tile_mode == 1 ? " direct" : " nearest"))
         << " base[" << describe_code(sp.desc, sp.desc.base_begin, sp.desc.base_end) << "]";
      for (uint32_t li = 0; li < sp.desc.n_layers; ++li) {
        const Layer& ly = sp.desc.layer[li];
        t2 << " layer" << li << "(" << (ly.mode == LAYER_BILINEAR_OVER ? "bilinear-over" : (ly.mode == LAYER_NEAREST_OVER ? "nearest-over" : "nearest-replace"))
           << " " << ly.src_w << "x" << ly.src_h << ")[" << describe_code(sp.desc, ly.prog_begin, ly.prog_end) << "]";
      }
      t2 << " post[" << describe_code(sp.desc, sp.desc.post_begin, sp.desc.post_end) << "]";
      const char* lds = sp.staged ? "_lds" : "";
      if (sp.desc.variant >= kJitWarpBase)
        t2 << " [kernel jitw#" << (sp.desc.variant - kJitWarpBase) / 2 << ((sp.desc.variant & 1) ? "_sep" : "") << lds
           << "]";
      else if (sp.desc.variant >= 0)
        t2 << " [kernel " << zh_warp_name(sp.desc.variant) << lds << "]";
      sp.text = t2.str();
    }
    plan.stages.push_back(sp);
  }

  for (size_t i = 0; i < plan.stages.size(); ++i) text << "stage " << i << ": " << plan.stages[i].text << "\n";
  plan.text = text.str();
  *out = std::move(plan);
  return ZH_OK;
}

}  // namespace zh
