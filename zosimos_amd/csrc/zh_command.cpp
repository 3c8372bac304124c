// zh_command.cpp -- C++ mirror of zosimos's CommandBuffer and Linker.
//
// Each constructor validates its operands the way lib/zosimos/src/command.rs
// does and records an SSA op; zs_command_compile emits the High list the way
// Linker::link_in does (command.rs:2345-2880), one texture per register.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/zosimos_command.h"
#include "zh_internal.h"

using namespace zh;

namespace {

enum OpKind {
  K_INPUT, K_OUTPUT, K_CROP, K_CONVERT, K_INSCRIBE, K_TRANSMUTE, K_AFFINE, K_BLEND, K_SOLID, K_ADJUST, K_BILINEAR,
  K_DERIVATIVE, K_BUFFER, K_FROM_BUFFER, K_NORMAL2D, K_NOISE, K_PALETTE, K_EXTRACT, K_INJECT,
  K_RESULT,  // Op::InvokedResult (command.rs:911-916): a result register of the Invoke that follows
  K_INVOKE,  // Op::Invoke (command.rs:918-923)
  K_DYNAMIC  // Op::DynamicImage (command.rs:180-185): a user stage kernel
};

// GenericDescriptor (command.rs:89-92): size and chroma (texel + colour),
// each concrete or a generic parameter of the command buffer.
struct GenDesc {
  int32_t size_var = -1;    // -1: desc.width / desc.height
  int32_t chroma_var = -1;  // -1: desc.texel / desc.color
  zh_descriptor desc{};     // the concrete parts (placeholders where generic)
};

// CommandSignature (command.rs:819-849): generic parameters, inputs, outputs.
struct Signature {
  uint32_t vars = 0;
  std::vector<GenDesc> inputs, outputs;
};

struct CmdOp {
  OpKind kind;
  zh_descriptor desc{};  // register descriptor (outputs: the source's)
  uint32_t src = 0, above = 0;
  zh_rect rect{};
  uint32_t shader = 0;                  // colour conversion shader key
  std::vector<uint8_t> blob;            // std140 parameters (buffers: their bytes)
  float matrix[9] = {0};
  uint32_t sampling = 0;
  int32_t knob = -1;
  int32_t gsize = -1, gchroma = -1;  // generic size / chroma of the register (GenDesc)
  // K_INVOKE: the function symbol, its arguments, result registers and the
  // generic arguments.
  uint32_t fn = 0;
  std::vector<uint32_t> args, results;
  std::vector<GenDesc> generics;
  // K_DYNAMIC: the shader (blob = zh_dynamic_call + the shader data).
  zh::ShaderRef dyn_shader;
};

// CIE 1931 2-degree chromaticities of the white points A .. F11 (the order of
// image-canvas Whitepoint); D65 and D50 use the 4-digit values of BT.709 /
// IEC 61966-2-1 and ICC.
const double kWhite[11][2] = {{0.44757, 0.40745}, {0.34840, 0.35160}, {0.31006, 0.31616}, {0.3457, 0.3585},
                              {0.33242, 0.34743}, {0.3127, 0.3290},   {0.29902, 0.31485}, {1.0 / 3, 1.0 / 3},
                              {0.37208, 0.37529}, {0.31292, 0.32933}, {0.38052, 0.37713}};

// XYZ of a white point at Y = 1.
void white_xyz(uint32_t wp, double W[3]) {
  const double wx = kWhite[wp][0], wy = kWhite[wp][1];
  W[0] = wx / wy;
  W[1] = 1.0;
  W[2] = (1.0 - wx - wy) / wy;
}

// RGB -> XYZ from chromaticities and white point (textbook derivation, as
// image-canvas Primaries::to_xyz_row_matrix provides it to command.rs:1023).
bool rgb_to_xyz(uint32_t primaries, uint32_t wp, Mat3* out) {
  static const double chroma[8][3][2] = {
      {{0, 0}, {0, 0}, {0, 0}},
      {{0.630, 0.340}, {0.310, 0.595}, {0.155, 0.070}},  // BT.601 525 (SMPTE C)
      {{0.640, 0.330}, {0.290, 0.600}, {0.150, 0.060}},  // BT.601 625 (EBU)
      {{0.640, 0.330}, {0.300, 0.600}, {0.150, 0.060}},  // BT.709 / sRGB
      {{0.630, 0.340}, {0.310, 0.595}, {0.155, 0.070}},  // SMPTE 240M
      {{0.708, 0.292}, {0.170, 0.797}, {0.131, 0.046}},  // BT.2020
      {{0.708, 0.292}, {0.170, 0.797}, {0.131, 0.046}},  // BT.2100
      {{0.680, 0.320}, {0.265, 0.690}, {0.150, 0.060}},  // Display P3 (SMPTE EG 432-1)
  };
  if (primaries > 7 || wp > 10) return false;
  if (primaries == 0) { *out = mat_identity(); return true; }
  Mat3 P;
  for (int c = 0; c < 3; ++c) {
    const double x = chroma[primaries][c][0], y = chroma[primaries][c][1];
    P[0 * 3 + c] = x / y;
    P[1 * 3 + c] = 1.0;
    P[2 * 3 + c] = (1.0 - x - y) / y;
  }
  double W[3];
  white_xyz(wp, W);
  Mat3 Pi;
  if (!mat_inv(P, &Pi)) return false;
  double s[3];
  for (int i = 0; i < 3; ++i) s[i] = Pi[i * 3] * W[0] + Pi[i * 3 + 1] * W[1] + Pi[i * 3 + 2] * W[2];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) P[r * 3 + c] *= s[c];
  *out = P;
  return true;
}

// RowMatrix::into_mat3x3_std140 (color_matrix.rs:151-161).
std::vector<uint8_t> std140(const Mat3& m) {
  const float f[12] = {(float)m[0], (float)m[3], (float)m[6], 0.0f, (float)m[1], (float)m[4],
                       (float)m[7], 0.0f,        (float)m[2], (float)m[5], (float)m[8], 0.0f};
  std::vector<uint8_t> b(sizeof(f));
  std::memcpy(b.data(), f, sizeof(f));
  return b;
}

bool chroma_equal(const zh_descriptor& a, const zh_descriptor& b) {
  if (a.texel.bits != b.texel.bits || a.texel.parts != b.texel.parts || a.color.model != b.color.model) return false;
  if (a.color.model == ZH_COLOR_OKLAB) return true;
  if (a.color.model == ZH_COLOR_SRLAB2) return a.color.whitepoint == b.color.whitepoint;
  return a.color.transfer == b.color.transfer &&
         (a.color.model != ZH_COLOR_RGB ||
          (a.color.primaries == b.color.primaries && a.color.whitepoint == b.color.whitepoint));
}

zh_rect with_size(uint32_t w, uint32_t h) { return zh_rect{0, 0, w, h}; }
uint32_t rw(zh_rect r) { return r.max_x > r.x ? r.max_x - r.x : 0; }
uint32_t rh(zh_rect r) { return r.max_y > r.y ? r.max_y - r.y : 0; }
bool rect_eq(zh_rect a, zh_rect b) { return a.x == b.x && a.y == b.y && a.max_x == b.max_x && a.max_y == b.max_y; }
// Rectangle::contains, command.rs:3560-3571.
bool contains(zh_rect s, zh_rect o) {
  if (!(s.x <= o.x && s.y <= o.y)) return false;
  const uint32_t ox = o.x - s.x, oy = o.y - s.y;
  if (rw(s) < ox || rh(s) < oy) return false;
  return rw(s) - ox >= rw(o) && rh(s) - oy >= rh(o);
}
// Rectangle::normalize as written at command.rs:3536-3543: the height of the
// result is the width (kept: inscribe's draws depend on it).
zh_rect normalize(zh_rect r) { return zh_rect{r.x, r.y, r.x + rw(r), r.y + rw(r)}; }

void quad_of_rect(zh_rect r, float q[8]) {
  const float x0 = (float)r.x, y0 = (float)r.y, x1 = (float)r.max_x, y1 = (float)r.max_y;
  const float v[8] = {x0, y0, x1, y0, x1, y1, x0, y1};
  std::memcpy(q, v, sizeof(v));
}

}  // namespace

bool zh::primaries_to_xyz(uint32_t primaries, uint32_t whitepoint, Mat3* out) {
  return rgb_to_xyz(primaries, whitepoint, out);
}

struct zs_command {
  std::vector<CmdOp> ops;
  int32_t next_knob = 0;
  uint32_t vars = 0;                // generic parameters declared (command.rs:772-791)
  std::vector<GenDesc> tys;         // DescriptorVar -> descriptor
  std::vector<Signature> symbols;   // FunctionVar -> declared signature
};

struct zs_signature {
  Signature sig;
};

struct zs_linked {
  std::vector<zh_descriptor> textures;
  std::vector<zh_op> ops;
  std::vector<std::vector<uint8_t>> blobs;
  std::vector<int64_t> input_slot, output_slot;  // per register of the entry buffer
  // RegisterKnob {link_idx, register} -> knob (command.rs:2128-2140).
  std::vector<std::array<int64_t, 3>> knobs;
  std::vector<zh::ShaderRef> shaders;  // held for the zh_dynamic_call blobs
};

// The register's descriptor, generic or not (describe_reg(..).as_texture()).
static zh_status reg_gen(const zs_command* c, uint32_t r, GenDesc* g) {
  if (!c || r >= c->ops.size()) return fail(ZH_ERR_OTHER, "bad register");
  const CmdOp& op = c->ops[r];
  if (op.kind == K_OUTPUT) return fail(ZH_ERR_TYPE, "register is an output, not a texture");
  if (op.kind == K_BUFFER) return fail(ZH_ERR_TYPE, "register is a buffer, not a texture");
  if (op.kind == K_INVOKE) return fail(ZH_ERR_TYPE, "register is an invocation, not a texture");
  g->desc = op.desc;
  g->size_var = op.gsize;
  g->chroma_var = op.gchroma;
  return ZH_OK;
}

// The register's concrete descriptor; a generic one is a type error
// (as_concrete(), CommandErrorKind::ConcreteDescriptorRequired).
static zh_status reg_desc(const zs_command* c, uint32_t r, zh_descriptor* d) {
  GenDesc g;
  zh_status st = reg_gen(c, r, &g);
  if (st != ZH_OK) return st;
  if (g.size_var >= 0 || g.chroma_var >= 0)
    return fail(ZH_ERR_TYPE, "operation requires a concrete descriptor; register " + std::to_string(r) + " is generic");
  *d = g.desc;
  return ZH_OK;
}

static bool gen_equal(const GenDesc& a, const GenDesc& b) {
  if (a.size_var != b.size_var || a.chroma_var != b.chroma_var) return false;
  if (a.size_var < 0 && (a.desc.width != b.desc.width || a.desc.height != b.desc.height)) return false;
  if (a.chroma_var < 0) {
    const zh_texel &ta = a.desc.texel, &tb = b.desc.texel;
    const zh_color &ca = a.desc.color, &cb = b.desc.color;
    if (ta.bits != tb.bits || ta.parts != tb.parts || ca.model != cb.model || ca.primaries != cb.primaries ||
        ca.transfer != cb.transfer || ca.whitepoint != cb.whitepoint || ca.luminance != cb.luminance)
      return false;
  }
  return true;
}

// GenericDescriptor::rewrite (command.rs:3143-3154): generic parts replaced
// by the corresponding parts of `decl`.
static GenDesc rewrite(const GenDesc& g, const std::vector<GenDesc>& decl) {
  GenDesc r = g;
  if (g.size_var >= 0) {
    const GenDesc& s = decl[(size_t)g.size_var];
    r.size_var = s.size_var;
    r.desc.width = s.desc.width;
    r.desc.height = s.desc.height;
  }
  if (g.chroma_var >= 0) {
    const GenDesc& s = decl[(size_t)g.chroma_var];
    r.chroma_var = s.chroma_var;
    r.desc.texel = s.desc.texel;
    r.desc.color = s.desc.color;
    r.desc.texel_stride = s.desc.texel_stride;
  }
  // A substituted part makes the placeholder's row layout meaningless: the
  // instance has tight rows.  A concrete descriptor keeps its own.
  if ((g.size_var >= 0 || g.chroma_var >= 0) && r.size_var < 0 && r.chroma_var < 0) r.desc.row_stride = 0;
  return r;
}

static uint32_t push(zs_command* c, CmdOp op) {
  c->ops.push_back(std::move(op));
  return (uint32_t)c->ops.size() - 1;
}

static zh_descriptor normalized(zh_descriptor d) {
  d.texel_stride = bits_bytes(d.texel.bits);
  return d;
}

extern "C" {

zh_status zs_command_create(zs_command** out) {
  if (!out) return fail(ZH_ERR_INVALID_ARGUMENT, "null output");
  *out = new zs_command();
  return ZH_OK;
}

zh_status zs_command_destroy(zs_command* cmd) {
  delete cmd;
  return ZH_OK;
}

zh_status zs_command_input(zs_command* cmd, const zh_descriptor* desc, uint32_t* reg) {
  if (!cmd || !desc || !reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  std::string why;
  if (!descriptor_valid(*desc, &why)) return fail(ZH_ERR_TYPE, "inconsistent input declared: " + why);
  CmdOp op;
  op.kind = K_INPUT;
  op.desc = normalized(*desc);
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_command_output(zs_command* cmd, uint32_t src, uint32_t* reg, zh_descriptor* outformat) {
  if (!cmd || !reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  GenDesc g;  // a function's output may be generic (command.rs:1707-1712)
  zh_status st = reg_gen(cmd, src, &g);
  if (st != ZH_OK) return st;
  CmdOp op;
  op.kind = K_OUTPUT;
  op.src = src;
  op.desc = g.desc;
  op.gsize = g.size_var;
  op.gchroma = g.chroma_var;
  *reg = push(cmd, op);
  if (outformat) *outformat = g.desc;
  return ZH_OK;
}

zh_status zs_command_render(zs_command* cmd, uint32_t src, uint32_t* reg, zh_descriptor* outformat) {
  if (!cmd || !reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  zh_descriptor d;
  zh_status st = reg_desc(cmd, src, &d);
  if (st != ZH_OK) return st;
  // A render target needs a texture form (ImageDescriptor::new,
  // program.rs:782-946): native RGBA8/BGRA8, or a staged texel of 1, 2, 4, 8
  // or 16 bytes (StageKind, shaders/stage.rs:63-72); 3- and 6-byte texels
  // have none (command.rs:1729-1731).
  const uint32_t b = bits_bytes(d.texel.bits);
  if (!is_native(d) && b != 1 && b != 2 && b != 4 && b != 8 && b != 16)
    return fail(ZH_ERR_TYPE, "render target has no texture format");
  return zs_command_output(cmd, src, reg, outformat);
}

zh_status zs_command_crop(zs_command* cmd, uint32_t src, zh_rect rect, uint32_t* reg) {
  if (!reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  GenDesc g;
  zh_status st = reg_gen(cmd, src, &g);
  if (st != ZH_OK) return st;
  CmdOp op;
  op.kind = K_CROP;
  op.src = src;
  op.rect = rect;
  op.desc = g.desc;  // command.rs:972-977 keeps the source descriptor, generic or not
  op.gsize = g.size_var;
  op.gchroma = g.chroma_var;
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_command_color_convert(zs_command* cmd, uint32_t src, const zh_color* color, const zh_texel* texel,
                                   uint32_t* reg) {
  if (!color || !texel || !reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  zh_descriptor s;
  zh_status st = reg_desc(cmd, src, &s);
  if (st != ZH_OK) return st;
  CmdOp op;
  op.kind = K_CONVERT;
  op.src = src;
  const zh_color& a = s.color;
  const zh_color& b = *color;
  Mat3 ps, pd;
  if (a.model == ZH_COLOR_RGB && b.model == ZH_COLOR_RGB && a.whitepoint == b.whitepoint) {
    // Linear RGB(src) -> XYZ -> linear RGB(dst): P_dst^-1 P_src.
    Mat3 pdi;
    if (!rgb_to_xyz(a.primaries, a.whitepoint, &ps) || !rgb_to_xyz(b.primaries, b.whitepoint, &pd) ||
        !mat_inv(pd, &pdi))
      return fail(ZH_ERR_TYPE, "unknown primaries or white point");
    op.shader = ZH_SHADER_LINEAR_COLOR_MATRIX;
    op.blob = std140(mat_mul(pdi, ps));
  } else if (a.model == ZH_COLOR_RGB && a.whitepoint == ZH_WHITEPOINT_D65 && b.model == ZH_COLOR_OKLAB) {
    if (!rgb_to_xyz(a.primaries, ZH_WHITEPOINT_D65, &ps)) return fail(ZH_ERR_TYPE, "unknown primaries");
    op.shader = ZH_SHADER_OKLAB_ENCODE;
    op.blob = std140(ps);
  } else if (a.model == ZH_COLOR_OKLAB && b.model == ZH_COLOR_RGB && b.whitepoint == ZH_WHITEPOINT_D65) {
    Mat3 pdi;
    if (!rgb_to_xyz(b.primaries, ZH_WHITEPOINT_D65, &pd) || !mat_inv(pd, &pdi))
      return fail(ZH_ERR_TYPE, "unknown primaries");
    op.shader = ZH_SHADER_OKLAB_DECODE;
    op.blob = std140(pdi);
  } else if ((a.model == ZH_COLOR_RGB && b.model == ZH_COLOR_SRLAB2) ||
             (a.model == ZH_COLOR_SRLAB2 && b.model == ZH_COLOR_RGB)) {
    // ColorConversion::XyzToSrLab2 / SrLab2ToXyz (command.rs:1052-1075,
    // 3252-3274): the RGB side's XYZ matrix at its own white point (inverted
    // for the decode) and the SrLab2 white point's XYZ.
    const bool enc = a.model == ZH_COLOR_RGB;
    const zh_color& rgb = enc ? a : b;
    const zh_color& lab = enc ? b : a;
    Mat3 m, mi;
    if (lab.whitepoint > 10 || !rgb_to_xyz(rgb.primaries, rgb.whitepoint, &m) || !mat_inv(m, &mi))
      return fail(ZH_ERR_TYPE, "unknown primaries or white point");
    double W[3];
    white_xyz(lab.whitepoint, W);
    const float wp[4] = {(float)W[0], (float)W[1], (float)W[2], 0.0f};
    op.shader = enc ? ZH_SHADER_SRLAB2_ENCODE : ZH_SHADER_SRLAB2_DECODE;
    op.blob.resize(16);
    std::memcpy(op.blob.data(), wp, 16);
    const std::vector<uint8_t> mb = std140(enc ? m : mi);
    op.blob.insert(op.blob.end(), mb.begin(), mb.end());
  } else {
    return fail(ZH_ERR_TYPE, "No conversion");
  }
  zh_descriptor d = s;
  d.color = *color;
  d.texel = *texel;
  d.row_stride = 0;
  d = normalized(d);
  std::string why;
  if (!descriptor_valid(d, &why)) return fail(ZH_ERR_TYPE, "conversion target: " + why);
  op.desc = d;
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_command_inscribe(zs_command* cmd, uint32_t below, zh_rect rect, uint32_t above, uint32_t* reg) {
  if (!reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  zh_descriptor db, da;
  zh_status st = reg_desc(cmd, below, &db);
  if (st == ZH_OK) st = reg_desc(cmd, above, &da);
  if (st != ZH_OK) return st;
  if (!chroma_equal(db, da)) return fail(ZH_ERR_TYPE, "ConflictingTypes: inscribe of a different texel/colour");
  if (!rect_eq(with_size(da.width, da.height), rect)) return fail(ZH_ERR_OTHER, "inscribe rect must be the above image");
  if (!contains(with_size(db.width, db.height), rect)) return fail(ZH_ERR_OTHER, "inscribe rect outside the image");
  CmdOp op;
  op.kind = K_INSCRIBE;
  op.src = below;
  op.above = above;
  op.rect = normalize(rect);
  op.desc = db;
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_command_transmute(zs_command* cmd, uint32_t src, const zh_descriptor* target, uint32_t* reg) {
  if (!target || !reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  zh_descriptor s;
  zh_status st = reg_desc(cmd, src, &s);
  if (st != ZH_OK) return st;
  if (s.width != target->width || s.height != target->height)
    return fail(ZH_ERR_TYPE, "invalid transmute with mismatched size");
  if (bits_bytes(s.texel.bits) != bits_bytes(target->texel.bits))
    return fail(ZH_ERR_TYPE, "ConflictingTypes: transmute between texels of different size");
  std::string why;
  if (!descriptor_valid(*target, &why)) return fail(ZH_ERR_TYPE, "invalid transmute with inconsistent result");
  CmdOp op;
  op.kind = K_TRANSMUTE;
  op.src = src;
  op.desc = normalized(*target);
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_command_affine(zs_command* cmd, uint32_t below, const float matrix[9], uint32_t sampling,
                            uint32_t above, uint32_t* reg) {
  if (!matrix || !reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  zh_descriptor db, da;
  zh_status st = reg_desc(cmd, below, &db);
  if (st == ZH_OK) st = reg_desc(cmd, above, &da);
  if (st != ZH_OK) return st;
  if (!chroma_equal(db, da)) return fail(ZH_ERR_TYPE, "affine of a different texel/colour");
  // RowMatrix::det in f32 against f32::EPSILON (command.rs:1648-1656).
  const float* m = matrix;
  const float det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                    m[2] * (m[3] * m[7] - m[4] * m[6]);
  if (!(std::fabs(det) >= 1.1920929e-07f)) return fail(ZH_ERR_OTHER, "singular affine transformation");
  if (sampling != ZS_AFFINE_NEAREST && sampling != ZS_AFFINE_BILINEAR)
    return fail(ZH_ERR_INVALID_ARGUMENT, "unknown sampling");
  CmdOp op;
  op.kind = K_AFFINE;
  op.src = below;
  op.above = above;
  std::memcpy(op.matrix, matrix, sizeof(op.matrix));
  op.sampling = sampling;
  op.desc = db;
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_command_blend(zs_command* cmd, uint32_t below, zh_rect rect, uint32_t above, uint32_t blend,
                           uint32_t* reg) {
  if (!reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  if (blend != ZS_BLEND_ALPHA) return fail(ZH_ERR_UNIMPLEMENTED, "unknown blend mode");
  zh_descriptor db, da;
  zh_status st = reg_desc(cmd, below, &db);
  if (st == ZH_OK) st = reg_desc(cmd, above, &da);
  if (st != ZH_OK) return st;
  if (!chroma_equal(db, da)) return fail(ZH_ERR_TYPE, "blend of a different texel/colour");
  if (rw(rect) != da.width || rh(rect) != da.height) return fail(ZH_ERR_OTHER, "blend rect must have the above size");
  if (!contains(with_size(db.width, db.height), rect)) return fail(ZH_ERR_OTHER, "blend rect outside the image");
  CmdOp op;
  op.kind = K_BLEND;
  op.src = below;
  op.above = above;
  op.rect = rect;
  op.desc = db;
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_command_solid_rgba(zs_command* cmd, const zh_descriptor* desc, const float color[4], uint32_t* reg) {
  if (!cmd || !desc || !color || !reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  std::string why;
  if (!descriptor_valid(*desc, &why)) return fail(ZH_ERR_TYPE, "inconsistent constant color image created");
  CmdOp op;
  op.kind = K_SOLID;
  op.desc = normalized(*desc);
  op.blob.resize(16);
  std::memcpy(op.blob.data(), color, 16);
  *reg = push(cmd, op);
  return ZH_OK;
}

// CommandBuffer::bilinear (command.rs:1615-1633): a generated image, one
// bilinear function of the pixel centre's normalised coordinates
// (bilinear.frag); params = u_min, u_max, v_min, v_max, uv_min, uv_max
// (shaders/bilinear.rs:33-45).
zh_status zs_command_bilinear(zs_command* cmd, const zh_descriptor* desc, const float params[24], uint32_t* reg) {
  if (!cmd || !desc || !params || !reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  std::string why;
  if (!descriptor_valid(*desc, &why)) return fail(ZH_ERR_TYPE, "inconsistent descriptor for bilinear");
  CmdOp op;
  op.kind = K_BILINEAR;
  op.desc = normalized(*desc);
  op.blob.resize(96);
  std::memcpy(op.blob.data(), params, 96);
  *reg = push(cmd, op);
  return ZH_OK;
}

// CommandBuffer::chromatic_adaptation (command.rs:1112-1174), lowered as one
// linear colour matrix (command.rs:2527-2548): RGB relative to the target
// white <- XYZ <- adaptation <- XYZ <- RGB relative to the source white.  The
// adaptation is the von Kries transform M^-1 diag(M w_dst / M w_src) M in the
// method's cone space: identity (XYZ scaling), Hunt-Pointer-Estevez (von
// Kries) or Bradford, the matrices the reference documents at
// command.rs:422-460.
zh_status zs_command_chromatic_adaptation(zs_command* cmd, uint32_t src, uint32_t method, uint32_t target,
                                          uint32_t* reg) {
  if (!reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  zh_descriptor s;
  zh_status st = reg_desc(cmd, src, &s);
  if (st != ZH_OK) return st;
  if (s.color.model != ZH_COLOR_RGB) return fail(ZH_ERR_TYPE, "non-rgb chromatic adaptation");
  if (target > 10 || s.color.whitepoint > 10) return fail(ZH_ERR_TYPE, "unknown white point");
  Mat3 M;
  switch (method) {
    case ZS_ADAPT_XYZ: M = mat_identity(); break;
    case ZS_ADAPT_VON_KRIES: M = {0.40024, 0.70760, -0.08081, -0.22630, 1.16532, 0.04570, 0.0, 0.0, 0.91822}; break;
    case ZS_ADAPT_BRADFORD_VON_KRIES: M = {0.8951, 0.2664, -0.1614, -0.7502, 1.7135, 0.0367, 0.0389, -0.0685, 1.0296}; break;
    case ZS_ADAPT_BRADFORD_NONLINEAR: return fail(ZH_ERR_UNIMPLEMENTED, "Bradford's non-linear adaptation");
    default: return fail(ZH_ERR_INVALID_ARGUMENT, "unknown adaptation method");
  }
  double ws[3], wd[3];
  white_xyz(s.color.whitepoint, ws);
  white_xyz(target, wd);
  Mat3 Mi, to_xyz, from_xyz, xyz_to_rgb;
  if (!mat_inv(M, &Mi) || !rgb_to_xyz(s.color.primaries, s.color.whitepoint, &to_xyz) ||
      !rgb_to_xyz(s.color.primaries, target, &from_xyz) || !mat_inv(from_xyz, &xyz_to_rgb))
    return fail(ZH_ERR_TYPE, "unknown primaries");
  Mat3 D = mat_identity();
  for (int i = 0; i < 3; ++i) {
    const double cs = M[i * 3] * ws[0] + M[i * 3 + 1] * ws[1] + M[i * 3 + 2] * ws[2];
    const double cd = M[i * 3] * wd[0] + M[i * 3 + 1] * wd[1] + M[i * 3 + 2] * wd[2];
    D[i * 4] = cd / cs;
  }
  const Mat3 adapt = mat_mul(Mi, mat_mul(D, M));
  CmdOp op;
  op.kind = K_CONVERT;
  op.src = src;
  op.shader = ZH_SHADER_LINEAR_COLOR_MATRIX;
  op.blob = std140(mat_mul(xyz_to_rgb, mat_mul(adapt, to_xyz)));
  zh_descriptor d = s;
  d.color.whitepoint = target;
  d.row_stride = 0;
  op.desc = normalized(d);
  *reg = push(cmd, op);
  return ZH_OK;
}

// CommandBuffer::derivative (command.rs:1487-1507): box3.frag with the
// outer product of a smoothing and a central-difference vector
// (command.rs:3344-3409), transposed for the height direction.  The output
// keeps the source's descriptor.
zh_status zs_command_derivative(zs_command* cmd, uint32_t src, uint32_t method, uint32_t direction, uint32_t* reg) {
  if (!reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  zh_descriptor s;
  zh_status st = reg_desc(cmd, src, &s);
  if (st != ZH_OK) return st;
  if (direction != ZS_DIRECTION_WIDTH && direction != ZS_DIRECTION_HEIGHT)
    return fail(ZH_ERR_INVALID_ARGUMENT, "unknown direction");
  double smooth[3];
  switch (method) {
    case ZS_DERIVATIVE_PREWITT: smooth[0] = smooth[1] = smooth[2] = 1.0 / 3.0; break;
    case ZS_DERIVATIVE_SOBEL: smooth[0] = 0.25; smooth[1] = 0.5; smooth[2] = 0.25; break;
    case ZS_DERIVATIVE_SCHARR3: smooth[0] = smooth[2] = 46.84 / 256.0; smooth[1] = 162.32 / 256.0; break;
    case ZS_DERIVATIVE_SCHARR3_TO_4BIT: smooth[0] = smooth[2] = 3.0 / 16.0; smooth[1] = 10.0 / 16.0; break;
    case ZS_DERIVATIVE_SCHARR3_TO_8BIT: smooth[0] = smooth[2] = 47.0 / 256.0; smooth[1] = 162.0 / 256.0; break;
    case ZS_DERIVATIVE_ROBERTS:
    case ZS_DERIVATIVE_SCHARR4:
    case ZS_DERIVATIVE_SCHARR5:
    case ZS_DERIVATIVE_SCHARR5_TAB: return fail(ZH_ERR_UNIMPLEMENTED, "derivative method not implemented");
    default: return fail(ZH_ERR_INVALID_ARGUMENT, "unknown derivative method");
  }
  const double diff[3] = {0.5, 0.0, -0.5};
  Mat3 m;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m[r * 3 + c] = smooth[r] * diff[c];
  if (direction == ZS_DIRECTION_HEIGHT) {
    Mat3 t;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) t[r * 3 + c] = m[c * 3 + r];
    m = t;
  }
  CmdOp op;
  op.kind = K_DERIVATIVE;
  op.src = src;
  op.blob = std140(m);
  s.row_stride = 0;
  op.desc = normalized(s);
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_command_oklab_adjust(zs_command* cmd, uint32_t src, float chroma, float hue_rad, uint32_t* reg) {
  if (!reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  zh_descriptor s;
  zh_status st = reg_desc(cmd, src, &s);
  if (st != ZH_OK) return st;
  if (s.color.model != ZH_COLOR_OKLAB) return fail(ZH_ERR_TYPE, "oklab_adjust needs an Oklab register");
  CmdOp op;
  op.kind = K_ADJUST;
  op.src = src;
  op.desc = s;
  const float p[4] = {chroma, hue_rad, 0.0f, 0.0f};
  op.blob.resize(16);
  std::memcpy(op.blob.data(), p, 16);
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_normal2d_with_diagonal(float var0, float var1, float out[8]) {
  if (!out) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  // ShaderData::with_diagonal (shaders/distribution_normal2d.rs): a zero
  // variance drops its axis from the inverse and the determinant.
  const float two_pi = 2.0f * 3.14159265358979323846f;
  const float d0 = var0 == 0.0f ? 0.0f : 1.0f / var0, d1 = var1 == 0.0f ? 0.0f : 1.0f / var1;
  const float f0 = var0 == 0.0f ? 1.0f : two_pi * var0, f1 = var1 == 0.0f ? 1.0f : two_pi * var1;
  const float v[8] = {0.0f, 0.0f, d0, 0.0f, 0.0f, d1, f0 * f1, 0.0f};
  std::memcpy(out, v, sizeof(v));
  return ZH_OK;
}

zh_status zs_normal2d_with_direction(float x, float y, float out[8]) {
  if (!out) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  // ShaderData::with_direction: the covariance d d^T of direction d has the
  // single eigenvalue |d|^2 (the pseudo-determinant, times 2 pi) and the
  // pseudo-inverse d d^T / |d|^4, whose entries are evaluated in f32 in the
  // same rearranged forms (hypot first) for stability.  The diagonal is
  // evaluated as the reference does, diag(x, x) and diag(y, y): 1 / (4 x^2)
  // and 1 / (4 y^2), which equal the pseudo-inverse's only when |x| = |y|.
  const double lsq = (double)x * x + (double)y * y;
  const float length_sq = (float)lsq;
  if (!std::isfinite(length_sq)) return fail(ZH_ERR_INVALID_ARGUMENT, "direction length is not finite");
  auto diag = [](float a, float b) {  // a a / (a a + b b)^2
    const float h = std::hypot(a, b);
    return ((1.0f / h) * (a / h)) / (a + b * (b / a));
  };
  auto off = [](float a, float b) {  // a b / (a a + b b)^2
    const float lo = std::min(a, b), hi = std::max(a, b);
    const float h = std::hypot(lo, hi);
    return ((lo / h) / std::fma(lo, lo / hi, hi)) / h;
  };
  const float two_pi = 2.0f * 3.14159265358979323846f;
  const float v[8] = {0.0f, 0.0f, diag(x, x), off(x, y), off(y, x), diag(y, y), two_pi * length_sq, 0.0f};
  std::memcpy(out, v, sizeof(v));
  return ZH_OK;
}

zh_status zs_command_distribution_normal2d(zs_command* cmd, const zh_descriptor* desc, const float params[8],
                                           uint32_t* reg) {
  if (!cmd || !desc || !params || !reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  std::string why;
  if (!descriptor_valid(*desc, &why)) return fail(ZH_ERR_TYPE, "inconsistent normal2d: " + why);
  // command.rs:1559-1583: one grey channel (Luma or LumaA parts).
  if (desc->texel.parts != 4 && desc->texel.parts != 5) return fail(ZH_ERR_TYPE, "normal2d for non-LumA texel");
  CmdOp op;
  op.kind = K_NORMAL2D;
  op.desc = normalized(*desc);
  op.blob.resize(32);
  std::memcpy(op.blob.data(), params, 32);
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_command_palette(zs_command* cmd, uint32_t palette, const zs_palette* config, uint32_t indices,
                             uint32_t* reg) {
  if (!config || !reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  GenDesc pg, ig;  // either may be generic (command.rs:1448-1470)
  zh_status st = reg_gen(cmd, palette, &pg);
  if (st == ZH_OK) st = reg_gen(cmd, indices, &ig);
  if (st != ZH_OK) return st;
  const zh_descriptor& pd = pg.desc;
  const zh_descriptor& id = ig.desc;
  // ChannelPosition::into_vec4: a one-hot row per axis, or zeros.
  float x[4] = {0, 0, 0, 0}, y[4] = {0, 0, 0, 0};
  for (int axis = 0; axis < 2; ++axis) {
    const int32_t ch = axis == 0 ? config->width : config->height;
    if (ch == ZS_CHANNEL_NONE) continue;
    if (ch < ZS_CHANNEL_R || ch > ZS_CHANNEL_B) return fail(ZH_ERR_TYPE, "palette channel must be R, G or B");
    (axis == 0 ? x : y)[ch] = 1.0f;
  }
  CmdOp op;
  op.kind = K_PALETTE;
  op.src = palette;
  op.above = indices;
  // GenericDescriptor { chroma: palette's, ..indices' } (command.rs:1466-1470).
  op.desc = pd;
  op.desc.width = id.width;
  op.desc.height = id.height;
  op.desc.row_stride = 0;
  op.gchroma = pg.chroma_var;
  op.gsize = ig.size_var;
  // binary_data: the mat4x2's columns (x_k, y_k) (shaders/palette.rs:57-66).
  const float m[8] = {x[0], y[0], x[1], y[1], x[2], y[2], x[3], y[3]};
  op.blob.resize(32);
  std::memcpy(op.blob.data(), m, 32);
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_command_resize(zs_command* cmd, uint32_t below, uint32_t width, uint32_t height, uint32_t* reg) {
  if (!cmd || !reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  if (width == 0 || height == 0) return fail(ZH_ERR_OTHER, "empty resize target");
  // Descriptor::with_texel(Texel::new_u8(RgbA)): RGBA8, Scalars, linear.
  zh_descriptor grid;
  std::memset(&grid, 0, sizeof(grid));
  grid.width = width;
  grid.height = height;
  grid.texel.bits = ZH_BITS_INT8X4;
  grid.texel.parts = ZH_PARTS_RGBA;
  grid.color.model = ZH_COLOR_SCALARS;
  grid.color.transfer = ZH_TRANSFER_LINEAR;
  // u_max = (1, 0, 0, 1), v_max = (0, 1, 0, 1), every alpha 1.
  float p[24] = {0};
  for (int k = 0; k < 6; ++k) p[4 * k + 3] = 1.0f;
  p[4 * 1 + 0] = 1.0f;
  p[4 * 3 + 1] = 1.0f;
  uint32_t g;
  zh_status st = zs_command_bilinear(cmd, &grid, p, &g);
  if (st != ZH_OK) return st;
  const zs_palette cfg = {ZS_CHANNEL_R, ZS_CHANNEL_G, 0, 0};
  return zs_command_palette(cmd, below, &cfg, g, reg);
}

// Texel::channel_texel (image-canvas): the single-channel texel holding one
// colour channel of `t`: 8- or 16-bit per channel, parts R (1), G (2), B (3).
static bool channel_texel(const zh_texel& t, int32_t channel, zh_texel* out) {
  if (channel < ZS_CHANNEL_R || channel > ZS_CHANNEL_B) return false;
  uint32_t bits;
  if (t.bits == ZH_BITS_INT8X4 || t.bits == 9 /* UInt8x3 */) bits = ZH_BITS_INT8;
  else if (t.bits == ZH_BITS_INT16X4 || t.bits == 12 /* UInt16x3 */) bits = ZH_BITS_INT16;
  else return false;
  // Colour-ordered parts (Rgb, Bgr, Rgba, Bgra and their x/A variants).
  if (t.parts < 6 || t.parts > 15) return false;
  out->bits = bits;
  out->parts = 1u + (uint32_t)channel;
  return true;
}

zh_status zs_command_extract(zs_command* cmd, uint32_t src, int32_t channel, uint32_t* reg) {
  if (!reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  zh_descriptor d;
  zh_status st = reg_desc(cmd, src, &d);
  if (st != ZH_OK) return st;
  zh_texel t;
  if (!channel_texel(d.texel, channel, &t)) return fail(ZH_ERR_OTHER, "no single-channel texel for that channel");
  CmdOp op;
  op.kind = K_EXTRACT;
  op.src = src;
  op.desc = d;
  op.desc.texel = t;
  op.desc.row_stride = 0;
  op.desc = normalized(op.desc);
  // The channel's position in the decoded colour (ChannelPosition).
  op.blob.assign(16, 0);
  const uint32_t pos = (uint32_t)channel;
  std::memcpy(op.blob.data(), &pos, 4);
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_command_inject(zs_command* cmd, uint32_t below, int32_t channel, uint32_t above, uint32_t* reg) {
  if (!reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  zh_descriptor b, a;
  zh_status st = reg_desc(cmd, below, &b);
  if (st == ZH_OK) st = reg_desc(cmd, above, &a);
  if (st != ZH_OK) return st;
  zh_texel want;
  if (!channel_texel(b.texel, channel, &want)) return fail(ZH_ERR_OTHER, "no single-channel texel for that channel");
  // One component above (A, R, G, B or Luma), with the expected bits and the
  // same colour; its own parts say where its value sits (channel_weight_vec4).
  float color[4] = {0, 0, 0, 0};
  switch (a.texel.parts) {
    case 0: color[3] = 1.0f; break;                                     // A
    case 1: case 4: color[0] = 1.0f; break;                             // R, Luma
    case 2: color[1] = 1.0f; break;                                     // G
    case 3: color[2] = 1.0f; break;                                     // B
    default: return fail(ZH_ERR_TYPE, "inject from a texel with more than one channel");
  }
  zh_descriptor cmp = a;
  cmp.texel.parts = want.parts;
  if (cmp.texel.bits != want.bits || cmp.texel.parts != want.parts || !chroma_equal(cmp, [&] {
        zh_descriptor e = b;
        e.texel = want;
        return e;
      }()) || a.width != b.width || a.height != b.height)
    return fail(ZH_ERR_TYPE, "inject of a mismatched channel image");
  float select[4] = {0, 0, 0, 0};
  select[channel] = 1.0f;
  CmdOp op;
  op.kind = K_INJECT;
  op.src = below;
  op.above = above;
  op.desc = b;
  op.blob.resize(32);
  std::memcpy(op.blob.data(), select, 16);
  std::memcpy(op.blob.data() + 16, color, 16);
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_fractal_noise_with_octaves(uint32_t num_octaves, zs_fractal_noise* out) {
  if (!out) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  out->num_octaves = num_octaves;
  out->initial_amplitude = (float)(1.0 / (double)num_octaves);
  out->amplitude_damping = 1.0f;
  out->grid_scale[0] = out->grid_scale[1] = 100.0f;
  return ZH_OK;
}

zh_status zs_fractal_noise_set_damping(zs_fractal_noise* noise, float damping) {
  if (!noise) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  // Geometric series: sum_{i<n} d^i = (1 - d^n) / (1 - d); a vanishing
  // 1 - d^n (d = 1) leaves amplitude 1, as the reference does.
  const float total = 1.0f - std::pow(damping, (float)noise->num_octaves);
  noise->initial_amplitude = std::fabs(total) < 1e-7f ? 1.0f : (1.0f - damping) / total;
  noise->amplitude_damping = damping;
  return ZH_OK;
}

zh_status zs_command_distribution_fractal_noise(zs_command* cmd, const zh_descriptor* desc,
                                                const zs_fractal_noise* noise, uint32_t* reg) {
  if (!cmd || !desc || !noise || !reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  std::string why;
  if (!descriptor_valid(*desc, &why)) return fail(ZH_ERR_TYPE, "inconsistent descriptor for fractal noise: " + why);
  CmdOp op;
  op.kind = K_NOISE;
  op.desc = normalized(*desc);
  // binary_data: grid scale, amplitude, damping, octaves; padded to 8 bytes.
  op.blob.assign(24, 0);
  std::memcpy(op.blob.data(), noise->grid_scale, 8);
  std::memcpy(op.blob.data() + 8, &noise->initial_amplitude, 4);
  std::memcpy(op.blob.data() + 12, &noise->amplitude_damping, 4);
  std::memcpy(op.blob.data() + 16, &noise->num_octaves, 4);
  *reg = push(cmd, op);
  return ZH_OK;
}

// A buffer register stands for its bytes; it has no texture of its own, so
// its slot in the texture table holds a one-byte placeholder.
static zh_descriptor buffer_placeholder() {
  zh_descriptor d;
  std::memset(&d, 0, sizeof(d));
  d.width = d.height = 1;
  d.texel.bits = ZH_BITS_INT8;
  d.texel_stride = 1;
  d.color.model = ZH_COLOR_SCALARS;
  return d;
}

zh_status zs_command_buffer_init(zs_command* cmd, const void* data, uint64_t len, uint32_t* reg) {
  if (!cmd || !reg || (len && !data)) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  if (len > 0xffffffffull) return fail(ZH_ERR_UNSUPPORTED, "buffers are limited to 4 GiB");
  CmdOp op;
  op.kind = K_BUFFER;
  op.desc = buffer_placeholder();
  const uint8_t* b = static_cast<const uint8_t*>(data);
  op.blob.assign(b, b + len);
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_command_buffer_zero(zs_command* cmd, uint64_t len, uint32_t* reg) {
  if (!cmd || !reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  if (len > 0xffffffffull) return fail(ZH_ERR_UNSUPPORTED, "buffers are limited to 4 GiB");
  CmdOp op;
  op.kind = K_BUFFER;
  op.desc = buffer_placeholder();
  op.blob.assign((size_t)len, 0);
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_command_buffer_size(zs_command* cmd, uint32_t buffer, uint64_t* len) {
  if (!cmd || !len || buffer >= cmd->ops.size()) return fail(ZH_ERR_OTHER, "bad register");
  if (cmd->ops[buffer].kind != K_BUFFER) return fail(ZH_ERR_TYPE, "register is not a buffer");
  *len = cmd->ops[buffer].blob.size();
  return ZH_OK;
}

zh_status zs_command_from_buffer(zs_command* cmd, uint32_t buffer, const zh_descriptor* desc, uint32_t* reg) {
  if (!cmd || !desc || !reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  if (buffer >= cmd->ops.size()) return fail(ZH_ERR_OTHER, "bad register");
  if (cmd->ops[buffer].kind != K_BUFFER) return fail(ZH_ERR_TYPE, "register is not a buffer");
  std::string why;
  if (!descriptor_valid(*desc, &why)) return fail(ZH_ERR_TYPE, "inconsistent image descriptor: " + why);
  // Descriptor::to_aligned (buffer.rs:121-134): rows padded to 256 bytes; the
  // buffer must hold height such rows (command.rs:946-962).
  const uint64_t pitch = (tight_stride(*desc) + 255) / 256 * 256;
  if (pitch * desc->height > cmd->ops[buffer].blob.size())
    return fail(ZH_ERR_INVALID_ARGUMENT, "buffer of " + std::to_string(cmd->ops[buffer].blob.size()) +
                                             " bytes is smaller than the image's " +
                                             std::to_string(pitch * desc->height));
  CmdOp op;
  op.kind = K_FROM_BUFFER;
  op.desc = normalized(*desc);
  op.desc.row_stride = 0;
  op.src = buffer;
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_command_invoke(zs_command* cmd, const zs_command* body, const uint32_t* args, uint32_t num_args,
                            uint32_t* results, uint32_t num_results) {
  if (!cmd || !body || (num_args && !args) || (num_results && !results) || cmd == body)
    return fail(ZH_ERR_INVALID_ARGUMENT, "null or self-referencing argument");
  uint32_t inputs = 0, outputs = 0;
  for (const CmdOp& op : body->ops) {
    inputs += op.kind == K_INPUT;
    outputs += op.kind == K_OUTPUT;
  }
  if (inputs != num_args) return fail(ZH_ERR_TYPE, "function takes " + std::to_string(inputs) + " arguments");
  if (outputs != num_results) return fail(ZH_ERR_TYPE, "function returns " + std::to_string(outputs) + " results");
  if (body->vars || !body->symbols.empty())
    return fail(ZH_ERR_TYPE, "a generic function or one that calls others is linked with zs_link");
  // Check every argument before touching the caller.
  uint32_t k = 0;
  for (const CmdOp& op : body->ops) {
    if (op.kind != K_INPUT) continue;
    // The argument's full descriptor (size, texel, colour) must be the
    // parameter's (command.rs:895-902).
    GenDesc have, want;
    zh_status st = reg_gen(cmd, args[k], &have);
    if (st != ZH_OK) return st;
    want.desc = op.desc;
    if (!gen_equal(want, have))
      return fail(ZH_ERR_TYPE, "argument " + std::to_string(k) + " does not match the function's input");
    ++k;
  }
  // Copy the body with its registers renumbered into the caller's.
  std::vector<uint32_t> remap(body->ops.size(), 0);
  uint32_t a = 0, r = 0;
  for (size_t i = 0; i < body->ops.size(); ++i) {
    const CmdOp& op = body->ops[i];
    if (op.kind == K_INPUT) {
      remap[i] = args[a++];
      continue;
    }
    if (op.kind == K_OUTPUT) {
      results[r++] = remap[op.src];
      continue;
    }
    CmdOp c = op;
    c.src = remap[op.src];
    c.above = remap[op.above];
    c.knob = -1;
    remap[i] = push(cmd, c);
  }
  return ZH_OK;
}

// construct/unary/binary_dynamic (command.rs:2933-3048): the operands must
// describe as textures; the shader's data() gives the result's descriptor.
static zh_status dynamic_op(zs_command* cmd, uint32_t nargs, uint32_t lhs, uint32_t rhs,
                            const zs_shader_command* dyn, uint32_t* reg) {
  if (!cmd || !dyn || !reg || !dyn->shader || (dyn->data_len && !dyn->data))
    return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  const uint32_t regs[2] = {lhs, rhs};
  for (uint32_t k = 0; k < nargs; ++k) {
    GenDesc g;
    zh_status st = reg_gen(cmd, regs[k], &g);
    if (st != ZH_OK) return st;
    if (g.size_var >= 0 || g.chroma_var >= 0)
      return fail(ZH_ERR_TYPE, "a dynamic shader's operand needs a concrete descriptor");
  }
  std::string why;
  if (!descriptor_valid(dyn->desc, &why)) return fail(ZH_ERR_TYPE, "dynamic shader result: " + why);
  CmdOp op;
  op.kind = K_DYNAMIC;
  op.src = lhs;
  op.above = rhs;
  op.desc = normalized(dyn->desc);
  op.dyn_shader = zh::ShaderRef(dyn->shader);
  zh_dynamic_call call;
  std::memset(&call, 0, sizeof(call));
  call.shader = dyn->shader;
  call.num_args = nargs;
  call.data_len = dyn->data_len;
  op.blob.resize(sizeof(call) + dyn->data_len);
  std::memcpy(op.blob.data(), &call, sizeof(call));
  if (dyn->data_len) std::memcpy(op.blob.data() + sizeof(call), dyn->data, dyn->data_len);
  *reg = push(cmd, std::move(op));
  return ZH_OK;
}

zh_status zs_command_construct_dynamic(zs_command* cmd, const zs_shader_command* dynamic, uint32_t* reg) {
  return dynamic_op(cmd, 0, 0, 0, dynamic, reg);
}

zh_status zs_command_unary_dynamic(zs_command* cmd, uint32_t src, const zs_shader_command* dynamic, uint32_t* reg) {
  return dynamic_op(cmd, 1, src, 0, dynamic, reg);
}

zh_status zs_command_binary_dynamic(zs_command* cmd, uint32_t lhs, uint32_t rhs, const zs_shader_command* dynamic,
                                    uint32_t* reg) {
  return dynamic_op(cmd, 2, lhs, rhs, dynamic, reg);
}

zh_status zs_command_knob(zs_command* cmd, uint32_t reg, int32_t* knob) {
  if (!cmd || !knob || reg >= cmd->ops.size()) return fail(ZH_ERR_OTHER, "bad register");
  CmdOp& op = cmd->ops[reg];
  if (op.kind == K_INPUT || op.kind == K_OUTPUT || op.kind == K_TRANSMUTE || op.kind == K_CROP ||
      op.kind == K_INSCRIBE || op.kind == K_BLEND || op.kind == K_FROM_BUFFER || op.kind == K_RESULT ||
      op.kind == K_INVOKE || op.kind == K_DYNAMIC)
    return fail(ZH_ERR_TYPE, "operation has no parameter data to override");
  if (op.knob < 0) op.knob = cmd->next_knob++;
  *knob = op.knob;
  return ZH_OK;
}

zh_status zs_command_describe(zs_command* cmd, uint32_t reg, zh_descriptor* desc) {
  if (!desc) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  return reg_desc(cmd, reg, desc);
}

}  // extern "C"

// The High list of a flat, concrete command buffer (Linker::link_in,
// command.rs:2345-2880): one texture per register.
static zh_status compile_flat(const zs_command* cmd, zs_linked** out) {
  zs_linked* L = new zs_linked();
  const size_t n = cmd->ops.size();
  L->textures.resize(n);
  L->input_slot.assign(n, -1);
  L->output_slot.assign(n, -1);
  uint32_t inputs = 0, outputs = 0;
  auto draw = [&](uint32_t dst, uint32_t target, uint32_t init, uint32_t shader, std::vector<uint32_t> operands,
                  std::vector<uint8_t> blob, int32_t knob) -> zh_op& {
    zh_op op;
    std::memset(&op, 0, sizeof(op));
    op.high = ZH_HIGH_DRAW;
    op.dst = dst;
    op.target = target;
    op.initializer = init;
    op.shader = shader;
    op.num_operands = (uint32_t)operands.size();
    for (size_t k = 0; k < operands.size() && k < 2; ++k) op.operands[k] = operands[k];
    op.knob = knob;
    L->ops.push_back(op);
    L->blobs.push_back(std::move(blob));
    return L->ops.back();
  };
  // A full copy of `src` into `dst` (the Discard draw that opens inscribe,
  // affine and blend, command.rs:2646-2661).
  auto copy_below = [&](uint32_t dst, uint32_t src) {
    const zh_descriptor& d = cmd->ops[src].desc;
    zh_op& op = draw(dst, ZH_TARGET_DISCARD, ZH_INIT_PAINT_TO_SELECTION, ZH_SHADER_PAINT_COPY, {src}, {}, -1);
    op.selection = with_size(d.width, d.height);
    quad_of_rect(op.selection, op.quad);
    op.viewport = op.selection;
  };
  for (uint32_t r = 0; r < n; ++r) {
    const CmdOp& c = cmd->ops[r];
    L->textures[r] = c.desc;
    switch (c.kind) {
      case K_INPUT: {
        zh_op op;
        std::memset(&op, 0, sizeof(op));
        op.high = ZH_HIGH_INPUT;
        op.dst = r;
        op.src = inputs;
        op.knob = -1;
        L->input_slot[r] = inputs++;
        L->ops.push_back(op);
        L->blobs.emplace_back();
        break;
      }
      case K_OUTPUT: {
        zh_op op;
        std::memset(&op, 0, sizeof(op));
        op.high = ZH_HIGH_OUTPUT;
        op.src = c.src;
        op.dst = outputs;
        op.knob = -1;
        L->output_slot[r] = outputs++;
        L->ops.push_back(op);
        L->blobs.emplace_back();
        break;
      }
      case K_CROP: {
        // command.rs:2507-2526: selection = region, target = viewport = the
        // region's size at the origin.
        const zh_rect t = with_size(rw(c.rect), rh(c.rect));
        zh_op& op = draw(r, ZH_TARGET_DISCARD, ZH_INIT_PAINT_TO_SELECTION, ZH_SHADER_PAINT_COPY, {c.src}, {}, c.knob);
        op.selection = c.rect;
        quad_of_rect(t, op.quad);
        op.viewport = t;
        break;
      }
      case K_CONVERT:
        draw(r, ZH_TARGET_DISCARD, ZH_INIT_PAINT_FULLSCREEN, c.shader, {c.src}, c.blob, c.knob);
        break;
      case K_ADJUST:
        draw(r, ZH_TARGET_DISCARD, ZH_INIT_PAINT_FULLSCREEN, ZH_SHADER_OKLAB_ADJUST, {c.src}, c.blob, c.knob);
        break;
      case K_SOLID:
        draw(r, ZH_TARGET_DISCARD, ZH_INIT_PAINT_FULLSCREEN, ZH_SHADER_SOLID_RGB, {}, c.blob, c.knob);
        break;
      case K_BILINEAR:
        draw(r, ZH_TARGET_DISCARD, ZH_INIT_PAINT_FULLSCREEN, ZH_SHADER_BILINEAR, {}, c.blob, c.knob);
        break;
      case K_DERIVATIVE:
        draw(r, ZH_TARGET_DISCARD, ZH_INIT_PAINT_FULLSCREEN, ZH_SHADER_BOX3, {c.src}, c.blob, c.knob);
        break;
      case K_NORMAL2D:
        draw(r, ZH_TARGET_DISCARD, ZH_INIT_PAINT_FULLSCREEN, ZH_SHADER_DISTRIBUTION_NORMAL2D, {}, c.blob, c.knob);
        break;
      case K_PALETTE:
        draw(r, ZH_TARGET_DISCARD, ZH_INIT_PAINT_FULLSCREEN, ZH_SHADER_PALETTE, {c.src, c.above}, c.blob, c.knob);
        break;
      case K_EXTRACT:
        draw(r, ZH_TARGET_DISCARD, ZH_INIT_PAINT_FULLSCREEN, ZH_SHADER_EXTRACT, {c.src}, c.blob, c.knob);
        break;
      case K_INJECT:
        draw(r, ZH_TARGET_DISCARD, ZH_INIT_PAINT_FULLSCREEN, ZH_SHADER_INJECT, {c.src, c.above}, c.blob, c.knob);
        break;
      case K_NOISE:
        draw(r, ZH_TARGET_DISCARD, ZH_INIT_PAINT_FULLSCREEN, ZH_SHADER_FRACTAL_NOISE, {}, c.blob, c.knob);
        break;
      case K_DYNAMIC: {
        // command.rs:2767-2830: the operands pushed, one full-screen draw.
        zh_dynamic_call call;
        std::memcpy(&call, c.blob.data(), sizeof(call));
        std::vector<uint32_t> operands;
        if (call.num_args > 0) operands.push_back(c.src);
        if (call.num_args > 1) operands.push_back(c.above);
        draw(r, ZH_TARGET_DISCARD, ZH_INIT_PAINT_FULLSCREEN, ZH_SHADER_DYNAMIC, operands, c.blob, -1);
        L->shaders.push_back(c.dyn_shader);
        break;
      }
      case K_BUFFER:
        break;  // read through the images constructed from it
      case K_FROM_BUFFER: {
        // The buffer's bytes as a texture of 256-byte rows, then a full copy
        // into the register's (tight) texture.
        const CmdOp& b = cmd->ops[c.src];
        zh_descriptor pitched = c.desc;
        pitched.row_stride = (tight_stride(c.desc) + 255) / 256 * 256;
        const uint32_t t = (uint32_t)L->textures.size();
        L->textures.push_back(pitched);
        zh_op op;
        std::memset(&op, 0, sizeof(op));
        op.high = ZH_HIGH_BUFFER;
        op.dst = t;
        op.knob = b.knob;
        L->ops.push_back(op);
        L->blobs.push_back(b.blob);
        draw(r, ZH_TARGET_DISCARD, ZH_INIT_PAINT_FULLSCREEN, ZH_SHADER_PAINT_COPY, {t}, {}, -1);
        break;
      }
      case K_TRANSMUTE: {
        zh_op op;
        std::memset(&op, 0, sizeof(op));
        op.high = ZH_HIGH_COPY;
        op.src = c.src;
        op.dst = r;
        op.knob = -1;
        L->ops.push_back(op);
        L->blobs.emplace_back();
        break;
      }
      case K_INSCRIBE: {
        copy_below(r, c.src);
        const zh_descriptor& a = cmd->ops[c.above].desc;
        const zh_descriptor& b = cmd->ops[c.src].desc;
        zh_op& op = draw(r, ZH_TARGET_LOAD, ZH_INIT_PAINT_TO_SELECTION, ZH_SHADER_PAINT_COPY, {c.above}, {}, c.knob);
        op.selection = with_size(a.width, a.height);
        quad_of_rect(c.rect, op.quad);
        op.viewport = with_size(b.width, b.height);
        break;
      }
      case K_BLEND: {
        copy_below(r, c.src);
        const zh_descriptor& a = cmd->ops[c.above].desc;
        const zh_descriptor& b = cmd->ops[c.src].desc;
        zh_op& op = draw(r, ZH_TARGET_LOAD, ZH_INIT_PAINT_TO_SELECTION, ZH_SHADER_BLEND_OVER, {c.above}, {}, c.knob);
        op.selection = with_size(a.width, a.height);
        quad_of_rect(c.rect, op.quad);
        op.viewport = with_size(b.width, b.height);
        break;
      }
      case K_AFFINE: {
        copy_below(r, c.src);
        const zh_descriptor& a = cmd->ops[c.above].desc;
        const zh_descriptor& b = cmd->ops[c.src].desc;
        Mat3 m;
        for (int k = 0; k < 9; ++k) m[k] = c.matrix[k];
        const bool bil = c.sampling == ZS_AFFINE_BILINEAR;
        zh_op& op = draw(r, ZH_TARGET_LOAD, ZH_INIT_PAINT_TO_SELECTION,
                         bil ? ZH_SHADER_AFFINE_BILINEAR : ZH_SHADER_PAINT_COPY, {c.above},
                         bil ? std140(m) : std::vector<uint8_t>(), c.knob);
        op.selection = with_size(a.width, a.height);
        // QuadTarget::affine (program.rs:1897-1906): the unit quad of the
        // above image in its pixels, each corner through multiply_point.
        const double cx[4] = {0.0, (double)a.width, (double)a.width, 0.0};
        const double cy[4] = {0.0, 0.0, (double)a.height, (double)a.height};
        for (int k = 0; k < 4; ++k) {
          const double q = m[6] * cx[k] + m[7] * cy[k] + m[8];
          op.quad[2 * k] = (float)((m[0] * cx[k] + m[1] * cy[k] + m[2]) / q);
          op.quad[2 * k + 1] = (float)((m[3] * cx[k] + m[4] * cy[k] + m[5]) / q);
        }
        op.viewport = with_size(b.width, b.height);
        break;
      }
      case K_RESULT:
      case K_INVOKE:
        delete L;
        return fail(ZH_ERR_OTHER, "internal: call left in a linked program");
    }
  }
  for (size_t i = 0; i < L->ops.size(); ++i) {
    L->ops[i].params = L->blobs[i].empty() ? nullptr : L->blobs[i].data();
    L->ops[i].params_len = (uint32_t)L->blobs[i].size();
  }
  *out = L;
  return ZH_OK;
}

namespace {

// Linker::link's monomorphization (command.rs:2083-2185), folded: every
// invocation is expanded in place with its function's body instantiated at
// the call's generic arguments, so the program handed to the backend has no
// calls and no generic descriptors left, and a call's stages fuse with the
// caller's.  Knobs of a function body get fresh knob ids per instantiation;
// the last one of each (buffer, register) is what zs_linked_knob reports, as
// the reference's map keeps the last insertion.
struct Expander {
  std::vector<const zs_command*> bufs;          // 0: the entry buffer, then the functions
  std::vector<std::vector<uint32_t>> links;     // per buffer: function symbol -> buffer index
  zs_command out;                               // the flat program
  std::vector<std::array<int64_t, 3>> knobs;    // (buffer, register, knob)
  std::vector<int64_t> entry_remap;             // entry register -> register of `out`
  int depth = 0;

  static Signature signature_of(const zs_command* c) {
    Signature s;
    s.vars = c->vars;
    for (const CmdOp& op : c->ops) {
      GenDesc g;
      g.desc = op.desc;
      g.size_var = op.gsize;
      g.chroma_var = op.gchroma;
      if (op.kind == K_INPUT) s.inputs.push_back(g);
      if (op.kind == K_OUTPUT) s.outputs.push_back(g);
    }
    return s;
  }

  zh_status expand(uint32_t b, const std::vector<GenDesc>& tys, const std::vector<uint32_t>* args,
                   std::vector<uint32_t>* results) {
    if (++depth > 64) return fail(ZH_ERR_OTHER, "function calls nest deeper than 64 (recursive links?)");
    const zs_command* c = bufs[b];
    if (tys.size() != c->vars)
      return fail(ZH_ERR_OTHER, "buffer " + std::to_string(b) + " takes " + std::to_string(c->vars) +
                                    " generic arguments, " + std::to_string(tys.size()) + " given");
    std::vector<uint32_t> remap(c->ops.size(), 0);
    uint32_t a = 0;
    for (uint32_t i = 0; i < c->ops.size(); ++i) {
      const CmdOp& op = c->ops[i];
      GenDesc g;
      g.desc = op.desc;
      g.size_var = op.gsize;
      g.chroma_var = op.gchroma;
      const GenDesc cg = rewrite(g, tys);
      const bool concrete = cg.size_var < 0 && cg.chroma_var < 0;
      if (op.kind == K_RESULT) continue;  // set by the Invoke that follows
      if (op.kind == K_INVOKE) {
        if (b >= links.size() || op.fn >= links[b].size())
          return fail(ZH_ERR_OTHER, "function symbol " + std::to_string(op.fn) + " of buffer " + std::to_string(b) +
                                        " is not linked");
        const uint32_t callee = links[b][op.fn];
        if (callee == 0 || callee >= bufs.size()) return fail(ZH_ERR_OTHER, "link target out of range");
        // The definition must have the declared signature (command.rs:2076-2081).
        const Signature want = c->symbols[op.fn], have = signature_of(bufs[callee]);
        bool same = want.vars == have.vars && want.inputs.size() == have.inputs.size() &&
                    want.outputs.size() == have.outputs.size();
        for (size_t k = 0; same && k < want.inputs.size(); ++k) same = gen_equal(want.inputs[k], have.inputs[k]);
        for (size_t k = 0; same && k < want.outputs.size(); ++k) same = gen_equal(want.outputs[k], have.outputs[k]);
        if (!same) return fail(ZH_ERR_OTHER, "linked function does not have the declared signature");
        std::vector<GenDesc> gens;
        for (const GenDesc& gg : op.generics) {
          GenDesc r = rewrite(gg, tys);
          if (r.size_var >= 0 || r.chroma_var >= 0) return fail(ZH_ERR_OTHER, "generic argument left unbound");
          gens.push_back(r);
        }
        std::vector<uint32_t> in, res;
        for (uint32_t r : op.args) in.push_back(remap[r]);
        zh_status st = expand(callee, gens, &in, &res);
        if (st != ZH_OK) return st;
        for (size_t k = 0; k < op.results.size() && k < res.size(); ++k) remap[op.results[k]] = res[k];
        continue;
      }
      if (!concrete) return fail(ZH_ERR_OTHER, "generic descriptor left unbound in buffer " + std::to_string(b));
      if (op.kind == K_INPUT && args) {
        remap[i] = (*args)[a++];
        continue;
      }
      if (op.kind == K_OUTPUT && args) {
        results->push_back(remap[op.src]);
        continue;
      }
      CmdOp n = op;
      n.src = op.src < remap.size() ? remap[op.src] : 0;
      n.above = op.above < remap.size() ? remap[op.above] : 0;
      n.desc = cg.desc;
      n.desc.texel_stride = bits_bytes(n.desc.texel.bits);
      n.gsize = n.gchroma = -1;
      n.args.clear();
      n.results.clear();
      n.generics.clear();
      if (op.knob >= 0) {
        // The entry buffer keeps its knob ids; function bodies get fresh ones.
        n.knob = b == 0 ? op.knob : out.next_knob++;
        knobs.push_back({(int64_t)b, (int64_t)i, (int64_t)n.knob});
      }
      remap[i] = push(&out, n);
    }
    if (b == 0) entry_remap.assign(remap.begin(), remap.end());
    --depth;
    return ZH_OK;
  }
};

}  // namespace

static zh_status link_impl(const zs_command* main, const zh_descriptor* tys, uint32_t num_tys,
                           const zs_command* const* functions, uint32_t num_functions, const uint32_t* link_counts,
                           const uint32_t* link_targets, zs_linked** out) {
  if (!main || !out || (num_tys && !tys) || (num_functions && !functions)) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  Expander E;
  E.bufs.push_back(main);
  for (uint32_t k = 0; k < num_functions; ++k) {
    if (!functions[k]) return fail(ZH_ERR_INVALID_ARGUMENT, "null function buffer");
    E.bufs.push_back(functions[k]);
  }
  E.links.resize(E.bufs.size());
  if (link_counts) {
    size_t at = 0;
    for (size_t bi = 0; bi < E.bufs.size(); ++bi) {
      for (uint32_t k = 0; k < link_counts[bi]; ++k) E.links[bi].push_back(link_targets[at++]);
      if (!E.links[bi].empty() && E.links[bi].size() != E.bufs[bi]->symbols.size())
        return fail(ZH_ERR_OTHER, "buffer " + std::to_string(bi) + " declares " +
                                      std::to_string(E.bufs[bi]->symbols.size()) + " functions, links name " +
                                      std::to_string(E.links[bi].size()));
    }
  }
  // Knob ids of function bodies start above the entry buffer's.
  E.out.next_knob = main->next_knob;
  std::vector<GenDesc> entry_tys;
  for (uint32_t k = 0; k < num_tys; ++k) {
    GenDesc g;
    g.desc = normalized(tys[k]);
    entry_tys.push_back(g);
  }
  zh_status st = E.expand(0, entry_tys, nullptr, nullptr);
  if (st != ZH_OK) return st;
  zs_linked* flat = nullptr;
  st = compile_flat(&E.out, &flat);
  if (st != ZH_OK) return st;
  // Slots per register of the entry buffer.
  std::vector<int64_t> in(main->ops.size(), -1), outs(main->ops.size(), -1);
  for (size_t r = 0; r < main->ops.size(); ++r) {
    const int64_t t = E.entry_remap[r];
    if (main->ops[r].kind == K_INPUT) in[r] = flat->input_slot[(size_t)t];
    if (main->ops[r].kind == K_OUTPUT) outs[r] = flat->output_slot[(size_t)t];
  }
  flat->input_slot = in;
  flat->output_slot = outs;
  flat->knobs = E.knobs;
  *out = flat;
  return ZH_OK;
}

extern "C" {

zh_status zs_command_compile(zs_command* cmd, zs_linked** out) {
  if (!cmd || !out) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  return link_impl(cmd, nullptr, 0, nullptr, 0, nullptr, nullptr, out);
}

zh_status zs_link(const zs_command* main, const zh_descriptor* tys, uint32_t num_tys, const zs_command* const* functions,
                  uint32_t num_functions, const uint32_t* link_counts, const uint32_t* link_targets, zs_linked** out) {
  return link_impl(main, tys, num_tys, functions, num_functions, link_counts, link_targets, out);
}

zh_status zs_linked_knob(zs_linked* L, uint32_t link_idx, uint32_t reg, int32_t* knob) {
  if (!L || !knob) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  for (size_t k = L->knobs.size(); k-- > 0;)
    if (L->knobs[k][0] == (int64_t)link_idx && L->knobs[k][1] == (int64_t)reg) {
      *knob = (int32_t)L->knobs[k][2];
      return ZH_OK;
    }
  return fail(ZH_ERR_INVALID_ARGUMENT, "no knob on that register");
}

// ---- generics (command.rs:753-923) --------------------------------------------
zh_status zs_command_generic(zs_command* cmd, uint32_t* var) {
  if (!cmd || !var) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  GenDesc g;
  g.size_var = g.chroma_var = (int32_t)cmd->vars++;
  cmd->tys.push_back(g);
  *var = (uint32_t)cmd->tys.size() - 1;
  return ZH_OK;
}

zh_status zs_command_input_generic(zs_command* cmd, uint32_t var, uint32_t* reg) {
  if (!cmd || !reg) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  if (var >= cmd->tys.size()) return fail(ZH_ERR_OTHER, "bad descriptor variable");
  const GenDesc& g = cmd->tys[var];
  CmdOp op;
  op.kind = K_INPUT;
  op.desc = g.desc;
  op.gsize = g.size_var;
  op.gchroma = g.chroma_var;
  *reg = push(cmd, op);
  return ZH_OK;
}

zh_status zs_command_register_descriptor(zs_command* cmd, uint32_t reg, uint32_t* var) {
  if (!var) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  GenDesc g;
  zh_status st = reg_gen(cmd, reg, &g);
  if (st != ZH_OK) return st;
  g.desc.row_stride = 0;
  cmd->tys.push_back(g);
  *var = (uint32_t)cmd->tys.size() - 1;
  return ZH_OK;
}

zh_status zs_command_derive_descriptor(zs_command* cmd, uint32_t base, const uint32_t* size, uint32_t* var) {
  if (!cmd || !var) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  if (base >= cmd->tys.size()) return fail(ZH_ERR_OTHER, "bad descriptor variable");
  GenDesc g = cmd->tys[base];
  if (size) {
    if (size[0] == 0 || size[1] == 0) return fail(ZH_ERR_TYPE, "empty size");
    g.size_var = -1;
    g.desc.width = size[0];
    g.desc.height = size[1];
  }
  cmd->tys.push_back(g);
  *var = (uint32_t)cmd->tys.size() - 1;
  return ZH_OK;
}

zh_status zs_command_signature(const zs_command* body, zs_signature** out) {
  if (!body || !out) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  zs_signature* s = new zs_signature();
  s->sig = Expander::signature_of(body);
  *out = s;
  return ZH_OK;
}

zh_status zs_signature_counts(const zs_signature* sig, uint32_t* vars, uint32_t* inputs, uint32_t* outputs) {
  if (!sig) return fail(ZH_ERR_INVALID_ARGUMENT, "null signature");
  if (vars) *vars = sig->sig.vars;
  if (inputs) *inputs = (uint32_t)sig->sig.inputs.size();
  if (outputs) *outputs = (uint32_t)sig->sig.outputs.size();
  return ZH_OK;
}

zh_status zs_signature_destroy(zs_signature* sig) {
  delete sig;
  return ZH_OK;
}

zh_status zs_command_function(zs_command* cmd, const zs_signature* sig, uint32_t* fn) {
  if (!cmd || !sig || !fn) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  cmd->symbols.push_back(sig->sig);
  *fn = (uint32_t)cmd->symbols.size() - 1;
  return ZH_OK;
}

zh_status zs_command_invoke_function(zs_command* cmd, uint32_t fn, const uint32_t* generics, uint32_t num_generics,
                                     const uint32_t* args, uint32_t num_args, uint32_t* results,
                                     uint32_t num_results) {
  if (!cmd || (num_generics && !generics) || (num_args && !args) || (num_results && !results))
    return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  if (fn >= cmd->symbols.size()) return fail(ZH_ERR_OTHER, "bad function variable");
  const Signature sig = cmd->symbols[fn];
  // CommandError::INVALID_CALL (command.rs:866-900) for arity and types.
  if (sig.inputs.size() != num_args) return fail(ZH_ERR_OTHER, "invalid call: wrong number of arguments");
  if (sig.vars != num_generics) return fail(ZH_ERR_OTHER, "invalid call: wrong number of generic arguments");
  if (sig.outputs.size() != num_results) return fail(ZH_ERR_OTHER, "invalid call: wrong number of results");
  std::vector<GenDesc> gens;
  for (uint32_t k = 0; k < num_generics; ++k) {
    if (generics[k] >= cmd->tys.size()) return fail(ZH_ERR_OTHER, "bad descriptor variable");
    gens.push_back(cmd->tys[generics[k]]);
  }
  for (uint32_t k = 0; k < num_args; ++k) {
    GenDesc have;
    zh_status st = reg_gen(cmd, args[k], &have);
    if (st != ZH_OK) return st;
    if (!gen_equal(rewrite(sig.inputs[k], gens), have))
      return fail(ZH_ERR_OTHER, "invalid call: argument " + std::to_string(k) + " does not match the parameter");
  }
  // Result registers first, then the invocation (command.rs:907-923).
  CmdOp inv;
  inv.kind = K_INVOKE;
  inv.fn = fn;
  inv.args.assign(args, args + num_args);
  inv.generics = gens;
  for (uint32_t k = 0; k < num_results; ++k) {
    const GenDesc r = rewrite(sig.outputs[k], gens);
    CmdOp res;
    res.kind = K_RESULT;
    res.desc = r.desc;
    res.gsize = r.size_var;
    res.gchroma = r.chroma_var;
    results[k] = push(cmd, res);
    inv.results.push_back(results[k]);
  }
  push(cmd, inv);
  return ZH_OK;
}

zh_status zs_linked_destroy(zs_linked* prog) {
  delete prog;
  return ZH_OK;
}

zh_status zs_linked_high(zs_linked* L, const zh_descriptor** textures, uint32_t* num_textures, const zh_op** ops,
                         uint32_t* num_ops) {
  if (!L) return fail(ZH_ERR_INVALID_ARGUMENT, "null program");
  if (textures) *textures = L->textures.data();
  if (num_textures) *num_textures = (uint32_t)L->textures.size();
  if (ops) *ops = L->ops.data();
  if (num_ops) *num_ops = (uint32_t)L->ops.size();
  return ZH_OK;
}

zh_status zs_linked_input_slot(zs_linked* L, uint32_t reg, uint32_t* slot) {
  if (!L || !slot || reg >= L->input_slot.size() || L->input_slot[reg] < 0)
    return fail(ZH_ERR_INVALID_ARGUMENT, "register is not an input");
  *slot = (uint32_t)L->input_slot[reg];
  return ZH_OK;
}

zh_status zs_linked_output_slot(zs_linked* L, uint32_t reg, uint32_t* slot) {
  if (!L || !slot || reg >= L->output_slot.size() || L->output_slot[reg] < 0)
    return fail(ZH_ERR_INVALID_ARGUMENT, "register is not an output");
  *slot = (uint32_t)L->output_slot[reg];
  return ZH_OK;
}

zh_status zs_linked_lower(zs_linked* L, zh_context* ctx, uint32_t flags, zh_program** out) {
  if (!L) return fail(ZH_ERR_INVALID_ARGUMENT, "null program");
  return zh_program_build(ctx, L->textures.data(), (uint32_t)L->textures.size(), L->ops.data(), (uint32_t)L->ops.size(),
                          flags, out);
}

}  // extern "C"
