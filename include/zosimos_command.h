/* zosimos_command.h -- C ABI of the C++ mirror of zosimos's front end
 * (CommandBuffer, Linker::compile, Program::lower_to).
 *
 * The reference builds programs in Rust (lib/zosimos/src/command.rs) and
 * hands the linked High IR to the backend.  No Rust toolchain exists in this
 * build's image, so the front end is mirrored here in C++ with the same
 * operation names, argument meaning and error classes; it emits exactly the
 * zh_op list of include/zosimos_hip.h that a Rust adapter would.  Register
 * handles are plain uint32_t indices (command.rs:31 `Register(usize)`).
 */
#ifndef ZOSIMOS_COMMAND_H
#define ZOSIMOS_COMMAND_H

#include "zosimos_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct zs_command zs_command; /* CommandBuffer, command.rs:60-68 */
typedef struct zs_linked zs_linked;   /* linked Program, program.rs:24-39 */
typedef struct zs_signature zs_signature; /* CommandSignature, command.rs:819-849 */

/* AffineSample, command.rs:313-327 */
#define ZS_AFFINE_NEAREST 0
#define ZS_AFFINE_BILINEAR 1
/* Blend, command.rs:287-291 */
#define ZS_BLEND_ALPHA 0
/* DerivativeMethod and Direction, command.rs:547-566 (declaration order). */
#define ZS_DERIVATIVE_ROBERTS 0
#define ZS_DERIVATIVE_PREWITT 1
#define ZS_DERIVATIVE_SOBEL 2
#define ZS_DERIVATIVE_SCHARR3 3
#define ZS_DERIVATIVE_SCHARR4 4
#define ZS_DERIVATIVE_SCHARR5 5
#define ZS_DERIVATIVE_SCHARR3_TO_4BIT 6
#define ZS_DERIVATIVE_SCHARR5_TAB 7
#define ZS_DERIVATIVE_SCHARR3_TO_8BIT 8
#define ZS_DIRECTION_HEIGHT 0
#define ZS_DIRECTION_WIDTH 1
/* ChromaticAdaptationMethod, command.rs:422-460 (declaration order). */
#define ZS_ADAPT_XYZ 0
#define ZS_ADAPT_VON_KRIES 1
#define ZS_ADAPT_BRADFORD_VON_KRIES 2
#define ZS_ADAPT_BRADFORD_NONLINEAR 3 /* not implemented, as in the reference (command.rs:3328-3332) */

zh_status zs_command_create(zs_command** out);
zh_status zs_command_destroy(zs_command* cmd);

/* CommandBuffer::input, command.rs:743-751 */
zh_status zs_command_input(zs_command* cmd, const zh_descriptor* desc, uint32_t* reg);
/* CommandBuffer::output, command.rs:1707-1712; `outformat` may be NULL */
zh_status zs_command_output(zs_command* cmd, uint32_t src, uint32_t* reg, zh_descriptor* outformat);
/* CommandBuffer::render, command.rs:1719-1736: an output meant for a render
 * target (Environment::bind_render); here an output slot like any other,
 * bound with zh_binding.is_output = 1.  ZH_ERR_TYPE when the register's
 * texel has no texture form in the reference (3- or 6-byte texels). */
zh_status zs_command_render(zs_command* cmd, uint32_t src, uint32_t* reg, zh_descriptor* outformat);
/* CommandBuffer::crop, command.rs:971-978 (result keeps the source size) */
zh_status zs_command_crop(zs_command* cmd, uint32_t src, zh_rect rect, uint32_t* reg);
/* CommandBuffer::color_convert, command.rs:986-1107 */
zh_status zs_command_color_convert(zs_command* cmd, uint32_t src, const zh_color* color, const zh_texel* texel,
                                   uint32_t* reg);
/* CommandBuffer::inscribe, command.rs:1177-1218 */
zh_status zs_command_inscribe(zs_command* cmd, uint32_t below, zh_rect rect, uint32_t above, uint32_t* reg);
/* CommandBuffer::transmute, command.rs:1276-1350 */
zh_status zs_command_transmute(zs_command* cmd, uint32_t src, const zh_descriptor* target, uint32_t* reg);
/* CommandBuffer::affine, command.rs:1636-1673.  `matrix` is Affine::transformation
 * (row-major, command.rs:287-296).  BiLinear (UNIMPLEMENTED in the reference)
 * samples premultiplied texels bilinearly with a transparent border and
 * composites them 'over' `below`. */
zh_status zs_command_affine(zs_command* cmd, uint32_t below, const float matrix[9], uint32_t sampling,
                            uint32_t above, uint32_t* reg);
/* CommandBuffer::blend, command.rs:1510-1519 (UNIMPLEMENTED in the reference):
 * Porter-Duff 'over' of `above` placed at `rect` onto `below`. */
zh_status zs_command_blend(zs_command* cmd, uint32_t below, zh_rect rect, uint32_t above, uint32_t blend,
                           uint32_t* reg);
/* CommandBuffer::solid_rgba, command.rs:1524-1551 */
zh_status zs_command_solid_rgba(zs_command* cmd, const zh_descriptor* desc, const float color[4], uint32_t* reg);
/* CommandBuffer::bilinear, command.rs:1615-1633: a generated image, per pixel
 * mix(u_min, u_max, u) + mix(v_min, v_max, v) + mix(uv_min, uv_max, u v) with
 * (u, v) the pixel centre over the image size (bilinear.frag).  params: the
 * six vec4 u_min, u_max, v_min, v_max, uv_min, uv_max (shaders/bilinear.rs). */
zh_status zs_command_bilinear(zs_command* cmd, const zh_descriptor* desc, const float params[24], uint32_t* reg);
/* Oklab chroma scale and hue rotation (radians) on an Oklab register; no
 * counterpart in the reference, where such an adjustment needs a user shader
 * (ShaderCommand, command/dynamic.rs:7-28). */
zh_status zs_command_oklab_adjust(zs_command* cmd, uint32_t src, float chroma, float hue_rad, uint32_t* reg);
/* CommandBuffer::chromatic_adaptation, command.rs:1112-1174: an RGB register
 * adapted to the target white point (ZH_WHITEPOINT_*) by a ZS_ADAPT_* method;
 * one linear colour matrix (command.rs:2527-2548). */
zh_status zs_command_chromatic_adaptation(zs_command* cmd, uint32_t src, uint32_t method, uint32_t target,
                                          uint32_t* reg);
/* CommandBuffer::derivative, command.rs:1487-1507: a 3x3 derivative filter
 * (ZS_DERIVATIVE_*) along a ZS_DIRECTION_*, lowered to box3.frag
 * (command.rs:2598-2609, 3344-3409).  Roberts and the Scharr 4/5 variants are
 * not implemented, as in the reference. */
zh_status zs_command_derivative(zs_command* cmd, uint32_t src, uint32_t method, uint32_t direction, uint32_t* reg);
/* DistributionNormal2d (shaders/distribution_normal2d.rs ShaderData): the 8
 * floats {expectation (2), inverse covariance (4, row-major),
 * pseudo-determinant, 0} of a centred normal distribution with variances
 * (var0, var1) on the axes (with_diagonal), or of a 1-D one along the
 * direction (x, y), whose squared length is the variance (with_direction). */
zh_status zs_normal2d_with_diagonal(float var0, float var1, float out[8]);
zh_status zs_normal2d_with_direction(float x, float y, float out[8]);
/* CommandBuffer::distribution_normal2d, command.rs:1559-1583: the density at
 * the pixel centres, in screen coordinates [-1, 1]^2, as a grey image;
 * `desc` must have Luma or LumaA parts (ZH_ERR_TYPE otherwise). */
zh_status zs_command_distribution_normal2d(zs_command* cmd, const zh_descriptor* desc, const float params[8],
                                           uint32_t* reg);
/* Colour channels that can index a palette (image-canvas ColorChannel; the
 * shader takes R, G or B, shaders/palette.rs:30-38). */
#define ZS_CHANNEL_NONE (-1)
#define ZS_CHANNEL_R 0
#define ZS_CHANNEL_G 1
#define ZS_CHANNEL_B 2
#define ZS_CHANNEL_A 3
/* Palette, command.rs:527-537. */
typedef struct zs_palette {
  int32_t width;   /* ZS_CHANNEL_*: the coordinate along the palette's width */
  int32_t height;  /* ZS_CHANNEL_*: the coordinate along its height */
  int32_t width_base, height_base; /* carried; the shader does not read them */
} zs_palette;
/* CommandBuffer::palette, command.rs:1442-1485: per pixel of `indices`, the
 * `palette` texel at the coordinates its channels select (nearest, clamped
 * to the edge, palette.frag).  The result has the palette's texel and colour
 * and the indices' size.  ZH_ERR_TYPE for a channel other than R, G, B. */
zh_status zs_command_palette(zs_command* cmd, uint32_t palette, const zs_palette* config, uint32_t indices,
                             uint32_t* reg);
/* CommandBuffer::resize, command.rs:1675-1705: `below` sampled onto a
 * width x height grid (a bilinear coordinate ramp in RGBA8 with a Scalars
 * colour, then a palette lookup of R across and G down). */
zh_status zs_command_resize(zs_command* cmd, uint32_t below, uint32_t width, uint32_t height, uint32_t* reg);
/* CommandBuffer::extract, command.rs:1221-1265: one colour channel (R, G or
 * B; ZH_ERR_OTHER otherwise, and for texels without an 8- or 16-bit
 * single-channel form) as a single-channel image of the same colour. */
zh_status zs_command_extract(zs_command* cmd, uint32_t src, int32_t channel, uint32_t* reg);
/* CommandBuffer::inject, command.rs:1360-1440: `below` with its `channel`
 * replaced by the single channel of `above`, which must have the texel and
 * colour `extract` of that channel would give (ZH_ERR_TYPE otherwise). */
zh_status zs_command_inject(zs_command* cmd, uint32_t below, int32_t channel, uint32_t above, uint32_t* reg);
/* FractalNoise (shaders/fractal_noise.rs ShaderData). */
typedef struct zs_fractal_noise {
  uint32_t num_octaves;
  float initial_amplitude;
  float amplitude_damping;
  float grid_scale[2];
} zs_fractal_noise;
/* ShaderData::with_octaves: amplitude 1 / octaves, damping 1, cells of
 * 1/100 of the image. */
zh_status zs_fractal_noise_with_octaves(uint32_t num_octaves, zs_fractal_noise* out);
/* ShaderData::set_damping: the damping, and the initial amplitude that makes
 * the octaves' amplitudes sum to 1. */
zh_status zs_fractal_noise_set_damping(zs_fractal_noise* noise, float damping);
/* CommandBuffer::distribution_fractal_noise, command.rs:1589-1607: fractal
 * value noise (fractal_noise.frag) at the pixel centres, all four channels;
 * at most 4096 octaves (ZH_ERR_UNSUPPORTED beyond). */
zh_status zs_command_distribution_fractal_noise(zs_command* cmd, const zh_descriptor* desc,
                                                const zs_fractal_noise* noise, uint32_t* reg);
/* CommandBuffer::buffer_init, command.rs:1777-1790: a buffer register holding
 * a copy of `len` bytes. */
zh_status zs_command_buffer_init(zs_command* cmd, const void* data, uint64_t len, uint32_t* reg);
/* CommandBuffer::buffer_zero, command.rs:1793-1801: `len` zero bytes. */
zh_status zs_command_buffer_zero(zs_command* cmd, uint64_t len, uint32_t* reg);
/* Byte size of a buffer register (GenericBuffer::size, command.rs:93-96). */
zh_status zs_command_buffer_size(zs_command* cmd, uint32_t buffer, uint64_t* len);
/* CommandBuffer::from_buffer, command.rs:937-968: the image `desc` read from a
 * buffer register whose rows are padded to 256 bytes (Descriptor::to_aligned,
 * buffer.rs:121-134).  ZH_ERR_TYPE unless `buffer` is a buffer;
 * ZH_ERR_INVALID_ARGUMENT (CommandError::INVALID_CALL) when the buffer is
 * shorter than height padded rows.  The register's own descriptor is tight. */
zh_status zs_command_from_buffer(zs_command* cmd, uint32_t buffer, const zh_descriptor* desc, uint32_t* reg);
/* A call of `body` folded into the caller at once (the shortcut of
 * zs_command_function + zs_command_invoke_function + zs_link for a body
 * without generics or calls of its own).  The inputs of `body`, in
 * declaration order, take the registers `args`, whose descriptors (size,
 * texel, colour) must equal them (ZH_ERR_TYPE otherwise); its outputs, in
 * order, come back in `results`.  Knobs inside `body` are not carried over
 * (zs_link carries them); a generic body or one that calls functions is
 * ZH_ERR_TYPE. */
zh_status zs_command_invoke(zs_command* cmd, const zs_command* body, const uint32_t* args, uint32_t num_args,
                            uint32_t* results, uint32_t num_results);
/* User stage kernels (CommandBuffer::{construct,unary,binary}_dynamic,
 * command.rs:2933-3048): what ShaderCommand::source and ShaderCommand::data
 * (command/dynamic.rs:7-28) return -- the shader, its data and the result's
 * descriptor.  The operands must be concrete texture registers (buffers and
 * generic registers are ZH_ERR_TYPE); the result may have any size.  The
 * command buffer holds a reference to the shader.  zh_dynamic_args in
 * zosimos_hip.h is what the kernel receives. */
typedef struct zs_shader_command {
  zh_shader* shader;  /* ShaderSource: a HIP code object and kernel */
  const void* data;   /* ShaderData::set_data bytes (NULL when data_len is 0) */
  uint32_t data_len;
  zh_descriptor desc; /* the descriptor data() returns */
} zs_shader_command;
zh_status zs_command_construct_dynamic(zs_command* cmd, const zs_shader_command* dynamic, uint32_t* reg);
zh_status zs_command_unary_dynamic(zs_command* cmd, uint32_t src, const zs_shader_command* dynamic, uint32_t* reg);
zh_status zs_command_binary_dynamic(zs_command* cmd, uint32_t lhs, uint32_t rhs, const zs_shader_command* dynamic,
                                    uint32_t* reg);
/* Mark the operation that defined `reg` as a knob (CommandBuffer::with_knob,
 * command.rs:1865-1905); `knob` receives its id for zh_program_set_knob.  A
 * buffer register's knob takes the buffer's bytes, of the same length
 * (WithKnob::buffer_init, command.rs:1937-1940).  Dynamic shaders take no
 * knob (WithKnob has no dynamic ops, command.rs:1865-2060): ZH_ERR_TYPE. */
zh_status zs_command_knob(zs_command* cmd, uint32_t reg, int32_t* knob);
/* Descriptor of a register (RegisterDescription::Texture). */
zh_status zs_command_describe(zs_command* cmd, uint32_t reg, zh_descriptor* desc);

/* ---- generics and functions (command.rs:753-923) ---------------------------
 * A generic parameter stands for a whole descriptor (size and chroma);
 * descriptor variables name generic or derived descriptors. */
/* CommandBuffer::generic (command.rs:772-791): a new generic parameter;
 * `var` is the descriptor variable standing for it. */
zh_status zs_command_generic(zs_command* cmd, uint32_t* var);
/* CommandBuffer::input_generic (command.rs:756-762). */
zh_status zs_command_input_generic(zs_command* cmd, uint32_t var, uint32_t* reg);
/* CommandBuffer::register_descriptor (command.rs:811-821): a variable
 * describing a register (generic or not). */
zh_status zs_command_register_descriptor(zs_command* cmd, uint32_t reg, uint32_t* var);
/* CommandBuffer::derive_descriptor (command.rs:794-808): `base` with its
 * size replaced by size[0] x size[1], or unchanged when `size` is NULL. */
zh_status zs_command_derive_descriptor(zs_command* cmd, uint32_t base, const uint32_t* size, uint32_t* var);
/* CommandBuffer::computed_signature (command.rs:824-852). */
zh_status zs_command_signature(const zs_command* body, zs_signature** out);
zh_status zs_signature_destroy(zs_signature* sig);
/* Generic parameters, inputs and outputs of a signature. */
zh_status zs_signature_counts(const zs_signature* sig, uint32_t* vars, uint32_t* inputs, uint32_t* outputs);
/* CommandBuffer::function (command.rs:855-860): declare a function symbol
 * of that signature, defined at link time. */
zh_status zs_command_function(zs_command* cmd, const zs_signature* sig, uint32_t* fn);
/* CommandBuffer::invoke (command.rs:863-923): call symbol `fn` with the
 * descriptor variables `generics` (one per generic parameter) and the
 * registers `args`, each equal to its parameter rewritten with the generics
 * (ZH_ERR_OTHER, CommandError::INVALID_CALL, otherwise); `results` receive
 * the result registers. */
zh_status zs_command_invoke_function(zs_command* cmd, uint32_t fn, const uint32_t* generics, uint32_t num_generics,
                                     const uint32_t* args, uint32_t num_args, uint32_t* results, uint32_t num_results);

/* Linker::compile, command.rs:2069-2071 and link_in's High emission.  A
 * buffer with unlinked function calls fails (ZH_ERR_OTHER). */
zh_status zs_command_compile(zs_command* cmd, zs_linked** out);
/* Linker::link (command.rs:2083-2185): `main` (buffer 0) instantiated with
 * the descriptors `tys` (one per generic parameter of main), and the function
 * buffers `functions` (buffers 1 ..).  For buffer b, link_counts[b] entries
 * of link_targets (consecutive, buffer by buffer) name the buffer defining
 * each of its function symbols in declaration order (counts of 0 for a
 * buffer without symbols; link_counts may be NULL when no buffer calls).
 * Every call is monomorphized at its generic arguments and folded into the
 * caller, so the High list has no calls; the definition must have the
 * declared signature. */
zh_status zs_link(const zs_command* main, const zh_descriptor* tys, uint32_t num_tys, const zs_command* const* functions,
                  uint32_t num_functions, const uint32_t* link_counts, const uint32_t* link_targets, zs_linked** out);
/* Executable::query_knob (run.rs:1012-1014) with RegisterKnob {link_idx,
 * register} (command.rs:2128-2140): the knob id of a knob register of buffer
 * `link_idx` (0 = main) in the linked program (the last instantiation's,
 * for a function called more than once). */
zh_status zs_linked_knob(zs_linked* prog, uint32_t link_idx, uint32_t reg, int32_t* knob);
zh_status zs_linked_destroy(zs_linked* prog);
/* Borrow the High list (valid until zs_linked_destroy). */
zh_status zs_linked_high(zs_linked* prog, const zh_descriptor** textures, uint32_t* num_textures, const zh_op** ops,
                         uint32_t* num_ops);
/* Binding slot of an input / output register. */
zh_status zs_linked_input_slot(zs_linked* prog, uint32_t reg, uint32_t* slot);
zh_status zs_linked_output_slot(zs_linked* prog, uint32_t reg, uint32_t* slot);
/* Program::lower_to (program.rs:1304): plan for `ctx` (NULL = plan only). */
zh_status zs_linked_lower(zs_linked* prog, zh_context* ctx, uint32_t flags, zh_program** out);

#ifdef __cplusplus
}
#endif

#endif /* ZOSIMOS_COMMAND_H */
