/* zosimos_hip.h -- C ABI of the MI355X execution backend for zosimos programs.
 *
 * This is the seam a `zosimos-hip-sys` crate binds.  It replaces, below the
 * High IR, the reference's lowering and wgpu execution:
 *   Program::lower_to                 lib/zosimos/src/program.rs:1304-1423
 *   Encoder (High -> Low for wgpu)    lib/zosimos/src/program/encoder.rs
 *   Executable/Environment/Execution  lib/zosimos/src/run.rs:38-1499
 *   SyncPoint::block_on               lib/zosimos/src/run.rs:3015-3029
 * The op list handed to zh_program_build is the High IR of program.rs:88-240
 * (one zh_op per High instruction), with the std140 parameter blobs that
 * FragmentShaderData::binary_data writes (shaders.rs:63-85).
 *
 * All functions return a zh_status (0 = ZH_OK).  Nothing throws or aborts
 * across the boundary; zh_last_error() returns the message of the last
 * failure on the calling thread.  Contexts are externally synchronised: one
 * host thread at a time per context, like stepping an Execution.
 */
#ifndef ZOSIMOS_HIP_H
#define ZOSIMOS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t zh_status;
#define ZH_OK 0
#define ZH_ERR_INVALID_ARGUMENT 1   /* bad pointer, size or id */
#define ZH_ERR_UNSUPPORTED 2        /* valid request the backend cannot run */
#define ZH_ERR_HIP 3                /* a HIP runtime call failed */
#define ZH_ERR_MISMATCHED_DESCRIPTOR 4 /* StartError::MismatchedDescriptor, run.rs:369-392 */
#define ZH_ERR_MISSING_KEY 5        /* StartError::MissingKey: an input is unbound */
#define ZH_ERR_OUT_OF_MEMORY 6
#define ZH_ERR_TYPE 7               /* CommandError type errors, command.rs:3617-3630 */
#define ZH_ERR_UNIMPLEMENTED 8      /* CommandError::UNIMPLEMENTED */
#define ZH_ERR_KNOB_LENGTH 9        /* knob length mismatch, run.rs:1295-1298 */
#define ZH_ERR_NO_DEVICE 10         /* a device call on a context-free program */
#define ZH_ERR_OTHER 11             /* CommandError::OTHER */

/* ---- identifiers ------------------------------------------------------
 * SAMPLE_BITS ids   = lib/std/src/stage.frag:151-170 (std140 XyzParameter.z)
 * SAMPLE_PARTS ids  = lib/std/src/stage.frag:125-145 (XyzParameter.y)
 * TRANSFER ids      = lib/std/src/stage.frag:107-119 (XyzParameter.x)
 */
#define ZH_BITS_INT8 0
#define ZH_BITS_INT16 3
#define ZH_BITS_INT8X4 10
#define ZH_BITS_INT16X4 13
#define ZH_BITS_FLOAT16X4 18
#define ZH_BITS_FLOAT32X4 19

#define ZH_PARTS_RGBA 8
#define ZH_PARTS_BGRA 10
#define ZH_PARTS_LCHA 20

#define ZH_TRANSFER_BT709 0
#define ZH_TRANSFER_LINEAR 4
#define ZH_TRANSFER_SRGB 5
#define ZH_TRANSFER_LABLCH 0x100

/* Color models (image-canvas Color, lib/lang/src/serde_bindings.rs:16-32). */
#define ZH_COLOR_RGB 0
#define ZH_COLOR_OKLAB 1
#define ZH_COLOR_SCALARS 2
#define ZH_COLOR_SRLAB2 3 /* Color::SrLab2 { whitepoint } (LabA / LchA texels) */

/* Primaries: image-canvas order (serde_bindings.rs:65-75) + Display-P3. */
#define ZH_PRIMARIES_XYZ 0
#define ZH_PRIMARIES_BT601_525 1
#define ZH_PRIMARIES_BT601_625 2
#define ZH_PRIMARIES_BT709 3
#define ZH_PRIMARIES_SMPTE240 4
#define ZH_PRIMARIES_BT2020 5
#define ZH_PRIMARIES_BT2100 6
#define ZH_PRIMARIES_DISPLAY_P3 7

/* White points (image-canvas Whitepoint incl. D50 used at tests/blend.rs:180). */
#define ZH_WHITEPOINT_D50 3
#define ZH_WHITEPOINT_D65 5

typedef struct zh_texel {
  uint32_t bits;   /* SAMPLE_BITS id */
  uint32_t parts;  /* SAMPLE_PARTS id */
} zh_texel;

typedef struct zh_color {
  uint32_t model;      /* ZH_COLOR_* */
  uint32_t primaries;  /* ZH_PRIMARIES_* (RGB only) */
  uint32_t transfer;   /* TRANSFER id (RGB and Scalars) */
  uint32_t whitepoint; /* ZH_WHITEPOINT_* (RGB only) */
  uint32_t luminance;  /* image-canvas Luminance, carried only */
} zh_color;

/* Descriptor, lib/zosimos/src/buffer.rs:13-30.  row_stride 0 = tight rows. */
typedef struct zh_descriptor {
  uint32_t width;
  uint32_t height;
  uint64_t row_stride;
  uint32_t texel_stride; /* bytes per texel (SampleBits::bytes) */
  zh_texel texel;
  zh_color color;
} zh_descriptor;

/* Rectangle, command.rs:279-285: [x, max_x) x [y, max_y). */
typedef struct zh_rect {
  uint32_t x, y, max_x, max_y;
} zh_rect;

/* ---- context and images ------------------------------------------------ */
typedef struct zh_context zh_context;
typedef struct zh_image zh_image;
typedef struct zh_program zh_program;

/* One HIP device and one stream.  `stream` may be NULL (the context creates
 * its own default-priority stream) or an existing hipStream_t to launch on. */
zh_status zh_context_create(int32_t device, void* stream, zh_context** out);
zh_status zh_context_destroy(zh_context* ctx);
/* Block until all work queued on the context's stream is done
 * (SyncPoint::block_on, run.rs:3019-3029). */
zh_status zh_sync(zh_context* ctx);
/* 1 if work is still queued, 0 if idle (Execution::is_running, run.rs:1352). */
zh_status zh_query(zh_context* ctx, int32_t* running);
/* The context's hipStream_t. */
void* zh_context_stream(zh_context* ctx);

/* Device image with the layout of `desc` (row_stride 0 = tight).  The
 * context owns the memory (pool.rs:122-156 ImageData::GpuBuffer analogue). */
zh_status zh_image_alloc(zh_context* ctx, const zh_descriptor* desc, zh_image** out);
/* Wrap caller-owned device memory (e.g. a torch tensor); not freed by us. */
zh_status zh_image_wrap(zh_context* ctx, const zh_descriptor* desc, void* device_ptr,
                        size_t bytes, zh_image** out);
zh_status zh_image_free(zh_image* img);
zh_status zh_image_device_ptr(zh_image* img, void** ptr, size_t* bytes);
/* Host <-> device copies with pitch (replaces the row-by-row memcpy of
 * copy_host_to_buffer, run.rs:3282-3309).  Asynchronous on the context
 * stream; host memory must stay valid until zh_sync. */
zh_status zh_image_upload(zh_image* img, const void* host, uint64_t host_row_stride);
zh_status zh_image_download(zh_image* img, void* host, uint64_t host_row_stride);

/* ---- High IR --------------------------------------------------------- */
/* High variants, program.rs:88-139. */
#define ZH_HIGH_INPUT 0   /* High::Input: texture `dst` <- binding slot `src` */
#define ZH_HIGH_OUTPUT 1  /* High::Output: binding slot `dst` <- texture `src` */
#define ZH_HIGH_DRAW 2    /* High::PushOperand* + High::DrawInto */
#define ZH_HIGH_COPY 3    /* High::Copy: byte copy `src` -> `dst` */
#define ZH_HIGH_UNINIT 4  /* High::Uninit */
/* A buffer register's bytes read as texture `dst` (Op::BufferInit and
 * ConstructOp::FromBuffer, command.rs:937-968, 1777-1801): `params` holds the
 * bytes, laid out by the texture's descriptor (its row_stride; 256-byte rows
 * for a device buffer, buffer.rs:121-134), at least row_stride x height of
 * them.  The program keeps them in device memory; a knob replaces them
 * (Environment::knob of a WithBuffer register, run.rs:1292-1306). */
#define ZH_HIGH_BUFFER 5

/* Target, program.rs:142-148. */
#define ZH_TARGET_DISCARD 0 /* cleared first (to opaque blue, program.rs:1495-1499) */
#define ZH_TARGET_LOAD 1

/* Initializer, program.rs:176-240. */
#define ZH_INIT_PAINT_FULLSCREEN 0
#define ZH_INIT_PAINT_TO_SELECTION 1

/* Shader keys: FragmentShaderKey (shaders.rs:116-143) and the ops the
 * reference lacks.  Parameter blobs (std140):
 *  PAINT_COPY            none (copy.frag)
 *  LINEAR_COLOR_MATRIX   mat3 std140, 48 B (color_matrix.rs:151-161)
 *  OKLAB_ENCODE/DECODE   mat3 std140, 48 B (shaders/oklab.rs:36-39)
 *  SOLID_RGB             vec4, 16 B (solid_rgb.frag)
 *  BILINEAR              6 vec4 {u_min, u_max, v_min, v_max, uv_min, uv_max},
 *                        96 B (shaders/bilinear.rs:33-45, bilinear.frag)
 *  BOX3                  mat3 std140, 48 B (shaders/box3.rs:25-27): row-major
 *                        weights, row = y offset + 1, column = x offset + 1;
 *                        operand = the sampled image (box3.frag)
 *  DISTRIBUTION_NORMAL2D 8 floats {expectation (2), inverse covariance (4,
 *                        row-major), pseudo-determinant, 0}, 32 B
 *                        (shaders/distribution_normal2d.rs binary_data)
 *  SRLAB2_ENCODE/DECODE  {white point XYZ vec4, mat3 std140}, 64 B
 *                        (shaders/srlab2.rs binary_data): encode takes the
 *                        RGB -> XYZ matrix, decode the XYZ -> RGB one
 *                        (srlab2.frag)
 *  PALETTE               mat4x2 of 8 floats {x0, y0, x1, y1, x2, y2, x3, y3}
 *                        (shaders/palette.rs binary_data); operands (palette,
 *                        indices): the palette texel nearest to
 *                        (sum x_k c_k, sum y_k c_k) + 0.5 / indices' size, c
 *                        the indices' colour, clamped to the palette's edge
 *                        (palette.frag)
 *  EXTRACT               vec4 {channel position, 0, 0, 0} as u32, 16 B: one
 *                        colour channel of the operand, decoded as a
 *                        single-channel texel is (the channel at its
 *                        position, 0 elsewhere, alpha 1)
 *  INJECT                2 vec4 {select, color}, 32 B (inject.frag); operands
 *                        (below, above): mix(below, dot(above, color), select)
 *  FRACTAL_NOISE         {grid_scale (2 f32), initial amplitude f32, amplitude
 *                        damping f32, octaves u32}, 20 B padded to 24
 *                        (shaders/fractal_noise.rs binary_data); at most 4096
 *                        octaves
 *  OKLAB_ADJUST          vec4 {chroma scale, hue rotation (rad), 0, 0}
 *  BLEND_OVER            none; operands (below, above)
 *  AFFINE_BILINEAR       mat3 std140 of the row-major Affine matrix, 48 B;
 *                        operand = layer, composited over the target (Load)
 *  PREMULTIPLY, UNPREMULTIPLY  none
 *  DYNAMIC               a zh_dynamic_call followed by its data_len bytes of
 *                        shader data: a user stage kernel (High::DrawInto of
 *                        FragmentShaderInvocation::Runtime, command.rs:
 *                        2767-2830); operands = its num_args sampled images
 */
#define ZH_SHADER_PAINT_COPY 0
#define ZH_SHADER_LINEAR_COLOR_MATRIX 1
#define ZH_SHADER_OKLAB_ENCODE 2
#define ZH_SHADER_OKLAB_DECODE 3
#define ZH_SHADER_SOLID_RGB 4
#define ZH_SHADER_BILINEAR 5
#define ZH_SHADER_BOX3 6
#define ZH_SHADER_DISTRIBUTION_NORMAL2D 7
#define ZH_SHADER_FRACTAL_NOISE 8
#define ZH_SHADER_SRLAB2_ENCODE 9
#define ZH_SHADER_SRLAB2_DECODE 10
#define ZH_SHADER_PALETTE 11
#define ZH_SHADER_EXTRACT 12
#define ZH_SHADER_INJECT 13
#define ZH_SHADER_OKLAB_ADJUST 16
#define ZH_SHADER_BLEND_OVER 17
#define ZH_SHADER_AFFINE_BILINEAR 18
#define ZH_SHADER_PREMULTIPLY 19
#define ZH_SHADER_UNPREMULTIPLY 20
#define ZH_SHADER_DYNAMIC 21

/* ---- user stage kernels (ShaderCommand, command/dynamic.rs:7-51) -------
 * The HIP form of ShaderSource (dynamic.rs:31-34, a #[non_exhaustive] enum
 * whose one variant is SPIR-V): a code object for gfx950 (`hipcc --genco
 * --offload-arch=gfx950`) and the name of a kernel in it declared
 *     extern "C" __global__ void name(zh_dynamic_args args)
 * It runs where the reference runs the fragment shader: over every texel of
 * the result, with 16 x 16 workgroups and a grid of ceil(width / 16) x
 * ceil(height / 16) (the kernel tests x < width, y < height itself).  Its
 * operands and its result are RGBA f32 in tight rows, straight alpha, in the
 * linear values of each register's colour space -- the values a fragment
 * shader reads through its sampler and writes to its target, before the
 * register's texel encoding.  The texture coordinate of texel (x, y) is
 * ((x + 0.5) / width, (y + 0.5) / height); the reference samples operands
 * with a nearest, clamp-to-edge sampler (encoder.rs:1533-1537). */
typedef struct zh_dynamic_args {
  uint32_t width, height;       /* the result's size in texels */
  uint32_t num_inputs;          /* operands: 0, 1 or 2 */
  uint32_t data_len;            /* bytes of shader data at `data` */
  float* out;                   /* width * height RGBA f32 texels */
  const float* in[2];           /* the operands, in_width[k] * in_height[k] texels each */
  uint32_t in_width[2], in_height[2];
  const void* data;             /* device copy of the shader data (ShaderData::set_data) */
} zh_dynamic_args;

typedef struct zh_shader zh_shader;
/* Copy a code object and name its kernel.  Nothing is loaded until a program
 * that runs the shader is built on a context (hipModuleLoadData on the
 * context's device, once per device).  The shader is reference counted:
 * command buffers, linked programs and programs that use it hold it, and
 * zh_shader_release drops the creator's reference. */
zh_status zh_shader_create(const void* code_object, size_t len, const char* kernel, zh_shader** out);
zh_status zh_shader_retain(zh_shader* shader);
zh_status zh_shader_release(zh_shader* shader);

/* The leading bytes of a ZH_SHADER_DYNAMIC parameter blob (the
 * ShaderInvocation of command.rs:2943-2958). */
typedef struct zh_dynamic_call {
  zh_shader* shader;
  uint32_t num_args;  /* operands of the draw: 0, 1 or 2 */
  uint32_t data_len;  /* shader data bytes following this header */
} zh_dynamic_call;

typedef struct zh_op {
  uint32_t high;        /* ZH_HIGH_* */
  uint32_t dst;         /* texture id written (OUTPUT: output binding slot) */
  uint32_t src;         /* INPUT: input binding slot; OUTPUT/COPY: texture id */
  uint32_t target;      /* ZH_TARGET_* (DRAW) */
  uint32_t initializer; /* ZH_INIT_* (DRAW) */
  uint32_t shader;      /* ZH_SHADER_* (DRAW) */
  uint32_t num_operands;
  uint32_t operands[2]; /* texture ids pushed by High::PushOperand */
  int32_t knob;         /* -1, or the knob id that may override `params` */
  const void* params;   /* std140 blob, copied at build time */
  uint32_t params_len;
  /* PaintToSelection (program.rs:191-205): source selection in texels of
   * operands[0], target quad corners in target pixels (clockwise from
   * top-left; QuadTarget::Absolute, program.rs:254-258) and viewport. */
  zh_rect selection;
  float quad[8];
  zh_rect viewport;
} zh_op;

/* Plan flags. */
#define ZH_PLAN_DEFAULT 0u
/* Quantise every register to its declared texel format, as the reference's
 * per-op staging does (encoder.rs:871-956), instead of keeping fused
 * intermediates in FP32. */
#define ZH_PLAN_STRICT_REGISTERS 1u
/* Do not capture the program's launches into a hipGraph. */
#define ZH_PLAN_NO_GRAPH 2u
/* Run pointwise chains that have no registered specialised kernel on the
 * general interpreter instead of compiling one for them at run time. */
#define ZH_PLAN_NO_JIT 4u
/* Resample separable one-layer stages from each tile's source box staged in
 * LDS (decoded and premultiplied once per texel) wherever the box fits,
 * instead of gathering every tap directly.  Measured slower than the direct
 * gathers on MI355X at every scale from 1x to 4x, so not the default. */
#define ZH_PLAN_STAGED 8u
/* Launch the two warp forms of a run of independent resample stages (as a
 * batch of RGBA8 and RGBA16F images) as one persistent grid that deals their
 * tiles alternately (a kernel compiled at run time for the pair of forms),
 * instead of two tile queues one after the other.  Measured slower on
 * MI355X, so not the default. */
#define ZH_PLAN_PAIR 16u

/* Lower a High program (Program::lower_to, program.rs:1304).  `textures`
 * holds the descriptor of every texture id.  `ctx` may be NULL: the plan is
 * then built for inspection only (zh_program_describe) and cannot launch. */
zh_status zh_program_build(zh_context* ctx, const zh_descriptor* textures, uint32_t num_textures,
                           const zh_op* ops, uint32_t num_ops, uint32_t flags, zh_program** out);
zh_status zh_program_destroy(zh_program* prog);
/* Human-readable fused plan (stages, kernels, micro-ops, bytes). */
zh_status zh_program_describe(zh_program* prog, char* buf, size_t len, size_t* needed);
/* Number of kernel launches per program launch, and algorithmic HBM bytes
 * (each input read once, each output written once). */
zh_status zh_program_stats(zh_program* prog, uint32_t* launches, uint64_t* algorithmic_bytes);

/* Algorithmic HBM bytes of each kernel launch of the program, in launch
 * order (the sum is zh_program_stats' figure); for measurement, no
 * counterpart in the reference. */
zh_status zh_program_launch_bytes(zh_program* prog, uint64_t* bytes, uint32_t max_launches, uint32_t* num_launches);

/* Knobs (Environment::knob, run.rs:1292-1306): override the parameter blob of
 * every op built with this knob id; the length must equal the original. */
zh_status zh_program_set_knob(zh_program* prog, int32_t knob, const void* data, uint32_t len);
/* The knob's current parameter bytes: *len receives their length and the
 * first min(cap, *len) bytes are copied to `out` (which may be NULL when cap
 * is 0).  A run's knob data (Environment::knob, run.rs:1292-1306) holds for
 * that run only; a caller restores the bytes read here once it finishes. */
zh_status zh_program_knob_data(zh_program* prog, int32_t knob, void* out, uint32_t cap, uint32_t* len);

typedef struct zh_binding {
  uint32_t slot;   /* input slot (ZH_HIGH_INPUT.src) or output slot (ZH_HIGH_OUTPUT.dst) */
  uint32_t is_output;
  zh_image* image;
} zh_binding;

/* Bind and launch asynchronously (Executable::launch + Execution::step
 * until done, run.rs:1016-1464).  Every input and output slot must be bound
 * to an image whose descriptor matches.  Successive launches may bind other
 * images without a zh_sync in between (double buffering): each launch keeps
 * the images it was bound with.  The host waits only when three launches
 * with changed bindings are still queued. */
zh_status zh_program_launch(zh_program* prog, const zh_binding* bindings, uint32_t num_bindings);
/* Executable::launch_once (run.rs:1050-1079): one launch that consumes the
 * program -- kernel by kernel (no graph is captured for a single run), then
 * the program is destroyed once its kernels are done; `prog` is invalid
 * after the call, whatever it returns. */
zh_status zh_program_launch_once(zh_program* prog, const zh_binding* bindings, uint32_t num_bindings);
/* Executable::dot (run.rs:490-1010): a Graphviz digraph of the plan -- input,
 * output and intermediate images as boxes, stages (with their kernels) as
 * ellipses, an edge per read and per write.  Same buffer protocol as
 * zh_program_describe. */
zh_status zh_program_dot(zh_program* prog, char* buf, size_t len, size_t* needed);
/* hipGraph state of a program (no counterpart in the reference, which
 * records a wgpu command buffer per launch, run.rs:1394-1464).  A program of
 * two or more kernel launches is captured into a hipGraph at its first
 * launch and replayed after; a rebinding that changes which kernels run
 * captures again.  A capture that fails is reported here (and in
 * zh_program_describe) and the program launches kernel by kernel. */
#define ZH_GRAPH_NOT_YET 0       /* several launches, not launched yet */
#define ZH_GRAPH_CAPTURED 1
#define ZH_GRAPH_FAILED 2
#define ZH_GRAPH_DISABLED 3      /* built with ZH_PLAN_NO_GRAPH */
#define ZH_GRAPH_SINGLE_LAUNCH 4 /* one kernel launch: launched directly */
typedef struct zh_graph_info {
  uint32_t state;    /* ZH_GRAPH_* */
  uint32_t nodes;    /* nodes of the captured graph */
  uint32_t captures; /* captures so far */
  uint32_t pad;
  uint64_t replays;  /* launches that replayed the graph */
} zh_graph_info;
zh_status zh_program_graph_info(zh_program* prog, zh_graph_info* info);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): lower the
 * High program (as zh_program_build) and launch it once (as
 * zh_program_launch) with every device table of the plan -- stage
 * descriptors, tile queues and their counters, tile-layer masks, buffer
 * registers and intermediates -- placed in the library's static scratch
 * array of ZH_STATIC_SCRATCH_BYTES per device instead of allocated.  The
 * call makes no hipMalloc, hipFree or other allocating call; inputs and
 * outputs are the caller's images (zh_image_wrap adopts existing device
 * memory).  It waits for the work already queued on the context's stream,
 * uploads the tables, and launches the kernels asynchronously, without a
 * hipGraph.  A program whose tables exceed the scratch fails with
 * ZH_ERR_OUT_OF_MEMORY and launches nothing.  Every call on a device shares
 * its scratch, which keeps the last call's tables: calls through one context
 * are ordered by its stream, but calls that overlap on one GPU through
 * different contexts or streams race. */
#define ZH_STATIC_SCRATCH_BYTES (32u << 20)
zh_status zh_program_launch_static(zh_context* ctx, const zh_descriptor* textures, uint32_t num_textures,
                                   const zh_op* ops, uint32_t num_ops, uint32_t flags, const zh_binding* bindings,
                                   uint32_t num_bindings);

/* Average device time (ms) of each stage over the last `zh_program_time`
 * call: launches the program `iterations` times and records HIP events
 * around every stage on the context stream. */
zh_status zh_program_time(zh_program* prog, const zh_binding* bindings, uint32_t num_bindings,
                          uint32_t iterations, float* stage_ms, uint32_t max_stages,
                          uint32_t* num_stages);

/* Autotuner (no counterpart in the reference, whose tiles are the
 * rasteriser's): for every warp-kernel launch of the program (affine layer
 * stacks and their batches), time each tile shape the kernel accepts
 * (64x16, 32x32, 128x8, 16x64 pixels) with each persistent-grid width
 * (all resident workgroups, or 4, 3, 2, 1 per CU) on the bound images over
 * `iterations` launches and keep the fastest.  The choice persists in the program until a
 * knob re-plans it; zh_program_describe shows the tiles. */
zh_status zh_program_autotune(zh_program* prog, const zh_binding* bindings, uint32_t num_bindings,
                              uint32_t iterations);

/* ---- learned colour surrogate ---------------------------------------------
 * No counterpart in the reference (it has no learned colour path).  A
 * pointwise stage that maps RGBA8 sRGB to RGBA8 sRGB through colour-only
 * micro-ops (3x3 matrices, cube roots, cubes, clamps) is an R^3 -> R^3
 * function of the encoded colour; zh_program_learn fits a 3-64-64-3 SiLU
 * network to it (bfloat16 weights, FP32 accumulation on the matrix cores):
 * training samples come from the stage's own kernel, training runs on the
 * GPU, and the network replaces the stage only if its max Oklab Delta-E
 * against the stage's program evaluated in FP64 on a lattice**3 grid of the
 * input cube stays below max_delta_e.  Otherwise the plain kernel keeps
 * running.  Alpha passes through unchanged either way. */
typedef struct zh_learn_options {
  uint32_t uniform_points; /* 0: 2^20 samples uniform in the cube */
  uint32_t face_points;    /* 0: 2^16 samples on the cube's faces */
  uint32_t steps;          /* 0: 20000 Adam steps */
  uint32_t batch;          /* 0: 8192 samples per step */
  float lr;                /* 0: 3e-3, cosine decay to lr / 100 */
  uint32_t seed;
  uint32_t lattice;        /* 0: 65 points per axis */
  float max_delta_e;       /* 0: 0.005 */
} zh_learn_options;

typedef struct zh_learn_report {
  uint32_t stage;             /* stage index (zh_program_describe) */
  uint32_t accepted;          /* 1: the surrogate runs instead of the stage's kernel */
  uint32_t steps, samples;
  float initial_loss;         /* MSE of the first step's batch */
  float final_loss;           /* mean MSE over the last 100 steps */
  double max_delta_e;         /* trained network on the lattice */
  double mean_delta_e;
  double max_delta_e_initial; /* the untrained network on the same lattice */
  float train_ms;             /* sample generation + training, wall clock */
} zh_learn_report;

/* Learn a surrogate for every eligible stage; fills up to max_reports
 * reports (num_reports = eligible stages).  Needs a context. */
zh_status zh_program_learn(zh_program* prog, const zh_learn_options* options, zh_learn_report* reports,
                           uint32_t max_reports, uint32_t* num_reports);
/* 1 (default): accepted surrogates replace their stages; 0: plain kernels;
 * 2: every fitted surrogate runs, accepted or not (for measurement). */
zh_status zh_program_use_surrogates(zh_program* prog, int32_t enable);
/* Master (FP32) weights of a stage's surrogate, 4,611 floats in the order
 * W1[3][64], b1[64], W2[64][64], b2[64], W3[64][3], b3[3]. */
zh_status zh_program_surrogate_params(zh_program* prog, uint32_t stage, float* params, uint32_t len);
/* Loss (mean squared error over n x 3 values) and its gradient for the given
 * parameters on n samples x[n][3] -> t[n][3], by the training kernels. */
zh_status zh_mlp_gradient(zh_context* ctx, const float* params, uint32_t n, const float* x, const float* t,
                          float* grad, float* loss);
/* The network on n inputs x[n][3] -> y[n][3] by the matrix-core inference
 * kernel. */
zh_status zh_mlp_evaluate(zh_context* ctx, const float* params, uint32_t n, const float* x, float* y);

/* Host-only: the 8-bit sRGB tables the kernels stage in LDS (decode values,
 * bucket codes and code thresholds; 6,164 bytes), for inspection and tests.
 * No counterpart in the reference, whose sRGB codec is the texture unit's. */
zh_status zh_srgb_tables(void* out, size_t len);

/* ---- image pool (replaces zosimos::pool::Pool, pool.rs:19-156) ----------
 * A pool holds images under keys: host bytes (ImageData::Host), a HIP device
 * image (ImageData::GpuBuffer) or only a descriptor (ImageData::LateBound).
 * Keys are never 0; 0 binds an output to no pool image. */
typedef struct zh_pool zh_pool;
typedef uint64_t zh_pool_key;
#define ZH_IMAGE_HOST 0
#define ZH_IMAGE_HIP_BUFFER 1
#define ZH_IMAGE_LATE_BOUND 2
zh_status zh_pool_create(zh_pool** out);                        /* Pool::new, pool.rs:182 */
zh_status zh_pool_destroy(zh_pool* pool);
/* Pool::insert / insert_srgb (pool.rs:244-280): copies `bytes` (rows
 * `row_stride` apart, 0 = tight) into the pool as host bytes. */
zh_status zh_pool_insert_host(zh_pool* pool, const zh_descriptor* desc, const void* bytes, uint64_t row_stride,
                              zh_pool_key* key);
/* Pool::declare (pool.rs:281-291): a LateBound image, a destination only. */
zh_status zh_pool_declare(zh_pool* pool, const zh_descriptor* desc, zh_pool_key* key);
/* ImageData::GpuBuffer (pool.rs:133-137): the pool takes ownership of `img`. */
zh_status zh_pool_insert_device(zh_pool* pool, zh_image* img, zh_pool_key* key);
/* PoolImage::descriptor (pool.rs:661-668). */
zh_status zh_pool_image_info(zh_pool* pool, zh_pool_key key, zh_descriptor* desc, int32_t* kind);
/* PoolImage::as_bytes (pool.rs:669-680): tight host rows, valid until the
 * image changes. */
zh_status zh_pool_host_bytes(zh_pool* pool, zh_pool_key key, const void** data, uint64_t* len);
/* The device image of a ZH_IMAGE_HIP_BUFFER entry; the pool keeps ownership. */
zh_status zh_pool_device_image(zh_pool* pool, zh_pool_key key, zh_image** img);
/* ImageMeta::no_read (pool.rs:104-109, 912-916): an input launched from this image
 * moves out of the pool until the execution retires it or finishes. */
zh_status zh_pool_set_no_read(zh_pool* pool, zh_pool_key key, int32_t no_read);
/* Pool::upload (pool.rs:292-370) and its reverse: move the image to the device of `ctx`, or
 * back to host bytes (synchronous). */
zh_status zh_pool_upload(zh_pool* pool, zh_pool_key key, zh_context* ctx);
zh_status zh_pool_download(zh_pool* pool, zh_pool_key key);
zh_status zh_pool_remove(zh_pool* pool, zh_pool_key key);      /* Pool::remove */
/* Pool::clear_cache (pool.rs:450): free the device buffers kept from
 * finished executions for reuse. */
zh_status zh_pool_clear_cache(zh_pool* pool);
zh_status zh_pool_cached_buffers(zh_pool* pool, uint32_t* count);

/* Launch from the pool (Executable::launch with a pool, run.rs:1016-1167):
 * binding {slot, 0, key} reads input `slot` from pool image `key`; {slot, 1,
 * key} writes output `slot` into `key` (0 = a new image at retirement).  All
 * checks run before anything moves: ZH_ERR_MISSING_KEY for an unbound used
 * slot or an input whose image holds no data, ZH_ERR_MISMATCHED_DESCRIPTOR
 * when the image's descriptor differs from the slot's, ZH_ERR_INVALID_ARGUMENT
 * for an unknown key or an image bound twice.  Device images are used in
 * place; host bytes are uploaded into device buffers the pool caches for the
 * next launch.  The launch is asynchronous; the pool must outlive the
 * execution. */
typedef struct zh_pool_binding {
  uint32_t slot;
  uint32_t is_output;
  zh_pool_key key;
} zh_pool_binding;
typedef struct zh_execution zh_execution;
zh_status zh_program_launch_pool(zh_program* prog, zh_pool* pool, const zh_pool_binding* bindings,
                                 uint32_t num_bindings, zh_execution** out);
/* Executable::launch without stepping (run.rs:1016-1050): everything
 * above is checked, traded and bound, but no kernel is queued until
 * zh_execution_step.  Until its last launch is issued the program runs no
 * other execution (ZH_ERR_INVALID_ARGUMENT), and retiring is refused. */
zh_status zh_program_start_pool(zh_program* prog, zh_pool* pool, const zh_pool_binding* bindings,
                                uint32_t num_bindings, zh_execution** out);
/* Execution::step_to(StepLimits { instructions }) (run.rs:1394-1464):
 * queue up to `max_launches` more kernel launches (kernel by kernel, no
 * hipGraph); `launched` receives how many (0 once all are queued). */
zh_status zh_execution_step(zh_execution* exec, uint32_t max_launches, uint32_t* launched);
/* Execution::is_running (run.rs:1381-1384): 1 while launches are left to
 * step or the context's stream still has work. */
zh_status zh_execution_is_running(zh_execution* exec, int32_t* running);
/* Execution::resources_used (ResourcesUsed, run.rs:212-221, 1495-1498). */
typedef struct zh_resources_used {
  uint64_t buffer_mem;     /* device bytes newly allocated for the launch's images */
  uint64_t buffer_reused;  /* device bytes reused: the pool's cache, images used in place */
  uint64_t program_mem;    /* the program's own device memory: intermediates, buffers, tables */
  uint32_t buffers_allocated;
  uint32_t buffers_reused;
} zh_resources_used;
zh_status zh_execution_resources(zh_execution* exec, zh_resources_used* used);
/* Retire::output (run.rs:2809-2875): move output `slot` into the pool, as
 * host bytes (`to_host`, waits for the launch) or as the device image itself
 * (no copy, no wait).  `key` receives the pool key. */
zh_status zh_execution_retire_output(zh_execution* exec, uint32_t slot, int32_t to_host, zh_pool_key* key);
/* Retire::input (run.rs:2791-2808): give a moved (no_read) input back. */
zh_status zh_execution_retire_input(zh_execution* exec, uint32_t slot, zh_pool_key* key);
/* Retire::prune (run.rs:2944-2952; a no-op there): the execution's own device buffers are
 * freed at finish instead of kept in the pool's cache. */
zh_status zh_execution_prune(zh_execution* exec);
/* Retire::finish (run.rs:2962-2998): wait, return what is not retired to the
 * pool and its cache, free the execution. */
zh_status zh_execution_finish(zh_execution* exec);
/* Retire::finish_by_discarding (run.rs:2989-2997): wait and drop everything
 * not retired -- moved inputs stay out of the pool, buffers are freed. */
zh_status zh_execution_discard(zh_execution* exec);

/* Message of the last failed call on this thread (never NULL). */
const char* zh_last_error(void);
/* Library version string. */
const char* zh_version(void);

#ifdef __cplusplus
}
#endif

#endif /* ZOSIMOS_HIP_H */
