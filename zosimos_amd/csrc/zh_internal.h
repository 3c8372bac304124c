// zh_internal.h -- host-side structures shared by the runtime, planner and
// command front end.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/zosimos_hip.h"
#include "zh_stage.h"

namespace zh {

zh_status fail(zh_status code, const std::string& msg);


// ---- texel format helpers ------------------------------------------------
uint32_t bits_bytes(uint32_t bits);  // SampleBits::bytes for a stage.frag id, 0 if unknown
// RGBA8/BGRA8 with sRGB or linear transfer on an RGB colour: the formats the
// reference hands to the texture unit (program.rs:794-838).
bool is_native(const zh_descriptor& d);
// Transfer applied when staging this descriptor (ImageDescriptor::new,
// program.rs:840-899): the RGB/Scalars transfer, or LabLch / Linear for Oklab.
uint32_t stage_transfer(const zh_descriptor& d);
uint32_t stage_parts(const zh_descriptor& d);
uint64_t tight_stride(const zh_descriptor& d);
uint64_t row_stride_of(const zh_descriptor& d);
bool descriptor_valid(const zh_descriptor& d, std::string* why);
bool same_layout(const zh_descriptor& a, const zh_descriptor& b);

using Mat3 = std::array<double, 9>;  // row-major
Mat3 mat_mul(const Mat3& a, const Mat3& b);
bool mat_inv(const Mat3& a, Mat3* out);
Mat3 mat_identity();
// RGB -> XYZ matrix of ZH_PRIMARIES_* under a ZH_WHITEPOINT_* (zh_command.cpp).
bool primaries_to_xyz(uint32_t primaries, uint32_t whitepoint, Mat3* out);
Mat3 mat_from_std140(const void* blob);  // 12 floats, column-major with padding

// ---- user stage kernels (zh_dynamic.cpp) ------------------------------------
// A counted reference to a zh_shader (retained on copy, released on
// destruction).
class ShaderRef {
 public:
  ShaderRef() = default;
  explicit ShaderRef(zh_shader* s);  // takes a new reference
  ShaderRef(const ShaderRef& o);
  ShaderRef& operator=(const ShaderRef& o);
  ~ShaderRef();
  zh_shader* get() const { return s_; }
 private:
  zh_shader* s_ = nullptr;
};
// The kernel of `shader` on `device` (the current device), its code object
// loaded on first use.
zh_status shader_function(zh_shader* shader, int device, hipFunction_t* fn);
// Is the code object already loaded on `device`?
bool shader_loaded(zh_shader* shader, int device);
const char* shader_kernel_name(zh_shader* shader);

// ---- kernels compiled at run time (zh_jit.cpp) -------------------------------
// Can the specialised bodies run this signature (zh_spec.h elements)?
bool jit_supported(const uint16_t* sig, int n);
// The process-wide id of a signature (nothing is compiled yet).
int jit_signature(const uint16_t* sig, int n);
// The stage and batch kernels of signature `id` on `device` (the current
// device), compiled on first use; `why` receives the reason of a failure.
zh_status jit_kernels(int id, int device, hipFunction_t* stage, hipFunction_t* batch, std::string* why);
double jit_compile_seconds(int id);
// Affine layer-stack forms without a registered warp kernel: a post program
// of operations the warp bodies run, and the process-wide id of a form.
bool jit_warp_supported(const uint16_t* post, int n);
int jit_warp_form(uint32_t src, uint32_t fmt, uint32_t mat, const uint16_t* post, int n);
// The general, separable or LDS-staged (separable, one source) stage kernel
// and batch kernel of form `id`.
zh_status jit_warp_kernels(int id, bool sep, bool staged, int device, hipFunction_t* stage, hipFunction_t* batch,
                           std::string* why);
// Does form `id` have LDS-staged kernels (one-source forms)?
bool jit_warp_has_staged(int id);
// An affine layer-stack form as the warp bodies take it.
struct WarpForm {
  uint32_t src = 0, fmt = 0, mat = 0;
  std::vector<uint16_t> post;
  bool sep = false;
};
// The form behind a StageDesc::variant (registered or run-time); false for none.
bool warp_form_of(int variant, WarpForm* f);
// Two warp forms dealt by one persistent grid (warp_pair): the process-wide
// id of the pair, and its kernel on `device`, compiled on first use.
int jit_warp_pair(const WarpForm& a, const WarpForm& b);
zh_status jit_pair_kernel(int id, int device, hipFunction_t* fn, std::string* why);
// StageDesc::variant of a run-time warp form: kJitWarpBase + 2 id (+ 1 for
// separable maps); registered forms are below it.
constexpr int kJitWarpBase = 1 << 16;

// A user stage of a plan: the shader and its data.
struct DynamicCall {
  ShaderRef shader;
  std::vector<uint8_t> data;
  uint32_t num_args = 0;
};

// ---- planned program -------------------------------------------------------
struct StagePlan {
  StageDesc desc;                 // host copy; slot ptrs patched per launch
  int tile_mode = 0;              // resample: TILE_* of zh_kernels.hip
  bool fast = false;              // pointwise: every slot has an in-line codec (k_pointwise<true>)
  int spec = -1;                  // pointwise: index of the specialised kernel (zh_spec.h), -1 = none
  int jit = -1;                   // pointwise: run-time kernel of the signature (zh_jit.cpp), -1 = none
  bool staged = false;            // warp: the LDS-staged kernel (staged_tile), the box fitting its tile
  uint32_t lds_bytes = 0;
  uint64_t alg_bytes = 0;         // algorithmic HBM bytes (reads once + writes once)
  // slot -> what it refers to: >=0 input slot, or ~output slot (<0), or
  // INTERMEDIATE_BASE + buffer id.
  std::vector<int64_t> slot_ref;
  int dynamic = -1;               // STAGE_DYNAMIC: index into Plan::dynamics
  std::string text;               // description
};

constexpr int64_t kIntermediateBase = 1ll << 40;

struct Plan {
  std::vector<StagePlan> stages;
  std::vector<zh_descriptor> intermediates;  // device buffers owned by the program
  std::vector<zh_descriptor> input_desc;     // per input slot (declared texture desc)
  std::vector<zh_descriptor> output_desc;    // per output slot
  std::vector<bool> input_used, output_used;
  // Input slots from user_inputs on are the program's own buffers
  // (ZH_HIGH_BUFFER): slot user_inputs + k holds the bytes of op buffer_ops[k].
  uint32_t user_inputs = 0;
  std::vector<uint32_t> buffer_ops;
  bool batched = false;                      // resample stages run as one tile queue
  std::vector<DynamicCall> dynamics;         // user stages (STAGE_DYNAMIC)
  std::string text;
};

// Does warp stage d get the LDS-staged kernel on tw x th tiles: a separable
// one-layer one-source form that has one, every tile's source box within
// kBoxTexels?
bool warp_staged_fits(const StageDesc& d, uint32_t tw, uint32_t th);

// Build a fused plan from a High op list.
zh_status plan_program(const std::vector<zh_descriptor>& textures, const std::vector<zh_op>& ops,
                       const std::vector<std::vector<uint8_t>>& params, uint32_t flags, Plan* out);

}  // namespace zh

struct zh_context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  void* d_lut = nullptr;         // zh::Luts (8-bit sRGB tables)
  int cus = 256;                 // compute units of the device
  uint32_t* d_error = nullptr;   // kernel error flag
  void* scratch = nullptr;       // the device's static scratch (zh_program_launch_static)
  size_t scratch_bytes = 0;
  // Images and programs hold the context; zh_context_destroy of a context
  // with live objects defers the teardown to the release of the last one.
  int refs = 0;
  bool closing = false;
};

struct zh_image {
  zh_context* ctx = nullptr;
  zh_descriptor desc{};
  void* ptr = nullptr;
  size_t bytes = 0;
  bool owned = false;
};

// ---- programs (zh_runtime.cpp, zh_surrogate.cpp) --------------------------
struct Unit {
  bool batch = false;
  bool warp = false;             // batch of one warp form (k_warp_batch<form>)
  bool spec = false;             // batch of one pointwise signature (k_spec_batch<signature>)
  bool staged = false;           // warp stages on the LDS-staged kernels
  // Two warp forms in one grid (warp_pair): ids[0, split) are the first
  // form's stages (the batch above), ids[split, ..) the second's (below).
  bool pair = false;
  uint32_t split = 0;
  zh::BatchDesc* d_batch2 = nullptr;
  uint32_t* d_prefix2 = nullptr;
  uint32_t* d_ids2 = nullptr;
  uint32_t total_tiles2 = 0;
  hipFunction_t pair_fn = nullptr;
  uint32_t pair_resident = 1;    // resident workgroups per CU of pair_fn
  std::vector<uint32_t> ids;     // stages of the unit (one unless batched)
  uint32_t* d_ids = nullptr;     // device copy of ids (batches)
  zh::BatchDesc* d_batch = nullptr;  // device
  uint32_t* d_counter = nullptr;
  uint32_t* d_prefix = nullptr;
  uint32_t grid = 0, lds = 0;
  uint32_t total_tiles = 0;
  uint32_t blocks_per_cu = 0;    // persistent warp grids: 0 = as many as are resident
  uint32_t strip_w = 0;          // warp units: tiles dealt in strips this many tiles wide (0: by rows)
};

struct zh_program {
  zh_context* ctx = nullptr;
  uint32_t flags = 0;
  std::vector<zh_descriptor> textures;
  std::vector<zh_op> ops;
  std::vector<std::vector<uint8_t>> params;
  zh::Plan plan;
  std::vector<Unit> units;
  zh::StageDesc* d_stages = nullptr;       // device copies, one per stage
  zh::StageDesc* h_stages = nullptr;       // host copy of the descriptors (pinned)
  // Descriptor uploads of successive launches go through a ring of pinned
  // staging copies, each reused only after the event recorded behind its
  // last upload has passed: a launch still queued keeps reading the
  // descriptors it was bound with, whatever later launches bind.
  static constexpr int kRing = 3;
  zh::StageDesc* h_ring[kRing] = {nullptr, nullptr, nullptr};
  hipEvent_t ring_ev[kRing] = {nullptr, nullptr, nullptr};
  bool ring_used[kRing] = {false, false, false};
  int ring_next = 0;
  std::vector<void*> d_intermediates;
  std::vector<zh_image> buffer_images;   // the program's buffers (Plan::buffer_ops), device memory owned here
  std::vector<uint32_t*> d_tile_layers;  // per stage: the tile-layer masks (warp stages)
  std::vector<uint32_t> tile_layer_words;  // their capacities in words
  std::vector<std::vector<uint32_t>> h_tile_layers;  // host copies of the masks uploaded
  std::vector<void*> bound_inputs, bound_outputs;
  bool uploaded = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool graph_failed = false;
  std::string graph_error;     // why the capture failed (graph_failed)
  uint32_t graph_nodes = 0;    // kernel nodes of the captured graph
  uint32_t graph_captures = 0; // captures so far (a rebinding that changes kernels captures again)
  uint64_t graph_replays = 0;  // launches that replayed the graph
  uint32_t max_lds = 0;
  // Kernel per stage for the bound images: 0 general interpreter, 1
  // interpreter with in-line codecs, 2 specialised pointwise, 3 warp, 5 the
  // learned surrogate (an image the faster kernel cannot address forces 0).
  std::vector<int> variant;
  // Learned colour surrogates per stage (zh_surrogate.cpp): packed matrix-core
  // operands (nullptr = none), the lattice check's verdict, master weights.
  std::vector<uint16_t*> d_surrogate;
  std::vector<uint8_t> surrogate_ok;
  std::vector<std::vector<float>> surrogate_params;
  int use_surrogates = 1;  // 0 plain kernels, 1 accepted surrogates, 2 every fitted surrogate
  // A program launched from the static scratch (zh_program_launch_static):
  // its device tables are carved from `arena` in 256-byte steps instead of
  // allocated, its host copy of the descriptors lives in h_stage_vec, and
  // nothing is freed.
  uint8_t* arena = nullptr;
  size_t arena_size = 0, arena_used = 0;
  std::vector<zh::StageDesc> h_stage_vec;
  // User stages: the loaded kernel and the device copy of the shader data,
  // per entry of plan.dynamics.
  std::vector<hipFunction_t> dyn_fn;
  std::vector<void*> dyn_data;
  // Run-time kernels per stage (StagePlan::jit): stage and batch forms and
  // their resident workgroups per CU.
  std::vector<hipFunction_t> jit_stage_fn, jit_batch_fn;
  std::vector<uint32_t> jit_resident, jit_batch_resident;
  // The stepped execution (zh_program_start_pool) still launching this
  // program's kernels; other launches wait for it to finish.
  zh_execution* stepping = nullptr;
};

namespace zh {
// Launch pieces for stepped executions (zh_runtime.cpp).
zh_status program_bind(zh_program* p, const zh_binding* bindings, uint32_t num_bindings);
size_t program_units(const zh_program* p);
zh_status program_launch_units(zh_program* p, size_t first, size_t count);
uint64_t program_device_bytes(const zh_program* p);
}  // namespace zh


