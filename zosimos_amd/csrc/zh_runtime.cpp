// zh_runtime.cpp -- contexts, images, programs and launches of the C ABI in
// include/zosimos_hip.h.
//
// A program owns its planned stages, one device copy of every stage
// descriptor, the intermediate buffers of the plan and (unless disabled) a
// hipGraph of its launches, captured at the first launch and replayed after.
// Launch-time binding patches image addresses into the host copies of the
// descriptors and re-uploads only when they changed, so steady-state launches
// are one graph launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <set>
#include <sstream>

#include "zh_internal.h"

extern "C" {
hipError_t zh_launch_pointwise(const zh::StageDesc* d_dev, uint32_t width, uint32_t height, int flat, int fast,
                               uint32_t cus, hipStream_t stream);
hipError_t zh_launch_spec(int index, const zh::StageDesc* d_dev, uint32_t width, uint32_t height,
                          uint32_t cus, hipStream_t stream);
hipError_t zh_launch_spec_batch(int index, const zh::BatchDesc* b_dev, uint32_t total_chunks, uint32_t cus,
                                hipStream_t stream);
hipError_t zh_launch_warp(int index, int staged, const zh::StageDesc* d_dev, uint32_t tiles, uint32_t cus,
                          uint32_t per_cu, hipStream_t stream);
hipError_t zh_launch_warp_batch(int index, int staged, const zh::BatchDesc* b_dev, uint32_t total_tiles,
                                uint32_t cus, uint32_t per_cu, hipStream_t stream);
hipError_t zh_launch_resample(const zh::StageDesc* d_dev, int mode, uint32_t tiles_x, uint32_t tiles_y,
                              uint32_t lds_bytes, hipStream_t stream);
hipError_t zh_launch_batch(const zh::BatchDesc* b_dev, uint32_t* counter, uint32_t grid, uint32_t lds_bytes,
                           hipStream_t stream);
hipError_t zh_launch_copy(const zh::StageDesc* d_dev, uint64_t row_bytes, uint32_t rows, hipStream_t stream);
hipError_t zh_set_resample_lds(uint32_t bytes);
hipError_t zh_scratch_setup(int device, void** ptr, size_t* bytes);
}
#include "zh_mlp.h"

namespace zh {

static thread_local std::string g_error = "no error";

zh_status fail(zh_status code, const std::string& msg) {
  g_error = msg;
  return code;
}

static zh_status hip_fail(hipError_t e, const char* what) {
  return fail(ZH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define ZH_HIP(call)                                 \
  do {                                               \
    hipError_t _e = (call);                          \
    if (_e != hipSuccess) return hip_fail(_e, #call); \
  } while (0)

// IEC 61966-2-1 in FP64.
static double srgb_eotf(double v) { return v <= 0.04045 ? v / 12.92 : std::pow((v + 0.055) / 1.055, 2.4); }
static double srgb_oetf(double l) { return l <= 0.0031308 ? l * 12.92 : 1.055 * std::pow(l, 1.0 / 2.4) - 0.055; }
// 8-bit code of a linear value: clamp, OETF, round to nearest (even).
static uint32_t srgb_code(double l) {
  if (!(l > 0.0)) return 0;
  if (l >= 1.0) return 255;
  return (uint32_t)std::nearbyint(255.0 * srgb_oetf(l));
}

// The context's 8-bit sRGB tables (zh_stage.h Luts).
static void srgb_tables(Luts* L) {
  std::memset(L, 0, sizeof(*L));
  for (int k = 0; k < 256; ++k) L->dec[k] = (float)srgb_eotf(k / 255.0);
  for (int i = 0; i <= 4096; ++i) L->idx[i] = (uint8_t)srgb_code(i / 4096.0);
  // thr[k]: smallest non-negative f32 whose code is >= k, by bisection over
  // the ordered bit patterns of non-negative floats.
  for (int k = 1; k < 256; ++k) {
    uint32_t lo = 0, hi = 0x3f800000u;  // code(0) < k <= code(1.0f)
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      float f;
      std::memcpy(&f, &mid, 4);
      if (srgb_code(f) >= (uint32_t)k) hi = mid; else lo = mid;
    }
    std::memcpy(&L->thr[k], &hi, 4);
  }
  for (int k = 256; k < 260; ++k) L->thr[k] = INFINITY;
}

}  // namespace zh

using namespace zh;

// One launch unit: a single stage, or a batch of independent staged
// resample stages run by one persistent tile-queue kernel.
extern "C" {

const char* zh_last_error(void) { return g_error.c_str(); }

zh_status zh_srgb_tables(void* out, size_t len) {
  if (!out || len < sizeof(Luts)) return fail(ZH_ERR_INVALID_ARGUMENT, "buffer smaller than the tables");
  Luts L;
  srgb_tables(&L);
  std::memcpy(out, &L, sizeof(L));
  return ZH_OK;
}
const char* zh_version(void) { return "zosimos-hip 0.1.0 (gfx950)"; }

zh_status zh_context_create(int32_t device, void* stream, zh_context** out) {
  if (!out) return fail(ZH_ERR_INVALID_ARGUMENT, "null output pointer");
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return fail(ZH_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= count) return fail(ZH_ERR_INVALID_ARGUMENT, "device ordinal out of range");
  ZH_HIP(hipSetDevice(device));
  zh_context* ctx = new zh_context();
  ctx->device = device;
  if (stream) {
    ctx->stream = (hipStream_t)stream;
  } else {
    // Default priority on purpose (streams of other priorities are avoided).
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete ctx; return hip_fail(e, "hipStreamCreate"); }
    ctx->own_stream = true;
  }
  Luts lut;
  srgb_tables(&lut);
  e = hipMalloc(&ctx->d_lut, sizeof(lut));
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess) e = hipMalloc(&ctx->d_error, sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemcpy(ctx->d_lut, &lut, sizeof(lut), hipMemcpyHostToDevice);
  uint32_t zero = 0;
  if (e == hipSuccess) e = hipMemcpy(ctx->d_error, &zero, sizeof(zero), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = zh_set_resample_lds(kTileLdsBytes);
  // The device address of the static scratch, resolved once per device
  // with the device current (zh_program_launch_static).
  if (e == hipSuccess) e = zh_scratch_setup(device, &ctx->scratch, &ctx->scratch_bytes);
  if (e != hipSuccess) {
    zh_context_destroy(ctx);
    return hip_fail(e, "context setup");
  }
  *out = ctx;
  return ZH_OK;
}

static void context_teardown(zh_context* ctx) {
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  if (ctx->d_lut) hipFree(ctx->d_lut);
  if (ctx->d_error) hipFree(ctx->d_error);
  if (ctx->own_stream) hipStreamDestroy(ctx->stream);
  delete ctx;
}

static void context_retain(zh_context* ctx) {
  if (ctx) ctx->refs++;
}

static void context_release(zh_context* ctx) {
  if (!ctx) return;
  if (--ctx->refs == 0 && ctx->closing) context_teardown(ctx);
}

zh_status zh_context_destroy(zh_context* ctx) {
  if (!ctx || ctx->closing) return ZH_OK;
  ctx->closing = true;
  if (ctx->refs == 0) context_teardown(ctx);
  return ZH_OK;
}

void* zh_context_stream(zh_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

zh_status zh_sync(zh_context* ctx) {
  if (!ctx) return fail(ZH_ERR_INVALID_ARGUMENT, "null context");
  ZH_HIP(hipSetDevice(ctx->device));
  ZH_HIP(hipStreamSynchronize(ctx->stream));
  uint32_t flag = 0;
  ZH_HIP(hipMemcpy(&flag, ctx->d_error, sizeof(flag), hipMemcpyDeviceToHost));
  if (flag) {
    uint32_t zero = 0;
    hipMemcpy(ctx->d_error, &zero, sizeof(zero), hipMemcpyHostToDevice);
    return fail(ZH_ERR_UNSUPPORTED, "a kernel met a tile larger than its planned LDS box");
  }
  return ZH_OK;
}

zh_status zh_query(zh_context* ctx, int32_t* running) {
  if (!ctx || !running) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  const hipError_t e = hipStreamQuery(ctx->stream);
  if (e == hipSuccess) { *running = 0; return ZH_OK; }
  if (e == hipErrorNotReady) { *running = 1; return ZH_OK; }
  return hip_fail(e, "hipStreamQuery");
}

// ---- images -----------------------------------------------------------------
static uint64_t image_bytes(const zh_descriptor& d) { return row_stride_of(d) * d.height; }

zh_status zh_image_alloc(zh_context* ctx, const zh_descriptor* desc, zh_image** out) {
  if (!ctx || !desc || !out) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  if (ctx->closing) return fail(ZH_ERR_INVALID_ARGUMENT, "context is being destroyed");
  std::string why;
  if (!descriptor_valid(*desc, &why)) return fail(ZH_ERR_INVALID_ARGUMENT, "descriptor: " + why);
  ZH_HIP(hipSetDevice(ctx->device));
  zh_image* img = new zh_image();
  img->ctx = ctx;
  img->desc = *desc;
  img->desc.texel_stride = bits_bytes(desc->texel.bits);
  img->bytes = image_bytes(*desc);
  hipError_t e = hipMalloc(&img->ptr, img->bytes);
  if (e != hipSuccess) {
    delete img;
    return fail(ZH_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  img->owned = true;
  context_retain(ctx);
  *out = img;
  return ZH_OK;
}

zh_status zh_image_wrap(zh_context* ctx, const zh_descriptor* desc, void* device_ptr, size_t bytes,
                        zh_image** out) {
  if (!ctx || !desc || !out || !device_ptr) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  std::string why;
  if (!descriptor_valid(*desc, &why)) return fail(ZH_ERR_INVALID_ARGUMENT, "descriptor: " + why);
  if (bytes < image_bytes(*desc)) return fail(ZH_ERR_INVALID_ARGUMENT, "wrapped buffer smaller than the image");
  zh_image* img = new zh_image();
  img->ctx = ctx;
  img->desc = *desc;
  img->desc.texel_stride = bits_bytes(desc->texel.bits);
  img->ptr = device_ptr;
  img->bytes = bytes;
  img->owned = false;
  context_retain(ctx);
  *out = img;
  return ZH_OK;
}

zh_status zh_image_free(zh_image* img) {
  if (!img) return ZH_OK;
  if (img->owned && img->ptr) {
    hipSetDevice(img->ctx->device);
    hipStreamSynchronize(img->ctx->stream);
    hipFree(img->ptr);
  }
  zh_context* ctx = img->ctx;
  delete img;
  context_release(ctx);
  return ZH_OK;
}

zh_status zh_image_device_ptr(zh_image* img, void** ptr, size_t* bytes) {
  if (!img) return fail(ZH_ERR_INVALID_ARGUMENT, "null image");
  if (ptr) *ptr = img->ptr;
  if (bytes) *bytes = img->bytes;
  return ZH_OK;
}

zh_status zh_image_upload(zh_image* img, const void* host, uint64_t host_row_stride) {
  if (!img || !host) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  const uint64_t row = tight_stride(img->desc);
  if (host_row_stride == 0) host_row_stride = row;
  if (host_row_stride < row) return fail(ZH_ERR_INVALID_ARGUMENT, "host row stride shorter than a row");
  ZH_HIP(hipSetDevice(img->ctx->device));
  const uint64_t dev_stride = row_stride_of(img->desc);
  if (host_row_stride == dev_stride) {
    // Same pitch on both sides: one linear copy (a pitched copy from
    // pageable memory moves row by row).
    const uint64_t n = dev_stride * (img->desc.height - 1) + row;
    ZH_HIP(hipMemcpyAsync(img->ptr, host, n, hipMemcpyHostToDevice, img->ctx->stream));
    return ZH_OK;
  }
  ZH_HIP(hipMemcpy2DAsync(img->ptr, dev_stride, host, host_row_stride, row, img->desc.height,
                          hipMemcpyHostToDevice, img->ctx->stream));
  return ZH_OK;
}

zh_status zh_image_download(zh_image* img, void* host, uint64_t host_row_stride) {
  if (!img || !host) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  const uint64_t row = tight_stride(img->desc);
  if (host_row_stride == 0) host_row_stride = row;
  if (host_row_stride < row) return fail(ZH_ERR_INVALID_ARGUMENT, "host row stride shorter than a row");
  ZH_HIP(hipSetDevice(img->ctx->device));
  const uint64_t dev_stride = row_stride_of(img->desc);
  if (host_row_stride == dev_stride) {
    const uint64_t n = dev_stride * (img->desc.height - 1) + row;
    ZH_HIP(hipMemcpyAsync(host, img->ptr, n, hipMemcpyDeviceToHost, img->ctx->stream));
    return ZH_OK;
  }
  ZH_HIP(hipMemcpy2DAsync(host, host_row_stride, img->ptr, dev_stride, row, img->desc.height,
                          hipMemcpyDeviceToHost, img->ctx->stream));
  return ZH_OK;
}

// ---- programs -----------------------------------------------------------------
// Device memory of a program's tables: allocated, or carved in 256-byte steps
// from the static scratch for a program of zh_program_launch_static.
static zh_status dev_alloc(zh_program* p, size_t bytes, void** out) {
  if (p->arena) {
    const size_t at = (p->arena_used + 255) & ~(size_t)255;
    if (at > p->arena_size || bytes > p->arena_size - at)
      return fail(ZH_ERR_OUT_OF_MEMORY, "the program's device tables need more than the static scratch of " +
                                            std::to_string(p->arena_size >> 20) + " MiB (ZH_STATIC_SCRATCH_BYTES)");
    *out = p->arena + at;
    p->arena_used = at + bytes;
    return ZH_OK;
  }
  const hipError_t e = hipMalloc(out, bytes);
  if (e != hipSuccess) return fail(ZH_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  return ZH_OK;
}
#define ZH_ALLOC(p, bytes, ptr)                                      \
  do {                                                                \
    zh_status _s = dev_alloc((p), (bytes), (void**)(ptr));            \
    if (_s != ZH_OK) return _s;                                       \
  } while (0)

static void drop_surrogates(zh_program* p) {
  for (uint16_t*& b : p->d_surrogate)
    if (b) { hipFree(b); b = nullptr; }
  p->d_surrogate.clear();
  p->surrogate_ok.clear();
  p->surrogate_params.clear();
}

static void free_device(zh_program* p) {
  drop_surrogates(p);
  if (p->arena) {
    // Static-scratch tables are never freed; only the host side is dropped.
    p->units.clear();
    p->d_intermediates.clear();
    p->dyn_data.clear();
    p->dyn_fn.clear();
    p->buffer_images.clear();
    p->d_tile_layers.clear();
    p->tile_layer_words.clear();
    p->h_tile_layers.clear();
    p->d_stages = nullptr;
    p->h_stages = nullptr;
    p->h_stage_vec.clear();
    return;
  }
  if (p->graph_exec) hipGraphExecDestroy(p->graph_exec);
  if (p->graph) hipGraphDestroy(p->graph);
  p->graph_exec = nullptr;
  p->graph = nullptr;
  for (Unit& u : p->units) {
    if (u.d_batch) hipFree(u.d_batch);
    if (u.d_counter) hipFree(u.d_counter);
    if (u.d_prefix) hipFree(u.d_prefix);
    if (u.d_ids) hipFree(u.d_ids);
    if (u.d_batch2) hipFree(u.d_batch2);
    if (u.d_prefix2) hipFree(u.d_prefix2);
    if (u.d_ids2) hipFree(u.d_ids2);
  }
  p->units.clear();
  for (void* b : p->d_intermediates) hipFree(b);
  p->d_intermediates.clear();
  for (void* b : p->dyn_data) hipFree(b);
  p->dyn_data.clear();
  p->dyn_fn.clear();
  for (zh_image& b : p->buffer_images)
    if (b.ptr) hipFree(b.ptr);
  p->buffer_images.clear();
  for (uint32_t* m : p->d_tile_layers)
    if (m) hipFree(m);
  p->d_tile_layers.clear();
  p->tile_layer_words.clear();
  p->h_tile_layers.clear();
  if (p->d_stages) hipFree(p->d_stages);
  if (p->h_stages) hipHostFree(p->h_stages);
  p->d_stages = nullptr;
  p->h_stages = nullptr;
  for (int k = 0; k < zh_program::kRing; ++k) {
    if (p->h_ring[k]) hipHostFree(p->h_ring[k]);
    if (p->ring_ev[k]) hipEventDestroy(p->ring_ev[k]);
    p->h_ring[k] = nullptr;
    p->ring_ev[k] = nullptr;
    p->ring_used[k] = false;
  }
  p->ring_next = 0;
}

// Group stages into launch units.  A run of independent resample stages
// (none reads what an earlier one of the run writes) is batched: warp-form
// stages by form, one persistent tile-queue launch per form, and LDS-staged
// interpreter stages together; everything else launches alone.
static void make_units(zh_program* p) {
  p->units.clear();
  const auto& st = p->plan.stages;
  // 1: warp stages (partitioned by form), 2: LDS-staged resample stages,
  // 3: flat pointwise stages with a specialised kernel (by signature).
  auto family = [&](const StagePlan& sp) {
    if (sp.desc.kind == STAGE_POINTWISE) return sp.spec >= 0 || sp.jit >= 0 ? 3 : 0;
    if (sp.desc.kind != STAGE_RESAMPLE) return 0;
    if (sp.desc.variant >= 0) return 1;
    return sp.tile_mode == 0 ? 2 : 0;
  };
  size_t i = 0;
  while (i < st.size()) {
    const int fam = family(st[i]);
    size_t j = i + 1;
    if (fam) {
      std::vector<int64_t> written;
      for (size_t k = 0; k < st[i].slot_ref.size(); ++k)
        if (st[i].slot_ref[k] < 0 || st[i].slot_ref[k] >= kIntermediateBase) written.push_back(st[i].slot_ref[k]);
      while (j < st.size() && family(st[j]) == fam) {
        bool dep = false;
        for (int64_t r : st[j].slot_ref)
          for (int64_t w : written)
            if (r == w && r >= kIntermediateBase) dep = true;
        if (dep) break;
        for (int64_t r : st[j].slot_ref)
          if (r < 0 || r >= kIntermediateBase) written.push_back(r);
        ++j;
      }
    }
    // Partition [i, j): by warp form, or as one group.
    std::vector<std::pair<int, std::vector<uint32_t>>> groups;
    for (size_t k = i; k < j; ++k) {
      const int key = fam == 1   ? 2 * st[k].desc.variant + (st[k].staged ? 1 : 0)
                      : fam == 3 ? (st[k].spec >= 0 ? st[k].spec : 100000 + st[k].jit)
                                 : -1;
      auto it = std::find_if(groups.begin(), groups.end(), [&](const auto& g) { return g.first == key; });
      if (it == groups.end()) groups.push_back({key, {(uint32_t)k}});
      else it->second.push_back((uint32_t)k);
    }
    // Under ZH_PLAN_PAIR, two warp forms of one run (say, a batch of RGBA8
    // and RGBA16F images) share one persistent grid that deals their tiles
    // alternately.
    const bool pair = fam == 1 && groups.size() == 2 && (p->flags & ZH_PLAN_PAIR) && !(p->flags & ZH_PLAN_NO_JIT) &&
                      !p->arena &&
                      !st[groups[0].second[0]].staged && !st[groups[1].second[0]].staged;
    if (pair) {
      Unit u;
      u.ids = groups[0].second;
      u.split = (uint32_t)u.ids.size();
      u.ids.insert(u.ids.end(), groups[1].second.begin(), groups[1].second.end());
      u.batch = u.warp = u.pair = true;
      p->units.push_back(u);
      i = j;
      continue;
    }
    for (auto& g : groups) {
      Unit u;
      u.ids = g.second;
      u.batch = fam && u.ids.size() >= 2;
      u.warp = u.batch && fam == 1;
      u.spec = u.batch && fam == 3;
      u.staged = fam == 1 && st[u.ids[0]].staged;
      p->units.push_back(u);
    }
    i = j;
  }
}

// The kernels' per-tile layer test (warp_tile in zh_warp.hip) on the host:
// does any bilinear tap of the tile's pixel centres land in the layer's
// source?  Widened by 1e-6 texel so that rounding never drops a tile the
// kernel's own test would keep (a kept tile whose taps all miss adds 0).
static bool tile_meets_layer(const Layer& L, double ex0, double ey0, double ex1, double ey1) {
  if (L.covers_all) return true;
  const double* m = L.inv;
  const double u[4] = {m[0] * ex0 + m[1] * ey0 + m[2], m[0] * ex1 + m[1] * ey0 + m[2], m[0] * ex0 + m[1] * ey1 + m[2],
                       m[0] * ex1 + m[1] * ey1 + m[2]};
  const double v[4] = {m[3] * ex0 + m[4] * ey0 + m[5], m[3] * ex1 + m[4] * ey0 + m[5], m[3] * ex0 + m[4] * ey1 + m[5],
                       m[3] * ex1 + m[4] * ey1 + m[5]};
  const double umin = std::min(std::min(u[0], u[1]), std::min(u[2], u[3])) - 0.5;
  const double umax = std::max(std::max(u[0], u[1]), std::max(u[2], u[3])) - 0.5;
  const double vmin = std::min(std::min(v[0], v[1]), std::min(v[2], v[3])) - 0.5;
  const double vmax = std::max(std::max(v[0], v[1]), std::max(v[2], v[3])) - 0.5;
  const double e = 1e-6;
  return umax > -1.0 - e && vmax > -1.0 - e && umin < (double)L.src_w + e && vmin < (double)L.src_h + e;
}

// The tile-layer masks of warp stage s for its current tile shape (host
// computation, uploaded once; the kernels then read one byte per tile).
static zh_status tile_layer_masks(zh_program* p, size_t s) {
  StageDesc& d = p->plan.stages[s].desc;
  if (d.kind != STAGE_RESAMPLE || d.variant < 0 || d.n_layers == 0 || d.n_layers > 8) return ZH_OK;
  if (p->d_tile_layers.size() < p->plan.stages.size()) {
    p->d_tile_layers.resize(p->plan.stages.size(), nullptr);
    p->tile_layer_words.resize(p->plan.stages.size(), 0);
  }
  const uint32_t tiles = d.tiles_x * d.tiles_y, words = (tiles + 3) / 4;
  std::vector<uint32_t> m(std::max(1u, words), 0u);
  const int TW = (int)d.tile_w, TH = (int)d.tile_h, W = (int)d.width, H = (int)d.height;
  for (uint32_t ty = 0; ty < d.tiles_y; ++ty)
    for (uint32_t tx = 0; tx < d.tiles_x; ++tx) {
      const double ex0 = (double)tx * TW + 0.5, ey0 = (double)ty * TH + 0.5;
      const double ex1 = std::min((int)tx * TW + TW, W) - 0.5, ey1 = std::min((int)ty * TH + TH, H) - 0.5;
      uint32_t bits = 0;
      for (uint32_t l = 0; l < d.n_layers; ++l)
        if (tile_meets_layer(d.layer[l], ex0, ey0, ex1, ey1)) bits |= 1u << l;
      const uint32_t t = ty * d.tiles_x + tx;
      m[t >> 2] |= bits << ((t & 3u) * 8u);
    }
  if (p->h_tile_layers.size() < p->plan.stages.size()) p->h_tile_layers.resize(p->plan.stages.size());
  if (p->d_tile_layers[s] && p->h_tile_layers[s] == m) {
    // Unchanged (a knob that moved no layer): nothing to upload.
    d.tile_layers = p->d_tile_layers[s];
    if (p->h_stages) p->h_stages[s].tile_layers = d.tile_layers;
    return ZH_OK;
  }
  // Launches still in flight may read the old masks.
  ZH_HIP(hipStreamSynchronize(p->ctx->stream));
  if (p->tile_layer_words[s] < m.size()) {
    if (p->d_tile_layers[s] && !p->arena) hipFree(p->d_tile_layers[s]);
    p->d_tile_layers[s] = nullptr;
    p->tile_layer_words[s] = 0;
    ZH_ALLOC(p, m.size() * sizeof(uint32_t), &p->d_tile_layers[s]);
    p->tile_layer_words[s] = (uint32_t)m.size();
  }
  ZH_HIP(hipMemcpy(p->d_tile_layers[s], m.data(), m.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  p->h_tile_layers[s] = std::move(m);
  d.tile_layers = p->d_tile_layers[s];
  if (p->h_stages) p->h_stages[s].tile_layers = d.tile_layers;
  return ZH_OK;
}

// Copy the bytes of the program's buffers to their device images (the
// stream is idle: at set-up, or after a knob change synchronised it).
static zh_status upload_buffers(zh_program* p) {
  const Plan& plan = p->plan;
  for (size_t k = 0; k < plan.buffer_ops.size(); ++k) {
    const zh_image& b = p->buffer_images[k];
    ZH_HIP(hipMemcpy(b.ptr, p->params[plan.buffer_ops[k]].data(), b.bytes, hipMemcpyHostToDevice));
  }
  return ZH_OK;
}

// The tile queue of stages ids[0, n): their tile (or chunk) prefix, the ids
// and the BatchDesc, in device memory.
static zh_status alloc_batch(zh_program* p, const uint32_t* ids, size_t n, bool spec, uint32_t* counter,
                             BatchDesc** d_batch, uint32_t** d_prefix, uint32_t** d_ids, uint32_t* total) {
  std::vector<uint32_t> prefix(n + 1, 0);
  for (size_t k = 0; k < n; ++k) {
    const StageDesc& d = p->plan.stages[ids[k]].desc;
    const uint64_t chunks = ((uint64_t)d.width * d.height + kSpecChunkPixels - 1) / kSpecChunkPixels;
    prefix[k + 1] = prefix[k] + (spec ? (uint32_t)chunks : d.tiles_x * d.tiles_y);
  }
  ZH_ALLOC(p, sizeof(BatchDesc), d_batch);
  ZH_ALLOC(p, prefix.size() * sizeof(uint32_t), d_prefix);
  ZH_ALLOC(p, std::max<size_t>(n, 1) * sizeof(uint32_t), d_ids);
  ZH_HIP(hipMemcpy(*d_prefix, prefix.data(), prefix.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  ZH_HIP(hipMemcpy(*d_ids, ids, n * sizeof(uint32_t), hipMemcpyHostToDevice));
  BatchDesc b{};
  b.n_stages = (uint32_t)n;
  b.total_tiles = prefix.back();
  b.counter = counter;
  b.stages = p->d_stages;
  b.stage_ids = *d_ids;
  b.tile_prefix = *d_prefix;
  ZH_HIP(hipMemcpy(*d_batch, &b, sizeof(b), hipMemcpyHostToDevice));
  *total = b.total_tiles;
  return ZH_OK;
}

// Unit ui, a pair of warp forms, as the two launches of its forms.
static void split_pair(zh_program* p, size_t ui) {
  const Unit u = p->units[ui];
  Unit a, b;
  a.ids.assign(u.ids.begin(), u.ids.begin() + u.split);
  b.ids.assign(u.ids.begin() + u.split, u.ids.end());
  for (Unit* x : {&a, &b}) {
    x->batch = x->ids.size() >= 2;
    x->warp = x->batch;
  }
  p->units[ui] = a;
  p->units.insert(p->units.begin() + (ptrdiff_t)ui + 1, b);
}

static zh_status alloc_device(zh_program* p) {
  zh_context* ctx = p->ctx;
  ZH_HIP(hipSetDevice(ctx->device));
  const size_t n = p->plan.stages.size();
  ZH_ALLOC(p, std::max<size_t>(1, n) * sizeof(StageDesc), &p->d_stages);
  if (p->arena) {
    // Host memory of the caller's process heap; uploads are synchronous.
    p->h_stage_vec.assign(std::max<size_t>(1, n), StageDesc{});
    p->h_stages = p->h_stage_vec.data();
  } else {
    ZH_HIP(hipHostMalloc(&p->h_stages, std::max<size_t>(1, n) * sizeof(StageDesc), hipHostMallocDefault));
    for (int k = 0; k < zh_program::kRing; ++k) {
      ZH_HIP(hipHostMalloc(&p->h_ring[k], std::max<size_t>(1, n) * sizeof(StageDesc), hipHostMallocDefault));
      ZH_HIP(hipEventCreateWithFlags(&p->ring_ev[k], hipEventDisableTiming));
    }
  }
  for (const zh_descriptor& d : p->plan.intermediates) {
    void* b = nullptr;
    ZH_ALLOC(p, image_bytes(d), &b);
    p->d_intermediates.push_back(b);
  }
  // Run-time kernels of unregistered pointwise signatures (none from the
  // static scratch: loading a code object allocates).  A kernel that cannot
  // be had leaves its stage on the interpreter, and the description says why.
  p->jit_stage_fn.assign(p->plan.stages.size(), nullptr);
  p->jit_batch_fn.assign(p->plan.stages.size(), nullptr);
  p->jit_resident.assign(p->plan.stages.size(), 1);
  p->jit_batch_resident.assign(p->plan.stages.size(), 1);
  for (size_t s = 0; s < p->plan.stages.size(); ++s) {
    StagePlan& sp = p->plan.stages[s];
    if (sp.desc.kind == STAGE_RESAMPLE && sp.desc.variant >= kJitWarpBase) {
      // A run-time warp form; without it the stage runs on the resample
      // interpreter, whose kernels accept the same tiles.
      std::string why;
      hipFunction_t f = nullptr, fb = nullptr;
      const int id = (sp.desc.variant - kJitWarpBase) / 2;
      const bool sep = (sp.desc.variant & 1) != 0;
      if (p->arena || jit_warp_kernels(id, sep, sp.staged, ctx->device, &f, &fb, &why) != ZH_OK) {
        sp.desc.variant = -1;
        sp.staged = false;
        const size_t k = sp.text.rfind(" [kernel jitw#");
        if (k != std::string::npos)
          sp.text = sp.text.substr(0, k) +
                    (p->arena ? std::string() : " (run-time kernel unavailable: " + why.substr(0, 120) + ")");
        continue;
      }
      int n = 0;
      if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, kBlock, 0) == hipSuccess && n > 0)
        p->jit_resident[s] = (uint32_t)n;
      n = 0;
      if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&n, fb, kBlock, 0) == hipSuccess && n > 0)
        p->jit_batch_resident[s] = (uint32_t)n;
      p->jit_stage_fn[s] = f;
      p->jit_batch_fn[s] = fb;
      continue;
    }
    if (sp.jit < 0) continue;
    std::string why;
    hipFunction_t f = nullptr, fb = nullptr;
    const size_t k = sp.text.rfind(" [kernel jit#");
    if (p->arena || jit_kernels(sp.jit, ctx->device, &f, &fb, &why) != ZH_OK) {
      sp.jit = -1;
      if (k != std::string::npos)
        sp.text = sp.text.substr(0, k) + " [kernel interp-inline]" +
                  (p->arena ? std::string() : " (run-time kernel unavailable: " + why.substr(0, 120) + ")");
      continue;
    }
    int n = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, kBlock, 0) == hipSuccess && n > 0)
      p->jit_resident[s] = (uint32_t)n;
    n = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&n, fb, kBlock, 0) == hipSuccess && n > 0)
      p->jit_batch_resident[s] = (uint32_t)n;
    p->jit_stage_fn[s] = f;
    p->jit_batch_fn[s] = fb;
  }
  // User stages: the kernel on this device and a device copy of the data.
  for (const DynamicCall& dc : p->plan.dynamics) {
    hipFunction_t fn = nullptr;
    zh_status st = shader_function(dc.shader.get(), ctx->device, &fn);
    if (st != ZH_OK) return st;
    void* b = nullptr;
    ZH_ALLOC(p, std::max<size_t>(dc.data.size(), 16), &b);
    if (!dc.data.empty()) ZH_HIP(hipMemcpy(b, dc.data.data(), dc.data.size(), hipMemcpyHostToDevice));
    p->dyn_fn.push_back(fn);
    p->dyn_data.push_back(b);
  }
  make_units(p);
  // The kernels of paired warp forms, compiled at run time; a pair whose
  // kernel cannot be had launches its two forms one after the other.
  for (size_t ui = 0; ui < p->units.size(); ++ui) {
    Unit& u = p->units[ui];
    if (!u.pair) continue;
    WarpForm fa, fb;
    std::string why;
    hipFunction_t fn = nullptr;
    const bool ok = warp_form_of(p->plan.stages[u.ids[0]].desc.variant, &fa) &&
                    warp_form_of(p->plan.stages[u.ids[u.split]].desc.variant, &fb) &&
                    jit_pair_kernel(jit_warp_pair(fa, fb), ctx->device, &fn, &why) == ZH_OK;
    if (!ok) {
      (void)hipGetLastError();
      split_pair(p, ui);
      continue;
    }
    int nb = 0;
    u.pair_fn = fn;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kBlock, 0) == hipSuccess && nb > 0)
      u.pair_resident = (uint32_t)nb;
  }
  for (size_t s = 0; s < n; ++s) {
    zh_status st = tile_layer_masks(p, s);
    if (st != ZH_OK) return st;
  }
  for (size_t k = 0; k < p->plan.buffer_ops.size(); ++k) {
    zh_image b;
    b.ctx = ctx;
    b.desc = p->plan.input_desc[p->plan.user_inputs + k];
    b.bytes = image_bytes(b.desc);
    ZH_ALLOC(p, b.bytes, &b.ptr);
    p->buffer_images.push_back(b);
  }
  {
    zh_status st = upload_buffers(p);
    if (st != ZH_OK) return st;
  }
  for (Unit& u : p->units) {
    if (!u.batch) continue;
    for (size_t k = 0; k < u.ids.size(); ++k) u.lds = std::max(u.lds, p->plan.stages[u.ids[k]].lds_bytes);
    ZH_ALLOC(p, sizeof(uint32_t), &u.d_counter);
    const size_t na = u.pair ? u.split : u.ids.size();
    zh_status st = alloc_batch(p, u.ids.data(), na, u.spec, u.d_counter, &u.d_batch, &u.d_prefix, &u.d_ids,
                               &u.total_tiles);
    if (st == ZH_OK && u.pair)
      st = alloc_batch(p, u.ids.data() + na, u.ids.size() - na, false, u.d_counter, &u.d_batch2, &u.d_prefix2,
                       &u.d_ids2, &u.total_tiles2);
    if (st != ZH_OK) return st;
    // Persistent grid: enough workgroups to fill every CU at the kernel's
    // occupancy (2 per CU with a 64 KiB tile), capped by the work.
    hipDeviceProp_t prop;
    ZH_HIP(hipGetDeviceProperties(&prop, ctx->device));
    u.grid = std::min<uint32_t>(u.total_tiles, (uint32_t)prop.multiProcessorCount * 2);
    if (u.grid == 0) u.grid = 1;
  }
  return ZH_OK;
}

// The program's own copy of the High list (parameter blobs included) and its plan.
static zh_status init_program(zh_program* p, zh_context* ctx, const zh_descriptor* textures, uint32_t num_textures,
                              const zh_op* ops, uint32_t num_ops, uint32_t flags) {
  p->ctx = ctx;
  p->flags = flags;
  p->textures.assign(textures, textures + num_textures);
  p->ops.assign(ops, ops + num_ops);
  p->params.resize(num_ops);
  for (uint32_t i = 0; i < num_ops; ++i) {
    if (ops[i].params_len) {
      if (!ops[i].params) return fail(ZH_ERR_INVALID_ARGUMENT, "params pointer is null");
      const uint8_t* b = static_cast<const uint8_t*>(ops[i].params);
      p->params[i].assign(b, b + ops[i].params_len);
    }
    p->ops[i].params = nullptr;
  }
  return plan_program(p->textures, p->ops, p->params, flags, &p->plan);
}

zh_status zh_program_build(zh_context* ctx, const zh_descriptor* textures, uint32_t num_textures, const zh_op* ops,
                           uint32_t num_ops, uint32_t flags, zh_program** out) {
  if (!out || (num_textures && !textures) || (num_ops && !ops)) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  zh_program* p = new zh_program();
  zh_status st = init_program(p, ctx, textures, num_textures, ops, num_ops, flags);
  if (st != ZH_OK) { delete p; return st; }
  if (ctx) {
    if (ctx->closing) { delete p; return fail(ZH_ERR_INVALID_ARGUMENT, "context is being destroyed"); }
    st = alloc_device(p);
    if (st != ZH_OK) { free_device(p); delete p; return st; }
    context_retain(ctx);
  } else {
    make_units(p);
  }
  *out = p;
  return ZH_OK;
}

zh_status zh_program_destroy(zh_program* p) {
  if (!p) return ZH_OK;
  if (p->ctx) {
    hipSetDevice(p->ctx->device);
    hipStreamSynchronize(p->ctx->stream);
    free_device(p);
  }
  zh_context* ctx = p->ctx;
  delete p;
  context_release(ctx);
  return ZH_OK;
}

zh_status zh_program_launch_bytes(zh_program* p, uint64_t* bytes, uint32_t max_launches, uint32_t* num_launches) {
  if (!p) return fail(ZH_ERR_INVALID_ARGUMENT, "null program");
  if (p->units.empty() && !p->plan.stages.empty()) make_units(p);
  if (num_launches) *num_launches = (uint32_t)p->units.size();
  for (size_t i = 0; i < p->units.size() && i < max_launches && bytes; ++i) {
    uint64_t b = 0;
    for (uint32_t id : p->units[i].ids) b += p->plan.stages[id].alg_bytes;
    bytes[i] = b;
  }
  return ZH_OK;
}

zh_status zh_program_describe(zh_program* p, char* buf, size_t len, size_t* needed) {
  if (!p) return fail(ZH_ERR_INVALID_ARGUMENT, "null program");
  std::ostringstream o;
  for (size_t i = 0; i < p->plan.stages.size(); ++i) o << "stage " << i << ": " << p->plan.stages[i].text << "\n";
  for (size_t i = 0; i < p->units.size(); ++i) {
    const Unit& u = p->units[i];
    const std::string& t = p->plan.stages[u.ids[0]].text;
    const size_t k = t.rfind("[kernel ");
    std::string tag = k == std::string::npos ? std::string("stage ") + std::to_string(u.ids[0]) : t.substr(k);
    if (u.pair) {
      // "[kernel pair <first form> + <second form>]"
      auto name = [&](uint32_t s) {
        const std::string& x = p->plan.stages[s].text;
        const size_t a = x.rfind("[kernel ");
        const size_t b = a == std::string::npos ? a : x.find(']', a);
        return b == std::string::npos ? std::string("?") : x.substr(a + 8, b - a - 8);
      };
      tag = "[kernel pair " + name(u.ids[0]) + " + " + name(u.ids[u.split]) + "]";
    }
    o << "launch " << i << ": " << tag << " over " << u.ids.size() << (u.ids.size() == 1 ? " stage" : " stages")
      << "\n";
    if (p->units[i].batch) {
      o << "batch: stages ";
      for (size_t k = 0; k < p->units[i].ids.size(); ++k) o << (k ? "," : "") << p->units[i].ids[k];
      o << (p->units[i].spec ? " as one launch\n"
                              : p->units[i].pair ? " as one grid dealing two tile queues alternately\n"
                                                 : " as one tile queue\n");
    }
    const StageDesc& d0 = p->plan.stages[u.ids[0]].desc;
    if (d0.kind == STAGE_RESAMPLE) o << "launch " << i << ": tiles " << d0.tile_w << "x" << d0.tile_h << "\n";
    if (p->units[i].blocks_per_cu)
      o << "launch " << i << ": " << p->units[i].blocks_per_cu << " workgroups per CU\n";
    if (p->units[i].strip_w)
      o << "launch " << i << ": tiles in strips " << p->units[i].strip_w << " tiles wide\n";
  }
  {
    std::set<int> shown;
    for (size_t i = 0; i < p->plan.stages.size(); ++i) {
      const int j = p->plan.stages[i].jit;
      if (j >= 0 && i < p->jit_stage_fn.size() && p->jit_stage_fn[i] && shown.insert(j).second)
        o << "jit#" << j << ": compiled at run time in " << jit_compile_seconds(j) << " s\n";
    }
  }
  for (size_t i = 0; i < p->plan.intermediates.size(); ++i) {
    const zh_descriptor& d = p->plan.intermediates[i];
    o << "intermediate " << i << ": " << d.width << "x" << d.height << " bits " << d.texel.bits << "\n";
  }
  if (p->flags & ZH_PLAN_NO_GRAPH) o << "graph: disabled\n";
  else if (p->graph_failed) o << "graph: capture failed (" << p->graph_error << "); launching kernel by kernel\n";
  else if (p->graph_exec) o << "graph: captured, " << p->graph_nodes << " nodes, " << p->graph_replays << " replays\n";
  else if (p->units.size() <= 1) o << "graph: none (one launch)\n";
  else o << "graph: captured at the first launch\n";
  const std::string s = o.str();
  if (needed) *needed = s.size() + 1;
  if (buf && len) {
    const size_t n = std::min(len - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  return ZH_OK;
}

zh_status zh_program_stats(zh_program* p, uint32_t* launches, uint64_t* algorithmic_bytes) {
  if (!p) return fail(ZH_ERR_INVALID_ARGUMENT, "null program");
  if (launches) *launches = (uint32_t)p->units.size();
  if (algorithmic_bytes) {
    uint64_t b = 0;
    for (const StagePlan& s : p->plan.stages) b += s.alg_bytes;
    *algorithmic_bytes = b;
  }
  return ZH_OK;
}

zh_status zh_program_knob_data(zh_program* p, int32_t knob, void* out, uint32_t cap, uint32_t* len) {
  if (!p || !len || (!out && cap)) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  for (size_t i = 0; i < p->ops.size(); ++i)
    if (p->ops[i].knob == knob) {
      const std::vector<uint8_t>& b = p->params[i];
      *len = (uint32_t)b.size();
      if (cap && !b.empty()) std::memcpy(out, b.data(), std::min<size_t>(cap, b.size()));
      return ZH_OK;
    }
  return fail(ZH_ERR_INVALID_ARGUMENT, "no op uses knob " + std::to_string(knob));
}

// A warp stage's kernel tag with or without the "_lds" suffix of the staged
// kernels.
static void set_staged_tag(std::string& text, bool staged) {
  const size_t k = text.rfind(" [kernel ");
  if (k == std::string::npos) return;
  const size_t e = text.find(']', k);
  if (e == std::string::npos) return;
  const bool has = e >= k + 4 && text.compare(e - 4, 4, "_lds") == 0;
  if (staged && !has) text.insert(e, "_lds");
  else if (!staged && has) text.erase(e - 4, 4);
}

zh_status zh_program_set_knob(zh_program* p, int32_t knob, const void* data, uint32_t len) {
  if (!p || (!data && len)) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  bool found = false;
  for (size_t i = 0; i < p->ops.size(); ++i) {
    if (p->ops[i].knob != knob) continue;
    if (p->params[i].size() != len)
      return fail(ZH_ERR_KNOB_LENGTH, "knob data length " + std::to_string(len) + " differs from " +
                                          std::to_string(p->params[i].size()));
    found = true;
  }
  if (!found) return fail(ZH_ERR_INVALID_ARGUMENT, "no op uses knob " + std::to_string(knob));
  std::vector<std::vector<uint8_t>> params = p->params;
  for (size_t i = 0; i < p->ops.size(); ++i)
    if (p->ops[i].knob == knob) {
      const uint8_t* b = static_cast<const uint8_t*>(data);
      params[i].assign(b, b + len);
    }
  Plan np;
  zh_status st = plan_program(p->textures, p->ops, params, p->flags, &np);
  if (st != ZH_OK) return st;
  // Warp stages keep the tiles they run on (the autotuner's choice): a warp
  // kernel's tiles do not depend on the layers' maps.
  if (np.stages.size() == p->plan.stages.size())
    for (size_t i = 0; i < np.stages.size(); ++i) {
      StagePlan& a = np.stages[i];
      const StagePlan& b = p->plan.stages[i];
      if (a.desc.kind != STAGE_RESAMPLE || b.desc.kind != STAGE_RESAMPLE || a.desc.variant < 0 ||
          a.desc.variant != b.desc.variant || a.desc.width != b.desc.width || a.desc.height != b.desc.height)
        continue;
      a.desc.tile_w = b.desc.tile_w;
      a.desc.tile_h = b.desc.tile_h;
      a.desc.band_h = b.desc.band_h;
      a.desc.tiles_x = b.desc.tiles_x;
      a.desc.tiles_y = b.desc.tiles_y;
      a.desc.lds_texels = b.desc.lds_texels;
      a.tile_mode = b.tile_mode;
      a.lds_bytes = b.lds_bytes;
      // The new maps decide whether the box still fits the kept tiles.
      a.staged = (p->flags & ZH_PLAN_STAGED) && warp_staged_fits(a.desc, a.desc.tile_w, a.desc.tile_h);
      a.text = b.text;
      set_staged_tag(a.text, a.staged);
    }
  // Same structure (stage kinds, tiles, slots): keep buffers and the graph and
  // refresh the descriptors at the next launch.
  bool same = np.stages.size() == p->plan.stages.size() && np.intermediates.size() == p->plan.intermediates.size();
  for (size_t i = 0; same && i < np.stages.size(); ++i) {
    const StagePlan& a = np.stages[i];
    const StagePlan& b = p->plan.stages[i];
    same = a.desc.kind == b.desc.kind && a.tile_mode == b.tile_mode && a.desc.tiles_x == b.desc.tiles_x &&
           a.desc.tiles_y == b.desc.tiles_y && a.slot_ref == b.slot_ref && a.lds_bytes <= b.lds_bytes &&
           a.spec == b.spec && a.jit == b.jit && a.fast == b.fast && a.desc.flat == b.desc.flat &&
           a.desc.variant == b.desc.variant && a.staged == b.staged;
  }
  p->params = params;
  // A surrogate was fitted to the old parameters; a buffer knob rewrites the
  // buffer in place: both wait for the launches in flight.  Parameters the
  // stage descriptors carry go through the descriptor ring instead.
  bool buffer_knob = false;
  for (size_t i = 0; i < p->ops.size(); ++i)
    buffer_knob = buffer_knob || (p->ops[i].knob == knob && p->ops[i].high == ZH_HIGH_BUFFER);
  if (p->ctx && (buffer_knob || !p->d_surrogate.empty() || !same)) hipStreamSynchronize(p->ctx->stream);
  drop_surrogates(p);
  if (same) {
    for (size_t i = 0; i < np.stages.size(); ++i) np.stages[i].lds_bytes = p->plan.stages[i].lds_bytes;
    p->plan = std::move(np);
    p->uploaded = false;
    // The new maps move the layers' footprints; buffers take the new bytes.
    if (p->ctx) {
      for (size_t s = 0; s < p->plan.stages.size(); ++s) {
        zh_status ms = tile_layer_masks(p, s);
        if (ms != ZH_OK) return ms;
      }
      if (buffer_knob) {
        zh_status bs = upload_buffers(p);
        if (bs != ZH_OK) return bs;
      }
    }
    return ZH_OK;
  }
  if (p->ctx) {
    hipStreamSynchronize(p->ctx->stream);
    free_device(p);
  }
  p->plan = std::move(np);
  p->uploaded = false;
  p->bound_inputs.clear();
  p->bound_outputs.clear();
  if (p->ctx) return alloc_device(p);
  make_units(p);
  return ZH_OK;
}

// Resolve bindings, patch slot addresses and upload descriptors if needed.
static zh_status bind(zh_program* p, const zh_binding* b, uint32_t n) {
  const Plan& plan = p->plan;
  std::vector<zh_image*> in(plan.input_desc.size(), nullptr), outs(plan.output_desc.size(), nullptr);
  for (uint32_t i = 0; i < n; ++i) {
    if (!b[i].image) return fail(ZH_ERR_INVALID_ARGUMENT, "binding without image");
    if (b[i].image->ctx != p->ctx) return fail(ZH_ERR_INVALID_ARGUMENT, "image belongs to another context");
    auto& vec = b[i].is_output ? outs : in;
    const auto& descs = b[i].is_output ? plan.output_desc : plan.input_desc;
    if (b[i].slot >= (b[i].is_output ? vec.size() : plan.user_inputs))
      return fail(ZH_ERR_INVALID_ARGUMENT, "binding slot out of range");
    const zh_descriptor& want = descs[b[i].slot];
    const zh_descriptor& have = b[i].image->desc;
    // Environment::bind requires the pool image's descriptor to equal the
    // register's (run.rs:1187-1193); row layout included.
    if (!same_layout(want, have))
      return fail(ZH_ERR_MISMATCHED_DESCRIPTOR, std::string(b[i].is_output ? "output" : "input") + " slot " +
                                                    std::to_string(b[i].slot) + " bound to a mismatched image");
    vec[b[i].slot] = b[i].image;
  }
  for (size_t k = 0; k < p->buffer_images.size(); ++k) in[plan.user_inputs + k] = &p->buffer_images[k];
  for (size_t i = 0; i < in.size(); ++i)
    if (plan.input_used[i] && !in[i]) return fail(ZH_ERR_MISSING_KEY, "input slot " + std::to_string(i) + " unbound");
  for (size_t i = 0; i < outs.size(); ++i)
    if (plan.output_used[i] && !outs[i]) return fail(ZH_ERR_MISSING_KEY, "output slot " + std::to_string(i) + " unbound");
  std::vector<void*> ip(in.size()), op(outs.size());
  for (size_t i = 0; i < in.size(); ++i) ip[i] = in[i] ? in[i]->ptr : nullptr;
  for (size_t i = 0; i < outs.size(); ++i) op[i] = outs[i] ? outs[i]->ptr : nullptr;
  if (p->uploaded && ip == p->bound_inputs && op == p->bound_outputs) return ZH_OK;

  std::vector<int> variant(plan.stages.size(), 0);
  for (size_t s = 0; s < plan.stages.size(); ++s) {
    StageDesc d = plan.stages[s].desc;
    d.srgb_lut = p->ctx->d_lut;
    d.error_flag = p->ctx->d_error;
    const auto& refs = plan.stages[s].slot_ref;
    auto resolve = [&](int64_t ref, uint64_t* stride) -> uint64_t {
      if (ref >= kIntermediateBase) {
        const size_t k = (size_t)(ref - kIntermediateBase);
        *stride = row_stride_of(plan.intermediates[k]);
        return (uint64_t)(uintptr_t)p->d_intermediates[k];
      }
      if (ref >= 0) { *stride = row_stride_of(in[ref]->desc); return (uint64_t)(uintptr_t)in[ref]->ptr; }
      if (ref > -1000000) {
        const size_t k = (size_t)(~ref);
        *stride = row_stride_of(outs[k]->desc);
        return (uint64_t)(uintptr_t)outs[k]->ptr;
      }
      *stride = 0;
      return 0;  // format-only slot (strict quantisation)
    };
    if (d.kind == STAGE_DYNAMIC) {
      // A user kernel reads program-owned intermediates, not a StageDesc.
    } else if (d.kind == STAGE_COPY) {
      uint64_t s0, s1;
      d.copy_src = resolve(refs[0], &s0);
      d.copy_dst = resolve(refs[1], &s1);
    } else {
      for (size_t k = 0; k < refs.size(); ++k) {
        Slot& sl = d.slot[k];
        uint64_t stride = 0;
        sl.ptr = resolve(refs[k], &stride);
        sl.vec_ok = ((sl.ptr & 15) == 0 && (d.flat || (sl.row_stride & 15) == 0)) ? 1u : 0u;
      }
      // Prefetch only slots whose addresses allow 16-byte loads; keep the
      // first entry filled so the kernel's "any prefetch" test holds.
      int32_t pre[2] = {-1, -1};
      int np = 0;
      for (int k = 0; k < 2; ++k) {
        const int32_t sl = d.pre_slot[k];
        if (sl >= 0 && d.slot[sl].vec_ok) pre[np++] = sl;
      }
      d.pre_slot[0] = pre[0];
      d.pre_slot[1] = pre[1];
      if (d.kind == STAGE_RESAMPLE && d.variant >= 0) {
        // The warp kernel reads and writes whole texels with 4/8/16-byte
        // accesses: every image must be aligned to its texel size.
        bool ok = true;
        for (size_t k = 0; k < refs.size(); ++k)
          ok = ok && d.slot[k].ptr % d.slot[k].bpp == 0 && d.slot[k].row_stride % d.slot[k].bpp == 0;
        variant[s] = ok ? 3 : 0;
      }
      if (d.kind == STAGE_POINTWISE && p->use_surrogates && s < p->d_surrogate.size() && p->d_surrogate[s] &&
          (p->surrogate_ok[s] || p->use_surrogates == 2)) {
        // The surrogate reads and writes whole RGBA8 texels.
        const Slot& si = d.slot[d.code[0].a];
        const Slot& so = d.slot[d.code[d.n_instr - 1].a];
        if (si.ptr % 4 == 0 && so.ptr % 4 == 0) {
          variant[s] = 5;
          p->h_stages[s] = d;
          continue;
        }
      }
      if (d.kind == STAGE_POINTWISE) {
        bool aligned = true;
        for (size_t k = 0; k < refs.size(); ++k) aligned = aligned && d.slot[k].vec_ok;
        const StagePlan& sp = plan.stages[s];
        const bool jit = sp.jit >= 0 && s < p->jit_stage_fn.size() && p->jit_stage_fn[s];
        variant[s] = !aligned ? 0 : sp.spec >= 0 ? 2 : jit ? 6 : sp.fast ? 1 : 0;
      }
    }
    p->h_stages[s] = d;
  }
  if (variant != p->variant) {
    // A captured graph holds the previous kernels: capture again.
    if (p->graph_exec) hipGraphExecDestroy(p->graph_exec);
    if (p->graph) hipGraphDestroy(p->graph);
    p->graph_exec = nullptr;
    p->graph = nullptr;
    p->variant = variant;
  }
  if (p->arena) {
    // Static-scratch programs upload synchronously (the stream is idle).
    ZH_HIP(hipMemcpy(p->d_stages, p->h_stages, plan.stages.size() * sizeof(StageDesc), hipMemcpyHostToDevice));
  } else {
    // Stage through the next ring copy; its previous upload must have left
    // the stream before the copy is overwritten.
    const int k = p->ring_next;
    p->ring_next = (k + 1) % zh_program::kRing;
    if (p->ring_used[k]) ZH_HIP(hipEventSynchronize(p->ring_ev[k]));
    const size_t bytes = plan.stages.size() * sizeof(StageDesc);
    std::memcpy(p->h_ring[k], p->h_stages, bytes);
    ZH_HIP(hipMemcpyAsync(p->d_stages, p->h_ring[k], bytes, hipMemcpyHostToDevice, p->ctx->stream));
    ZH_HIP(hipEventRecord(p->ring_ev[k], p->ctx->stream));
    p->ring_used[k] = true;
  }
  p->bound_inputs = ip;
  p->bound_outputs = op;
  p->uploaded = true;
  return ZH_OK;
}

static zh_status launch_unit(zh_program* p, const Unit& u) {
  hipStream_t s = p->ctx->stream;
  if (u.batch) {
    bool all_spec = u.spec, all_jit = u.spec;
    for (uint32_t id : u.ids) {
      all_spec = all_spec && p->variant[id] == 2;
      all_jit = all_jit && p->variant[id] == 6;
    }
    if (all_spec) {
      ZH_HIP(zh_launch_spec_batch(p->plan.stages[u.ids[0]].spec, u.d_batch, u.total_tiles, (uint32_t)p->ctx->cus, s));
      return ZH_OK;
    }
    if (all_jit) {
      const uint32_t first = u.ids[0];
      const uint32_t blocks = std::min<uint32_t>(u.total_tiles, p->jit_batch_resident[first] * (uint32_t)p->ctx->cus);
      if (blocks == 0) return ZH_OK;
      const BatchDesc* bd = u.d_batch;
      void* args[] = {&bd};
      ZH_HIP(hipModuleLaunchKernel(p->jit_batch_fn[first], blocks, 1, 1, kBlock, 1, 1, 0, s, args, nullptr));
      return ZH_OK;
    }
    bool all_warp = u.warp;
    for (uint32_t id : u.ids) all_warp = all_warp && p->variant[id] == 3;
    if (all_warp && u.pair) {
      // Both forms' tiles in one persistent grid.
      const uint32_t bpc = u.blocks_per_cu ? std::min(u.blocks_per_cu, u.pair_resident) : u.pair_resident;
      const uint32_t blocks = std::min<uint32_t>(u.total_tiles + u.total_tiles2, bpc * (uint32_t)p->ctx->cus);
      if (blocks == 0) return ZH_OK;
      const BatchDesc* ba = u.d_batch;
      const BatchDesc* bb = u.d_batch2;
      void* args[] = {&ba, &bb};
      ZH_HIP(hipModuleLaunchKernel(u.pair_fn, blocks, 1, 1, kBlock, 1, 1, 0, s, args, nullptr));
      return ZH_OK;
    }
    if (all_warp) {
      if (p->plan.stages[u.ids[0]].desc.variant >= kJitWarpBase) {
        const uint32_t first = u.ids[0], res = p->jit_batch_resident[first];
        const uint32_t bpc = u.blocks_per_cu ? std::min(u.blocks_per_cu, res) : res;
        const uint32_t blocks = std::min<uint32_t>(u.total_tiles, bpc * (uint32_t)p->ctx->cus);
        if (blocks == 0) return ZH_OK;
        const BatchDesc* bd = u.d_batch;
        void* args[] = {&bd};
        ZH_HIP(hipModuleLaunchKernel(p->jit_batch_fn[first], blocks, 1, 1, kBlock, 1, 1, 0, s, args, nullptr));
        return ZH_OK;
      }
      ZH_HIP(zh_launch_warp_batch(p->plan.stages[u.ids[0]].desc.variant, u.staged, u.d_batch, u.total_tiles,
                                  (uint32_t)p->ctx->cus, u.blocks_per_cu, s));
      return ZH_OK;
    }
    bool all_staged = !u.spec;  // pointwise stages have no staged form
    for (uint32_t id : u.ids) all_staged = all_staged && p->plan.stages[id].tile_mode == 0;
    if (all_staged) {
      ZH_HIP(zh_launch_batch(u.d_batch, u.d_counter, u.grid, u.lds, s));
      return ZH_OK;
    }
    // Mixed fallback kernels: one launch per stage.
    for (uint32_t id : u.ids) {
      Unit one;
      one.ids = {id};
      zh_status st = launch_unit(p, one);
      if (st != ZH_OK) return st;
    }
    return ZH_OK;
  }
  const uint32_t first = u.ids[0];
  const StagePlan& sp = p->plan.stages[first];
  const StageDesc* d = p->d_stages + first;
  switch (sp.desc.kind) {
    case STAGE_POINTWISE: {
      const int v = first < p->variant.size() ? p->variant[first] : 0;
      if (v == 5)
        ZH_HIP(zh_mlp_infer_stage(p->d_surrogate[first], d, (uint64_t)sp.desc.width * sp.desc.height,
                                  (uint32_t)p->ctx->cus, s));
      else if (v == 2)
        ZH_HIP(zh_launch_spec(sp.spec, d, sp.desc.width, sp.desc.height, (uint32_t)p->ctx->cus, s));
      else if (v == 6) {
        // A run-time kernel: the persistent grid of zh_launch_spec.
        const uint64_t groups = ((uint64_t)sp.desc.width * sp.desc.height + kPix - 1) / kPix;
        const uint64_t blocks =
            std::min<uint64_t>((groups + kBlock - 1) / kBlock, (uint64_t)p->jit_resident[first] * p->ctx->cus);
        if (blocks == 0) break;
        const StageDesc* dp = d;
        void* args[] = {&dp};
        ZH_HIP(hipModuleLaunchKernel(p->jit_stage_fn[first], (uint32_t)blocks, 1, 1, kBlock, 1, 1, 0, s, args,
                                     nullptr));
      }
      else
        ZH_HIP(zh_launch_pointwise(d, sp.desc.width, sp.desc.height, (int)sp.desc.flat, v, (uint32_t)p->ctx->cus,
                                   s));
      break;
    }
    case STAGE_RESAMPLE:
      if (first < p->variant.size() && p->variant[first] == 3 && sp.desc.variant >= kJitWarpBase) {
        // A run-time warp form: the persistent grid of zh_launch_warp.
        const uint32_t res = p->jit_resident[first];
        const uint32_t bpc = u.blocks_per_cu ? std::min(u.blocks_per_cu, res) : res;
        const uint32_t blocks = std::min<uint32_t>(sp.desc.tiles_x * sp.desc.tiles_y, bpc * (uint32_t)p->ctx->cus);
        if (blocks == 0) break;
        const StageDesc* dp = d;
        void* args[] = {&dp};
        ZH_HIP(hipModuleLaunchKernel(p->jit_stage_fn[first], blocks, 1, 1, kBlock, 1, 1, 0, s, args, nullptr));
      } else if (first < p->variant.size() && p->variant[first] == 3) {
        ZH_HIP(zh_launch_warp(sp.desc.variant, sp.staged, d, sp.desc.tiles_x * sp.desc.tiles_y, (uint32_t)p->ctx->cus,
                              u.blocks_per_cu, s));
      } else {
        ZH_HIP(zh_launch_resample(d, sp.tile_mode, sp.desc.tiles_x, sp.desc.tiles_y, sp.lds_bytes, s));
      }
      break;
    case STAGE_COPY:
      ZH_HIP(zh_launch_copy(d, sp.desc.copy_row_bytes, sp.desc.copy_rows, s));
      break;
    case STAGE_DYNAMIC: {
      // The user kernel over the result, 16 x 16 texels per workgroup.
      const DynamicCall& dc = p->plan.dynamics[sp.dynamic];
      zh_dynamic_args a;
      std::memset(&a, 0, sizeof(a));
      a.width = sp.desc.width;
      a.height = sp.desc.height;
      a.num_inputs = dc.num_args;
      a.data_len = (uint32_t)dc.data.size();
      for (uint32_t k = 0; k < dc.num_args; ++k) {
        const size_t b = (size_t)(sp.slot_ref[k] - kIntermediateBase);
        a.in[k] = static_cast<const float*>(p->d_intermediates[b]);
        a.in_width[k] = p->plan.intermediates[b].width;
        a.in_height[k] = p->plan.intermediates[b].height;
      }
      a.out = static_cast<float*>(p->d_intermediates[(size_t)(sp.slot_ref[dc.num_args] - kIntermediateBase)]);
      a.data = p->dyn_data[sp.dynamic];
      void* args[] = {&a};
      ZH_HIP(hipModuleLaunchKernel(p->dyn_fn[sp.dynamic], (a.width + 15) / 16, (a.height + 15) / 16, 1, 16, 16, 1, 0,
                                   s, args, nullptr));
      break;
    }
  }
  return ZH_OK;
}

static zh_status launch_all(zh_program* p) {
  for (const Unit& u : p->units) {
    zh_status st = launch_unit(p, u);
    if (st != ZH_OK) return st;
  }
  return ZH_OK;
}

}  // extern "C"

// ---- stepped executions (Execution::step_to, run.rs:1394-1464) -------------
namespace zh {

zh_status program_bind(zh_program* p, const zh_binding* bindings, uint32_t num_bindings) {
  if (!p->ctx) return fail(ZH_ERR_NO_DEVICE, "program was built without a context");
  ZH_HIP(hipSetDevice(p->ctx->device));
  return bind(p, bindings, num_bindings);
}

size_t program_units(const zh_program* p) { return p->units.size(); }

zh_status program_launch_units(zh_program* p, size_t first, size_t count) {
  ZH_HIP(hipSetDevice(p->ctx->device));
  for (size_t i = first; i < first + count && i < p->units.size(); ++i) {
    zh_status st = launch_unit(p, p->units[i]);
    if (st != ZH_OK) return st;
  }
  return ZH_OK;
}

uint64_t program_device_bytes(const zh_program* p) {
  uint64_t b = (uint64_t)p->plan.stages.size() * sizeof(StageDesc);
  for (const zh_descriptor& d : p->plan.intermediates) b += image_bytes(d);
  for (const zh_image& img : p->buffer_images) b += img.bytes;
  for (uint32_t w : p->tile_layer_words) b += (uint64_t)w * 4;
  for (const DynamicCall& dc : p->plan.dynamics) b += std::max<size_t>(dc.data.size(), 16);
  return b;
}

}  // namespace zh

extern "C" {

zh_status zh_program_launch(zh_program* p, const zh_binding* bindings, uint32_t num_bindings) {
  if (!p) return fail(ZH_ERR_INVALID_ARGUMENT, "null program");
  if (!p->ctx) return fail(ZH_ERR_NO_DEVICE, "program was built without a context");
  if (p->stepping) return fail(ZH_ERR_INVALID_ARGUMENT, "a stepped execution of this program is unfinished");
  ZH_HIP(hipSetDevice(p->ctx->device));
  zh_status st = bind(p, bindings, num_bindings);
  if (st != ZH_OK) return st;
  // A graph pays off from two launches on: one kernel launches directly.
  const bool use_graph = !(p->flags & ZH_PLAN_NO_GRAPH) && !p->graph_failed && p->units.size() > 1;
  if (use_graph && !p->graph_exec) {
    // Capture once; descriptor contents change through their device copies,
    // never through the graph, so the executable graph stays valid.
    const char* step = "hipStreamBeginCapture";
    hipError_t e = hipStreamBeginCapture(p->ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      st = launch_all(p);
      hipGraph_t g = nullptr;
      step = "hipStreamEndCapture";
      e = hipStreamEndCapture(p->ctx->stream, &g);
      if (st != ZH_OK) { if (g) hipGraphDestroy(g); return st; }
      if (e == hipSuccess && !g) e = hipErrorUnknown;
      if (e == hipSuccess) {
        size_t nodes = 0;
        hipGraphGetNodes(g, nullptr, &nodes);
        hipGraphExec_t ge = nullptr;
        step = "hipGraphInstantiate";
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (e == hipSuccess) {
          p->graph = g;
          p->graph_exec = ge;
          p->graph_nodes = (uint32_t)nodes;
          p->graph_captures++;
        } else {
          hipGraphDestroy(g);
        }
      }
    }
    if (!p->graph_exec) {
      // Reported through zh_program_graph_info and zh_program_describe; the
      // program launches its kernels one by one from now on.
      (void)hipGetLastError();
      p->graph_failed = true;
      p->graph_error = std::string(step) + ": " + hipGetErrorString(e);
    }
  }
  if (p->graph_exec) {
    ZH_HIP(hipGraphLaunch(p->graph_exec, p->ctx->stream));
    p->graph_replays++;
    return ZH_OK;
  }
  return launch_all(p);
}

zh_status zh_program_launch_once(zh_program* p, const zh_binding* bindings, uint32_t num_bindings) {
  if (!p) return fail(ZH_ERR_INVALID_ARGUMENT, "null program");
  if (!p->ctx) return fail(ZH_ERR_NO_DEVICE, "program was built without a context");
  if (p->stepping) return fail(ZH_ERR_INVALID_ARGUMENT, "a stepped execution of this program is unfinished");
  ZH_HIP(hipSetDevice(p->ctx->device));
  zh_status st = bind(p, bindings, num_bindings);
  // One run: kernel by kernel (a graph pays off only when replayed).
  if (st == ZH_OK) st = launch_all(p);
  const std::string msg = st == ZH_OK ? std::string() : std::string(zh_last_error());
  zh_program_destroy(p);  // waits for the queued kernels first
  return st == ZH_OK ? ZH_OK : fail(st, msg);
}

// Graphviz text of the plan: images and stages, an edge per read and write.
zh_status zh_program_dot(zh_program* p, char* buf, size_t len, size_t* needed) {
  if (!p) return fail(ZH_ERR_INVALID_ARGUMENT, "null program");
  if (p->units.empty() && !p->plan.stages.empty()) make_units(p);
  const Plan& plan = p->plan;
  std::ostringstream o;
  o << "digraph program {\n  rankdir=LR;\n";
  auto image = [&](int64_t ref) -> std::string {
    if (ref >= kIntermediateBase) return "t" + std::to_string(ref - kIntermediateBase);
    if (ref >= 0) return "in" + std::to_string(ref);
    return "out" + std::to_string(~ref);
  };
  auto label = [&](int64_t ref) -> std::string {
    const zh_descriptor& d = ref >= kIntermediateBase ? plan.intermediates[(size_t)(ref - kIntermediateBase)]
                             : ref >= 0               ? plan.input_desc[(size_t)ref]
                                                      : plan.output_desc[(size_t)~ref];
    const char* kind = ref >= kIntermediateBase ? "intermediate" : ref >= 0 ? "input" : "output";
    return std::string(kind) + " " + std::to_string(ref >= kIntermediateBase ? ref - kIntermediateBase
                                                    : ref >= 0               ? ref
                                                                             : ~ref) +
           "\\n" + std::to_string(d.width) + "x" + std::to_string(d.height) + " bits " + std::to_string(d.texel.bits);
  };
  std::set<int64_t> seen;
  auto node = [&](int64_t ref) {
    if (seen.insert(ref).second)
      o << "  " << image(ref) << " [shape=box, label=\"" << label(ref) << "\"];\n";
  };
  for (size_t s = 0; s < plan.stages.size(); ++s) {
    const StagePlan& sp = plan.stages[s];
    std::string kernel = "stage " + std::to_string(s);
    const size_t k = sp.text.rfind("[kernel ");
    if (k != std::string::npos) kernel += "\\n" + sp.text.substr(k + 8, sp.text.size() - k - 9);
    o << "  s" << s << " [shape=ellipse, label=\"" << kernel << "\"];\n";
    std::vector<bool> read(sp.slot_ref.size(), false), written(sp.slot_ref.size(), false);
    if (sp.desc.kind == STAGE_COPY) {
      read[0] = written[1] = true;
    } else if (sp.desc.kind == STAGE_DYNAMIC) {
      for (size_t i = 0; i + 1 < sp.slot_ref.size(); ++i) read[i] = true;
      written.back() = true;
    } else {
      for (uint32_t i = 0; i < sp.desc.n_instr; ++i) {
        const Instr& in = sp.desc.code[i];
        if (in.op == OP_STORE) written[in.a] = true;
        if (in.op == OP_LOAD || in.op == OP_BOX3 || in.op == OP_PALETTE) read[in.a] = true;
      }
    }
    for (size_t i = 0; i < sp.slot_ref.size(); ++i) {
      const int64_t ref = sp.slot_ref[i];
      if (ref < 0 && ref <= -1000000) continue;  // quantisation-only slot
      if (!read[i] && !written[i]) continue;
      node(ref);
      if (read[i]) o << "  " << image(ref) << " -> s" << s << ";\n";
      if (written[i]) o << "  s" << s << " -> " << image(ref) << ";\n";
    }
  }
  o << "}\n";
  const std::string t = o.str();
  if (needed) *needed = t.size() + 1;
  if (buf && len) {
    const size_t n = std::min(len - 1, t.size());
    std::memcpy(buf, t.data(), n);
    buf[n] = 0;
  }
  return ZH_OK;
}

zh_status zh_program_launch_static(zh_context* ctx, const zh_descriptor* textures, uint32_t num_textures,
                                   const zh_op* ops, uint32_t num_ops, uint32_t flags, const zh_binding* bindings,
                                   uint32_t num_bindings) {
  if (!ctx || (num_textures && !textures) || (num_ops && !ops) || (num_bindings && !bindings))
    return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  if (ctx->closing) return fail(ZH_ERR_INVALID_ARGUMENT, "context is being destroyed");
  if (!ctx->scratch || ctx->scratch_bytes == 0)
    return fail(ZH_ERR_HIP, "the device's static scratch was not resolved at context creation");
  ZH_HIP(hipSetDevice(ctx->device));
  zh_program prog;
  zh_program* p = &prog;
  // No hipGraph: instantiating one allocates.
  zh_status st = init_program(p, ctx, textures, num_textures, ops, num_ops, flags | ZH_PLAN_NO_GRAPH);
  if (st != ZH_OK) return st;
  // Loading a user kernel's code object allocates.
  if (!p->plan.dynamics.empty())
    return fail(ZH_ERR_UNSUPPORTED, "user stage kernels cannot run from the static scratch (loading them allocates)");
  // An earlier call's kernels on this stream may still read the scratch:
  // let them finish before its tables are overwritten.
  ZH_HIP(hipStreamSynchronize(ctx->stream));
  p->arena = static_cast<uint8_t*>(ctx->scratch);
  p->arena_size = ctx->scratch_bytes;
  p->arena_used = 0;
  st = alloc_device(p);
  if (st == ZH_OK) st = bind(p, bindings, num_bindings);
  if (st == ZH_OK) st = launch_all(p);
  free_device(p);  // host side only: the tables stay in the scratch for the queued kernels
  return st;
}

zh_status zh_program_graph_info(zh_program* p, zh_graph_info* info) {
  if (!p || !info) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  std::memset(info, 0, sizeof(*info));
  if (p->flags & ZH_PLAN_NO_GRAPH) info->state = ZH_GRAPH_DISABLED;
  else if (p->graph_failed) info->state = ZH_GRAPH_FAILED;
  else if (p->graph_exec) info->state = ZH_GRAPH_CAPTURED;
  else if (p->units.size() <= 1) info->state = ZH_GRAPH_SINGLE_LAUNCH;
  else info->state = ZH_GRAPH_NOT_YET;
  info->nodes = p->graph_nodes;
  info->captures = p->graph_captures;
  info->replays = p->graph_replays;
  return ZH_OK;
}

// ---- autotuner ---------------------------------------------------------------
// Tile shapes the warp kernels accept (tile_w / kPix * tile_h == kBlock).
// (Narrower ones, 8x128 and 4x256, measured 1.5-3x slower on the 256-image
// batch and the 8-layer 4K stack, and no faster on the small two-layer one.)
static const uint32_t kWarpTiles[][2] = {{64, 16}, {32, 32}, {128, 8}, {16, 64}};

static bool warp_unit(const zh_program* p, const Unit& u) {
  for (uint32_t s : u.ids) {
    if (p->plan.stages[s].desc.kind != STAGE_RESAMPLE || s >= p->variant.size() || p->variant[s] != 3) return false;
  }
  return true;
}

// Give every stage of a warp unit the tile shape tw x th (descriptors, device
// copies and, for a batch, the tile prefix).  The interpreter fallback of a
// re-tiled stage gathers its taps directly (TILE_DIRECT accepts any tile).
static zh_status set_warp_tile(zh_program* p, Unit& u, uint32_t tw, uint32_t th) {
  // The warp kernels split a tile's lanes by shifts: tw / 4 lanes per row,
  // a power of two, and 256 lanes per tile.
  if (tw < 4 || (tw & (tw - 1)) || tw * th != 1024) return fail(ZH_ERR_INVALID_ARGUMENT, "unsupported warp tile shape");
  std::vector<uint32_t> tiles(u.ids.size(), 0);
  for (size_t k = 0; k < u.ids.size(); ++k) {
    const size_t s = u.ids[k];
    StagePlan& sp = p->plan.stages[s];
    StageDesc& d = sp.desc;
    if (d.tile_w != tw || d.tile_h != th) {
      d.tile_w = tw;
      d.tile_h = th;
      d.band_h = th;
      d.tiles_x = (d.width + tw - 1) / tw;
      d.tiles_y = (d.height + th - 1) / th;
      sp.tile_mode = 1;  // TILE_DIRECT
      sp.lds_bytes = 0;
      const size_t a = sp.text.find(" tile "), b = a == std::string::npos ? a : sp.text.find(' ', a + 6);
      if (b != std::string::npos) sp.text.replace(a + 6, b - (a + 6), std::to_string(tw) + "x" + std::to_string(th));
    }
    zh_status ms = tile_layer_masks(p, s);
    if (ms != ZH_OK) return ms;
    StageDesc& h = p->h_stages[s];
    h.tile_w = d.tile_w;
    h.tile_h = d.tile_h;
    h.band_h = d.band_h;
    h.tiles_x = d.tiles_x;
    h.tiles_y = d.tiles_y;
    h.tile_layers = d.tile_layers;
    tiles[k] = d.tiles_x * d.tiles_y;
  }
  hipStream_t st = p->ctx->stream;
  for (uint32_t s : u.ids)
    ZH_HIP(hipMemcpyAsync(p->d_stages + s, p->h_stages + s, sizeof(StageDesc), hipMemcpyHostToDevice, st));
  // The tile prefix of each queue (a pair's two forms have one each).
  const size_t na = u.pair ? u.split : u.ids.size();
  std::vector<uint32_t> first(na + 1, 0), second(u.ids.size() - na + 1, 0);
  for (size_t k = 0; k < na; ++k) first[k + 1] = first[k] + tiles[k];
  for (size_t k = na; k < u.ids.size(); ++k) second[k - na + 1] = second[k - na] + tiles[k];
  if (u.batch) {
    ZH_HIP(hipMemcpyAsync(u.d_prefix, first.data(), first.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    u.total_tiles = first.back();
    ZH_HIP(hipMemcpyAsync((char*)u.d_batch + offsetof(BatchDesc, total_tiles), &u.total_tiles, sizeof(uint32_t),
                          hipMemcpyHostToDevice, st));
    if (u.pair) {
      ZH_HIP(hipMemcpyAsync(u.d_prefix2, second.data(), second.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
      u.total_tiles2 = second.back();
      ZH_HIP(hipMemcpyAsync((char*)u.d_batch2 + offsetof(BatchDesc, total_tiles), &u.total_tiles2,
                            sizeof(uint32_t), hipMemcpyHostToDevice, st));
    }
  }
  // The host vectors and totals must outlive the copies.
  ZH_HIP(hipStreamSynchronize(st));
  return ZH_OK;
}

// Deal every stage's tiles of a warp unit in vertical strips sw tiles wide
// (0: by rows).
static zh_status set_warp_strip(zh_program* p, Unit& u, uint32_t sw) {
  hipStream_t st = p->ctx->stream;
  for (uint32_t s : u.ids) {
    p->plan.stages[s].desc.strip_w = sw;
    p->h_stages[s].strip_w = sw;
    ZH_HIP(hipMemcpyAsync((char*)(p->d_stages + s) + offsetof(StageDesc, strip_w), &p->h_stages[s].strip_w,
                          sizeof(uint32_t), hipMemcpyHostToDevice, st));
  }
  ZH_HIP(hipStreamSynchronize(st));
  u.strip_w = sw;
  return ZH_OK;
}

zh_status zh_program_autotune(zh_program* p, const zh_binding* bindings, uint32_t num_bindings, uint32_t iterations) {
  if (!p) return fail(ZH_ERR_INVALID_ARGUMENT, "null program");
  if (!p->ctx) return fail(ZH_ERR_NO_DEVICE, "program was built without a context");
  if (iterations == 0) iterations = 5;
  ZH_HIP(hipSetDevice(p->ctx->device));
  zh_status st = bind(p, bindings, num_bindings);
  if (st != ZH_OK) return st;
  hipEvent_t e0, e1;
  ZH_HIP(hipEventCreate(&e0));
  ZH_HIP(hipEventCreate(&e1));
  bool changed = false;
  for (Unit& u : p->units) {
    if (!warp_unit(p, u)) continue;
    const uint32_t tw0 = p->plan.stages[u.ids[0]].desc.tile_w, th0 = p->plan.stages[u.ids[0]].desc.tile_h;
    float best = 1e30f;
    uint32_t bw = tw0, bh = th0, bb = u.blocks_per_cu;
    static const uint32_t kPerCu[] = {0, 4, 3, 2, 1};
    for (const auto& c : kWarpTiles) {
      // Staged units keep to the tile shapes whose boxes fit every stage.
      bool fits = true;
      for (uint32_t s : u.ids) fits = fits && (!u.staged || warp_staged_fits(p->plan.stages[s].desc, c[0], c[1]));
      if (!fits) continue;
      st = set_warp_tile(p, u, c[0], c[1]);
      for (uint32_t bpc : kPerCu) {
        if (st != ZH_OK) break;
        u.blocks_per_cu = bpc;
        st = launch_unit(p, u);  // warm-up
        if (st != ZH_OK) break;
        hipEventRecord(e0, p->ctx->stream);
        for (uint32_t i = 0; i < iterations && st == ZH_OK; ++i) st = launch_unit(p, u);
        hipEventRecord(e1, p->ctx->stream);
        if (st != ZH_OK) break;
        hipEventSynchronize(e1);
        float ms = 0.0f;
        hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) {
          best = ms;
          bw = c[0];
          bh = c[1];
          bb = bpc;
        }
      }
      if (st != ZH_OK) break;
    }
    u.blocks_per_cu = bb;
    if (st == ZH_OK) st = set_warp_tile(p, u, bw, bh);
    if (st != ZH_OK) break;
    changed = changed || bw != tw0 || bh != th0;
    // Then the dealing order for that tile and grid.
    static const uint32_t kStrips[] = {0, 2, 4, 8, 16};
    uint32_t bs = u.strip_w;
    for (uint32_t sw : kStrips) {
      st = set_warp_strip(p, u, sw);
      if (st == ZH_OK) st = launch_unit(p, u);  // warm-up
      if (st != ZH_OK) break;
      hipEventRecord(e0, p->ctx->stream);
      for (uint32_t i = 0; i < iterations && st == ZH_OK; ++i) st = launch_unit(p, u);
      hipEventRecord(e1, p->ctx->stream);
      if (st != ZH_OK) break;
      hipEventSynchronize(e1);
      float ms = 0.0f;
      hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) {
        best = ms;
        bs = sw;
      }
    }
    if (st == ZH_OK) st = set_warp_strip(p, u, bs);
    if (st != ZH_OK) break;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (st != ZH_OK) return st;
  // Tiles live in the descriptors, but a captured graph holds the grid
  // sizes: capture again at the next launch.
  if (p->graph_exec) hipGraphExecDestroy(p->graph_exec);
  if (p->graph) hipGraphDestroy(p->graph);
  p->graph_exec = nullptr;
  p->graph = nullptr;
  (void)changed;
  return ZH_OK;
}

zh_status zh_program_time(zh_program* p, const zh_binding* bindings, uint32_t num_bindings, uint32_t iterations,
                          float* stage_ms, uint32_t max_stages, uint32_t* num_stages) {
  if (!p) return fail(ZH_ERR_INVALID_ARGUMENT, "null program");
  if (!p->ctx) return fail(ZH_ERR_NO_DEVICE, "program was built without a context");
  if (iterations == 0) iterations = 1;
  ZH_HIP(hipSetDevice(p->ctx->device));
  zh_status st = bind(p, bindings, num_bindings);
  if (st != ZH_OK) return st;
  const size_t n = p->units.size();
  if (num_stages) *num_stages = (uint32_t)n;
  // Every iteration's launches and events are queued before one wait, so
  // the intervals measure the GPU back to back, not the host's launch latency.
  std::vector<hipEvent_t> ev((n + 1) * iterations);
  for (auto& e : ev) ZH_HIP(hipEventCreate(&e));
  st = launch_all(p);  // warm-up
  for (uint32_t it = 0; it < iterations && st == ZH_OK; ++it) {
    hipEvent_t* e = &ev[(size_t)it * (n + 1)];
    ZH_HIP(hipEventRecord(e[0], p->ctx->stream));
    for (size_t u = 0; u < n && st == ZH_OK; ++u) {
      st = launch_unit(p, p->units[u]);
      if (st == ZH_OK) ZH_HIP(hipEventRecord(e[u + 1], p->ctx->stream));
    }
  }
  std::vector<double> acc(n, 0.0);
  if (st == ZH_OK) {
    ZH_HIP(hipEventSynchronize(ev.back()));
    for (uint32_t it = 0; it < iterations; ++it)
      for (size_t u = 0; u < n; ++u) {
        float ms = 0;
        hipEventElapsedTime(&ms, ev[(size_t)it * (n + 1) + u], ev[(size_t)it * (n + 1) + u + 1]);
        acc[u] += ms;
      }
  }
  for (auto& e : ev) hipEventDestroy(e);
  if (st != ZH_OK) return st;
  for (size_t u = 0; u < n && u < max_stages; ++u) stage_ms[u] = (float)(acc[u] / iterations);
  return ZH_OK;
}

}  // extern "C"
