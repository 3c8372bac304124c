// zh_pool.cpp -- the image pool and its handoff to executions
// (include/zosimos_hip.h "pool"), the HIP counterpart of
//   Pool / Image / ImageData / ImageMeta        lib/zosimos/src/pool.rs:19-156
//   PoolImageMut::trade / swap_image            pool.rs:900-941
//   Executable::check_satisfiable               run.rs:1085-1167
//   Retire::{input, output, finish}             run.rs:2791-2998
//   recover_buffers / retire_buffers caching    run.rs:1312-1347, 2876-2960
//
// An image of the pool holds host bytes (ImageData::Host), a device image
// (ImageData::GpuBuffer; here HIP device memory), or only its descriptor
// (ImageData::LateBound).  Launching a program from the pool trades each
// bound input into the execution -- host bytes are uploaded into a device
// buffer, a device image is used in place -- and each keyed output is written
// into the pool's own device image when it has one of the output's layout.
// Retiring an output moves the result into the pool: as host bytes (one
// download), or as a device image handed over without a copy.  Device
// buffers the execution allocated return to the pool's cache at finish and
// are reused by the next launch with the same layout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "zh_internal.h"

using namespace zh;

namespace {

// A device image shared between the pool and executions; freed with the last
// holder.
struct DeviceBuf {
  zh_image* img = nullptr;
  explicit DeviceBuf(zh_image* i) : img(i) {}
  ~DeviceBuf() {
    if (img) zh_image_free(img);
  }
  DeviceBuf(const DeviceBuf&) = delete;
  DeviceBuf& operator=(const DeviceBuf&) = delete;
};
typedef std::shared_ptr<DeviceBuf> DevPtr;
typedef std::shared_ptr<std::vector<uint8_t>> HostPtr;

struct PoolImage {
  zh_descriptor desc{};
  int32_t kind = ZH_IMAGE_LATE_BOUND;
  bool no_read = false;  // ImageMeta::no_read (pool.rs:104-109)
  HostPtr host;          // tight rows
  DevPtr dev;
};

// The descriptor with tight rows (host images are tight).
zh_descriptor tight(zh_descriptor d) {
  d.row_stride = 0;
  d.texel_stride = bits_bytes(d.texel.bits);
  return d;
}

// Descriptor equality (Descriptor: PartialEq, buffer.rs:13-30): layout and
// every colour field, as check_satisfiable compares them (run.rs:1099-1108).
bool same_descriptor(const zh_descriptor& a, const zh_descriptor& b) {
  return same_layout(a, b) && a.color.model == b.color.model && a.color.primaries == b.color.primaries &&
         a.color.transfer == b.color.transfer && a.color.whitepoint == b.color.whitepoint &&
         a.color.luminance == b.color.luminance;
}

}  // namespace

struct zh_pool {
  std::map<zh_pool_key, PoolImage> items;
  zh_pool_key next = 1;
  std::vector<DevPtr> cache;  // device buffers of finished executions, by layout
};

struct zh_execution {
  zh_program* prog = nullptr;
  zh_pool* pool = nullptr;
  struct Slot {
    bool bound = false, retired = false;
    zh_pool_key key = 0;
    DevPtr dev;            // the device image the program reads or writes
    bool in_place = false; // the pool image's own device buffer
    HostPtr host_keep;     // host bytes being uploaded (kept until finish)
    bool moved = false;    // a no_read input moved out of the pool (pool.rs:912-916)
    PoolImage moved_image;
  };
  std::vector<Slot> in, out;
  // Stepped executions (Execution::step_to): launch units [next, units) are
  // still to be issued.
  bool stepped = false;
  size_t next = 0, units = 0;
  bool prune = false;  // Retire::prune: free the execution's buffers at finish
  zh_resources_used usage{};
};

// A device buffer of layout d: from the pool's cache, else newly allocated
// (ResourcesUsed::buffer_mem / buffer_reused, run.rs:212-221).
static DevPtr take_buffer(zh_execution* e, const zh_descriptor& d, zh_status* st) {
  zh_pool* pool = e->pool;
  zh_context* ctx = e->prog->ctx;
  for (size_t k = 0; k < pool->cache.size(); ++k) {
    const DevPtr& b = pool->cache[k];
    if (b->img->ctx == ctx && same_layout(b->img->desc, d)) {
      DevPtr r = b;
      pool->cache.erase(pool->cache.begin() + (long)k);
      e->usage.buffer_reused += r->img->bytes;
      e->usage.buffers_reused++;
      *st = ZH_OK;
      return r;
    }
  }
  zh_image* img = nullptr;
  *st = zh_image_alloc(ctx, &d, &img);
  if (*st != ZH_OK) return nullptr;
  e->usage.buffer_mem += img->bytes;
  e->usage.buffers_allocated++;
  return std::make_shared<DeviceBuf>(img);
}

static zh_status bad_key(zh_pool_key key) {
  return fail(ZH_ERR_INVALID_ARGUMENT, "no pool image under key " + std::to_string(key));
}

extern "C" {

zh_status zh_pool_create(zh_pool** out) {
  if (!out) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  *out = new zh_pool();
  return ZH_OK;
}

zh_status zh_pool_destroy(zh_pool* pool) {
  delete pool;
  return ZH_OK;
}

zh_status zh_pool_insert_host(zh_pool* pool, const zh_descriptor* desc, const void* bytes, uint64_t row_stride,
                              zh_pool_key* key) {
  if (!pool || !desc || !key || !bytes) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  std::string why;
  if (!descriptor_valid(*desc, &why)) return fail(ZH_ERR_INVALID_ARGUMENT, "descriptor: " + why);
  PoolImage im;
  im.desc = tight(*desc);
  const uint64_t row = tight_stride(im.desc);
  if (row_stride == 0) row_stride = row;
  if (row_stride < row) return fail(ZH_ERR_INVALID_ARGUMENT, "row stride shorter than a row");
  im.kind = ZH_IMAGE_HOST;
  im.host = std::make_shared<std::vector<uint8_t>>(row * im.desc.height);
  for (uint32_t y = 0; y < im.desc.height; ++y)
    std::memcpy(im.host->data() + row * y, static_cast<const uint8_t*>(bytes) + row_stride * y, row);
  *key = pool->next++;
  pool->items[*key] = std::move(im);
  return ZH_OK;
}

zh_status zh_pool_declare(zh_pool* pool, const zh_descriptor* desc, zh_pool_key* key) {
  if (!pool || !desc || !key) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  std::string why;
  if (!descriptor_valid(*desc, &why)) return fail(ZH_ERR_INVALID_ARGUMENT, "descriptor: " + why);
  PoolImage im;
  im.desc = tight(*desc);
  *key = pool->next++;
  pool->items[*key] = std::move(im);
  return ZH_OK;
}

zh_status zh_pool_insert_device(zh_pool* pool, zh_image* img, zh_pool_key* key) {
  if (!pool || !img || !key) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  PoolImage im;
  im.desc = img->desc;
  im.kind = ZH_IMAGE_HIP_BUFFER;
  im.dev = std::make_shared<DeviceBuf>(img);
  *key = pool->next++;
  pool->items[*key] = std::move(im);
  return ZH_OK;
}

zh_status zh_pool_image_info(zh_pool* pool, zh_pool_key key, zh_descriptor* desc, int32_t* kind) {
  if (!pool) return fail(ZH_ERR_INVALID_ARGUMENT, "null pool");
  auto it = pool->items.find(key);
  if (it == pool->items.end()) return bad_key(key);
  if (desc) *desc = it->second.desc;
  if (kind) *kind = it->second.kind;
  return ZH_OK;
}

zh_status zh_pool_host_bytes(zh_pool* pool, zh_pool_key key, const void** data, uint64_t* len) {
  if (!pool || !data) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  auto it = pool->items.find(key);
  if (it == pool->items.end()) return bad_key(key);
  if (it->second.kind != ZH_IMAGE_HOST) return fail(ZH_ERR_INVALID_ARGUMENT, "pool image holds no host bytes");
  *data = it->second.host->data();
  if (len) *len = it->second.host->size();
  return ZH_OK;
}

zh_status zh_pool_device_image(zh_pool* pool, zh_pool_key key, zh_image** img) {
  if (!pool || !img) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  auto it = pool->items.find(key);
  if (it == pool->items.end()) return bad_key(key);
  if (it->second.kind != ZH_IMAGE_HIP_BUFFER) return fail(ZH_ERR_INVALID_ARGUMENT, "pool image is not on a device");
  *img = it->second.dev->img;
  return ZH_OK;
}

zh_status zh_pool_set_no_read(zh_pool* pool, zh_pool_key key, int32_t no_read) {
  if (!pool) return fail(ZH_ERR_INVALID_ARGUMENT, "null pool");
  auto it = pool->items.find(key);
  if (it == pool->items.end()) return bad_key(key);
  it->second.no_read = no_read != 0;
  return ZH_OK;
}

zh_status zh_pool_upload(zh_pool* pool, zh_pool_key key, zh_context* ctx) {
  if (!pool || !ctx) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  auto it = pool->items.find(key);
  if (it == pool->items.end()) return bad_key(key);
  PoolImage& im = it->second;
  if (im.kind == ZH_IMAGE_HIP_BUFFER) return ZH_OK;
  if (im.kind != ZH_IMAGE_HOST) return fail(ZH_ERR_INVALID_ARGUMENT, "pool image has no data to upload");
  zh_image* img = nullptr;
  zh_status st = zh_image_alloc(ctx, &im.desc, &img);
  if (st != ZH_OK) return st;
  DevPtr dev = std::make_shared<DeviceBuf>(img);
  st = zh_image_upload(img, im.host->data(), 0);
  if (st == ZH_OK) st = zh_sync(ctx);
  if (st != ZH_OK) return st;
  im.kind = ZH_IMAGE_HIP_BUFFER;
  im.dev = dev;
  im.host.reset();
  return ZH_OK;
}

zh_status zh_pool_download(zh_pool* pool, zh_pool_key key) {
  if (!pool) return fail(ZH_ERR_INVALID_ARGUMENT, "null pool");
  auto it = pool->items.find(key);
  if (it == pool->items.end()) return bad_key(key);
  PoolImage& im = it->second;
  if (im.kind == ZH_IMAGE_HOST) return ZH_OK;
  if (im.kind != ZH_IMAGE_HIP_BUFFER) return fail(ZH_ERR_INVALID_ARGUMENT, "pool image has no data to download");
  HostPtr host = std::make_shared<std::vector<uint8_t>>(tight_stride(im.desc) * im.desc.height);
  zh_status st = zh_image_download(im.dev->img, host->data(), 0);
  if (st == ZH_OK) st = zh_sync(im.dev->img->ctx);
  if (st != ZH_OK) return st;
  im.desc = tight(im.desc);
  im.kind = ZH_IMAGE_HOST;
  im.host = host;
  im.dev.reset();
  return ZH_OK;
}

zh_status zh_pool_remove(zh_pool* pool, zh_pool_key key) {
  if (!pool) return fail(ZH_ERR_INVALID_ARGUMENT, "null pool");
  if (!pool->items.erase(key)) return bad_key(key);
  return ZH_OK;
}

zh_status zh_pool_clear_cache(zh_pool* pool) {
  if (!pool) return fail(ZH_ERR_INVALID_ARGUMENT, "null pool");
  pool->cache.clear();
  return ZH_OK;
}

zh_status zh_pool_cached_buffers(zh_pool* pool, uint32_t* count) {
  if (!pool || !count) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  *count = (uint32_t)pool->cache.size();
  return ZH_OK;
}

// Return what the execution still holds to the pool: moved inputs go back,
// device buffers it allocated go to the cache.
// Return what the execution still holds: moved inputs go back to the pool
// (unless discarded), device buffers it allocated go to the cache (unless
// pruned or discarded).
static void release(zh_execution* e, bool discard) {
  for (auto* slots : {&e->in, &e->out})
    for (zh_execution::Slot& s : *slots) {
      if (s.moved && !discard) {
        auto it = e->pool->items.find(s.key);
        if (it != e->pool->items.end()) it->second = std::move(s.moved_image);
      }
      s.moved = false;
      if (s.dev && !s.in_place && !discard && !e->prune) e->pool->cache.push_back(s.dev);
      s.dev.reset();
      s.host_keep.reset();
    }
  if (e->prog && e->prog->stepping == e) e->prog->stepping = nullptr;
}

// Executable::launch from a pool (run.rs:1016-1167): check, trade, bind, and
// either launch everything or leave the launches to zh_execution_step.
static zh_status start(zh_program* prog, zh_pool* pool, const zh_pool_binding* bindings, uint32_t num_bindings,
                       bool launch, zh_execution** out) {
  if (!prog || !pool || !out || (num_bindings && !bindings)) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  const Plan& plan = prog->plan;
  std::unique_ptr<zh_execution> e(new zh_execution());
  e->prog = prog;
  e->pool = pool;
  e->in.resize(plan.user_inputs);
  e->out.resize(plan.output_desc.size());
  std::set<zh_pool_key> used;
  // Check everything before anything is traded (run.rs:1085-1131).
  for (uint32_t i = 0; i < num_bindings; ++i) {
    const zh_pool_binding& b = bindings[i];
    auto& slots = b.is_output ? e->out : e->in;
    if (b.slot >= slots.size()) return fail(ZH_ERR_INVALID_ARGUMENT, "binding slot out of range");
    if (slots[b.slot].bound) return fail(ZH_ERR_INVALID_ARGUMENT, "slot bound twice");
    const zh_descriptor want = tight(b.is_output ? plan.output_desc[b.slot] : plan.input_desc[b.slot]);
    if (b.key == 0) {
      if (!b.is_output) return fail(ZH_ERR_MISSING_KEY, "input slot " + std::to_string(b.slot) + " has no pool key");
    } else {
      auto it = pool->items.find(b.key);
      if (it == pool->items.end()) return bad_key(b.key);
      if (!same_descriptor(tight(it->second.desc), want))
        return fail(ZH_ERR_MISMATCHED_DESCRIPTOR, std::string(b.is_output ? "output" : "input") + " slot " +
                                                      std::to_string(b.slot) + " bound to a pool image of another descriptor");
      if (!b.is_output) {
        if (it->second.kind == ZH_IMAGE_LATE_BOUND)
          return fail(ZH_ERR_MISSING_KEY, "input slot " + std::to_string(b.slot) + ": the pool image has no data");
        if (!used.insert(b.key).second) return fail(ZH_ERR_INVALID_ARGUMENT, "one pool image bound to two inputs");
      }
    }
    slots[b.slot].bound = true;
    slots[b.slot].key = b.key;
  }
  // An image is written by one output at most and never read by the same
  // launch: the pool hands an image to one owner at a time (pool.rs:900-941).
  std::set<zh_pool_key> written;
  for (const zh_execution::Slot& s : e->out)
    if (s.bound && s.key && (used.count(s.key) || !written.insert(s.key).second))
      return fail(ZH_ERR_INVALID_ARGUMENT, "pool image " + std::to_string(s.key) + " bound twice");
  for (size_t i = 0; i < e->in.size(); ++i)
    if (plan.input_used[i] && !e->in[i].bound) return fail(ZH_ERR_MISSING_KEY, "input slot " + std::to_string(i) + " unbound");
  for (size_t i = 0; i < e->out.size(); ++i)
    if (plan.output_used[i] && !e->out[i].bound)
      return fail(ZH_ERR_MISSING_KEY, "output slot " + std::to_string(i) + " unbound");

  if (!prog->ctx) return fail(ZH_ERR_NO_DEVICE, "program was built without a context");
  if (prog->stepping) return fail(ZH_ERR_INVALID_ARGUMENT, "a stepped execution of this program is unfinished");

  // Trade (pool.rs:912-941): device images are used in place, host bytes are
  // uploaded into a (cached) device buffer; a no_read image moves out of the
  // pool until it is retired.
  std::vector<zh_binding> binds;
  zh_status st = ZH_OK;
  for (uint32_t s = 0; s < e->in.size() && st == ZH_OK; ++s) {
    zh_execution::Slot& sl = e->in[s];
    if (!sl.bound) continue;
    PoolImage& im = pool->items[sl.key];
    if (im.kind == ZH_IMAGE_HIP_BUFFER && same_layout(im.dev->img->desc, plan.input_desc[s]) &&
        im.dev->img->ctx == prog->ctx) {
      sl.dev = im.dev;
      sl.in_place = true;
      e->usage.buffer_reused += im.dev->img->bytes;
      e->usage.buffers_reused++;
    } else {
      sl.dev = take_buffer(e.get(), tight(plan.input_desc[s]), &st);
      if (st != ZH_OK) break;
      if (im.kind == ZH_IMAGE_HOST) {
        sl.host_keep = im.host;
        st = zh_image_upload(sl.dev->img, im.host->data(), 0);
      } else {
        st = hipMemcpy2DAsync(sl.dev->img->ptr, row_stride_of(sl.dev->img->desc), im.dev->img->ptr,
                              row_stride_of(im.dev->img->desc), tight_stride(im.desc), im.desc.height,
                              hipMemcpyDeviceToDevice, prog->ctx->stream) == hipSuccess
                 ? ZH_OK
                 : fail(ZH_ERR_HIP, "device copy of a pool image failed");
      }
    }
    if (im.no_read) {
      sl.moved = true;
      sl.moved_image = im;
      im = PoolImage();
      im.desc = sl.moved_image.desc;
    }
    binds.push_back(zh_binding{s, 0, sl.dev->img});
  }
  for (uint32_t s = 0; s < e->out.size() && st == ZH_OK; ++s) {
    zh_execution::Slot& sl = e->out[s];
    if (!sl.bound) continue;
    auto it = sl.key ? pool->items.find(sl.key) : pool->items.end();
    if (it != pool->items.end() && it->second.kind == ZH_IMAGE_HIP_BUFFER &&
        same_layout(it->second.dev->img->desc, plan.output_desc[s]) && it->second.dev->img->ctx == prog->ctx) {
      sl.dev = it->second.dev;
      sl.in_place = true;
      e->usage.buffer_reused += sl.dev->img->bytes;
      e->usage.buffers_reused++;
    } else {
      sl.dev = take_buffer(e.get(), tight(plan.output_desc[s]), &st);
      if (st != ZH_OK) break;
    }
    binds.push_back(zh_binding{s, 1, sl.dev->img});
  }
  e->usage.program_mem = program_device_bytes(prog);
  if (st == ZH_OK) {
    if (launch) {
      st = zh_program_launch(prog, binds.data(), (uint32_t)binds.size());
    } else {
      st = program_bind(prog, binds.data(), (uint32_t)binds.size());
      e->stepped = true;
      e->units = program_units(prog);
    }
  }
  if (st != ZH_OK) {
    const std::string msg = zh_last_error();
    hipStreamSynchronize(prog->ctx->stream);
    release(e.get(), false);
    return fail(st, msg);
  }
  if (e->stepped) prog->stepping = e.get();
  *out = e.release();
  return ZH_OK;
}

zh_status zh_program_launch_pool(zh_program* prog, zh_pool* pool, const zh_pool_binding* bindings,
                                 uint32_t num_bindings, zh_execution** out) {
  return start(prog, pool, bindings, num_bindings, true, out);
}

zh_status zh_program_start_pool(zh_program* prog, zh_pool* pool, const zh_pool_binding* bindings,
                                uint32_t num_bindings, zh_execution** out) {
  return start(prog, pool, bindings, num_bindings, false, out);
}

zh_status zh_execution_step(zh_execution* e, uint32_t max_launches, uint32_t* launched) {
  if (!e) return fail(ZH_ERR_INVALID_ARGUMENT, "null execution");
  if (max_launches == 0) return fail(ZH_ERR_INVALID_ARGUMENT, "a step launches at least one kernel");
  uint32_t n = 0;
  if (e->stepped && e->next < e->units) {
    n = (uint32_t)std::min<size_t>(max_launches, e->units - e->next);
    zh_status st = program_launch_units(e->prog, e->next, n);
    if (st != ZH_OK) return st;
    e->next += n;
    if (e->next == e->units && e->prog->stepping == e) e->prog->stepping = nullptr;
  }
  if (launched) *launched = n;
  return ZH_OK;
}

zh_status zh_execution_is_running(zh_execution* e, int32_t* running) {
  if (!e || !running) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  if (e->stepped && e->next < e->units) {
    *running = 1;
    return ZH_OK;
  }
  return zh_query(e->prog->ctx, running);
}

zh_status zh_execution_resources(zh_execution* e, zh_resources_used* used) {
  if (!e || !used) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  *used = e->usage;
  return ZH_OK;
}

zh_status zh_execution_prune(zh_execution* e) {
  if (!e) return fail(ZH_ERR_INVALID_ARGUMENT, "null execution");
  e->prune = true;
  return ZH_OK;
}

zh_status zh_execution_discard(zh_execution* e) {
  if (!e) return ZH_OK;
  zh_status st = zh_sync(e->prog->ctx);
  release(e, true);
  delete e;
  return st;
}

zh_status zh_execution_retire_output(zh_execution* e, uint32_t slot, int32_t to_host, zh_pool_key* key) {
  if (!e || !key) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  if (slot >= e->out.size() || !e->out[slot].bound) return fail(ZH_ERR_INVALID_ARGUMENT, "no such output");
  zh_execution::Slot& sl = e->out[slot];
  if (sl.retired) return fail(ZH_ERR_INVALID_ARGUMENT, "output already retired");
  if (e->stepped && e->next < e->units) return fail(ZH_ERR_INVALID_ARGUMENT, "the execution has launches left to step");
  zh_pool* pool = e->pool;
  zh_context* ctx = e->prog->ctx;
  // The image the output goes to: its key, or a new one (run.rs:2809-2875).
  zh_pool_key k = sl.key;
  if (!k || pool->items.find(k) == pool->items.end()) {
    k = pool->next++;
    PoolImage im;
    im.desc = tight(e->prog->plan.output_desc[slot]);
    pool->items[k] = im;
  }
  PoolImage& im = pool->items[k];
  if (sl.in_place && !to_host) {
    // Already written into the pool's own device image.
  } else if (!to_host) {
    im.kind = ZH_IMAGE_HIP_BUFFER;
    im.dev = sl.dev;  // handed over without a copy
    im.host.reset();
    im.desc = sl.dev->img->desc;
    sl.dev.reset();
  } else {
    HostPtr host = std::make_shared<std::vector<uint8_t>>(tight_stride(im.desc) * im.desc.height);
    zh_status st = zh_image_download(sl.dev->img, host->data(), 0);
    if (st == ZH_OK) st = zh_sync(ctx);
    if (st != ZH_OK) return st;
    im.kind = ZH_IMAGE_HOST;
    im.host = host;
    im.dev.reset();
    im.desc = tight(im.desc);
    if (!sl.in_place && !e->prune) pool->cache.push_back(sl.dev);
    sl.dev.reset();
  }
  sl.retired = true;
  *key = k;
  return ZH_OK;
}

zh_status zh_execution_retire_input(zh_execution* e, uint32_t slot, zh_pool_key* key) {
  if (!e || !key) return fail(ZH_ERR_INVALID_ARGUMENT, "null argument");
  if (slot >= e->in.size() || !e->in[slot].bound) return fail(ZH_ERR_INVALID_ARGUMENT, "no such input");
  zh_execution::Slot& sl = e->in[slot];
  if (e->stepped && e->next < e->units) return fail(ZH_ERR_INVALID_ARGUMENT, "the execution has launches left to step");
  if (sl.moved) {
    zh_status st = zh_sync(e->prog->ctx);
    if (st != ZH_OK) return st;
    auto it = e->pool->items.find(sl.key);
    if (it != e->pool->items.end()) it->second = std::move(sl.moved_image);
    sl.moved = false;
  }
  sl.retired = true;
  *key = sl.key;
  return ZH_OK;
}

zh_status zh_execution_finish(zh_execution* e) {
  if (!e) return ZH_OK;
  zh_status st = zh_sync(e->prog->ctx);
  release(e, false);
  delete e;
  return st;
}

}  // extern "C"
