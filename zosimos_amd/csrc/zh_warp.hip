// zh_warp.hip -- the registered affine layer-stack kernels (zh_spec.h
// ZH_WARP_LIST): instantiations of the bodies in zh_warp_kernels.h, their
// registry and launchers.  Forms not registered here are compiled at run time
// from the same bodies (zh_jit.cpp).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstring>

#include "zh_warp_kernels.h"

namespace zh {
namespace spec {

template <class V>
__global__ __launch_bounds__(kBlock) void k_warp(const StageDesc* dd) {
  warp_stage<V>(dd);
}

template <class V>
__global__ __launch_bounds__(kBlock) void k_warp_batch(const BatchDesc* bb) {
  warp_batch<V>(bb);
}

// LDS-staged tiles of one-source separable forms (staged_tile).
template <class V>
__global__ __launch_bounds__(kBlock) void k_warp_staged(const StageDesc* dd) {
  warp_staged_stage<V>(dd);
}

template <class V>
__global__ __launch_bounds__(kBlock) void k_warp_staged_batch(const BatchDesc* bb) {
  warp_staged_batch<V>(bb);
}

template <class V>
constexpr void (*staged_fn())(const StageDesc*) {
  if constexpr (V::SEP && V::SRC == WSRC_ONE) return k_warp_staged<V>;
  else return nullptr;
}
template <class V>
constexpr void (*staged_batch_fn())(const BatchDesc*) {
  if constexpr (V::SEP && V::SRC == WSRC_ONE) return k_warp_staged_batch<V>;
  else return nullptr;
}

// ---- registry ------------------------------------------------------------------
#define ZH_WARP_TYPE(name, src, f, mat, ...)                         \
  typedef Warp<src, f, mat != 0, Prog<__VA_ARGS__>, false> V_##name;    \
  typedef Warp<src, f, mat != 0, Prog<__VA_ARGS__>, true> VS_##name;    \
  static const uint16_t WP_##name[] = {__VA_ARGS__};
ZH_WARP_LIST(ZH_WARP_TYPE)
#undef ZH_WARP_TYPE

// Kernel index 2 i: form i with general affine maps; 2 i + 1: separable.
enum {
#define ZH_WARP_ENUM(name, ...) WI_##name, WIS_##name,
  ZH_WARP_LIST(ZH_WARP_ENUM)
#undef ZH_WARP_ENUM
  kNumWarps
};

struct WarpEntry {
  const char* name;
  uint32_t src, fmt, mat;
  const uint16_t* post;
  int n;
  void (*fn)(const StageDesc*);
  void (*batch)(const BatchDesc*);
  void (*staged)(const StageDesc*);  // separable one-source forms: LDS-staged tiles
  void (*staged_batch)(const BatchDesc*);
};

#define ZH_WARP_ENTRY(name, src, f, mat, ...)                                                          \
  {#name, src, f, mat, WP_##name, (int)(sizeof(WP_##name) / sizeof(uint16_t)), k_warp<V_##name>,         \
   k_warp_batch<V_##name>, nullptr, nullptr},                                                            \
  {#name "_sep", src, f, mat, WP_##name, (int)(sizeof(WP_##name) / sizeof(uint16_t)), k_warp<VS_##name>, \
   k_warp_batch<VS_##name>, staged_fn<VS_##name>(), staged_batch_fn<VS_##name>()},
static const WarpEntry kWarps[] = {ZH_WARP_LIST(ZH_WARP_ENTRY)};
#undef ZH_WARP_ENTRY

}  // namespace spec
}  // namespace zh

extern "C" {

int zh_warp_lookup(uint32_t src, uint32_t fmt, uint32_t mat, const uint16_t* post, int n) {
  for (int i = 0; i < zh::spec::kNumWarps; i += 2) {
    const auto& e = zh::spec::kWarps[i];
    if (e.src == src && e.fmt == fmt && e.mat == mat && e.n == n &&
        std::memcmp(e.post, post, sizeof(uint16_t) * (size_t)n) == 0)
      return i;
  }
  return -1;
}

const char* zh_warp_name(int index) {
  return index >= 0 && index < zh::spec::kNumWarps ? zh::spec::kWarps[index].name : "";
}

int zh_warp_form_of(int index, uint32_t* src, uint32_t* fmt, uint32_t* mat, const uint16_t** post, int* n) {
  if (index < 0 || index >= zh::spec::kNumWarps) return 0;
  const auto& e = zh::spec::kWarps[index];
  *src = e.src;
  *fmt = e.fmt;
  *mat = e.mat;
  *post = e.post;
  *n = e.n;
  return 1;
}

int zh_warp_has_staged(int index) {
  return index >= 0 && index < zh::spec::kNumWarps && zh::spec::kWarps[index].staged != nullptr;
}

// per_cu: workgroups per compute unit of the persistent grid (0: as many as
// are resident; fewer keeps more of each workgroup's taps in the L1).
// staged: the LDS-staged kernel of a form that has one.
hipError_t zh_launch_warp(int index, int staged, const zh::StageDesc* d_dev, uint32_t tiles, uint32_t cus,
                          uint32_t per_cu, hipStream_t stream) {
  if (index < 0 || index >= zh::spec::kNumWarps) return hipErrorInvalidValue;
  const auto& e = zh::spec::kWarps[index];
  if (staged && !e.staged) return hipErrorInvalidValue;
  void (*fn)(const zh::StageDesc*) = staged ? e.staged : e.fn;
  static uint32_t resident[2][zh::spec::kNumWarps] = {};
  uint32_t& res = resident[staged ? 1 : 0][index];
  if (!res) res = zh::resident_blocks((const void*)fn, 1);
  const uint32_t bpc = per_cu ? std::min(per_cu, res) : res;
  const uint32_t blocks = std::min(tiles, bpc * cus);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(zh::kBlock), 0, stream, d_dev);
  return hipGetLastError();
}

hipError_t zh_launch_warp_batch(int index, int staged, const zh::BatchDesc* b_dev, uint32_t total_tiles,
                                uint32_t cus, uint32_t per_cu, hipStream_t stream) {
  if (index < 0 || index >= zh::spec::kNumWarps) return hipErrorInvalidValue;
  const auto& e = zh::spec::kWarps[index];
  if (staged && !e.staged_batch) return hipErrorInvalidValue;
  void (*fn)(const zh::BatchDesc*) = staged ? e.staged_batch : e.batch;
  static uint32_t resident[2][zh::spec::kNumWarps] = {};
  uint32_t& res = resident[staged ? 1 : 0][index];
  if (!res) res = zh::resident_blocks((const void*)fn, 1);
  const uint32_t bpc = per_cu ? std::min(per_cu, res) : res;
  const uint32_t blocks = std::min(total_tiles, bpc * cus);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(zh::kBlock), 0, stream, b_dev);
  return hipGetLastError();
}

}  // extern "C"
