// zh_spec.hip -- compile-time specialised pointwise kernels.
//
// A flat pointwise stage (every slot tight, 16-byte aligned and of the output
// size) whose micro-op sequence matches a registered signature (zh_spec.h)
// runs here as straight-line code.  The per-stage arguments -- slot base
// addresses and matrix coefficients -- are loaded once from the descriptor
// into scalar registers before the pixel loop; the loop itself is loads,
// colour math and stores with no dispatch (zh_spec_kernels.h).  Signatures
// not registered here are compiled at run time from the same bodies
// (zh_jit.cpp).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstring>

#include "zh_spec_kernels.h"

namespace zh {
namespace spec {

template <class P>
__global__ __launch_bounds__(kBlock) void k_spec(const StageDesc* dd) {
  spec_stage<P>(dd);
}

template <class P>
__global__ __launch_bounds__(kBlock) void k_spec_batch(const BatchDesc* bb) {
  spec_batch<P>(bb);
}

// ---- registry ---------------------------------------------------------------
typedef void (*KernelFn)(const StageDesc*);
typedef void (*BatchFn)(const BatchDesc*);

struct Entry {
  const char* name;
  const uint16_t* sig;
  int n;
  KernelFn fn;
  BatchFn batch;
};

#define ZH_SPEC_PROG(name, ...)                                        \
  typedef Prog<__VA_ARGS__> P_##name;                                  \
  static const uint16_t S_##name[] = {__VA_ARGS__};
ZH_SPEC_LIST(ZH_SPEC_PROG)
#undef ZH_SPEC_PROG

#define ZH_SPEC_ENTRY(name, ...) \
  {#name, S_##name, (int)(sizeof(S_##name) / sizeof(uint16_t)), k_spec<P_##name>, k_spec_batch<P_##name>},
static const Entry kEntries[] = {ZH_SPEC_LIST(ZH_SPEC_ENTRY)};
#undef ZH_SPEC_ENTRY

constexpr int kNumEntries = (int)(sizeof(kEntries) / sizeof(kEntries[0]));

}  // namespace spec
}  // namespace zh

extern "C" {

int zh_spec_lookup(const uint16_t* sig, int n) {
  for (int i = 0; i < zh::spec::kNumEntries; ++i) {
    const auto& e = zh::spec::kEntries[i];
    if (e.n == n && std::memcmp(e.sig, sig, sizeof(uint16_t) * (size_t)n) == 0) return i;
  }
  return -1;
}

const char* zh_spec_name(int index) {
  return index >= 0 && index < zh::spec::kNumEntries ? zh::spec::kEntries[index].name : "";
}

hipError_t zh_launch_spec(int index, const zh::StageDesc* d_dev, uint32_t width, uint32_t height, uint32_t cus,
                          hipStream_t stream) {
  if (index < 0 || index >= zh::spec::kNumEntries) return hipErrorInvalidValue;
  static uint32_t resident[zh::spec::kNumEntries] = {};
  if (!resident[index]) resident[index] = zh::resident_blocks((const void*)zh::spec::kEntries[index].fn, 1);
  const uint64_t groups = ((uint64_t)width * height + zh::kPix - 1) / zh::kPix;
  const uint64_t blocks = std::min<uint64_t>((groups + zh::kBlock - 1) / zh::kBlock, (uint64_t)resident[index] * cus);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(zh::spec::kEntries[index].fn, dim3((uint32_t)blocks), dim3(zh::kBlock), 0, stream, d_dev);
  return hipGetLastError();
}

// A batch of stages of signature `index`: total_chunks chunks of
// kSpecChunkPixels (the BatchDesc prefix counts chunks).
hipError_t zh_launch_spec_batch(int index, const zh::BatchDesc* b_dev, uint32_t total_chunks, uint32_t cus,
                                hipStream_t stream) {
  if (index < 0 || index >= zh::spec::kNumEntries) return hipErrorInvalidValue;
  static uint32_t resident[zh::spec::kNumEntries] = {};
  if (!resident[index]) resident[index] = zh::resident_blocks((const void*)zh::spec::kEntries[index].batch, 1);
  const uint32_t blocks = std::min(total_chunks, resident[index] * cus);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(zh::spec::kEntries[index].batch, dim3(blocks), dim3(zh::kBlock), 0, stream, b_dev);
  return hipGetLastError();
}

}  // extern "C"
