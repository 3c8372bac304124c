// zh_spec.hip -- compile-time specialised pointwise kernels.
//
// A flat pointwise stage (every slot tight, 16-byte aligned and of the output
// size) whose micro-op sequence matches a registered signature (zh_spec.h)
// runs here as straight-line code.  The per-stage arguments -- slot base
// addresses and matrix coefficients -- are loaded once from the descriptor
// into scalar registers before the pixel loop; the loop itself is loads,
// colour math and stores with no dispatch.
//
// Each lane owns kPix = 4 consecutive pixels per step and walks the image in a
// persistent grid-stride loop with two raw-texel buffers in ping-pong: the
// loads of step k+1 are issued before the colour math of step k.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstring>

#include "zh_prog.h"

namespace zh {
namespace spec {

// Groups g, g + stride, ... below `end` of one flat image of `total` pixels
// (nfull whole groups of kPix), with the next group's texels in flight while
// the current one is computed.  The clamp keeps the look-ahead load of a
// lane's last step inside the span (it re-reads the last group) instead of
// branching around it.  The image's partial last group, if any, is group
// nfull: the lane that reaches it with nfull < end writes it texel by texel.
template <class P>
__device__ __forceinline__ void spec_span(const Uni<P>& u, const Luts* L, uint64_t g, uint64_t end, uint64_t stride,
                                          uint64_t nfull, uint64_t total) {
  const uint64_t last = end < nfull ? end : nfull;
  uint4 ra[P::nwords], rb[P::nwords];
  if (g < last) fetch<P, true>(u, g * kPix, kPix, ra);
  while (g < last) {
    fetch<P, true>(u, std::min(g + stride, last - 1) * kPix, kPix, rb);
    {
      Px X, Y, Z;
      exec<P, true>(u, g * kPix, kPix, ra, X, Y, Z, L);
    }
    g += stride;
    if (g >= last) break;
    fetch<P, true>(u, std::min(g + stride, last - 1) * kPix, kPix, ra);
    {
      Px X, Y, Z;
      exec<P, true>(u, g * kPix, kPix, rb, X, Y, Z, L);
    }
    g += stride;
  }
  if (g == nfull && nfull < end && nfull * kPix < total) {
    const int n = (int)(total - nfull * kPix);
    fetch<P, false>(u, g * kPix, n, ra);
    Px X, Y, Z;
    exec<P, false>(u, g * kPix, n, ra, X, Y, Z, L);
  }
}

// One stage: a persistent grid striding over the image's groups.
template <class P>
__global__ __launch_bounds__(kBlock) void k_spec(const StageDesc* dd) {
  const CDesc d = as_const(dd);
  __shared__ Luts luts;
  load_lut(d, &luts);
  Uni<P> u;
  resolve<P>(d, u);
  __syncthreads();
  const uint64_t total = (uint64_t)d->width * d->height;
  const uint64_t nfull = total / kPix;
  spec_span<P>(u, &luts, (uint64_t)blockIdx.x * kBlock + threadIdx.x, nfull + 1, (uint64_t)gridDim.x * kBlock, nfull,
               total);
}

// Independent stages of one signature (the images of a batch) as one launch:
// each stage is cut into chunks of kSpecChunkPixels consecutive pixels, chunk
// j of the batch belongs to the last stage s with prefix[s] <= j, and a
// workgroup takes chunks grid-stride, resolving the stage's uniform
// arguments once per chunk.  One launch instead of one per image removes the
// per-launch ramp and drain of the grid.
template <class P>
__global__ __launch_bounds__(kBlock) void k_spec_batch(const BatchDesc* bb) {
  const CBatch b = (CBatch)(uintptr_t)bb;
  __shared__ Luts luts;
  load_lut(as_const(b->stages), &luts);  // one context: every stage holds the same tables
  __syncthreads();
  const uint32_t total = b->total_tiles;
  const uint32_t c0 = blockIdx.x;
  if (c0 >= total) return;
  uint32_t lo = 0, hi = b->n_stages;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (cword(b->tile_prefix, mid) <= c0) lo = mid; else hi = mid;
  }
  constexpr uint64_t CG = (uint64_t)kSpecChunkGroups * kBlock;  // groups per chunk
  for (uint32_t c = c0; c < total; c += gridDim.x) {
    while (cword(b->tile_prefix, lo + 1) <= c) ++lo;  // uniform
    const CDesc d = as_const(b->stages + cword(b->stage_ids, lo));
    Uni<P> u;
    resolve<P>(d, u);
    const uint64_t n = (uint64_t)d->width * d->height, nfull = n / kPix;
    const uint64_t g0 = (uint64_t)(c - cword(b->tile_prefix, lo)) * CG;
    spec_span<P>(u, &luts, g0 + threadIdx.x, std::min(g0 + CG, nfull + 1), kBlock, nfull, n);
  }
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
