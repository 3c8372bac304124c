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
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  uint4 ra[P::nwords], rb[P::nwords];
  // The clamp keeps the look-ahead load of a lane's last step inside the
  // image (it re-reads the last group) instead of branching around it.
  if (g < nfull) fetch<P, true>(u, g * kPix, kPix, ra);
  while (g < nfull) {
    fetch<P, true>(u, std::min(g + stride, nfull - 1) * kPix, kPix, rb);
    {
      Px X, Y, Z;
      exec<P, true>(u, g * kPix, kPix, ra, X, Y, Z, &luts);
    }
    g += stride;
    if (g >= nfull) break;
    fetch<P, true>(u, std::min(g + stride, nfull - 1) * kPix, kPix, ra);
    {
      Px X, Y, Z;
      exec<P, true>(u, g * kPix, kPix, rb, X, Y, Z, &luts);
    }
    g += stride;
  }
  // The partial group at the end of the image (one lane of the grid).
  if (g == nfull && nfull * kPix < total) {
    const int n = (int)(total - nfull * kPix);
    fetch<P, false>(u, g * kPix, n, ra);
    Px X, Y, Z;
    exec<P, false>(u, g * kPix, n, ra, X, Y, Z, &luts);
  }
}

// ---- registry ---------------------------------------------------------------
typedef void (*KernelFn)(const StageDesc*);

struct Entry {
  const char* name;
  const uint16_t* sig;
  int n;
  KernelFn fn;
};

#define ZH_SPEC_PROG(name, ...)                                        \
  typedef Prog<__VA_ARGS__> P_##name;                                  \
  static const uint16_t S_##name[] = {__VA_ARGS__};
ZH_SPEC_LIST(ZH_SPEC_PROG)
#undef ZH_SPEC_PROG

#define ZH_SPEC_ENTRY(name, ...) {#name, S_##name, (int)(sizeof(S_##name) / sizeof(uint16_t)), k_spec<P_##name>},
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

}  // extern "C"
