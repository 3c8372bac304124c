// zh_spec_kernels.h -- the bodies of the specialised pointwise kernels,
// device code only: included by zh_spec.hip for the registered signatures
// and by the kernels zh_jit.cpp compiles at run time for the others.
//
// Each lane owns kPix = 4 consecutive pixels per step and walks the image in a
// persistent grid-stride loop with two raw-texel buffers in ping-pong: the
// loads of step k+1 are issued before the colour math of step k.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "zh_prog.h"

namespace zh {
namespace spec {

__device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }


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
    fetch<P, true>(u, umin64(g + stride, last - 1) * kPix, kPix, rb);
    {
      Px X, Y, Z;
      exec<P, true>(u, g * kPix, kPix, ra, X, Y, Z, L);
    }
    g += stride;
    if (g >= last) break;
    fetch<P, true>(u, umin64(g + stride, last - 1) * kPix, kPix, ra);
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
__device__ __forceinline__ void spec_stage(const StageDesc* dd) {
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
__device__ __forceinline__ void spec_batch(const BatchDesc* bb) {
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
    spec_span<P>(u, &luts, g0 + threadIdx.x, umin64(g0 + CG, nfull + 1), kBlock, nfull, n);
  }
}

}  // namespace spec
}  // namespace zh
