// Device side of TermEpi (bsrg_internal.h): the band term of one element of a
// channel block, added to the exchange product in a product kernel's
// epilogue.  Included by the HIP sources only.
#pragma once

#include <hip/hip_runtime.h>

#include "bsrg_internal.h"

namespace bsrg {

// The band term of element (x, y) of block blk: true, with its value in v,
// when the element lies within ks-1 of the diagonal and the block has band
// terms (the block's summed band at (min(x, y), |x - y|)); called before
// the product so that the load's latency hides behind it, and added after.
__device__ __forceinline__ bool epi_band(const TermEpi& te, int blk, int x, int y, double& v) {
  const int d = x - y, ad = d < 0 ? -d : d;
  if (!(ad < te.ks && te.bptr[blk + 1] > te.bptr[blk])) return false;
  v = te.bsum[static_cast<size_t>(blk) * te.ns * te.ks + static_cast<size_t>(min(x, y)) +
              static_cast<size_t>(te.ns) * ad];
  return true;
}

}  // namespace bsrg
