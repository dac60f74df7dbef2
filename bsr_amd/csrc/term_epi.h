// Device side of TermEpi (bsrg_internal.h): the band term of one element of a
// channel block, added to the exchange product in a product kernel's
// epilogue.  Included by the HIP sources only.
#pragma once

#include <hip/hip_runtime.h>

#include "bsrg_internal.h"

namespace bsrg {

// Device pointers in the address spaces that select the load instructions:
// constant (read-only during the kernel; a wave-uniform load becomes a
// scalar load) and global.  Kernel arguments are known to be global; these
// also hold for pointers read from memory (a batch's descriptors), whose
// address space the compiler cannot infer.
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T* p) {
  return (cptr<T>)(p);
}
template <typename T>
__device__ __forceinline__ gptr<T> as_global(T* p) {
  return (gptr<T>)(p);
}

// The band term of element (x, y) of block blk: true, with its value in v,
// when the element lies within ks-1 of the diagonal and the block has band
// terms (the block's summed band at (min(x, y), |x - y|)); called before
// the product so that the load's latency hides behind it, and added after.
__device__ __forceinline__ bool epi_band(const TermEpi& te, int blk, int x, int y, double& v) {
  const int d = x - y, ad = d < 0 ? -d : d;
  const cptr<int> bptr = as_const(te.bptr);
  if (!(ad < te.ks && bptr[blk + 1] > bptr[blk])) return false;
  v = as_global(const_cast<double*>(te.bsum))[static_cast<size_t>(blk) * te.ns * te.ks +
                                               static_cast<size_t>(min(x, y)) + static_cast<size_t>(te.ns) * ad];
  return true;
}

}  // namespace bsrg
