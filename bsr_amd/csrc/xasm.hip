// The exchange contraction over each block's nonzero coefficients (engine 4
// of the FP64 wave).
//
// H[blk](x,y) = sum_t C(t, blk) E_t(x,y)              exchange (jtype 8, 9)
//
// The coefficient matrix C (nt x nblk, nt = nk nb^2) is sparse: a record of
// jtype 8/9 touches one multipole and one ordered bound pair of one channel
// pair, so a block holds a few tens of the nt possible t (under a tenth at
// the production shape).  A dense product spends most of its FLOPs on
// zeros; here each block runs over its own nonzeros only.
//
// Layout: a workgroup owns 64 consecutive positions p = x + ns y of the
// ns^2 block (one 512-byte row tile of every block) and a range of blocks.
// It stages the tile of every used basis E_t, [t][64] doubles, in LDS once
// and then runs over its blocks, one wavefront per block: lane l sums
// c_j E_{t_j}(p0 + l) over the block's nonzeros and writes its element:
// 512 coalesced bytes per block and tile, and no read of H.
//
// Two forms:
//   launch_xprod  the product alone, in the GEMM's place: a one-shot wave
//                 starts it as soon as the exchange records are decoded and
//                 the term pass adds the band and rank-1 terms afterwards;
//   launch_xasm   a wave whose product was not started early (prepared or
//                 captured waves): the band terms join the product in the
//                 same pass (H written once), then the term pass adds only
//                 the rank-1 terms to the blocks that have them.
//
// Sums: the exchange part in ascending t (four interleaved partial sums
// added pairwise), then the band term, then the rank-1 terms in list order
// (the order of the term pass after a separate product, so both forms give
// the same bits): a fixed order, so a block's values do not depend on what
// else runs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "bsrg_internal.h"
#include "term_epi.h"

namespace bsrg {

namespace {

constexpr int kXWaves = 16;  // wavefronts per workgroup (4 per SIMD)
constexpr int kXChunk = 8;  // list entries per scalar-load round (lists are padded to a multiple)

// Per block (one wavefront): the nonzeros of column blk of C (rows 0..nr-1,
// leading dimension ldc) in ascending t, stored as (byte offset t*rowb of
// E_t's tile in LDS, c) in slots blk*nrs.. (nrs = nr rounded up to
// kXChunk), padded with (0, 0) up to the next multiple of kXChunk; their
// count in xn[blk].
__global__ __launch_bounds__(256) void k_coef_pack(int nblk, int nr, int nrs, const double* __restrict__ C, size_t ldc,
                                                   int rowb, int* __restrict__ xo, double* __restrict__ xc,
                                                   int* __restrict__ xn) {
  const int blk = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (blk >= nblk) return;
  const double* col = C + static_cast<size_t>(blk) * ldc;
  int* ol = xo + static_cast<size_t>(blk) * nrs;
  double* cl = xc + static_cast<size_t>(blk) * nrs;
  int base = 0;
  for (int t0 = 0; t0 < nr; t0 += 64) {
    const int t = t0 + lane;
    const double v = t < nr ? col[t] : 0.0;
    const bool nz = v != 0.0;
    const unsigned long long m = __ballot(nz);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (nz) {
      ol[base + before] = t * rowb;
      cl[base + before] = v;
    }
    base += __popcll(m);
  }
  const int pad = (base + kXChunk - 1) / kXChunk * kXChunk;
  if (base + lane < pad) {
    ol[base + lane] = 0;
    cl[base + lane] = 0.0;
  }
  if (lane == 0) xn[blk] = base;
}

// One work item: tile `tile` (64 positions) of the blocks of group `grp`
// ([grp*per, min(nblk, (grp+1)*per)), per = ceil(nblk/ngroup)), by the
// whole 1024-thread workgroup.  E: the used bases, column t at E + t*nss.
// LDS: nr * 64 doubles.  The lists of a block are wave-uniform: kXChunk
// entries at a time come in through scalar loads and serve as scalar
// operands, so the inner loop is one address add, one LDS read and one FMA
// per nonzero (the padding adds zeros).  The band term (one read of the
// block's summed band, for the elements within ks-1 of the diagonal) is
// requested before the sum and added after it in the fused form.
template <bool FUSED>
__device__ __forceinline__ void xasm_item(int ns, int nblk, int nr, int nrs, int ngroup, int tile, int grp,
                                          const double* __restrict__ E, const int* __restrict__ xo,
                                          const double* __restrict__ xc, const int* __restrict__ xn,
                                          const TermEpi& te, const int* __restrict__ rptr,
                                          const unsigned char* __restrict__ live, double* __restrict__ hcc,
                                          double* El) {
  const size_t nss = static_cast<size_t>(ns) * ns;
  const size_t p0 = static_cast<size_t>(tile) * 64;
  const gptr<double> Eg = as_global(const_cast<double*>(E));
  const gptr<double> hg = as_global(hcc);
  const cptr<int> xoc = as_const(xo), xnc = as_const(xn), rpc = as_const(rptr), bpc = as_const(te.bptr);
  const cptr<double> xcc = as_const(xc);
  const cptr<unsigned char> livec = as_const(live);
  {
    // stage the tile: eight loads in flight per thread
    const int tot = nr * 64;
    for (int q0 = threadIdx.x; q0 < tot; q0 += 8 * 64 * kXWaves) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int q = q0 + u * 64 * kXWaves, t = q >> 6;
        const size_t pp = p0 + (q & 63);
        v[u] = (q < tot && pp < nss) ? Eg[static_cast<size_t>(t) * nss + pp] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int q = q0 + u * 64 * kXWaves;
        if (q < tot) El[q] = v[u];
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const size_t p = p0 + lane;
  const bool valid = p < nss;
  const int x = valid ? static_cast<int>(p % ns) : 0, y = valid ? static_cast<int>(p / ns) : 0;
  const char* Eb = reinterpret_cast<const char*>(El + lane);
  const int per = (nblk + ngroup - 1) / ngroup, blo = grp * per, bhi = min(nblk, blo + per);
  for (int blk = blo + wv; blk < bhi; blk += kXWaves) {
    const int n = xnc[blk];
    // a block without terms that already holds zeros stays as it is
    if (FUSED && n == 0 && bpc[blk] == bpc[blk + 1] && rpc[blk] == rpc[blk + 1] && live && !livec[blk])
      continue;
    double band = 0.0;
    const bool has_band = FUSED && epi_band(te, blk, x, y, band);
    const cptr<int> ol = xoc + static_cast<size_t>(blk) * nrs;
    const cptr<double> cl = xcc + static_cast<size_t>(blk) * nrs;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    // two chunks (16 LDS reads in flight) per round over the padded list,
    // then the last chunk
    const int npad = (n + kXChunk - 1) / kXChunk * kXChunk;
    int j0 = 0;
    for (; j0 + 2 * kXChunk <= npad; j0 += 2 * kXChunk) {
      int oo[2 * kXChunk];
      double cc[2 * kXChunk];
#pragma unroll
      for (int i = 0; i < 2 * kXChunk; ++i) {
        oo[i] = ol[j0 + i];
        cc[i] = cl[j0 + i];
      }
#pragma unroll
      for (int i = 0; i < 2 * kXChunk; ++i) a[i & 3] += cc[i] * *reinterpret_cast<const double*>(Eb + oo[i]);
    }
    if (j0 < npad) {
      int oo[kXChunk];
      double cc[kXChunk];
#pragma unroll
      for (int i = 0; i < kXChunk; ++i) {
        oo[i] = ol[j0 + i];
        cc[i] = cl[j0 + i];
      }
#pragma unroll
      for (int i = 0; i < kXChunk; ++i) a[i & 3] += cc[i] * *reinterpret_cast<const double*>(Eb + oo[i]);
    }
    double acc = (a[0] + a[1]) + (a[2] + a[3]);
    if (has_band) acc += band;
    if (valid) hg[static_cast<size_t>(blk) * nss + p] = acc;
  }
}

// grid ntile * ngroup: workgroup -> (tile = id % ntile, group = id / ntile).
template <bool FUSED>
__global__ __launch_bounds__(64 * kXWaves) void k_xasm(
    int ns, int nblk, int nr, int nrs, int ntile, int ngroup, const double* __restrict__ E,
    const int* __restrict__ xo, const double* __restrict__ xc, const int* __restrict__ xn, TermEpi te,
    const int* __restrict__ rptr, const unsigned char* __restrict__ live, double* __restrict__ hcc) {
  extern __shared__ double El[];
  xasm_item<FUSED>(ns, nblk, nr, nrs, ngroup, static_cast<int>(blockIdx.x % ntile),
                   static_cast<int>(blockIdx.x / ntile), E, xo, xc, xn, te, rptr, live, hcc, El);
}

// The fused products of a batch of waves in one persistent launch: each
// workgroup takes the next work item (a tile of a group of blocks of one
// wave) from a counter until the counter passes the last item.  The waves'
// items are numbered in order (wave w owns [item0, item0 + ntile*ngroup));
// the host puts the waves with the most work first, so the small items of
// the small waves fill the end.  Every element is computed by xasm_item as
// in the wave's own launch, so the results are the same bits.  counter is
// zero at the start (the launcher clears it with a kernel).
constexpr int kXBatchMax = 256;  // waves in one batched launch
__global__ __launch_bounds__(64 * kXWaves) void k_xasm_batch(const XWave* __restrict__ waves, int nw, int total,
                                                             int* __restrict__ counter) {
  extern __shared__ double El[];
  __shared__ int s_it[2], s_first[kXBatchMax];
  for (int w = threadIdx.x; w < nw; w += blockDim.x) s_first[w] = waves[w].item0;
  if (threadIdx.x == 0) s_it[0] = atomicAdd(counter, 1);
  __syncthreads();
  // the item index is the same in every lane; taken as a scalar, everything
  // derived from it (tile, group, block numbers, list addresses) stays in
  // scalar registers and the lists come in through scalar loads
  int it = __builtin_amdgcn_readfirstlane(s_it[0]);
  // the next item's index is taken while the current item runs: item k
  // reads slot k&1, and thread 0 fills slot (k+1)&1 before the barrier that
  // ends item k (every thread read that slot before the barrier of item k-1)
  for (int k = 0; it < total; ++k) {  // `it` is the same in every thread
    if (threadIdx.x == 0) s_it[(k + 1) & 1] = atomicAdd(counter, 1);
    // the wave of the item: the last one whose first item is <= it
    int lo = 0, hi = nw - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_first[mid] <= it) lo = mid;
      else hi = mid - 1;
    }
    const XWave X = waves[__builtin_amdgcn_readfirstlane(lo)];  // held in scalar registers for the item
    const int local = it - X.item0;
    xasm_item<true>(X.ns, X.nblk, X.nr, X.nrs, X.ngroup, local % X.ntile, local / X.ntile, X.E, X.xo, X.xc, X.xn,
                    X.te, X.rptr, X.live, X.hcc, El);
    __syncthreads();  // the tile is free, and the next index is in its slot
    it = __builtin_amdgcn_readfirstlane(s_it[(k + 1) & 1]);
  }
}

// The Dirac form (one family, one quadrant): H_q[blk](x,y) = sum_t C(t,blk)
// E_f,t(x,y) written to quadrant (qa, qb) of channel block blk (order 2ns,
// leading dimension 2ns).  A family may use more bases than a 64-position
// tile of them fits in LDS (325 at ks 9, k <= 12, five bound orbitals: 166
// KB), so the tile is wt positions, the widest of 64, 48 and 32 that fits
// ([t][wt] doubles, all nr rows at once); lanes past wt idle.  As in
// k_xasm, one wavefront per block with its list wave-uniform (scalar loads,
// scalar operands): a lane's work per nonzero is one address add, one LDS
// read and one FMA.  Lists from the slots of k_coef_pack (offsets t*wt*8).
// grid ntile * ngroup, ntile = ceil(ns^2/wt).
__global__ __launch_bounds__(64 * kXWaves) void k_xquad(int ns, int nblk, int nr, int nrs, int wt, int ntile,
                                                        int ngroup, const double* __restrict__ E,
                                                        const int* __restrict__ xo, const double* __restrict__ xc,
                                                        const int* __restrict__ xn, int qa, int qb,
                                                        double* __restrict__ hch) {
  extern __shared__ double El[];
  const size_t nss = static_cast<size_t>(ns) * ns, ms = 2 * static_cast<size_t>(ns);
  // workgroups go to the eight XCDs in turn (id mod 8); numbering them XCD
  // by XCD puts neighbouring tiles on one XCD, so the 128-byte lines that two
  // tiles share (a tile starts at any 8-byte alignment in the quadrant) are
  // merged in one L2 before they are written back
  unsigned g = blockIdx.x;
  const unsigned per8 = gridDim.x / 8;
  if (g < per8 * 8) g = (g % 8) * per8 + g / 8;
  const int tile = static_cast<int>(g % ntile), grp = static_cast<int>(g / ntile);
  const size_t p0 = static_cast<size_t>(tile) * wt;
  {
    const int tot = nr * wt;
    for (int q0 = threadIdx.x; q0 < tot; q0 += 8 * 64 * kXWaves) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        // clamped addresses, so all eight loads issue without branches
        const int q = min(q0 + u * 64 * kXWaves, tot - 1), t = q / wt;
        const size_t pp = p0 + (q - t * wt);
        const double e = E[static_cast<size_t>(t) * nss + min(pp, nss - 1)];
        v[u] = pp < nss ? e : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int q = q0 + u * 64 * kXWaves;
        if (q < tot) El[q] = v[u];
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const size_t p = p0 + lane;
  const bool valid = lane < wt && p < nss;
  const size_t x = valid ? p % ns : 0, y = valid ? p / ns : 0;
  const size_t off = (static_cast<size_t>(qa) * ns + x) + ms * (static_cast<size_t>(qb) * ns + y);
  // idle lanes read the tile's last column (in bounds) and write nothing
  const char* Eb = reinterpret_cast<const char*>(El + min(lane, wt - 1));
  const int per = (nblk + ngroup - 1) / ngroup, blo = grp * per, bhi = min(nblk, blo + per);
  for (int blk = blo + wv; blk < bhi; blk += kXWaves) {
    const int n = xn[blk];
    const int* ol = xo + static_cast<size_t>(blk) * nrs;
    const double* cl = xc + static_cast<size_t>(blk) * nrs;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int j0 = 0; j0 < n; j0 += kXChunk) {
      int oo[kXChunk];
      double cc[kXChunk];
#pragma unroll
      for (int i = 0; i < kXChunk; ++i) {
        oo[i] = ol[j0 + i];
        cc[i] = cl[j0 + i];
      }
#pragma unroll
      for (int i = 0; i < kXChunk; ++i) a[i & 3] += cc[i] * *reinterpret_cast<const double*>(Eb + oo[i]);
    }
    if (valid) hch[static_cast<size_t>(blk) * ms * ms + off] = (a[0] + a[1]) + (a[2] + a[3]);
  }
}

}  // namespace

int xquad_width(int nr, int device) {
  const int rows = xasm_max_t(device);  // rows of a 64-position tile
  if (nr <= rows) return 64;
  if (3 * nr <= 4 * rows) return 48;
  if (nr <= 2 * rows) return 32;
  return 0;
}

int xquad_max_t(int device) { return 2 * xasm_max_t(device); }

void launch_coef_pack_quad(int nblk, int nr, int wt, const double* C, size_t ldc, int* xo, double* xc, int* xn,
                           hipStream_t s) {
  if (nblk <= 0 || nr <= 0) return;
  hipLaunchKernelGGL(k_coef_pack, dim3((nblk + 3) / 4), dim3(256), 0, s, nblk, nr, xasm_slot(nr), C, ldc,
                     wt * static_cast<int>(sizeof(double)), xo, xc, xn);
}

int launch_xquad(int ns, int nblk, int nr, int wt, const double* E, const int* xo, const double* xc, const int* xn,
                 int qa, int qb, double* hch, hipStream_t s) {
  if (nblk <= 0) return 0;
  if (nr < 1 || (wt != 64 && wt != 48 && wt != 32)) return -1;
  const size_t nss = static_cast<size_t>(ns) * ns;
  const int ntile = static_cast<int>((nss + wt - 1) / wt);
  // about eight workgroups per CU in all, at least one block per wavefront
  int ngroup = std::max(1, (8 * 256 + ntile - 1) / ntile);
  ngroup = std::min(ngroup, std::max(1, nblk / kXWaves));
  hipLaunchKernelGGL(k_xquad, dim3(static_cast<unsigned>(ntile) * ngroup), dim3(64 * kXWaves),
                     static_cast<size_t>(nr) * wt * sizeof(double), s, ns, nblk, nr, xasm_slot(nr), wt, ntile, ngroup,
                     E, xo, xc, xn, qa, qb, hch);
  return 0;
}

int xasm_max_t(int device) {
  // the E tile of every used basis must fit the LDS a workgroup may use
  static std::mutex mu;
  static int lds_bytes[64] = {};
  if (device < 0 || device >= 64) return 0;
  std::lock_guard<std::mutex> g(mu);
  if (!lds_bytes[device]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || v <= 0) {
      (void)hipGetLastError();
      v = 64 * 1024;
    }
    lds_bytes[device] = v;
  }
  return lds_bytes[device] / (64 * static_cast<int>(sizeof(double)));
}

int xasm_slot(int nr) { return (nr + kXChunk - 1) / kXChunk * kXChunk; }

void launch_coef_pack(int nblk, int nr, const double* C, size_t ldc, int* xo, double* xc, int* xn, hipStream_t s) {
  if (nblk <= 0) return;
  hipLaunchKernelGGL(k_coef_pack, dim3((nblk + 3) / 4), dim3(256), 0, s, nblk, nr, xasm_slot(nr), C, ldc, 512, xo,
                     xc, xn);
}

void xasm_grid(int ns, int nblk, int& ntile, int& ngroup) {
  const size_t nss = static_cast<size_t>(ns) * ns;
  ntile = static_cast<int>((nss + 63) / 64);
  // about six workgroups per CU in all (one resident at a time when the
  // tile needs most of the LDS), at least one block per wavefront
  ngroup = std::max(1, (6 * 256 + ntile - 1) / ntile);
  ngroup = std::min(ngroup, std::max(1, nblk / kXWaves));
}

int launch_xasm_batch(const XWave* waves, int nw, int total, int max_nr, int nwg, int* counter, hipStream_t s) {
  if (nw < 1 || total < 1) return 0;
  if (nw > kXBatchMax || max_nr < 1 || nwg < 1) return -1;
  launch_fill_int(1, 0, counter, s);
  hipLaunchKernelGGL(k_xasm_batch, dim3(nwg), dim3(64 * kXWaves), static_cast<size_t>(max_nr) * 64 * sizeof(double),
                     s, waves, nw, total, counter);
  return 0;
}

int launch_xprod(int ns, int nblk, int nr, const double* E, const int* xo, const double* xc, const int* xn,
                 double* hcc, hipStream_t s) {
  if (nblk <= 0) return 0;
  if (nr < 1) return -1;
  int ntile, ngroup;
  xasm_grid(ns, nblk, ntile, ngroup);
  hipLaunchKernelGGL(k_xasm<false>, dim3(static_cast<unsigned>(ntile) * ngroup), dim3(64 * kXWaves),
                     static_cast<size_t>(nr) * 64 * sizeof(double), s, ns, nblk, nr, xasm_slot(nr), ntile, ngroup,
                     E, xo, xc, xn, TermEpi{}, nullptr, nullptr, hcc);
  return 0;
}

int launch_xasm(int ns, int ks, int nblk, int nr, const double* E, const int* xo, const double* xc, const int* xn,
                const int* bptr, const int* bid, const double* bc, const double* bands, const int* rptr,
                const int* ru, const int* rw, const int* rs, const double* rc, const double* stab, const double* V,
                double* hcc, double* bsum, double* rf, const unsigned char* live, const int* nobands,
                hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  if (nblk <= 0) return 0;
  if (nr < 1) return -1;
  launch_term_prep(ns, ks, nblk, bptr, bid, bc, bands, bsum, rptr, rs, rc, stab, rf, s);
  int ntile, ngroup;
  xasm_grid(ns, nblk, ntile, ngroup);
  const TermEpi te{ns, ks, bptr, bsum};
  if (e0) (void)hipEventRecord(e0, s);
  hipLaunchKernelGGL(k_xasm<true>, dim3(static_cast<unsigned>(ntile) * ngroup), dim3(64 * kXWaves),
                     static_cast<size_t>(nr) * 64 * sizeof(double), s, ns, nblk, nr, xasm_slot(nr), ntile, ngroup, E,
                     xo, xc, xn, te, rptr, live, hcc);
  if (e1) (void)hipEventRecord(e1, s);
  // the rank-1 terms: added by the term pass to the blocks that have them
  // (no band terms: nobands is an all-zero pointer array)
  launch_assemble(ns, ks, nblk, 1, nobands, bid, bc, bands, rptr, ru, rw, rs, rc, stab, V, hcc, bsum, rf, s, nullptr,
                  0, 0, nullptr, /*prep=*/false);
  return 0;
}

}  // namespace bsrg
