// The R^k lists of a partial wave decoded on the device: what I_data
// (src/bsr/BSR_MAT3/I_data.f90:2-224) and its structure-1 updates
// (UPDATE_HB / HV / HW / CF, called at I_data.f90:60-95) do per record,
// and GET_V (src/bsr/BSR_MAT3/get_V.f90:2-27) for the channel-orbital
// vectors the records name.  The host decoder in capi.hip (add_idata,
// add_struct1, get_v) does the same work on the host; both give the same
// term arrays, so the matrices agree bit for bit.
//
// Each record becomes up to two items (an item: a category, an index
// within the category and a payload): the term it adds (a band, rank-1,
// vector or exchange-coefficient term, or a host scalar) and, for the last
// record of a direct group, the group's ACF sum.  One stable radix sort by
// (category, index) puts the items of each output together in item order,
// which is list order: the CSR arrays of the term pass are gathered from
// it, and the sums over one output (exchange coefficients, scalars, ACF)
// are added run by run in that order -- the order of the host decoder.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "bsrg_internal.h"

namespace bsrg {
namespace {

constexpr unsigned kNone = 0xFFFFFFFFu;

// The stable LSD radix sort in rocPRIM's onesweep form at every size (its
// default sorts fewer than 2^20 items by block sorts and some twenty merge
// passes, one launch each).
using SortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                              rocprim::default_config, 0>;

__device__ __forceinline__ unsigned cat_key(int cat, int kb, long idx) {
  return (static_cast<unsigned>(cat) << kb) | static_cast<unsigned>(idx);
}

__device__ __forceinline__ int blk_of(int ich, int jch) {
  const int i = max(ich, jch), j = min(ich, jch);
  return i * (i - 1) / 2 + j - 1;
}

// pair index of two orbitals of the B-spline expansion (pair_index), -1
// with the message set otherwise
__device__ __forceinline__ int dl_pair(const DecodeArgs& a, int o1, int o2, int& msg) {
  if (o1 < 1 || o1 > a.nbf || o2 < 1 || o2 > a.nbf) { msg = kDlOrbRange; return -1; }
  int x = a.bidx[o1 - 1], y = a.bidx[o2 - 1];
  if (x < 0 || y < 0) { msg = kDlChanDensity; return -1; }
  if (x > y) { const int t = x; x = y; y = t; }
  return x * a.nb - x * (x - 1) / 2 + (y - x);
}

// GET_V: the vector-table id of the (ii, jj) vector, marking it as needed
__device__ __forceinline__ int dl_getv(const DecodeArgs& a, int ii, int jj, int& msg) {
  if (ii < 1 || ii > a.nbf || jj < 1 || jj > a.nbf) { msg = kDlGetvRange; return -1; }
  const int pair = (ii - 1) + a.nbf * (jj - 1);
  if (a.ibort[pair] == 0) { msg = kDlGetvZero; return -1; }
  a.need[pair] = 1;
  return a.w0 + pair;
}

__device__ __forceinline__ bool dl_chan(const DecodeArgs& a, int ich, int jch) {
  return ich >= 1 && ich <= a.nch && jch >= 1 && jch <= a.nch;
}

__device__ __forceinline__ bool same_group(const DecodeArgs& a, int r, int q) {
  return a.jt[q] == a.jt[r] && a.kp[q] == a.kp[r] && a.k1[q] == a.k1[r];
}

// One thread per record: its items base + 2r (the term) and base + 2r + 1
// (the ACF sum of a direct group, at the group's last record).
__global__ __launch_bounds__(256) void k_dl_items(DecodeArgs a) {
  // products and sums rounded one by one, as the host decoder rounds them
  #pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.nrec) return;
  const int jt = a.jt[r], k = a.kp[r], K1 = a.k1[r], K2 = a.k2[r], K3 = a.k3[r], K4 = a.k4[r];
  const double c = a.cd[r];
  unsigned key0 = kNone, key1 = kNone;
  int A = 0, B = 0, sidx = -1, msg = 0;
  double C = 0.0, ca = 0.0;
  do {
    if (jt < 1 || jt > 9) { msg = kDlItype; break; }
    if (k < 0 || k > a.mk) { msg = kDlMultipole; break; }
    if (k >= a.basis_nk) { msg = kDlPrepared; break; }
    if (jt == 1) {  // structure 1 (I_data.f90:52-96)
      const int p1 = dl_pair(a, K1 / kIbi, K1 % kIbi, msg);
      if (p1 < 0) break;
      const int p2 = dl_pair(a, K2 / kIbi, K2 % kIbi, msg);
      if (p2 < 0) break;
      sidx = (k * a.np + p1) * a.np + p2;
      C = c;
      const int ic = -K3, jc = -K4;
      if (ic > 0 && jc > 0) {  // UPDATE_HB
        if (ic > a.ncp || jc > a.ncp) { msg = kDlHb; break; }
        const int i = max(ic, jc), j = min(ic, jc);
        key0 = cat_key(kDlCatHbb, a.kb, i * (i - 1) / 2 + j - 1);
      } else if (ic < 0 && jc > 0) {  // UPDATE_HV
        const int io = -ic, i1 = io / kIbo, i2 = io % kIbo;
        const int w = dl_getv(a, i1, i2, msg);
        if (w < 0) break;
        const int ich = a.iech[i1 - 1];
        if (ich < 1 || ich > a.nch) { msg = kDlHvChannel; break; }
        if (jc > a.ncp) { msg = kDlHvBound; break; }
        key0 = cat_key(kDlCatVec, a.kb, static_cast<long>(jc - 1) * a.nch + (ich - 1));
        A = w;
      } else if (ic < 0 && jc < 0) {  // UPDATE_HW
        const int io = -ic, i1 = io / kIbo, i2 = io % kIbo, jo = -jc, j1 = jo / kIbo, j2 = jo % kIbo;
        const int v = dl_getv(a, i1, i2, msg);
        if (v < 0) break;
        const int w = dl_getv(a, j1, j2, msg);
        if (w < 0) break;
        const int ich = a.iech[i1 - 1], jch = a.iech[j1 - 1];
        if (!dl_chan(a, ich, jch)) { msg = kDlHwChannel; break; }
        key0 = cat_key(kDlCatR1, a.kb, blk_of(ich, jch));
        A = ich >= jch ? v : w;
        B = ich >= jch ? w : v;
      } else if (ic < 0 && jc == 0) {  // UPDATE_CF (target matrix)
        const int io = -ic, i1 = io / kIbo, i2 = io % kIbo;
        if (i1 < 1 || i1 > a.nbf || i2 < 1 || i2 > a.nbf) { msg = kDlOrbRange; break; }
        const int ich = a.iech[i1 - 1], jch = a.iech[i2 - 1];
        if (!dl_chan(a, ich, jch)) { msg = kDlCfChannel; break; }
        key0 = cat_key(kDlCatHtarg, a.kb, blk_of(ich, jch));
      } else {
        msg = kDlStruct1;
        break;
      }
    } else if (jt <= 5) {  // structures 2-5 (I_data.f90:98-140)
      const int p = dl_pair(a, K1 / kIbi, K1 % kIbi, msg);
      if (p < 0) break;
      if (K2 < 1 || K2 > a.nbf || a.bidx[K2 - 1] < 0) { msg = kDlBav; break; }
      const int vid = (k * a.np + p) * a.nb + a.bidx[K2 - 1];
      const int ich = K3, io = K4;
      C = c;
      if (io > 0) {
        const int j1 = io / kIbo, j2 = io % kIbo;
        const int w = dl_getv(a, j1, j2, msg);
        if (w < 0) break;
        const int jch = a.iech[j1 - 1];
        if (!dl_chan(a, ich, jch)) { msg = kDlHwChannel; break; }
        key0 = cat_key(kDlCatR1, a.kb, blk_of(ich, jch));
        A = ich >= jch ? vid : w;
        B = ich >= jch ? w : vid;
      } else if (io < 0) {
        if (ich < 1 || ich > a.nch) { msg = kDlHvChannel; break; }
        if (-io > a.ncp) { msg = kDlHvBound; break; }
        key0 = cat_key(kDlCatVec, a.kb, static_cast<long>(-io - 1) * a.nch + (ich - 1));
        A = vid;
      } else {
        msg = kDlStruct25;
        break;
      }
    } else {  // direct (6, 7) and exchange (8, 9) channel blocks (I_data.f90:142-222)
      const int o1 = K2, o2 = K3;
      if (o1 < 1 || o1 > a.nbf || o2 < 1 || o2 > a.nbf || a.bidx[o1 - 1] < 0 || a.bidx[o2 - 1] < 0) {
        msg = kDlChanDensity;
        break;
      }
      const int ich = K1 / kIbi, jch = K1 % kIbi;
      if (!dl_chan(a, ich, jch)) { msg = kDlHxChannel; break; }
      const int x = a.bidx[o1 - 1], y = a.bidx[o2 - 1];
      const long blk = blk_of(ich, jch);
      C = c;
      if (jt <= 7) {
        const int lo = min(x, y), hi = max(x, y);
        key0 = cat_key(kDlCatBand, a.kb, blk);
        A = 1 + k * a.np + (lo * a.nb - lo * (lo - 1) / 2 + (hi - lo));
        if (r == a.nrec - 1 || !same_group(a, r, r + 1)) {
          // the group's ACF sum: cd * <p_o1|r^k|p_o2> over its records, in order
          int s = r;
          while (s > 0 && same_group(a, r, s - 1)) --s;
          for (int q = s; q <= r; ++q) {
            const int q1 = a.k2[q], q2 = a.k3[q];
            if (q1 < 1 || q1 > a.nbf || q2 < 1 || q2 > a.nbf || a.bidx[q1 - 1] < 0 || a.bidx[q2 - 1] < 0) continue;
            const double qv = a.Qb[(static_cast<size_t>(k) * a.nb + a.bidx[q1 - 1]) * a.nb + a.bidx[q2 - 1]];
            ca += a.cd[q] * qv;
          }
          key1 = cat_key(kDlCatAcf, a.kb, static_cast<long>(k) * a.nch * a.nch + (ich - 1) +
                                              static_cast<long>(a.nch) * (jch - 1));
        }
      } else {
        const long t = static_cast<long>(k) * a.nb * a.nb + (ich >= jch ? x * a.nb + y : y * a.nb + x);
        key0 = cat_key(kDlCatGemm, a.kb, t + static_cast<long>(a.nt) * blk);
      }
    }
  } while (false);
  if (msg) {
    atomicMin(a.err, (static_cast<unsigned long long>(r) << 8) | static_cast<unsigned long long>(msg));
    key0 = key1 = kNone;
  }
  const size_t it = static_cast<size_t>(a.base) + 2 * static_cast<size_t>(r);
  a.key[it] = key0;
  a.ia[it] = A;
  a.ib[it] = B;
  a.is[it] = sidx;
  a.ic[it] = C;
  a.key[it + 1] = key1;
  a.ic[it + 1] = ca;
}

// GET_V vectors of the needed pairs: qbs(:, jj) minus the projections on
// the bound orbitals of ii's l that ii must be orthogonal to; one workgroup
// per pair.
__global__ __launch_bounds__(256) void k_dl_getv(DecodeArgs a, double* __restrict__ VT) {
  // products and sums rounded one by one, as the host decoder rounds them
  #pragma clang fp contract(off)
  const int pair = blockIdx.x;
  if (!a.need[pair]) return;
  const int nbf = a.nbf, ns = a.ns, ii = pair % nbf + 1, jj = pair / nbf + 1;
  for (int r = threadIdx.x; r < ns; r += blockDim.x) {
    double v = a.qbs[static_cast<size_t>(ns) * (jj - 1) + r];
    for (int i = 1; i <= nbf; ++i) {
      if (i == jj || a.iech[i - 1] != 0 || a.lbs[i - 1] != a.lbs[ii - 1]) continue;
      if (a.ibort[(i - 1) + nbf * (ii - 1)] != 0) continue;
      const double o = a.obs[(i - 1) + static_cast<size_t>(nbf) * (jj - 1)];
      if (fabs(o) < a.eps_ovl) continue;
      v -= o * a.qbs[static_cast<size_t>(ns) * (i - 1) + r];
    }
    VT[static_cast<size_t>(a.w0 + pair) * ns + r] = v;
  }
}

__device__ __forceinline__ int lower_bound_u(const unsigned* k, int lo, int hi, unsigned x) {
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (k[m] < x) lo = m + 1;
    else hi = m;
  }
  return lo;
}

// cb[c] = first sorted position of category c (c = 0..kDlNumCat), cb[kDlNumCat]
// the first item without a category
__global__ void k_dl_bounds(const unsigned* __restrict__ skey, int n, int kb, int* __restrict__ cb) {
  const int c = threadIdx.x;
  if (c > kDlNumCat) return;
  cb[c] = lower_bound_u(skey, 0, n, static_cast<unsigned>(c) << kb);
}

// CSR row pointers of the band (by block), rank-1 (by block) and vector (by
// channel-perturber position) terms: blockIdx.y selects the set.
__global__ void k_dl_ptr(const unsigned* __restrict__ skey, const int* __restrict__ cb, int kb, int nblk, int nq,
                         int* __restrict__ bptr, int* __restrict__ rptr, int* __restrict__ vptr) {
  const int set = blockIdx.y, cat = set == 0 ? kDlCatBand : (set == 1 ? kDlCatR1 : kDlCatVec);
  const int nkeys = set == 2 ? nq : nblk;
  int* ptr = set == 0 ? bptr : (set == 1 ? rptr : vptr);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nkeys) return;
  const int lo = cb[cat], hi = cb[cat + 1];
  ptr[b] = lower_bound_u(skey, lo, hi, (static_cast<unsigned>(cat) << kb) | static_cast<unsigned>(b)) - lo;
}

// Payloads of the band, rank-1 and vector terms into the CSR arrays.
__global__ __launch_bounds__(256) void k_dl_gather(const unsigned* __restrict__ sval, const int* __restrict__ cb,
                                                   const int* __restrict__ ia, const int* __restrict__ ib,
                                                   const int* __restrict__ is, const double* __restrict__ ic,
                                                   DecodeCsr o) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int c1 = cb[kDlCatR1], c2 = cb[kDlCatVec], c3 = cb[kDlCatVec + 1];
  if (i < cb[kDlCatBand] || i >= c3) return;
  const unsigned it = sval[i];
  if (i < c1) {
    const int j = i - cb[kDlCatBand];
    o.bc[j] = ic[it];
    o.bid[j] = ia[it];
  } else if (i < c2) {
    const int j = i - c1;
    o.rc[j] = ic[it];
    o.ru[j] = ia[it];
    o.rw[j] = ib[it];
    o.rs[j] = is[it];
  } else {
    const int j = i - c2;
    o.vc[j] = ic[it];
    o.vid[j] = ia[it];
    o.vs[j] = is[it];
  }
}

// Sums over one output in item order, one thread per run: exchange
// coefficients into the dense coefficient matrix, host scalars (times the
// scalar R^k table entry where the term has one) and ACF sums into their
// dense arrays.
__global__ __launch_bounds__(256) void k_dl_runs(const unsigned* __restrict__ skey, const unsigned* __restrict__ sval,
                                                 const int* __restrict__ cb, int kb, const int* __restrict__ is,
                                                 const double* __restrict__ ic, const double* __restrict__ St,
                                                 double* __restrict__ gemm, double* __restrict__ hbb,
                                                 double* __restrict__ htarg, double* __restrict__ acf) {
  #pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cb[kDlCatGemm] || i >= cb[kDlNumCat]) return;
  const unsigned key = skey[i];
  if (i > cb[kDlCatGemm] && skey[i - 1] == key) return;
  const int cat = static_cast<int>(key >> kb);
  const unsigned idx = key & ((1u << kb) - 1u);
  const int end = cb[cat + 1];
  const bool scal = cat == kDlCatHbb || cat == kDlCatHtarg;
  double v = 0.0;
  for (int j = i; j < end && skey[j] == key; ++j) {
    const unsigned it = sval[j];
    double f = ic[it];
    if (scal && is[it] >= 0) f = f * St[is[it]];
    v += f;
  }
  double* out = cat == kDlCatGemm ? gemm : (cat == kDlCatHbb ? hbb : (cat == kDlCatHtarg ? htarg : acf));
  if (out) out[idx] = v;
}

// The host terms of H as items: nh0 before the records' items, nh1 after
// them (from item tail on); staged as ic (nh doubles), then key, ia, ib, is
// (nh each).
__global__ void k_dl_host_items(int nh0, int nh1, size_t tail, const char* __restrict__ staged, unsigned* key,
                                int* ia, int* ib, int* is, double* ic) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x, nh = nh0 + nh1;
  if (t >= nh) return;
  const double* sc = reinterpret_cast<const double*>(staged);
  const unsigned* sk = reinterpret_cast<const unsigned*>(sc + nh);
  const int* sa = reinterpret_cast<const int*>(sk + nh);
  const int* sb = sa + nh;
  const int* ss = sb + nh;
  const size_t p = t < nh0 ? static_cast<size_t>(t) : tail + (t - nh0);
  key[p] = sk[t];
  ia[p] = sa[t];
  ib[p] = sb[t];
  is[p] = ss[t];
  ic[p] = sc[t];
}

__global__ void k_dl_iota(size_t n, unsigned* __restrict__ v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) v[i] = static_cast<unsigned>(i);
}

__global__ void k_dl_fill_u8(size_t n, unsigned char* __restrict__ p) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_dl_fill_u64(unsigned long long* __restrict__ p, unsigned long long v) { *p = v; }

unsigned grid_of(size_t n, unsigned b) { return static_cast<unsigned>((n + b - 1) / b); }

}  // namespace

size_t dl_sort_bytes(int nitems, int end_bit) {
  size_t bytes = 0;
  (void)rocprim::radix_sort_pairs<SortConfig>(nullptr, bytes, static_cast<const unsigned*>(nullptr),
                                              static_cast<unsigned*>(nullptr), static_cast<const unsigned*>(nullptr),
                                              static_cast<unsigned*>(nullptr), nitems, 0u,
                                              static_cast<unsigned>(end_bit));
  return bytes;
}

size_t dl_host_item_bytes(int nh) { return static_cast<size_t>(nh) * (8 + 4 * 4); }

void launch_dl_host_items(int nh0, int nh1, size_t tail, const void* staged, const DecodeArgs& a, hipStream_t s) {
  const int nh = nh0 + nh1;
  if (nh > 0)
    hipLaunchKernelGGL(k_dl_host_items, dim3(grid_of(nh, 256)), dim3(256), 0, s, nh0, nh1, tail,
                       static_cast<const char*>(staged), a.key, a.ia, a.ib, a.is, a.ic);
}

void launch_dl_iota(size_t n, unsigned* v, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_dl_iota, dim3(grid_of(n, 256)), dim3(256), 0, s, n, v);
}

int launch_dl_decode(const DecodeArgs& a, const DecodeSort& so, const DecodeCsr& o, const DecodeOut& out,
                     hipStream_t s) {
  const size_t npair = static_cast<size_t>(a.nbf) * a.nbf;
  hipLaunchKernelGGL(k_dl_fill_u8, dim3(grid_of(npair, 256)), dim3(256), 0, s, npair, a.need);
  hipLaunchKernelGGL(k_dl_fill_u64, dim3(1), dim3(1), 0, s, a.err, ~0ull);
  if (a.nrec > 0) hipLaunchKernelGGL(k_dl_items, dim3(grid_of(a.nrec, 256)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_dl_getv, dim3(static_cast<unsigned>(npair)), dim3(256), 0, s, a, out.VT);
  size_t bytes = so.tmp_bytes;
  if (rocprim::radix_sort_pairs<SortConfig>(so.tmp, bytes, a.key, so.skey, so.iota, so.sval, so.nitems, 0u,
                                             static_cast<unsigned>(a.kb + 3), s) != hipSuccess)
    return -1;
  hipLaunchKernelGGL(k_dl_bounds, dim3(1), dim3(64), 0, s, so.skey, so.nitems, a.kb, out.cb);
  const int nmax = std::max(o.nblk, o.nq) + 1;
  hipLaunchKernelGGL(k_dl_ptr, dim3(grid_of(nmax, 256), 3), dim3(256), 0, s, so.skey, out.cb, a.kb, o.nblk, o.nq,
                     o.bptr, o.rptr, o.vptr);
  hipLaunchKernelGGL(k_dl_gather, dim3(grid_of(so.nitems, 256)), dim3(256), 0, s, so.sval, out.cb, a.ia, a.ib, a.is,
                     a.ic, o);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_dl_runs(const DecodeArgs& a, const DecodeSort& so, const DecodeOut& out, hipStream_t s) {
  hipLaunchKernelGGL(k_dl_runs, dim3(grid_of(so.nitems, 256)), dim3(256), 0, s, so.skey, so.sval, out.cb, a.kb, a.is,
                     a.ic, a.St, out.gemm, out.hbb, out.htarg, out.acf);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace bsrg
