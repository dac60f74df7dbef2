// The steps of DBSR_MAT3's SUB1 that act on whole channel blocks of a Dirac
// partial wave, on the GPU:
//
//   overlap pass      S: fppqq on the channel diagonal blocks, then O_data
//                     over the overlap lists (sub1.f90:77-79, O_data.f90:2-90
//                     with UPDATE_HX/HW/HV/HB of updates.f90 and Target_h of
//                     target_subs.f90:2-31)
//   UPDATE_HX         c * D on channel blocks of S or H for host-given ms x ms
//                     matrices D (updates.f90:79-106), the update that the
//                     L(ij) records of L_data.f90:112-128 make
//   energy shift      H_ii += Ecore S_ii + (Etarg(iptar(i)) - Ecore) fppqq
//                     (sub1.f90:96 and :128-130)
//   Diag_channels     per channel the generalized problem H_ii c = E S_ii c
//                     on the B-splines Zero_cond keeps (diag.f90:2-81,
//                     zero_cond.f90:2-56), solutions in the energy window
//   transform_matrix  off-diagonal channel blocks to C_i^T H_ij C_j, channel-
//                     perturber vectors to C_i^T v (transform_matrix.f90:2-66)
//
// Layout as the Dirac assembly: channel blocks ms x ms (ms = 2 ns, P first),
// diagonal blocks first, then the lower triangle row by row; channel-
// perturber vectors hcp(ms, (ich-1)*npert + ip).  The channel solutions are
// the reference's diag(ms,ms,nch): column k < ipsol(ich) solution k, energy
// k in row k of column ms, other columns as S_ii left them.
//
// Updates are summed in list order, one thread per output element, with
// unfused multiplies and adds in the reference's association, so a block
// gets the bits the reference's loops give for the same inputs.  The
// eigenproblems go to rocSOLVER's batched sygvd, grouped by reduced size;
// the transforms to rocBLAS FP64 GEMMs, one row of blocks at a time.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <map>
#include <string>
#include <vector>

#include "../../include/bsr_gpu.h"
#include "ctx.h"

namespace {

// term kinds: rank-1 products of V_ch columns a, b (UPDATE_HW) and scaled
// matrices (UPDATE_HX)
enum : int { kHW = 0, kHWt = 1, kHWs = 2, kHX = 3, kHXt = 4 };

// ---- kernels -----------------------------------------------------------------

// fppqq = blockdiag(fpbs, fqbs): <B_i|B_j> of the P and of the Q B-splines
// (ZINTYm with the Gauss weights, zinty.f90:2-27; DBS_gauss.f90:256-280):
// element (i,j) sums over the cells iv it spans in ascending iv, each cell
// over its Gauss points in order, (w b_i) b_j.
__global__ __launch_bounds__(256) void k_fppqq(DevDbs g, double* __restrict__ fpq) {
  const int ns = g.ns, ms = 2 * ns, ks = g.ks;
  const size_t e = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= static_cast<size_t>(ms) * ms) return;
  const int x = static_cast<int>(e % ms), y = static_cast<int>(e / ms);
  double acc = 0.0;
  if ((x < ns) == (y < ns)) {
    const int q = x < ns ? 0 : 1, i = x - q * ns, j = y - q * ns, kb = g.kb[q];
    const double* v = g.val[q];
    const int lo = max(max(i, j) - kb + 1, 0), hi = min(min(i, j), g.nv - 1);
    for (int iv = lo; iv <= hi; ++iv) {
      double s = 0.0;
      for (int m = 0; m < ks; ++m) {
        const int c = iv * ks + m;
        s = __dadd_rn(s, __dmul_rn(__dmul_rn(g.grw[c], v[c * kb + (i - iv)]), v[c * kb + (j - iv)]));
      }
      acc = __dadd_rn(acc, s);
    }
  }
  fpq[e] = acc;
}

// Channel blocks with term lists: output o is block oblk[o] (o itself when
// oblk is null), its terms tk/tc[tptr[o] .. tptr[o+1]) in list order, added
// to the block's current values (load) or to zero.  Term (kind, a, b, c):
//   kHW   c (v_a v_b^T)            UPDATE_HW, ich > jch
//   kHWt  c (v_a v_b^T)^T                     ich < jch
//   kHWs  (c/2)(x + x^T), x = v_a v_b^T       ich = jch
//   kHX   c M_a                    UPDATE_HX, ich >= jch
//   kHXt  c M_a^T                             ich < jch
// v_a = column a of V (ms rows), M_a = mats[a] (ms x ms).
// grid (ceil(ms^2/256), nout)
__global__ __launch_bounds__(256) void k_blk_terms(int ms, const int* __restrict__ oblk, const int* __restrict__ tptr,
                                                   const int4* __restrict__ tk, const double* __restrict__ tc,
                                                   const double* __restrict__ V, const double* const* __restrict__ mats,
                                                   int load, double* __restrict__ h) {
  const size_t mss = static_cast<size_t>(ms) * ms;
  const size_t p = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (p >= mss) return;
  const int o = blockIdx.y;
  const size_t blk = oblk ? static_cast<size_t>(oblk[o]) : static_cast<size_t>(o);
  const int x = static_cast<int>(p % ms), y = static_cast<int>(p / ms);
  double* out = h + blk * mss + p;
  double acc = load ? *out : 0.0;
  const int t1 = tptr[o + 1];
  for (int t = tptr[o]; t < t1; ++t) {
    const int4 k = tk[t];
    const double c = tc[t];
    double xx;
    if (k.x == kHW) {
      xx = __dmul_rn(c, __dmul_rn(V[static_cast<size_t>(k.y) * ms + x], V[static_cast<size_t>(k.z) * ms + y]));
    } else if (k.x == kHWt) {
      xx = __dmul_rn(c, __dmul_rn(V[static_cast<size_t>(k.y) * ms + y], V[static_cast<size_t>(k.z) * ms + x]));
    } else if (k.x == kHWs) {
      const double a = __dmul_rn(V[static_cast<size_t>(k.y) * ms + x], V[static_cast<size_t>(k.z) * ms + y]);
      const double b = __dmul_rn(V[static_cast<size_t>(k.y) * ms + y], V[static_cast<size_t>(k.z) * ms + x]);
      xx = __dmul_rn(c * 0.5, __dadd_rn(a, b));
    } else if (k.x == kHX) {
      xx = __dmul_rn(c, mats[k.y][p]);
    } else {
      xx = __dmul_rn(c, mats[k.y][static_cast<size_t>(y) + static_cast<size_t>(ms) * x]);
    }
    acc = __dadd_rn(acc, xx);
  }
  *out = acc;
}

// Channel-perturber vectors from zero: column o gets c v_a for its terms in
// list order (UPDATE_HV).  grid (ceil(ms/256), ncol)
__global__ __launch_bounds__(256) void k_vec_terms(int ms, const int* __restrict__ tptr, const int* __restrict__ ta,
                                                   const double* __restrict__ tc, const double* __restrict__ V,
                                                   double* __restrict__ hcp) {
  const int x = blockIdx.x * 256 + threadIdx.x, o = blockIdx.y;
  if (x >= ms) return;
  double acc = 0.0;
  for (int t = tptr[o]; t < tptr[o + 1]; ++t)
    acc = __dadd_rn(acc, __dmul_rn(tc[t], V[static_cast<size_t>(ta[t]) * ms + x]));
  hcp[static_cast<size_t>(o) * ms + x] = acc;
}

// largest |element| of each of the n-element pieces of a (one workgroup per
// piece)
__global__ __launch_bounds__(256) void k_piece_max(size_t n, const double* __restrict__ a, double* __restrict__ out) {
  const double* p = a + static_cast<size_t>(blockIdx.x) * n;
  double m = 0.0;
  for (size_t i = threadIdx.x; i < n; i += 256) m = fmax(m, fabs(p[i]));
  __shared__ double red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// Record_matrix staging: off-diagonal block b (block number bk[b], rows ni
// = isol[bi[b]-1], columns nj = isol[bj[b]-1] of its top-left corner) packed
// column by column at out + off[b], its largest |element| in cmax[b].
// grid (nb), block 256.
__global__ __launch_bounds__(256) void k_corner_pack(int ms, const int* __restrict__ isol, const int* __restrict__ bi,
                                                     const int* __restrict__ bj, const int* __restrict__ bk,
                                                     const size_t* __restrict__ off, const double* __restrict__ H,
                                                     double* __restrict__ out, double* __restrict__ cmax) {
  const int b = blockIdx.x, ni = isol[bi[b] - 1], nj = isol[bj[b] - 1];
  const double* h = H + static_cast<size_t>(bk[b]) * ms * ms;
  double* o = out + off[b];
  const size_t n = static_cast<size_t>(ni) * nj;
  double m = 0.0;
  for (size_t q = threadIdx.x; q < n; q += 256) {
    const double v = h[q % ni + static_cast<size_t>(ms) * (q / ni)];
    o[q] = v;
    m = fmax(m, fabs(v));
  }
  __shared__ double red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) cmax[b] = red[0];
}

// The reduced problems of one group of channels (same kept count n): member
// g is channel chans[g] (0-based), keep[ich*ms + r] its r-th kept B-spline;
// A from the H diagonal block, B from the S diagonal block, both n x n.
// grid (ceil(n^2/256), count)
__global__ __launch_bounds__(256) void k_chan_gather(int ms, int n, const int* __restrict__ chans,
                                                     const int* __restrict__ keep, const double* __restrict__ H,
                                                     const double* __restrict__ S, double* __restrict__ A,
                                                     double* __restrict__ B) {
  const size_t nn = static_cast<size_t>(n) * n, e = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= nn) return;
  const int g = blockIdx.y, ich = chans[g];
  const int r = static_cast<int>(e % n), c = static_cast<int>(e / n);
  const size_t mss = static_cast<size_t>(ms) * ms;
  const size_t src = static_cast<size_t>(ich) * mss + keep[static_cast<size_t>(ich) * ms + r] +
                     static_cast<size_t>(ms) * keep[static_cast<size_t>(ich) * ms + c];
  A[g * nn + e] = H[src];
  B[g * nn + e] = S[src];
}

// Selection and storage of the solutions of one group (diag.f90:44-69): the
// eigenvalues W (ascending) of member g in the window (not below edmin, not
// above edmax, |E| not below egap) are numbered k = 0, 1, .. in order; the
// vector of each is spread back to the ms B-splines (zero at the deleted
// ones), its sign made positive at its first largest |component|, stored as
// column k of the channel's solutions, its energy in row k of column ms-1.
// A solver failure (info != 0) leaves the channel without solutions.  One
// workgroup of four wavefronts per member; wavefront w takes the eigenvalues
// i = w (mod 4).  grid (count)
__global__ __launch_bounds__(256) void k_chan_select(int ms, int n, const int* __restrict__ chans,
                                                     const int* __restrict__ pos, const double* __restrict__ A,
                                                     const double* __restrict__ Wv, const int* __restrict__ info,
                                                     double edmin, double edmax, double egap, double* __restrict__ sol,
                                                     int* __restrict__ ksol_out) {
  const int g = blockIdx.x, ich = chans[g];
  __shared__ int rank[1024];
  __shared__ int nsel;
  const double* w = Wv + static_cast<size_t>(g) * n;
  if (threadIdx.x == 0) {
    int k = 0;
    if (info[g] == 0)
      for (int i = 0; i < n; ++i) {
        const double e = w[i];
        const bool out = e < edmin || e > edmax || fabs(e) < egap;
        rank[i] = out ? -1 : k;
        k += out ? 0 : 1;
      }
    nsel = k;
    ksol_out[ich] = k;
  }
  __syncthreads();
  const int ns = nsel;
  if (ns == 0) return;
  const size_t mss = static_cast<size_t>(ms) * ms;
  double* D = sol + static_cast<size_t>(ich) * mss;
  const int* ps = pos + static_cast<size_t>(ich) * ms;
  const double* Ag = A + static_cast<size_t>(g) * n * n;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int i = wv; i < n; i += 4) {
    const int k = rank[i];
    if (k < 0) continue;
    const double* v = Ag + static_cast<size_t>(i) * n;
    // first largest |v_r| over the kept rows (in B-spline order)
    double bm = -1.0;
    int bi = n;
    for (int r = lane; r < n; r += 64) {
      const double a = fabs(v[r]);
      if (a > bm) { bm = a; bi = r; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double om = __shfl_xor(bm, off);
      const int oi = __shfl_xor(bi, off);
      if (om > bm || (om == bm && oi < bi)) { bm = om; bi = oi; }
    }
    const bool flip = bi < n && v[bi] < 0.0;
    double* col = D + static_cast<size_t>(k) * ms;
    for (int x = lane; x < ms; x += 64) {
      const int r = ps[x];
      double val = r >= 0 ? (flip ? -v[r] : v[r]) : 0.0;
      if (k == ms - 1 && x == ms - 1) val = w[i];  // its own energy, written after it
      col[x] = val;
    }
  }
  __syncthreads();
  // energies: row k of column ms-1 (a solution stored in column ms-1 holds
  // that column, written after the energies of the others)
  if (ns < ms)
    for (int i = threadIdx.x; i < n; i += 256)
      if (rank[i] >= 0) D[static_cast<size_t>(ms - 1) * ms + rank[i]] = w[i];
}

// Row ich of off-diagonal blocks after the GEMMs: block j of the row (0-based
// j < ich, block number bofs + j) becomes T_j (isol x jsol[j], leading
// dimension isol, T_j at T + j*isol*nmax) with zeros elsewhere, unless its
// largest |element| bmax was below 1e-20 (left as it is); ov[j] = largest
// |element| of T_j (-DBL_MAX for an empty one, 0 where skipped).
// grid (ich), block 1024
__global__ __launch_bounds__(1024) void k_row_scatter(int ms, int isol, int nmax, const int* __restrict__ jsol,
                                                      const double* __restrict__ T, const double* __restrict__ bmax,
                                                      double* __restrict__ hrow, double* __restrict__ ov) {
  const int j = blockIdx.x;
  if (bmax[j] < 1e-20) {
    if (threadIdx.x == 0) ov[j] = 0.0;
    return;
  }
  const int js = jsol[j];
  const size_t mss = static_cast<size_t>(ms) * ms;
  double* h = hrow + static_cast<size_t>(j) * mss;
  const double* t = T + static_cast<size_t>(j) * isol * nmax;
  double m = 0.0;
  for (size_t e = threadIdx.x; e < mss; e += 1024) {
    const int x = static_cast<int>(e % ms), y = static_cast<int>(e / ms);
    double v = 0.0;
    if (x < isol && y < js) {
      v = t[x + static_cast<size_t>(isol) * y];
      m = fmax(m, fabs(v));
    }
    h[e] = v;
  }
  __shared__ double red[1024];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) ov[j] = (isol > 0 && js > 0) ? red[0] : -DBL_MAX;
}

// Channel-perturber vectors after the GEMM: column q = ich*npert + ip becomes
// the first isol[ich] rows of Wv (ms x npert per channel, leading dimension
// ms) and zeros, unless its largest |element| vmax was below 1e-20; ov[q]
// as above.  grid (nch*npert)
__global__ __launch_bounds__(256) void k_vec_scatter(int ms, int npert, const int* __restrict__ isol,
                                                     const double* __restrict__ Wv, const double* __restrict__ vmax,
                                                     double* __restrict__ hcp, double* __restrict__ ov) {
  const int q = blockIdx.x, ich = q / npert;
  if (vmax[q] < 1e-20) {
    if (threadIdx.x == 0) ov[q] = 0.0;
    return;
  }
  const int is = isol[ich];
  double m = 0.0;
  for (int x = threadIdx.x; x < ms; x += 256) {
    const double v = x < is ? Wv[static_cast<size_t>(q) * ms + x] : 0.0;
    if (x < is) m = fmax(m, fabs(v));
    hcp[static_cast<size_t>(q) * ms + x] = v;
  }
  __shared__ double red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) ov[q] = is > 0 ? red[0] : -DBL_MAX;
}

// ---- host --------------------------------------------------------------------

int dcheck(bsrg_ctx* c) {
  if (!c) return fail(-2, "null context");
  if (!c->dbs.ok) return fail(-3, "Dirac grid not set (bsrg_dbs_set_grid)");
  return 0;
}

// block number of channel pair (ich, jch), 1-based channels (diagonal
// blocks first, then the lower triangle row by row)
inline size_t dblk(int nch, int ich, int jch) {
  if (ich == jch) return static_cast<size_t>(ich - 1);
  const int i = std::max(ich, jch), j = std::min(ich, jch);
  return static_cast<size_t>(nch) + static_cast<size_t>(i - 1) * (i - 2) / 2 + (j - 1);
}

// term lists of some channel blocks, in list order per block
struct BlockTerms {
  std::map<size_t, std::vector<std::pair<int4, double>>> by;
  void add(size_t blk, int kind, int a, int b, double c) { by[blk].push_back({make_int4(kind, a, b, 0), c}); }
};

// Runs the terms: every block of 0..nblk-1 from zero (all = true) or the
// listed blocks from their values.  V: V_ch columns on the device; mats: the
// device matrices the kHX terms name.
int run_block_terms(bsrg_ctx* c, int ms, size_t nblk, const BlockTerms& bt, bool all, const double* V,
                    const std::vector<const double*>& mats, double* h) {
  auto& W = c->dw;
  hipStream_t s = c->stream;
  const size_t nout = all ? nblk : bt.by.size();
  if (nout == 0) return 0;
  std::vector<int> oblk, tptr{0};
  std::vector<int4> tk;
  std::vector<double> tc;
  if (all) {
    for (size_t b = 0; b < nblk; ++b) {
      auto it = bt.by.find(b);
      if (it != bt.by.end())
        for (const auto& t : it->second) { tk.push_back(t.first); tc.push_back(t.second); }
      tptr.push_back(static_cast<int>(tk.size()));
    }
  } else {
    for (const auto& kv : bt.by) {
      oblk.push_back(static_cast<int>(kv.first));
      for (const auto& t : kv.second) { tk.push_back(t.first); tc.push_back(t.second); }
      tptr.push_back(static_cast<int>(tk.size()));
    }
  }
  HIPCHK(W.ti.upload(tptr.data(), tptr.size(), s));
  HIPCHK(W.tb.upload(oblk.data(), oblk.size(), s));
  HIPCHK(W.tk.upload(tk.data(), tk.size(), s));
  HIPCHK(W.tcd.upload(tc.data(), tc.size(), s));
  HIPCHK(W.mp.upload(mats.data(), mats.size(), s));
  const unsigned gx = static_cast<unsigned>((static_cast<size_t>(ms) * ms + 255) / 256);
  // one launch per 65535 outputs (grid y)
  for (size_t o0 = 0; o0 < nout; o0 += 65535) {
    const unsigned gy = static_cast<unsigned>(std::min<size_t>(65535, nout - o0));
    hipLaunchKernelGGL(k_blk_terms, dim3(gx, gy), dim3(256), 0, s, ms, all ? nullptr : W.tb.p + o0, W.ti.p + o0,
                       W.tk.p, W.tcd.p, V, W.mp.p, all ? 0 : 1, all ? h + o0 * ms * ms : h);
  }
  HIPCHK(hipGetLastError());
  // the host arrays the copies read go out of scope here
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int ensure_fpq(bsrg_ctx* c) {
  auto& W = c->dw;
  const int ms = 2 * c->dbs.T.ns;
  HIPCHK(W.fpq.resize(static_cast<size_t>(ms) * ms));
  hipLaunchKernelGGL(k_fppqq, dim3(static_cast<unsigned>((static_cast<size_t>(ms) * ms + 255) / 256)), dim3(256), 0,
                     c->stream, devdbs(c), W.fpq.p);
  HIPCHK(hipGetLastError());
  return 0;
}

int upload_vch(bsrg_ctx* c) {
  auto& W = c->dw;
  const size_t ms = 2 * static_cast<size_t>(c->dbs.T.ns);
  HIPCHK(W.vchd.resize(std::max<size_t>(W.vch.size(), ms)));
  if (!W.vch.empty())
    HIPCHK(hipMemcpyAsync(W.vchd.p, W.vch.data(), W.vch.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return 0;
}

// KBORT (src/libs/DBS/DBS_orbitals_pq.f90:248-277): 0-based V_ch column of
// (channel orbital, bound orbital), -1 if none
int vcol0(const DiracWave& W, int i, int j) {
  if (i < 1 || i > W.nbf || j < 1 || j > W.nbf) return -1;
  return W.kcol[(i - 1) + static_cast<size_t>(W.nbf) * (j - 1)] - 1;
}

int chan(const DiracWave& W, int o) { return (o >= 1 && o <= W.nbf) ? W.ipbs[o - 1] : 0; }

float elapsed(hipEvent_t a, hipEvent_t b) {
  float e = 0.f;
  (void)hipEventElapsedTime(&e, a, b);
  return e;
}

}  // namespace

extern "C" {

int bsrg_dirac_overlap(bsrg_ctx* c, int nch, int npert, int ntarg, const int* iptar, int nrec, const int* itype,
                       const int* k1, const int* k2, const double* cdata, double eps_c, int iitar) {
  if (int rc = dcheck(c)) return rc;
  auto& W = c->dw;
  if (!W.orb_ok) return fail(-4, "Dirac orbitals not set (bsrg_dirac_set_orbitals)");
  if (nch < 1 || npert < 0 || ntarg < 1 || !iptar) return fail(-1, "bsrg_dirac_overlap: bad sizes");
  if (nrec < 0 || (nrec > 0 && (!itype || !k1 || !k2 || !cdata)))
    return fail(-1, "bsrg_dirac_overlap: missing list arrays");
  for (int i = 0; i < nch; ++i)
    if (iptar[i] < 1 || iptar[i] > ntarg) return fail(-1, "bsrg_dirac_overlap: iptar out of 1..ntarg");
  HIPCHK(hipSetDevice(c->device));
  const int ms = 2 * c->dbs.T.ns;
  const size_t mss = static_cast<size_t>(ms) * ms, nblk = static_cast<size_t>(nch) * (nch + 1) / 2;
  const size_t ncol = static_cast<size_t>(nch) * npert;
  hipStream_t s = c->stream;
  W.s_ok = W.sol_ok = W.ovl_ok = W.tr[0] = false;
  W.s_nch = nch;
  W.s_npert = npert;
  W.s_ntarg = ntarg;
  W.s_iptar.assign(iptar, iptar + nch);
  W.shp.assign(static_cast<size_t>(npert) * (npert + 1) / 2, 0.0);
  W.sotarg.assign(static_cast<size_t>(ntarg) * (ntarg + 1) / 2, 0.0);
  // decode in list order (O_data.f90:33-80)
  BlockTerms bt;
  for (int ich = 1; ich <= nch; ++ich) bt.add(dblk(nch, ich, ich), kHX, 0, 0, 1.0);  // sub1.f90:78
  std::vector<std::vector<std::pair<int, double>>> vt(ncol);
  for (int r = 0; r < nrec; ++r) {
    const double cc = cdata[r];
    if (std::fabs(cc) < eps_c) continue;
    switch (itype[r]) {
      case 1: {  // <ic|jc>: UPDATE_HB
        const int ic = k1[r], jc = k2[r];
        if (ic > npert || jc > npert || ic <= 0 || jc <= 0) return fail(-10, "UPDATE_HB: index is out of range");
        const int i = std::max(ic, jc), j = std::min(ic, jc);
        W.shp[static_cast<size_t>(i) * (i - 1) / 2 + j - 1] += cc;
        break;
      }
      case 2: {  // <i|.> jc: UPDATE_HV
        const int io = k1[r], jc = k2[r], i1 = io / kIbo, i2 = io % kIbo;
        const int kv = vcol0(W, i1, i2);
        if (kv < 0) return fail(-12, "O: KBORT=0");
        const int ich = chan(W, i1);
        if (ich < 1 || ich > nch) return fail(-10, "UPDATE_HV: channel index out is of range");
        if (jc < 1 || jc > npert) return fail(-10, "UPDATE_HV: bound index out is of range");
        vt[static_cast<size_t>(ich - 1) * npert + (jc - 1)].push_back({kv, cc});
        break;
      }
      case 3: {  // <i|.> <.|j>: UPDATE_HW
        const int io = k1[r], jo = k2[r];
        const int kv = vcol0(W, io / kIbo, io % kIbo), kw = vcol0(W, jo / kIbo, jo % kIbo);
        if (kv < 0 || kw < 0) return fail(-12, "O: KBORT=0");
        const int ich = chan(W, io / kIbo), jch = chan(W, jo / kIbo);
        if (ich < 1 || jch < 1 || ich > nch || jch > nch) return fail(-10, "UPDATE_HW: channel index out of range");
        bt.add(dblk(nch, ich, jch), ich > jch ? kHW : (ich < jch ? kHWt : kHWs), kv, kw, cc);
        break;
      }
      case 4: {  // <i|j>: Target_h(ich,jch,0,C), UPDATE_HX(ich,jch,fppqq,C) if iitar > 1
        const int io = k1[r], ich = chan(W, io / kIbo), jch = chan(W, io % kIbo);
        if (ich < 1 || jch < 1 || ich > nch || jch > nch) return fail(-10, "target_h: channel index out of range");
        const int it = iptar[ich - 1], jt = iptar[jch - 1], i = std::max(it, jt), j = std::min(it, jt);
        W.sotarg[static_cast<size_t>(i - 1) * i / 2 + j - 1] += cc;
        if (iitar > 1 && ich != jch) bt.add(dblk(nch, ich, jch), ich >= jch ? kHX : kHXt, 0, 0, cc);
        break;
      }
      default:
        return fail(-21, "O_data: unknown structure (itype) ");
    }
  }
  (void)hipEventRecord(c->ev[0], s);
  if (int rc = ensure_fpq(c)) return rc;
  if (int rc = upload_vch(c)) return rc;
  HIPCHK(W.shch.resize(nblk * mss));
  HIPCHK(W.shcp.resize(std::max<size_t>(ncol * ms, 1)));
  std::vector<const double*> mats{W.fpq.p};
  if (int rc = run_block_terms(c, ms, nblk, bt, true, W.vchd.p, mats, W.shch.p)) return rc;
  if (ncol) {
    std::vector<int> vptr{0}, va;
    std::vector<double> vc;
    for (const auto& col : vt) {
      for (const auto& t : col) { va.push_back(t.first); vc.push_back(t.second); }
      vptr.push_back(static_cast<int>(va.size()));
    }
    HIPCHK(W.ti.upload(vptr.data(), vptr.size(), s));
    HIPCHK(W.tb.upload(va.data(), va.size(), s));
    HIPCHK(W.tcd.upload(vc.data(), vc.size(), s));
    for (size_t q0 = 0; q0 < ncol; q0 += 65535) {
      const unsigned gy = static_cast<unsigned>(std::min<size_t>(65535, ncol - q0));
      hipLaunchKernelGGL(k_vec_terms, dim3((ms + 255) / 256, gy), dim3(256), 0, s, ms, W.ti.p + q0, W.tb.p, W.tcd.p,
                         W.vchd.p, W.shcp.p + q0 * ms);
    }
    HIPCHK(hipGetLastError());
  }
  (void)hipEventRecord(c->ev[1], s);
  HIPCHK(hipEventSynchronize(c->ev[1]));
  W.dms[0] = elapsed(c->ev[0], c->ev[1]);
  W.s_ok = true;
  return 0;
}

int bsrg_dirac_overlap_get(bsrg_ctx* c, double* hch, double* hcp, double* hp, double* otarg) {
  if (int rc = dcheck(c)) return rc;
  auto& W = c->dw;
  if (!W.s_ok) return fail(-6, "no overlap matrix (bsrg_dirac_overlap)");
  HIPCHK(hipSetDevice(c->device));
  const size_t ms = 2 * static_cast<size_t>(c->dbs.T.ns), nblk = static_cast<size_t>(W.s_nch) * (W.s_nch + 1) / 2;
  const size_t ncol = static_cast<size_t>(W.s_nch) * W.s_npert;
  if (hch) HIPCHK(hipMemcpyAsync(hch, W.shch.p, nblk * ms * ms * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (hcp && ncol) HIPCHK(hipMemcpyAsync(hcp, W.shcp.p, ncol * ms * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (hp) std::copy(W.shp.begin(), W.shp.end(), hp);
  if (otarg) std::copy(W.sotarg.begin(), W.sotarg.end(), otarg);
  return 0;
}

int bsrg_dirac_update_hx(bsrg_ctx* c, int which, int nmat, const double* d, int nrec, const int* ich, const int* jch,
                         const int* imat, const double* cc) {
  if (int rc = dcheck(c)) return rc;
  auto& W = c->dw;
  if (which != 0 && which != 1) return fail(-1, "bsrg_dirac_update_hx: which must be 0 (S) or 1 (H)");
  if (which == 0 ? !W.s_ok : !W.ok)
    return fail(-6, which == 0 ? "no overlap matrix (bsrg_dirac_overlap)" : "no assembled Dirac wave (bsrg_dirac_assemble)");
  if (nmat < 1 || !d || nrec < 0 || (nrec > 0 && (!ich || !jch || !imat || !cc)))
    return fail(-1, "bsrg_dirac_update_hx: missing arguments");
  W.tr[which] = false;
  const int nch = which == 0 ? W.s_nch : W.nch;
  const int ms = 2 * c->dbs.T.ns;
  const size_t mss = static_cast<size_t>(ms) * ms, nblk = static_cast<size_t>(nch) * (nch + 1) / 2;
  BlockTerms bt;
  for (int r = 0; r < nrec; ++r) {
    if (ich[r] < 0) return fail(-10, "UPDATE_HX: ich < 0");
    if (jch[r] < 0) return fail(-10, "UPDATE_HX: jch < 0");
    if (ich[r] > nch) return fail(-10, "UPDATE_HX: ich > nch");
    if (jch[r] > nch) return fail(-10, "UPDATE_HX: jch > nch");
    if (ich[r] == 0 || jch[r] == 0) return fail(-10, "UPDATE_HX: i out of range");
    if (imat[r] < 1 || imat[r] > nmat) return fail(-1, "bsrg_dirac_update_hx: matrix index out of 1..nmat");
    bt.add(dblk(nch, ich[r], jch[r]), ich[r] >= jch[r] ? kHX : kHXt, imat[r] - 1, 0, cc[r]);
  }
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(W.dmat.upload(d, static_cast<size_t>(nmat) * mss, c->stream));
  std::vector<const double*> mats(nmat);
  for (int m = 0; m < nmat; ++m) mats[m] = W.dmat.p + static_cast<size_t>(m) * mss;
  return run_block_terms(c, ms, nblk, bt, false, nullptr, mats, which == 0 ? W.shch.p : W.hch.p);
}

int bsrg_dirac_energy_shift(bsrg_ctx* c, double ecore, const double* etarg) {
  if (int rc = dcheck(c)) return rc;
  auto& W = c->dw;
  if (!W.ok) return fail(-6, "no assembled Dirac wave (bsrg_dirac_assemble)");
  if (!W.s_ok) return fail(-6, "no overlap matrix (bsrg_dirac_overlap)");
  if (W.s_nch != W.nch) return fail(-1, "bsrg_dirac_energy_shift: S and H of different channel counts");
  W.tr[1] = false;
  if (!etarg) return fail(-2, "null argument");
  HIPCHK(hipSetDevice(c->device));
  const int ms = 2 * c->dbs.T.ns, nch = W.nch;
  const size_t mss = static_cast<size_t>(ms) * ms;
  if (int rc = ensure_fpq(c)) return rc;
  BlockTerms bt;
  std::vector<const double*> mats{W.fpq.p};
  for (int ich = 1; ich <= nch; ++ich) {
    mats.push_back(W.shch.p + static_cast<size_t>(ich - 1) * mss);
    bt.add(dblk(nch, ich, ich), kHX, ich, 0, ecore);
    bt.add(dblk(nch, ich, ich), kHX, 0, 0, etarg[W.iptar[ich - 1] - 1] - ecore);
  }
  return run_block_terms(c, ms, static_cast<size_t>(nch) * (nch + 1) / 2, bt, false, nullptr, mats, W.hch.p);
}

int bsrg_dirac_diag_channels(bsrg_ctx* c, const int* lch, const int* kch, int ilzero, int jlzero, int ibzero,
                             int jbzero, double edmin, double edmax, double egap, int* ipsol, int* nsol) {
  if (int rc = dcheck(c)) return rc;
  auto& W = c->dw;
  if (!W.ok) return fail(-6, "no assembled Dirac wave (bsrg_dirac_assemble)");
  if (!W.s_ok) return fail(-6, "no overlap matrix (bsrg_dirac_overlap)");
  if (W.s_nch != W.nch) return fail(-1, "bsrg_dirac_diag_channels: S and H of different channel counts");
  if (!lch || !kch) return fail(-2, "null argument");
  HIPCHK(hipSetDevice(c->device));
  const auto& T = c->dbs.T;
  const int ns = T.ns, ms = 2 * ns, nch = W.nch, ksp = T.kb[0], ksq = T.kb[1], nsp = T.nsb[0], nsq = T.nsb[1];
  const size_t mss = static_cast<size_t>(ms) * ms;
  if (ms > 1024) return fail(-1, "bsrg_dirac_diag_channels: ms above 1024");
  // Zero_cond (zero_cond.f90:17-53)
  W.iprm.assign(static_cast<size_t>(nch) * ms, 1);
  std::vector<int> keep(static_cast<size_t>(nch) * ms, 0), pos(static_cast<size_t>(nch) * ms, -1), nkeep(nch);
  for (int ich = 0; ich < nch; ++ich) {
    int il, jl = 0, mi = 0, mj = 0;
    if (ilzero >= 0) {
      il = ilzero;
    } else {
      il = lch[ich] + 1;
      if (il > ksp - 1) { il = 1; mi = 1; }
    }
    if (jlzero >= 0) {
      jl = jlzero;
    } else {
      if (kch[ich] < 0) jl = lch[ich] + 2;
      if (kch[ich] >= 1) jl = lch[ich];
      if (jl > ksq - 1) { jl = 1; mj = 1; }
    }
    if (mi == 1 || mj == 1) { il = 1; jl = 1; }
    int n = 0;
    for (int half = 0; half < 2; ++half) {
      const int lim = half ? jl : il, m = half ? nsq + 1 - jbzero : nsp + 1 - ibzero;
      for (int i = 1; i <= ns; ++i) {
        const size_t ip = static_cast<size_t>(ich) * ms + half * ns + (i - 1);
        if (i <= lim || i >= m) W.iprm[ip] = 0;
        if (W.iprm[ip]) {
          keep[static_cast<size_t>(ich) * ms + n] = half * ns + (i - 1);
          pos[ip] = n++;
        }
      }
    }
    nkeep[ich] = n;
  }
  hipStream_t s = c->stream;
  (void)hipEventRecord(c->ev[0], s);
  // the solutions start as the S diagonal blocks (sub1.f90:84)
  HIPCHK(W.sol.resize(static_cast<size_t>(nch) * mss));
  HIPCHK(hipMemcpyAsync(W.sol.p, W.shch.p, static_cast<size_t>(nch) * mss * sizeof(double), hipMemcpyDeviceToDevice, s));
  HIPCHK(W.keepd.upload(keep.data(), keep.size(), s));
  HIPCHK(W.posd.upload(pos.data(), pos.size(), s));
  HIPCHK(W.ksold.resize(nch));
  // groups of channels by kept count
  std::map<int, std::vector<int>> groups;
  for (int ich = 0; ich < nch; ++ich) groups[nkeep[ich]].push_back(ich);
  size_t need = 0, nmem = 0;
  for (const auto& g : groups) {
    const size_t n = g.first, cnt = g.second.size();
    need = std::max(need, cnt * (2 * n * n + 2 * n));
    nmem = std::max(nmem, cnt);
  }
  HIPCHK(W.dwork.resize(std::max<size_t>(need, 1)));
  HIPCHK(W.iwork.resize(2 * std::max<size_t>(nmem, 1)));
  std::vector<int> zero_ks(nch, 0);
  HIPCHK(hipMemcpyAsync(W.ksold.p, zero_ks.data(), nch * sizeof(int), hipMemcpyHostToDevice, s));
  for (const auto& g : groups) {
    const int n = g.first, cnt = static_cast<int>(g.second.size());
    if (n == 0) continue;
    const size_t nn = static_cast<size_t>(n) * n;
    double* A = W.dwork.p;
    double* B = A + cnt * nn;
    double* D = B + cnt * nn;
    double* E = D + static_cast<size_t>(cnt) * n;
    int* chans = W.iwork.p;
    int* info = W.iwork.p + cnt;
    HIPCHK(hipMemcpyAsync(chans, g.second.data(), cnt * sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_chan_gather, dim3(static_cast<unsigned>((nn + 255) / 256), cnt), dim3(256), 0, s, ms, n,
                       chans, W.keepd.p, W.hch.p, W.shch.p, A, B);
    HIPCHK(hipGetLastError());
    const rocblas_status st = rocsolver_dsygvd_strided_batched(
        c->blas, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_lower, n, A, n, nn, B, n, nn, D, n, E, n, info,
        cnt);
    if (st != rocblas_status_success)
      return fail(-100, std::string("rocsolver_dsygvd_strided_batched failed: ") + rocblas_status_to_string(st));
    hipLaunchKernelGGL(k_chan_select, dim3(cnt), dim3(256), 0, s, ms, n, chans, W.posd.p, A, D, info, edmin, edmax,
                       egap, W.sol.p, W.ksold.p);
    HIPCHK(hipGetLastError());
    // chans/info are reused by the next group: its copy waits for this one
    HIPCHK(hipStreamSynchronize(s));
  }
  (void)hipEventRecord(c->ev[1], s);
  W.ipsol.assign(nch, 0);
  HIPCHK(hipMemcpyAsync(W.ipsol.data(), W.ksold.p, nch * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  W.dms[1] = elapsed(c->ev[0], c->ev[1]);
  int tot = 0;
  for (int ich = 0; ich < nch; ++ich) {
    if (ipsol) ipsol[ich] = W.ipsol[ich];
    tot += W.ipsol[ich];
  }
  if (nsol) *nsol = tot;
  W.sol_ok = true;
  W.tr[0] = W.tr[1] = false;
  W.ovl_ok = false;
  return 0;
}

int bsrg_dirac_solutions(bsrg_ctx* c, double* diag, double* eval, int* iprm) {
  if (int rc = dcheck(c)) return rc;
  auto& W = c->dw;
  if (!W.sol_ok) return fail(-6, "no channel solutions (bsrg_dirac_diag_channels)");
  HIPCHK(hipSetDevice(c->device));
  const size_t ms = 2 * static_cast<size_t>(c->dbs.T.ns), nch = W.nch;
  std::vector<double> last;
  if (eval) last.resize(nch * ms);
  if (diag) HIPCHK(hipMemcpyAsync(diag, W.sol.p, nch * ms * ms * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (eval)  // column ms of every channel (sub1.f90:148-152)
    HIPCHK(hipMemcpy2DAsync(last.data(), ms * sizeof(double), W.sol.p + (ms - 1) * ms, ms * ms * sizeof(double),
                            ms * sizeof(double), nch, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (eval) {
    size_t k = 0;
    for (size_t ich = 0; ich < nch; ++ich)
      for (int is = 0; is < W.ipsol[ich]; ++is) eval[k++] = last[ich * ms + is];
  }
  if (iprm) std::copy(W.iprm.begin(), W.iprm.end(), iprm);
  return 0;
}

int bsrg_dirac_transform(bsrg_ctx* c, int which) {
  if (int rc = dcheck(c)) return rc;
  auto& W = c->dw;
  if (which != 0 && which != 1) return fail(-1, "bsrg_dirac_transform: which must be 0 (S) or 1 (H)");
  if (!W.sol_ok) return fail(-6, "no channel solutions (bsrg_dirac_diag_channels)");
  if (which == 0 ? !W.s_ok : !W.ok)
    return fail(-6, which == 0 ? "no overlap matrix (bsrg_dirac_overlap)" : "no assembled Dirac wave (bsrg_dirac_assemble)");
  const int nch = which == 0 ? W.s_nch : W.nch, npert = which == 0 ? W.s_npert : W.npert;
  if (nch != static_cast<int>(W.ipsol.size())) return fail(-1, "bsrg_dirac_transform: solutions of another channel count");
  HIPCHK(hipSetDevice(c->device));
  const int ms = 2 * c->dbs.T.ns;
  const size_t mss = static_cast<size_t>(ms) * ms, nblk = static_cast<size_t>(nch) * (nch + 1) / 2;
  const size_t ncol = static_cast<size_t>(nch) * npert;
  double* H = which == 0 ? W.shch.p : W.hch.p;
  double* V = which == 0 ? W.shcp.p : W.hcp.p;
  hipStream_t s = c->stream;
  (void)hipEventRecord(c->ev[0], s);
  // largest |element| of every block and vector before the transform
  HIPCHK(W.bmax.resize(nblk + ncol + 2 * (nblk + ncol)));
  double* bmax = W.bmax.p;
  double* vmax = bmax + nblk;
  double* ovb = vmax + ncol;       // per block
  double* ovv = ovb + nblk;        // per vector
  hipLaunchKernelGGL(k_piece_max, dim3(static_cast<unsigned>(nblk)), dim3(256), 0, s, mss, H, bmax);
  if (ncol) hipLaunchKernelGGL(k_piece_max, dim3(static_cast<unsigned>(ncol)), dim3(256), 0, s, static_cast<size_t>(ms), V, vmax);
  HIPCHK(hipGetLastError());
  std::vector<int> isol(W.ipsol);
  HIPCHK(W.isold.upload(isol.data(), isol.size(), s));
  // workspace: W_i = C_i^T [H_i1 .. H_i,i-1] and the products T
  HIPCHK(W.dwork.resize(std::max<size_t>(2 * static_cast<size_t>(ms) * ms * std::max(nch - 1, 1),
                                         ncol * static_cast<size_t>(ms))));
  double* Wb = W.dwork.p;
  double* Tb = Wb + static_cast<size_t>(ms) * ms * std::max(nch - 1, 1);
  const double one = 1.0, zero = 0.0;
  for (int ich = 2; ich <= nch; ++ich) {
    const int is = isol[ich - 1], nj = ich - 1;
    double* hrow = H + dblk(nch, ich, 1) * mss;
    int nmax = 0;
    for (int j = 0; j < nj; ++j) nmax = std::max(nmax, isol[j]);
    if (is > 0 && nmax > 0) {
      // W = C_i^T Hrow (is x ms*nj)
      rocblas_status st = rocblas_dgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, is, ms * nj, ms,
                                        &one, W.sol.p + static_cast<size_t>(ich - 1) * mss, ms, hrow, ms, &zero, Wb, is);
      if (st != rocblas_status_success) return fail(-101, "rocblas_dgemm failed");
      // T_j = W_j C_j (is x nmax), j = 0..nj-1
      st = rocblas_dgemm_strided_batched(c->blas, rocblas_operation_none, rocblas_operation_none, is, nmax, ms, &one,
                                         Wb, is, static_cast<rocblas_stride>(is) * ms, W.sol.p, ms,
                                         static_cast<rocblas_stride>(mss), &zero, Tb, is,
                                         static_cast<rocblas_stride>(is) * nmax, nj);
      if (st != rocblas_status_success) return fail(-101, "rocblas_dgemm_strided_batched failed");
    }
    hipLaunchKernelGGL(k_row_scatter, dim3(nj), dim3(1024), 0, s, ms, is, nmax, W.isold.p, Tb,
                       bmax + dblk(nch, ich, 1), hrow, ovb + dblk(nch, ich, 1));
    HIPCHK(hipGetLastError());
  }
  if (ncol) {
    // Wv_i = C_i^T V_i (ms x npert per channel; the first isol rows are used)
    const rocblas_status st = rocblas_dgemm_strided_batched(
        c->blas, rocblas_operation_transpose, rocblas_operation_none, ms, npert, ms, &one, W.sol.p, ms,
        static_cast<rocblas_stride>(mss), V, ms, static_cast<rocblas_stride>(ms) * npert, &zero, Wb, ms,
        static_cast<rocblas_stride>(ms) * npert, nch);
    if (st != rocblas_status_success) return fail(-101, "rocblas_dgemm_strided_batched failed");
    hipLaunchKernelGGL(k_vec_scatter, dim3(static_cast<unsigned>(ncol)), dim3(256), 0, s, ms, npert, W.isold.p, Wb,
                       vmax, V, ovv);
    HIPCHK(hipGetLastError());
  }
  (void)hipEventRecord(c->ev[1], s);
  // the overlaps matrix of the S transform (transform_matrix.f90:19-22, :39, :52, :56-60)
  if (which == 0) {
    std::vector<double> ob(nblk + ncol);
    HIPCHK(hipMemcpyAsync(ob.data(), ovb, (nblk + ncol) * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const int n = nch + npert;
    W.ovl.assign(static_cast<size_t>(n) * n, 0.0);
    W.ovl_n = n;
    for (int ich = 2; ich <= nch; ++ich)
      for (int jch = 1; jch < ich; ++jch)
        W.ovl[(ich - 1) + static_cast<size_t>(n) * (jch - 1)] = ob[dblk(nch, ich, jch)];
    for (int ich = 1; ich <= nch; ++ich)
      for (int ip = 1; ip <= npert; ++ip)
        W.ovl[(nch + ip - 1) + static_cast<size_t>(n) * (ich - 1)] = ob[nblk + static_cast<size_t>(ich - 1) * npert + ip - 1];
    for (int i = 1; i <= npert; ++i)
      for (int j = 1; j <= i; ++j) {
        const double x = W.shp[static_cast<size_t>(i) * (i - 1) / 2 + j - 1];
        if (x != 0.0) W.ovl[(nch + i - 1) + static_cast<size_t>(n) * (nch + j - 1)] = std::fabs(x);
      }
    W.ovl_ok = true;
  }
  HIPCHK(hipStreamSynchronize(s));
  W.dms[2 + which] = elapsed(c->ev[0], c->ev[1]);
  W.tr[which] = true;
  return 0;
}

int bsrg_dirac_overlaps(bsrg_ctx* c, double* ov) {
  if (!c || !ov) return fail(-2, "null argument");
  if (!c->dw.ovl_ok) return fail(-6, "no transformed overlap matrix (bsrg_dirac_transform(ctx, 0))");
  std::copy(c->dw.ovl.begin(), c->dw.ovl.end(), ov);
  return 0;
}

int bsrg_dirac_diag_timings(bsrg_ctx* c, double* ms4) {
  if (!c || !ms4) return fail(-2, "null argument");
  std::copy(c->dw.dms, c->dw.dms + 4, ms4);
  return 0;
}

int bsrg_dirac_record(bsrg_ctx* c, const char* path, int which, int nelc_core) {
  if (int rc = dcheck(c)) return rc;
  if (!path) return fail(-2, "null argument");
  auto& W = c->dw;
  if (which != 0 && which != 1) return fail(-1, "bsrg_dirac_record: which must be 0 (S) or 1 (H)");
  if (!W.sol_ok || !W.tr[which])
    return fail(-6, which == 0 ? "no transformed overlap matrix (bsrg_dirac_transform(ctx, 0))"
                               : "no transformed Hamiltonian matrix (bsrg_dirac_transform(ctx, 1))");
  HIPCHK(hipSetDevice(c->device));
  const auto& T = c->dbs.T;
  const int ns = T.ns, ms = 2 * ns, nch = which == 0 ? W.s_nch : W.nch, npert = which == 0 ? W.s_npert : W.npert;
  const std::vector<int>& isol = W.ipsol;
  if (static_cast<int>(isol.size()) != nch) return fail(-1, "bsrg_dirac_record: solutions of another channel count");
  hipStream_t s = c->stream;
  // the off-diagonal blocks' corners and maxima, in the order they are written
  const int nb = nch * (nch - 1) / 2;
  std::vector<int> bi(std::max(nb, 1)), bj(std::max(nb, 1)), bk(std::max(nb, 1));
  std::vector<size_t> off(std::max(nb, 1) + 1, 0);
  {
    int b = 0;
    for (int ich = 2; ich <= nch; ++ich)
      for (int jch = 1; jch < ich; ++jch, ++b) {
        bi[b] = ich;
        bj[b] = jch;
        bk[b] = static_cast<int>(dblk(nch, ich, jch));
        off[b + 1] = off[b] + static_cast<size_t>(isol[ich - 1]) * isol[jch - 1];
      }
  }
  const size_t tot = off[nb];
  std::vector<double> corners(tot), cmax(nb);
  if (nb > 0) {
    HIPCHK(W.rec.resize(tot + nb));
    HIPCHK(W.reci.resize(3 * static_cast<size_t>(nb)));
    HIPCHK(W.reco.resize(nb));
    HIPCHK(W.isold.upload(isol.data(), isol.size(), s));
    HIPCHK(hipMemcpyAsync(W.reci.p, bi.data(), nb * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(W.reci.p + nb, bj.data(), nb * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(W.reci.p + 2 * nb, bk.data(), nb * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(W.reco.p, off.data(), nb * sizeof(size_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_corner_pack, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, ms, W.isold.p, W.reci.p,
                       W.reci.p + nb, W.reci.p + 2 * nb, W.reco.p, which == 0 ? W.shch.p : W.hch.p, W.rec.p,
                       W.rec.p + tot);
    HIPCHK(hipGetLastError());
    if (tot) HIPCHK(hipMemcpyAsync(corners.data(), W.rec.p, tot * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(cmax.data(), W.rec.p + tot, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  const size_t ncol = static_cast<size_t>(nch) * npert;
  std::vector<double> hcp(ncol * ms);
  if (ncol)
    HIPCHK(hipMemcpyAsync(hcp.data(), which == 0 ? W.shcp.p : W.hcp.p, hcp.size() * sizeof(double),
                          hipMemcpyDeviceToHost, s));
  std::vector<double> sol;
  if (which == 0) {
    sol.resize(static_cast<size_t>(nch) * ms * ms);
    HIPCHK(hipMemcpyAsync(sol.data(), W.sol.p, sol.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  const std::vector<double>& hp = which == 0 ? W.shp : W.hp;
  std::FILE* f = std::fopen(path, which == 0 ? "wb" : "ab");
  if (!f) return fail(-40, std::string("cannot open ") + path);
  bool ok = true;
  if (which == 0) {
    // sub1.f90:154-158: sizes, nsol, jpsol(0:nch), eval(nsol), the solution vectors
    std::vector<int32_t> jpsol(nch + 1, 0);
    for (int i = 1; i <= nch; ++i) jpsol[i] = jpsol[i - 1] + isol[i - 1];
    const int32_t nsol = jpsol[nch];
    std::vector<double> eval, vecs;
    eval.reserve(nsol);
    vecs.reserve(static_cast<size_t>(nsol) * ms);
    for (int ich = 0; ich < nch; ++ich) {
      const double* d = sol.data() + static_cast<size_t>(ich) * ms * ms;
      for (int js = 0; js < isol[ich]; ++js) {
        eval.push_back(d[js + static_cast<size_t>(ms) * (ms - 1)]);
        vecs.insert(vecs.end(), d + static_cast<size_t>(js) * ms, d + static_cast<size_t>(js + 1) * ms);
      }
    }
    const int32_t head[5] = {ns, nch, npert, T.nsb[0], T.nsb[1]};
    ok = put_record(f, head, sizeof(head)) && put_record(f, &nsol, 4) &&
         put_record(f, jpsol.data(), jpsol.size() * 4) && put_record(f, eval.data(), eval.size() * 8) &&
         put_record(f, vecs.data(), vecs.size() * 8);
  }
  // record_matrix.f90:12-49: channel blocks with an element of 1e-20 or more,
  // channel-perturber vectors likewise, non-zero perturber elements, (0,0)
  for (int b = 0; b < nb && ok; ++b) {
    if (cmax[b] < 1e-20) continue;
    const int32_t hdr[2] = {bi[b], bj[b]};
    ok = put_record(f, hdr, 8) && put_record(f, corners.data() + off[b], (off[b + 1] - off[b]) * 8);
  }
  for (int ich = 1; ich <= nch && ok; ++ich)
    for (int ip = 1; ip <= npert && ok; ++ip) {
      const double* v = hcp.data() + (static_cast<size_t>(ich - 1) * npert + ip - 1) * ms;
      const int ni = isol[ich - 1];
      double m = 0.0;
      for (int q = 0; q < ni; ++q) m = std::max(m, std::fabs(v[q]));
      if (m < 1e-20) continue;
      const int32_t hdr[2] = {ip + nch, ich};
      ok = put_record(f, hdr, 8) && put_record(f, v, static_cast<size_t>(ni) * 8);
    }
  for (int ip = 1; ip <= npert && ok; ++ip)
    for (int jp = 1; jp <= ip && ok; ++jp) {
      const double v = hp[static_cast<size_t>(ip) * (ip - 1) / 2 + jp - 1];
      if (v == 0.0) continue;
      const int32_t hdr[2] = {ip + nch, jp + nch};
      ok = put_record(f, hdr, 8) && put_record(f, &v, 8);
    }
  const int32_t end[2] = {0, 0};
  ok = ok && put_record(f, end, 8);
  if (which == 1 && ok) {
    // sub1.f90:249-253, 277-284: the core electrons on the diagonal k = 0
    // coefficients, |acf| < 1e-5 cut to zero, then mk and the nch x nch
    // matrix of every k
    const size_t nblk = static_cast<size_t>(nch) * (nch + 1) / 2;
    std::vector<double> a(W.acf);
    for (int ich = 1; ich <= nch; ++ich) a[dblk(nch, ich, ich)] += 2.0 * nelc_core;
    for (double& x : a)
      if (std::fabs(x) < 0.00001) x = 0.0;
    std::vector<double> full(static_cast<size_t>(nch) * nch * (W.mk + 1));
    size_t q = 0;
    for (int k = 0; k <= W.mk; ++k)
      for (int jch = 1; jch <= nch; ++jch)
        for (int ich = 1; ich <= nch; ++ich) full[q++] = a[dblk(nch, ich, jch) + nblk * k];
    const int32_t mk = W.mk;
    ok = put_record(f, &mk, 4) && put_record(f, full.data(), full.size() * 8);
  }
  if (std::fclose(f) != 0) ok = false;
  if (!ok) return fail(-41, std::string("write failed: ") + path);
  return 0;
}

}  // extern "C"
