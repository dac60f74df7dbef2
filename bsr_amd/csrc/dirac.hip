// Dirac partial waves on the GPU: the R^k assembly of DBSR_MAT3, i.e. R_data
// (src/dbsr/DBSR_MAT3_MPI/R_data.f90:2-290) over the four integral families
// for every (itype, kpol) list, as Gen_matrix runs it (gen_matrix.f90:11-14),
// with the accumulators of updates.f90 and Target_h (target_subs.f90:2-31).
//
// Same design as the BS partial wave: the family bands stay resident, their
// contractions with the bound-orbital pairs are made once per orbital set
// ("bases"), and every record becomes terms on the channel blocks.  Family f
// pairs component b1(f) on r1 with b2(f) on r2 -- 0 PPPP (P,P), 1 QQQQ (Q,Q),
// 2 PQPQ (P,Q), 3 QPQP (Q,P) -- and its bases are
//   D[f][k][p]      direct band of the r2 pair density of components b2
//   St[f][k][p][q]  = sum over the band of D[f][k][p] * (r1 pair density q, components b1)
//   V[f][k][p][b]   = D[f][k][p] (symmetric band) * pq(:, b1, b)
//   E[f][k][a][b]   exchange matrix of the density pq(:,b1,a) (x) pq(:,b2,b)
// and a record lands in quadrant (a, b) of its channel block:
//   itype 3  band term c*D in quadrant (b1, b1)              (UPDATE_HS 's')
//   itype 4  coefficient of family f's GEMM, quadrant (b1,b2) (UPDATE_HS 'x')
//   itype 2  rank-1 c*V (x) V_ch half, quadrants (b1, h)     (UPDATE_HW, HV)
//   itype 1  S*c with S from St: V_ch (x) V_ch halves (HW), V_ch halves into
//            hcp (HV), hp (HB), htarg (Target_h)
// A record applies one coefficient to all four families, so the families
// that land in the same place are summed once per orbital set: the direct
// bands and vectors of PPPP + PQPQ (r1 = P) and QQQQ + QPQP (r1 = Q), and the
// scalars of all four.
// Diagonal channel blocks collect the raw sums and are symmetrised at the
// end, (X + X^T)/2, which is what the halved updates of the reference
// (updates.f90:166-173, :204-207) add up to.  Every output element is summed
// by one thread in list order (no atomics).
#include <rocblas/rocblas.h>

#include <chrono>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

#include "../../include/bsr_gpu.h"
#include "ctx.h"

namespace {

constexpr int kB1[4] = {0, 1, 0, 1};  // component on r1 of family f (0 P, 1 Q)
constexpr int kB2[4] = {0, 1, 1, 0};  // component on r2

// ---- kernels -----------------------------------------------------------------

// Exchange coefficients, column blk of coef (nt rows): the entries of blk in
// list order.  Thread tid owns the rows r = tid (mod 256).
__global__ __launch_bounds__(256) void k_dirac_coef(int nt, const int* __restrict__ ptr, const int* __restrict__ tt,
                                                    const double* __restrict__ cc, double* __restrict__ coef) {
  const size_t blk = blockIdx.x;
  double* col = coef + blk * nt;
  const int e0 = ptr[blk], e1 = ptr[blk + 1];
  __shared__ int st[256];
  __shared__ double sc[256];
  for (int r = threadIdx.x; r < nt; r += 256) col[r] = 0.0;
  for (int e = e0; e < e1; e += 256) {
    const int n = min(256, e1 - e);
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < n) {
      st[threadIdx.x] = tt[e + threadIdx.x];
      sc[threadIdx.x] = cc[e + threadIdx.x];
    }
    __syncthreads();
    for (int q = 0; q < n; ++q) {
      const int r = st[q];
      if ((r & 255) == static_cast<int>(threadIdx.x)) col[r] += sc[q];
    }
  }
}

// Diagonal channel blocks (the first nch blocks of hch, order ms) ->
// (X + X^T)/2 in place, through 32 x 32 LDS tiles: the workgroup of tile
// pair (ti >= tj) reads both tiles before it writes them.  grid
// (nt*(nt+1)/2, nch), block 256 = 32 rows x 8 column groups.
__global__ __launch_bounds__(256) void k_dirac_symm(int ms, double* __restrict__ hch) {
  int p = blockIdx.x, ti = 0;
  while (p > ti) { p -= ti + 1; ++ti; }
  const int tj = p;
  double* X = hch + static_cast<size_t>(blockIdx.y) * ms * ms;
  __shared__ double A[32][33], B[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int ri = ti * 32, cj = tj * 32;
  for (int j = ty; j < 32; j += 8) {
    const int r = ri + tx, c = cj + j, r2 = cj + tx, c2 = ri + j;
    A[j][tx] = (r < ms && c < ms) ? X[r + static_cast<size_t>(ms) * c] : 0.0;      // X(ri+tx, cj+j)
    B[j][tx] = (r2 < ms && c2 < ms) ? X[r2 + static_cast<size_t>(ms) * c2] : 0.0;  // X(cj+tx, ri+j)
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int r = ri + tx, c = cj + j, r2 = cj + tx, c2 = ri + j;
    if (r < ms && c < ms) X[r + static_cast<size_t>(ms) * c] = 0.5 * (A[j][tx] + B[tx][j]);
    if (ti != tj && r2 < ms && c2 < ms) X[r2 + static_cast<size_t>(ms) * c2] = 0.5 * (B[j][tx] + A[tx][j]);
  }
}

// out = a + b (n doubles)
__global__ void k_sum2(size_t n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + b[i];
}

// out = ((x0 + x1) + x2) + x3 with xf = in + f*stride (n doubles)
__global__ void k_sum4(size_t n, size_t stride, const double* __restrict__ in, double* __restrict__ out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = ((in[i] + in[i + stride]) + in[i + 2 * stride]) + in[i + 3 * stride];
}

// hcp(:, q) (ms) from the component-major hv[a][q](ns).  grid (nq), block 256
__global__ void k_dirac_vfinalize(int ns, int nq, const double* __restrict__ hv, double* __restrict__ hcp) {
  const int q = blockIdx.x, ms = 2 * ns;
  for (int r = threadIdx.x; r < ms; r += blockDim.x) {
    const int a = r >= ns;
    hcp[static_cast<size_t>(q) * ms + r] = hv[(static_cast<size_t>(a) * nq + q) * ns + (r - a * ns)];
  }
}

// ---- host --------------------------------------------------------------------

int check(bsrg_ctx* c) {
  if (!c) return fail(-2, "null context");
  if (!c->dbs.ok) return fail(-3, "Dirac grid not set (bsrg_dbs_set_grid)");
  return 0;
}

// block number of channel pair (ich, jch), 0-based: alloc_dbsr_matrix on one
// process (updates.f90:17-30) gives the diagonal blocks 1..nch, then the lower
// triangle row by row
inline size_t dblk(int nch, int ich, int jch) {
  if (ich == jch) return static_cast<size_t>(ich - 1);
  const int i = std::max(ich, jch), j = std::min(ich, jch);
  return static_cast<size_t>(nch) + static_cast<size_t>(i - 1) * (i - 2) / 2 + (j - 1);
}

// quadr_pq(a, b, k) for the bound orbitals (src/libs/DBS/quadr_pq.f90:2-38),
// from the orbital values at the Gauss points
const std::vector<double>& quadr_table(bsrg_ctx* c, int k) {
  auto& W = c->dw;
  const auto& T = c->dbs.T;
  if (static_cast<int>(W.Q.size()) <= k) W.Q.resize(k + 1);
  if (W.Q[k].empty()) {
    const int nb = W.nb, nv = T.nv, ks = T.ks, ns = T.ns;
    std::vector<double> f(static_cast<size_t>(2) * nb * nv * ks, 0.0);  // [cpt][b][iv*ks+m]
    int b = 0;
    for (int o = 0; o < W.nbf; ++o) {
      if (W.bidx[o] < 0) continue;
      for (int cpt = 0; cpt < 2; ++cpt) {
        const int kb = T.kb[cpt];
        const double* p = &W.pq[(static_cast<size_t>(o) * 2 + cpt) * ns];
        double* out = &f[(static_cast<size_t>(cpt) * nb + b) * nv * ks];
        for (int iv = 0; iv < nv; ++iv)
          for (int m = 0; m < ks; ++m) {
            double s = 0.0;
            for (int ith = 0; ith < kb; ++ith) s += T.val[cpt][(static_cast<size_t>(iv) * ks + m) * kb + ith] * p[iv + ith];
            out[iv * ks + m] = s;
          }
      }
      ++b;
    }
    std::vector<double> wq(static_cast<size_t>(nv) * ks);
    for (size_t q = 0; q < wq.size(); ++q) wq[q] = powi(T.gr[q], k) * T.grw[q];
    W.Q[k].assign(static_cast<size_t>(nb) * nb, 0.0);
    for (int a = 0; a < nb; ++a)
      for (int bb = 0; bb < nb; ++bb) {
        double s = 0.0;
        for (int cpt = 0; cpt < 2; ++cpt) {
          const double* fa = &f[(static_cast<size_t>(cpt) * nb + a) * nv * ks];
          const double* fb = &f[(static_cast<size_t>(cpt) * nb + bb) * nv * ks];
          for (size_t q = 0; q < wq.size(); ++q) s += wq[q] * fa[q] * fb[q];
        }
        W.Q[k][a * nb + bb] = s;
      }
  }
  return W.Q[k];
}

// Bases of the four families for k < nk (exchange matrices too when exch).
int dirac_bases(bsrg_ctx* c, int nk, bool exch) {
  auto& W = c->dw;
  const auto& T = c->dbs.T;
  const int ns = T.ns, ks = T.ks, nb = W.nb, np = W.npair;
  hipStream_t s = c->stream;
  for (int f = 0; f < 4; ++f)
    if (int rc = ensure_dbs_band(c, f, nk)) return rc;
  const size_t nsk = static_cast<size_t>(ns) * ks;
  if (W.basis_nk < nk) {
    const size_t nD = static_cast<size_t>(nk) * np * nsk, nS = static_cast<size_t>(nk) * np * np;
    const size_t nV = static_cast<size_t>(nk) * np * nb * ns;
    HIPCHK(W.D.resize(4 * nD + 1));
    HIPCHK(W.Dc.resize(2 * nD + 1));
    HIPCHK(W.VT.resize(2 * nV + (2 * static_cast<size_t>(W.nvch) + 1) * ns));
    if (np > 0) {
      HIPCHK(W.dpart.resize(static_cast<size_t>(nk) * np * ks * nsk));
      HIPCHK(W.St.resize(4 * nS));
      HIPCHK(W.Sc.resize(nS));
      HIPCHK(W.Vf.resize(4 * nV));
      for (int f = 0; f < 4; ++f) {
        double* Df = W.D.p + f * nD;
        launch_direct_basis(ns, ks, nk, np, c->dbs.rkb[f].p, W.dens.p + kB2[f] * np * nsk, W.dpart.p, Df, s);
        launch_stab(ns, ks, nk, np, Df, W.dens.p + kB1[f] * np * nsk, W.St.p + f * nS, s);
        launch_vtab(ns, ks, nk, np, nb, Df, W.P.p + static_cast<size_t>(kB1[f]) * nb * ns, W.Vf.p + f * nV, s);
      }
      // families with the same r1 component land in the same quadrant with
      // one coefficient: PPPP + PQPQ (r1 = P), QQQQ + QPQP (r1 = Q)
      for (int h = 0; h < 2; ++h) {
        hipLaunchKernelGGL(k_sum2, dim3(static_cast<unsigned>((nD + 255) / 256)), dim3(256), 0, s, nD,
                           W.D.p + h * nD, W.D.p + (h + 2) * nD, W.Dc.p + h * nD);
        hipLaunchKernelGGL(k_sum2, dim3(static_cast<unsigned>((nV + 255) / 256)), dim3(256), 0, s, nV,
                           W.Vf.p + h * nV, W.Vf.p + (h + 2) * nV, W.VT.p + h * nV);
      }
      hipLaunchKernelGGL(k_sum4, dim3(static_cast<unsigned>((nS + 255) / 256)), dim3(256), 0, s, nS, nS, W.St.p,
                         W.Sc.p);
    }
    // the <.|p> vectors after the family vectors; half h of column m is
    // vector 2*nk*np*nb + 2m + h, which is V_ch's own layout
    if (W.nvch)
      HIPCHK(hipMemcpyAsync(W.VT.p + 2 * nV, W.vch.data(), W.vch.size() * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipGetLastError());
    if (np > 0) {
      if (nS > W.St_host.cap) {
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(W.St_host.reserve(nS));
      }
      HIPCHK(hipMemcpyAsync(W.St_host.p, W.Sc.p, nS * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    W.basis_nk = nk;
  }
  if (exch && W.exch_nk < nk && nb > 0) {
    const size_t per = static_cast<size_t>(nk) * nb * nb * ns * ns;
    HIPCHK(W.E.resize(4 * per));
    for (int f = 0; f < 4; ++f)
      launch_exchange_basis(ns, ks, nk, nb, c->dbs.rkb[f].p, W.P.p + static_cast<size_t>(kB1[f]) * nb * ns,
                            W.P.p + static_cast<size_t>(kB2[f]) * nb * ns, W.E.p + f * per, s);
    HIPCHK(hipGetLastError());
    W.exch_nk = nk;
    ++W.e_ver;
  }
  return 0;
}

int pair_of(const DiracWave& W, int o1, int o2, int& p) {
  if (o1 < 1 || o1 > W.nbf || o2 < 1 || o2 > W.nbf) return fail(-30, "orbital index out of range");
  int a = W.bidx[o1 - 1], b = W.bidx[o2 - 1];
  if (a < 0 || b < 0) return fail(-31, "R^k density of a channel orbital (no B-spline expansion)");
  if (a > b) std::swap(a, b);
  p = a * W.nb - a * (a - 1) / 2 + (b - a);
  return 0;
}

// KBORT (src/libs/DBS/DBS_orbitals_pq.f90:248-277): column of V_ch, 0 if none
int vcol(const DiracWave& W, int i, int j) {
  if (i < 1 || i > W.nbf || j < 1 || j > W.nbf) return 0;
  return W.kcol[(i - 1) + static_cast<size_t>(W.nbf) * (j - 1)];
}

int chan_of(const DiracWave& W, int o) { return (o >= 1 && o <= W.nbf) ? W.ipbs[o - 1] : -1; }

int chk_pair(const DiracWave& W, int ich, int jch, const char* who) {
  if (ich < 1 || jch < 1 || ich > W.nch || jch > W.nch)
    return fail(-10, std::string(who) + ": channel index out of range");
  return 0;
}

// One (itype, kpol) list -> terms, for the four families (R_data.f90:88-281).
int dirac_decode(bsrg_ctx* c, DiracWave::Part& P, int itype, int k, int n, const int* K1, const int* K2,
                 const int* K3, const int* K4, const double* cd) {
  auto& W = c->dw;
  Terms& T = P.T;
  const int nb = W.nb, np = W.npair, nk = W.basis_nk, nkE = W.exch_nk, nch = W.nch, npert = W.npert;
  const size_t nblk = W.nblk;
  const int nq1 = nch * npert;
  const int nVc = 2 * nk * np * nb;  // combined vectors before the V_ch halves
  auto qblk = [nblk](int a, int b, size_t bl) { return static_cast<int>((2 * a + b) * nblk + bl); };
  auto vid = [=](int h, int p, int b) { return ((h * nk + k) * np + p) * nb + b; };
  auto sid = [=](int pB, int pA) { return (k * np + pB) * np + pA; };
  int rc = 0;
  if (itype == 1) {  // R_data.f90:92-165
    for (int q = 0; q < n; ++q) {
      int pA, pB;
      if ((rc = pair_of(W, K1[q] / kIbo, K1[q] % kIbo, pA))) return rc;
      if ((rc = pair_of(W, K2[q] / kIbo, K2[q] % kIbo, pB))) return rc;
      const int ic = -K3[q], jc = -K4[q], io = -ic, jo = -jc;
      const double cc = cd[q];
      if (ic > 0 && jc > 0) {
        if (ic > npert || jc > npert) return fail(-12, "UPDATE_HB: index is out of range");
        const int i = std::max(ic, jc), j = std::min(ic, jc);
        T.scal.push_back({0, i * (i - 1) / 2 + j - 1, sid(pB, pA), cc});
      } else if (ic < 0 && jc > 0) {
        const int io1 = io / kIbo, io2 = io % kIbo, m = vcol(W, io1, io2);
        if (m == 0) return fail(-19, "RK: KBORT =0");
        const int ich = chan_of(W, io1);
        if (ich < 1 || ich > nch) return fail(-13, "UPDATE_HV: channel index out is of range");
        if (jc > npert) return fail(-14, "UPDATE_HV: bound index out is of range");
        for (int h = 0; h < 2; ++h)
          T.vec.push_back({h * nq1 + (ich - 1) * npert + (jc - 1), nVc + 2 * (m - 1) + h, sid(pB, pA), cc});
      } else if (ic < 0 && jc < 0) {
        const int io1 = io / kIbo, io2 = io % kIbo, jo1 = jo / kIbo, jo2 = jo % kIbo;
        const int mv = vcol(W, io1, io2), mw = vcol(W, jo1, jo2);
        if (mv == 0 || mw == 0) return fail(-19, "RK: KBORT =0");
        const int ich = chan_of(W, io1), jch = chan_of(W, jo1);
        if ((rc = chk_pair(W, ich, jch, "UPDATE_HW"))) return rc;
        const size_t bl = dblk(nch, ich, jch);
        const int u0 = nVc + 2 * (mv - 1), w0 = nVc + 2 * (mw - 1);
        for (int a = 0; a < 2; ++a)
          for (int b = 0; b < 2; ++b) {
            if (ich >= jch) T.r1.push_back({qblk(a, b, bl), u0 + a, w0 + b, sid(pB, pA), cc});
            else T.r1.push_back({qblk(a, b, bl), w0 + a, u0 + b, sid(pB, pA), cc});
          }
      } else if (ic < 0 && jc == 0) {
        const int io1 = io / kIbo, io2 = io % kIbo;
        const int ich = chan_of(W, io1), jch = chan_of(W, io2);
        if (ich < 1 || jch < 1 || ich > nch || jch > nch)
          return fail(-10, "target_h: channel index out of range");
        const int it = W.iptar[ich - 1], jt = W.iptar[jch - 1];
        const int i = std::max(it, jt), j = std::min(it, jt);
        T.scal.push_back({1, (i - 1) * i / 2 + j - 1, sid(pB, pA), cc});
      } else {
        return fail(-22, " R_data: unknown structure for itype=1");
      }
    }
  } else if (itype == 2) {  // R_data.f90:167-212
    for (int q = 0; q < n; ++q) {
      int pB;
      if ((rc = pair_of(W, K1[q] / kIbo, K1[q] % kIbo, pB))) return rc;
      const int ii = K2[q], ich = K3[q], io = K4[q];
      if (ii < 1 || ii > W.nbf || W.bidx[ii - 1] < 0) return fail(-31, "BAV with a channel orbital");
      const int bi = W.bidx[ii - 1];
      const double cc = cd[q];
      if (io > 0) {
        const int io1 = io / kIbo, io2 = io % kIbo, m = vcol(W, io1, io2);
        if (m == 0) return fail(-19, "RK: KBORT = 0");
        const int jch = chan_of(W, io1);
        if ((rc = chk_pair(W, ich, jch, "UPDATE_HW"))) return rc;
        const size_t bl = dblk(nch, ich, jch);
        const int w0 = nVc + 2 * (m - 1);
        for (int a = 0; a < 2; ++a) {  // r1 component: P from PPPP + PQPQ, Q from QQQQ + QPQP
          const int v = vid(a, pB, bi);
          for (int h = 0; h < 2; ++h) {
            if (ich >= jch) T.r1.push_back({qblk(a, h, bl), v, w0 + h, -1, cc});
            else T.r1.push_back({qblk(h, a, bl), w0 + h, v, -1, cc});
          }
        }
      } else if (io < 0) {
        if (ich < 1 || ich > nch) return fail(-13, "UPDATE_HV: channel index out is of range");
        if (-io > npert) return fail(-14, "UPDATE_HV: bound index out is of range");
        for (int a = 0; a < 2; ++a) T.vec.push_back({a * nq1 + (ich - 1) * npert + (-io - 1), vid(a, pB, bi), -1, cc});
      } else {
        return fail(-23, " R_data: uknown structure for itype=2 ");
      }
    }
  } else if (itype == 3) {  // R_data.f90:214-248
    const std::vector<double>* Qk = nullptr;
    double CA = 0.0;
    for (int q = 0; q < n; ++q) {
      const int ich = K1[q] / kIbo, jch = K1[q] % kIbo;
      if ((rc = chk_pair(W, ich, jch, "UPDATE_HS"))) return rc;
      int p;
      if ((rc = pair_of(W, K2[q], K3[q], p))) return rc;
      const size_t bl = dblk(nch, ich, jch);
      for (int h = 0; h < 2; ++h) T.band.push_back({qblk(h, h, bl), (h * nk + k) * np + p, cd[q]});
      if (!Qk) Qk = &W.Q[k];  // built before the parallel pass
      CA += cd[q] * (*Qk)[W.bidx[K2[q] - 1] * nb + W.bidx[K3[q] - 1]];
      if (q == n - 1 || K1[q + 1] != K1[q]) {
        if (k > W.mk) return fail(-17, "Update_cf: k > kmk ");
        const size_t at = bl + static_cast<size_t>(W.nblk) * k;
        P.ax.push_back({at, CA});  // atype 0
        P.ax.push_back({at, CA});  // atype 1
        CA = 0.0;
      }
    }
  } else if (itype == 4) {  // R_data.f90:250-281
    const int nb2 = nb * nb;
    for (int q = 0; q < n; ++q) {
      const int ich = K1[q] / kIbo, jch = K1[q] % kIbo;
      if ((rc = chk_pair(W, ich, jch, "UPDATE_HS"))) return rc;
      const int o1 = K2[q], o2 = K3[q];
      if (o1 < 1 || o1 > W.nbf || o2 < 1 || o2 > W.nbf) return fail(-30, "orbital index out of range");
      const int a = W.bidx[o1 - 1], b = W.bidx[o2 - 1];
      if (a < 0 || b < 0) return fail(-31, "R^k density of a channel orbital (no B-spline expansion)");
      if (ich < jch)
        return fail(-25, " R_data: exchange record with ich < jch (lists pack the larger channel first, "
                         "idef_type.f90:268-296)");
      const int bl = static_cast<int>(dblk(nch, ich, jch));
      for (int f = 0; f < 4; ++f) {
        P.xc.push_back({bl, (f * nkE + k) * nb2 + a * nb + b, cd[q]});
        P.klo[f] = std::min(P.klo[f], k);
        P.khi[f] = std::max(P.khi[f], k);
      }
    }
  } else {
    return fail(-21, " R_data: unknown itype ");
  }
  return 0;
}

// CSR of the exchange coefficients over the blocks (list order kept within a
// block), in one pinned arena, uploaded in one copy.
int stage_exchange(bsrg_ctx* c) {
  auto& W = c->dw;
  size_t n = 0;
  for (int t = 0; t < W.nparts; ++t) n += W.parts[t]->xc.size();
  const size_t nblk = W.nblk;
  size_t o = 0;
  auto put = [&o](size_t count, size_t elem) { const size_t at = o; o += ((count * elem + 255) / 256) * 256; return at; };
  W.o_xc = put(n, 8);
  W.o_xptr = put(nblk + 1, 4);
  W.o_xt = put(n, 4);
  W.xbytes = o;
  if (int rc = wait_staged(W.st)) return rc;  // the arena copy of the previous wave has completed
  if (o > W.xh.cap) HIPCHK(W.xh.reserve(o));
  int* ptr = reinterpret_cast<int*>(W.xh.p + W.o_xptr);
  int* tt = reinterpret_cast<int*>(W.xh.p + W.o_xt);
  double* cv = reinterpret_cast<double*>(W.xh.p + W.o_xc);
  W.xcnt.assign(nblk + 1, 0);
  for (int t = 0; t < W.nparts; ++t)
    for (const auto& x : W.parts[t]->xc) ++W.xcnt[x.blk + 1];
  for (size_t b = 0; b < nblk; ++b) W.xcnt[b + 1] += W.xcnt[b];
  for (size_t b = 0; b <= nblk; ++b) ptr[b] = W.xcnt[b];
  for (int t = 0; t < W.nparts; ++t)
    for (const auto& x : W.parts[t]->xc) {
      const int at = W.xcnt[x.blk]++;
      tt[at] = x.t;
      cv[at] = x.c;
    }
  return 0;
}

// Exchange GEMM of family f into its quadrant: hq[b1,b2] (+)= E_f * coef_f.
// engine 0 library FP64, 1-3 FP64 MFMA kernel, 10+S bf16 and 20+S int8 split
// precision with S slices (the int8 slices of E_f prepared once per version
// of the exchange matrices, over every multipole).
int dirac_gemm(bsrg_ctx* c, int f, int engine, bool accumulate, bool into_output = false) {
  auto& W = c->dw;
  const int ns = c->dbs.T.ns, nb = W.nb, nkE = W.exch_nk;
  const size_t nss = static_cast<size_t>(ns) * ns, nb2 = static_cast<size_t>(nb) * nb;
  const size_t nt = 4 * static_cast<size_t>(nkE) * nb2;
  const size_t r0 = (static_cast<size_t>(f) * nkE + W.klo[f]) * nb2;
  const int Kt = (W.khi[f] - W.klo[f] + 1) * static_cast<int>(nb2);
  const double* A = W.E.p + r0 * nss;
  const double* B = W.coef.p + r0;
  double* C = W.hq.p + static_cast<size_t>(2 * kB1[f] + kB2[f]) * W.nblk * nss;
  const int m = static_cast<int>(nss), n = static_cast<int>(W.nblk);
  if (engine == 4) {
    // family f's quadrant (b1, b2) of every channel block from the packed
    // nonzeros of the coefficients (identical for the four families)
    const double* Ef = W.E.p + r0 * nss;
    if (launch_xquad(ns, n, Kt, xquad_width(Kt, c->device), Ef, W.qxo.p, W.qxc.p, W.qxn.p, kB1[f], kB2[f], W.hch.p,
                     c->stream))
      return fail(-1, "launch_xquad: no coefficient rows");
    return 0;
  }
  if (engine >= 20) {
    const int S = engine - 20, K = nkE * static_cast<int>(nb2);
    const double* Af = W.E.p + static_cast<size_t>(f) * nkE * nb2 * nss;
    if (W.oz8a_ver[f] != W.e_ver || W.oz8a_S[f] != S) {
      HIPCHK(W.oz8a[f].resize(ozaki8_a_bytes(m, K, S)));
      if (int rc = ozaki8_prepare_a(m, K, Af, nss, S, W.oz8a[f].p, c->stream)) return rc;
      W.oz8a_ver[f] = W.e_ver;
      W.oz8a_S[f] = S;
    }
    HIPCHK(W.oz8b.resize(ozaki8_b_bytes(n, K, S)));
    if (into_output) {
      // straight into quadrant (b1, b2) of the channel blocks
      const size_t ms = 2 * static_cast<size_t>(ns);
      double* Q = W.hch.p + kB1[f] * static_cast<size_t>(ns) + kB2[f] * static_cast<size_t>(ns) * ms;
      return ozaki8_dgemm(m, n, K, W.oz8a[f].p, W.coef.p + static_cast<size_t>(f) * nkE * nb2, nt, Q, nss,
                          accumulate, S, W.oz8b.p, c->stream, ns, static_cast<int>(ms));
    }
    return ozaki8_dgemm(m, n, K, W.oz8a[f].p, W.coef.p + static_cast<size_t>(f) * nkE * nb2, nt, C, nss, accumulate, S,
                        W.oz8b.p, c->stream);
  }
  if (engine >= 10) {
    const int S = engine - 10;
    HIPCHK(W.ozk.resize(ozaki_workspace_bytes(m, n, Kt, S)));
    return ozaki_dgemm(m, n, Kt, A, nss, B, nt, C, nss, accumulate, S, W.ozk.p, c->stream);
  }
  if (engine == 0) {
    const double alpha = 1.0, beta = accumulate ? 1.0 : 0.0;
    rocblas_status rs = rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, m, n, Kt, &alpha, A,
                                      m, B, static_cast<int>(nt), &beta, C, m);
    return rs == rocblas_status_success ? 0 : fail(-101, "rocblas_dgemm failed");
  }
  return mfma_dgemm(m, n, Kt, A, nss, B, nt, C, nss, accumulate, engine == 1 ? 1 : (engine == 2 ? 2 : 4), c->stream);
}

// FP64 engine of family f: the sparse exchange product (4) unless an engine
// is forced (0 library GEMM, 1-3 FP64 MFMA GEMM kernel); split precision:
// 10+S bf16, 20+S int8 slices
int dirac_engine(bsrg_ctx* c, int f) {
  if (c->precision_mode != 0) return 10 * c->precision_mode + c->nslices;
  if (c->forced_engine >= 0 && c->forced_engine != 4) return c->forced_engine;
  const int nr = (c->dw.khi[f] - c->dw.klo[f] + 1) * c->dw.nb * c->dw.nb;
  return nr <= xquad_max_t(c->device) ? 4 : 0;  // the library GEMM when the tile does not fit
}

}  // namespace

int bsrg_dirac_set_orbitals(bsrg_ctx* c, int nbf, const double* pq, const int* ipbs, int nv_ch, const int* i_ch,
                            const int* j_ch, const double* v_ch) {
  if (int rc = check(c)) return rc;
  if (nbf < 1 || !pq || !ipbs) return fail(-1, "bsrg_dirac_set_orbitals: no orbitals");
  if (nv_ch < 0 || (nv_ch > 0 && (!i_ch || !j_ch || !v_ch))) return fail(-1, "bsrg_dirac_set_orbitals: bad V_ch");
  HIPCHK(hipSetDevice(c->device));
  auto& W = c->dw;
  const auto& T = c->dbs.T;
  const int ns = T.ns, ks = T.ks, ms = 2 * ns;
  for (int m = 1; m < nv_ch; ++m)
    if (i_ch[m] < i_ch[m - 1] || (i_ch[m] == i_ch[m - 1] && j_ch[m] <= j_ch[m - 1]))
      return fail(-1, "bsrg_dirac_set_orbitals: (i_ch, j_ch) must be strictly increasing (KBORT searches them)");
  for (int m = 0; m < nv_ch; ++m)
    if (i_ch[m] < 1 || i_ch[m] > nbf || j_ch[m] < 1 || j_ch[m] > nbf)
      return fail(-1, "bsrg_dirac_set_orbitals: V_ch key out of range");
  // bases depend on the bound orbitals only; V_ch goes into the vector table
  std::vector<int> bidx(nbf, -1);
  int nb = 0;
  for (int o = 0; o < nbf; ++o)
    if (ipbs[o] == 0) bidx[o] = nb++;
  bool same_bound = W.orb_ok && W.nbf == nbf && W.bidx == bidx;
  if (same_bound)
    for (int o = 0; o < nbf && same_bound; ++o)
      if (bidx[o] >= 0)
        same_bound = std::equal(pq + static_cast<size_t>(o) * 2 * ns, pq + static_cast<size_t>(o + 1) * 2 * ns,
                                W.pq.begin() + static_cast<size_t>(o) * 2 * ns);
  W.nbf = nbf;
  W.pq.assign(pq, pq + static_cast<size_t>(ns) * 2 * nbf);
  W.ipbs.assign(ipbs, ipbs + nbf);
  W.bidx = bidx;
  W.nvch = nv_ch;
  W.vch.assign(v_ch ? v_ch : pq, v_ch ? v_ch + static_cast<size_t>(ms) * nv_ch : pq);
  W.kcol.assign(static_cast<size_t>(nbf) * nbf, 0);
  for (int m = 0; m < nv_ch; ++m) W.kcol[(i_ch[m] - 1) + static_cast<size_t>(nbf) * (j_ch[m] - 1)] = m + 1;
  if (same_bound) {
    if (W.basis_nk > 0) {  // new V_ch into the vector table, after the combined vectors
      const size_t nVc = 2 * static_cast<size_t>(W.basis_nk) * W.npair * W.nb;
      if ((nVc + 2 * static_cast<size_t>(nv_ch) + 1) * ns > W.VT.n) {
        W.basis_nk = 0;  // no room: the next assembly rebuilds the table
      } else if (nv_ch) {
        HIPCHK(hipMemcpyAsync(W.VT.p + nVc * ns, W.vch.data(), W.vch.size() * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
      }
    }
    W.orb_ok = true;
    return 0;
  }
  W.nb = nb;
  W.npair = nb * (nb + 1) / 2;
  W.basis_nk = W.exch_nk = 0;
  W.Q.clear();
  std::vector<double> P(static_cast<size_t>(2) * std::max(nb, 1) * ns, 0.0);  // [cpt][b](x)
  for (int o = 0; o < nbf; ++o)
    if (bidx[o] >= 0)
      for (int cpt = 0; cpt < 2; ++cpt)
        std::copy(pq + (static_cast<size_t>(o) * 2 + cpt) * ns, pq + (static_cast<size_t>(o) * 2 + cpt + 1) * ns,
                  P.begin() + (static_cast<size_t>(cpt) * nb + bidx[o]) * ns);
  HIPCHK(W.P.upload(P.data(), P.size(), c->stream));
  const size_t nsk = static_cast<size_t>(ns) * ks;
  HIPCHK(W.dens.resize(2 * static_cast<size_t>(std::max(W.npair, 1)) * nsk));
  if (nb > 0)
    for (int cpt = 0; cpt < 2; ++cpt)
      launch_pair_dens(ns, ks, nb, W.P.p + static_cast<size_t>(cpt) * nb * ns, W.dens.p + cpt * W.npair * nsk,
                       c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  W.orb_ok = true;
  return 0;
}

int bsrg_dirac_assemble(bsrg_ctx* c, int nch, int npert, int mk, int ntarg, const int* iptar, int nrec,
                        const int* itype, const int* kpol, const int* k1, const int* k2, const int* k3, const int* k4,
                        const double* cdata) {
  if (int rc = check(c)) return rc;
  auto& W = c->dw;
  if (!W.orb_ok) return fail(-4, "Dirac orbitals not set (bsrg_dirac_set_orbitals)");
  if (nch < 1 || npert < 0 || mk < 0 || ntarg < 1 || !iptar) return fail(-1, "bsrg_dirac_assemble: bad sizes");
  if (nrec < 0 || (nrec > 0 && (!itype || !kpol || !k1 || !k2 || !k3 || !k4 || !cdata)))
    return fail(-1, "bsrg_dirac_assemble: missing list arrays");
  for (int i = 0; i < nch; ++i)
    if (iptar[i] < 1 || iptar[i] > ntarg) return fail(-1, "bsrg_dirac_assemble: iptar out of 1..ntarg");
  HIPCHK(hipSetDevice(c->device));
  const auto hs = std::chrono::steady_clock::now();
  const auto& T = c->dbs.T;
  const int ns = T.ns, ks = T.ks, ms = 2 * ns;
  hipStream_t s = c->stream;
  int nk = 1;
  bool exch = false;
  for (int r = 0; r < nrec; ++r) {
    if (itype[r] < 1 || itype[r] > 4) return fail(-21, " R_data: unknown itype ");
    if (kpol[r] < 0) return fail(-1, "bsrg_dirac_assemble: kpol < 0");
    nk = std::max(nk, kpol[r] + 1);
    exch = exch || itype[r] == 4;
  }
  if (nk > 64) return fail(-1, "bsrg_dirac_assemble: kpol > 63");
  W.ok = W.tr[1] = false;
  (void)hipEventRecord(c->ev[0], s);
  if (int rc = dirac_bases(c, nk, exch)) return rc;
  (void)hipEventRecord(c->ev[1], s);
  // matrices
  W.nch = nch;
  W.npert = npert;
  W.mk = mk;
  W.ntarg = ntarg;
  W.iptar.assign(iptar, iptar + nch);
  W.nblk = static_cast<size_t>(nch) * (nch + 1) / 2;
  const size_t nss = static_cast<size_t>(ns) * ns, nq1 = static_cast<size_t>(nch) * npert;
  // the quadrant-major GEMM output: only the library and MFMA GEMMs and
  // the bf16 split form need it (the sparse product and the int8 split
  // GEMMs write the output layout directly)
  if (c->precision_mode == 1 || (c->precision_mode == 0 && c->forced_engine >= 0 && c->forced_engine <= 3))
    HIPCHK(W.hq.resize(4 * W.nblk * nss));
  HIPCHK(W.hv.resize(std::max<size_t>(2 * nq1 * ns, 1)));
  HIPCHK(W.hch.resize(W.nblk * ms * ms));
  HIPCHK(W.hcp.resize(std::max<size_t>(nq1 * ms, 1)));
  W.hp.assign(static_cast<size_t>(npert) * (npert + 1) / 2, 0.0);
  W.acf.assign(W.nblk * (mk + 1), 0.0);
  W.htarg.assign(static_cast<size_t>(ntarg) * (ntarg + 1) / 2, 0.0);
  // lists -> terms: records cut into ranges at list boundaries or where K1
  // changes (the ACF group sums of itype 3 stay whole); each host worker
  // decodes its ranges into its own part; parts stay in list order
  const auto h0 = std::chrono::steady_clock::now();
  if (int rc = wait_staged(W.st)) return rc;
  for (int r = 0; r < nrec; ++r)
    if (itype[r] == 3) (void)quadr_table(c, kpol[r]);  // tables the workers read
  if (!c->pool) {
    c->pool.reset(new Pool(host_workers()));
  }
  const int nth = static_cast<int>(std::min<long>(4L * c->pool->size(), nrec / 8192L + 1));
  std::vector<int> cuts{0};
  for (int t = 1; t < nth; ++t) {
    int b = static_cast<int>(static_cast<long>(nrec) * t / nth);
    if (b <= cuts.back()) continue;
    while (b < nrec && itype[b] == itype[b - 1] && kpol[b] == kpol[b - 1] && k1[b] == k1[b - 1]) ++b;
    if (b < nrec) cuts.push_back(b);
  }
  cuts.push_back(nrec);
  W.nparts = static_cast<int>(cuts.size()) - 1;
  while (static_cast<int>(W.parts.size()) < W.nparts) W.parts.emplace_back(new DiracWave::Part());
  const std::function<void(int)> work = [&](int t) {
    DiracWave::Part& P = *W.parts[t];
    P.T.clear();
    P.xc.clear();
    P.ax.clear();
    P.err = 0;
    for (int f = 0; f < 4; ++f) { P.klo[f] = 1 << 30; P.khi[f] = -1; }
    for (int r = cuts[t]; r < cuts[t + 1];) {
      int e = r + 1;
      while (e < cuts[t + 1] && itype[e] == itype[r] && kpol[e] == kpol[r]) ++e;
      if (int rc = dirac_decode(c, P, itype[r], kpol[r], e - r, k1 + r, k2 + r, k3 + r, k4 + r, cdata + r)) {
        P.err = rc;
        P.msg = bsrg_last_error();
        return;
      }
      r = e;
    }
  };
  c->pool->run(W.nparts, work);
  for (int t = 0; t < W.nparts; ++t)
    if (W.parts[t]->err) return fail(W.parts[t]->err, W.parts[t]->msg);
  bool has_x = false;
  for (int f = 0; f < 4; ++f) { W.klo[f] = 1 << 30; W.khi[f] = -1; }
  std::vector<Terms*> parts;
  for (int t = 0; t < W.nparts; ++t) {
    DiracWave::Part& P = *W.parts[t];
    for (const auto& a : P.ax) W.acf[a.first] += a.second;
    for (int f = 0; f < 4; ++f) { W.klo[f] = std::min(W.klo[f], P.klo[f]); W.khi[f] = std::max(W.khi[f], P.khi[f]); }
    has_x = has_x || !P.xc.empty();
    parts.push_back(&P.T);
  }
  const auto h1 = std::chrono::steady_clock::now();
  // the exchange coefficients first: the GEMMs need only them, and compute
  // while the other terms are sorted on the host and uploaded beside them
  if (int rc = stage_exchange(c)) return rc;
  const size_t nt = 4 * static_cast<size_t>(W.exch_nk) * W.nb * W.nb;
  if (has_x) {
    HIPCHK(W.xd.resize(W.xbytes));
    HIPCHK(hipMemcpyAsync(W.xd.p, W.xh.p, W.xbytes, hipMemcpyHostToDevice, s));
    HIPCHK(W.coef.resize(nt * W.nblk));
    hipLaunchKernelGGL(k_dirac_coef, dim3(static_cast<unsigned>(W.nblk)), dim3(256), 0, s, static_cast<int>(nt),
                       reinterpret_cast<const int*>(W.xd.p + W.o_xptr), reinterpret_cast<const int*>(W.xd.p + W.o_xt),
                       reinterpret_cast<const double*>(W.xd.p + W.o_xc), W.coef.p);
    HIPCHK(hipGetLastError());
  }
  size_t nr1 = 0;
  for (const Terms* P : parts) nr1 += P->r1.size();
  HIPCHK(c->bsum.resize(4 * W.nblk * ns * ks));
  HIPCHK(c->rfac.resize(std::max<size_t>(nr1, 1)));
  int eng[4] = {-1, -1, -1, -1};
  for (int f = 0; f < 4; ++f)
    if (W.khi[f] >= 0) {
      eng[f] = dirac_engine(c, f);
      if (eng[f] < 0) return eng[f];
    }
  // device work.  FP64 (and bf16) GEMMs write their quadrants of hq; one
  // pass then adds the band and rank-1 terms and writes the output layout.
  // int8 split GEMMs place their results in the output layout themselves
  // and the pass adds the terms in place, so no quadrant-major copy exists.
  // The diagonal channel blocks are symmetrised in place at the end.
  bool fused = true, sparse = false;
  for (int f = 0; f < 4; ++f) {
    fused = fused && (eng[f] < 0 || eng[f] == 4 || eng[f] >= 20);
    sparse = sparse || eng[f] == 4;
  }
  W.fused = fused;
  if (sparse) {
    // per-block nonzero lists of the coefficients of family 0 (the others'
    // are the same)
    const size_t nb2 = static_cast<size_t>(W.nb) * W.nb;
    const int nr = (W.khi[0] - W.klo[0] + 1) * static_cast<int>(nb2);
    const size_t slots = W.nblk * static_cast<size_t>(xasm_slot(nr));
    HIPCHK(W.qxo.resize(slots));
    HIPCHK(W.qxc.resize(slots));
    HIPCHK(W.qxn.resize(W.nblk));
    launch_coef_pack_quad(static_cast<int>(W.nblk), nr, xquad_width(nr, c->device),
                          W.coef.p + static_cast<size_t>(W.klo[0]) * nb2,
                          4 * static_cast<size_t>(W.exch_nk) * nb2, W.qxo.p, W.qxc.p, W.qxn.p, s);
    HIPCHK(hipGetLastError());
  }
  (void)hipEventRecord(c->ev[2], s);
  int qmask = 0;
  for (int f = 0; f < 4; ++f)
    if (eng[f] >= 0) {
      if (int rc = dirac_gemm(c, f, eng[f], false, fused)) return rc;
      qmask |= 1 << (2 * kB1[f] + kB2[f]);
    }
  (void)hipEventRecord(c->ev[3], s);
  // band, rank-1 and vector terms: sorted and uploaded while the GEMMs run
  if (int rc = stage_terms(parts, W.st, static_cast<int>(4 * W.nblk), static_cast<int>(2 * nq1), c->pool.get()))
    return rc;
  Staged& st = W.st;
  if (st.bytes) {
    HIPCHK(st.d.resize(st.bytes));
    HIPCHK(hipMemcpyAsync(st.d.p, st.h.p, st.bytes, hipMemcpyHostToDevice, c->copy_stream));
  }
  HIPCHK(hipEventRecord(st.done, c->copy_stream));
  st.pending = true;
  HIPCHK(hipStreamWaitEvent(s, st.done, 0));
  const auto h2 = std::chrono::steady_clock::now();
  // in place on the quadrants the sparse or int8 products wrote, else from hq
  launch_assemble(ns, ks, static_cast<int>(4 * W.nblk), 0, st.dp<int>(st.o_bptr), st.dp<int>(st.o_bid),
                  st.dp<double>(st.o_bc), W.Dc.p, st.dp<int>(st.o_rptr), st.dp<int>(st.o_ru), st.dp<int>(st.o_rw),
                  st.dp<int>(st.o_rs), st.dp<double>(st.o_rc), W.Sc.p, W.VT.p, W.hch.p, c->bsum.p, c->rfac.p, s,
                  fused ? W.hch.p : W.hq.p, static_cast<int>(W.nblk), qmask);
  W.engine = -1;
  for (int f = 0; f < 4 && W.engine < 0; ++f) W.engine = eng[f];
  HIPCHK(hipGetLastError());
  (void)hipEventRecord(c->ev[4], s);
  {
    const int nt = (ms + 31) / 32;
    hipLaunchKernelGGL(k_dirac_symm, dim3(nt * (nt + 1) / 2, nch), dim3(256), 0, s, ms, W.hch.p);
  }
  launch_vec_assemble(ns, static_cast<int>(2 * nq1), 0, st.dp<int>(st.o_vptr), st.dp<int>(st.o_vid),
                      st.dp<int>(st.o_vs), st.dp<double>(st.o_vc), W.Sc.p, W.VT.p, W.hv.p, s);
  if (nq1) hipLaunchKernelGGL(k_dirac_vfinalize, dim3(static_cast<unsigned>(nq1)), dim3(256), 0, s, ns, static_cast<int>(nq1), W.hv.p, W.hcp.p);
  HIPCHK(hipGetLastError());
  (void)hipEventRecord(c->ev[5], s);
  HIPCHK(hipEventSynchronize(c->ev[5]));
  st.pending = false;
  // host scalars (HB, Target_h) in list order
  for (int t = 0; t < W.nparts; ++t)
    for (const auto& x : W.parts[t]->T.scal) {
      const double v = x.c * (x.sidx >= 0 ? W.St_host.p[x.sidx] : 1.0);
      if (x.kind == 0) W.hp[x.idx] += v;
      else W.htarg[x.idx] += v;
    }
  float e01 = 0, e23 = 0, e34 = 0, e45 = 0, e05 = 0;  // bases, GEMMs, terms + layout, symmetrise + vectors
  (void)hipEventElapsedTime(&e01, c->ev[0], c->ev[1]);
  (void)hipEventElapsedTime(&e23, c->ev[2], c->ev[3]);
  (void)hipEventElapsedTime(&e34, c->ev[3], c->ev[4]);
  (void)hipEventElapsedTime(&e45, c->ev[4], c->ev[5]);
  (void)hipEventElapsedTime(&e05, c->ev[0], c->ev[5]);
  W.ms[0] = e01; W.ms[1] = e23; W.ms[2] = e34; W.ms[3] = e45; W.ms[4] = e05;
  auto msd = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  W.host_ms[0] = msd(h0, h1);
  W.host_ms[1] = msd(h1, h2);
  W.host_ms[2] = msd(hs, std::chrono::steady_clock::now());
  W.ok = true;
  return 0;
}

int bsrg_dirac_get(bsrg_ctx* c, double* hch, double* hcp, double* hp, double* acf, double* htarg) {
  if (int rc = check(c)) return rc;
  auto& W = c->dw;
  if (!W.ok) return fail(-6, "no assembled Dirac wave (bsrg_dirac_assemble)");
  HIPCHK(hipSetDevice(c->device));
  const size_t ms = 2 * static_cast<size_t>(c->dbs.T.ns), nq1 = static_cast<size_t>(W.nch) * W.npert;
  if (hch) HIPCHK(hipMemcpyAsync(hch, W.hch.p, W.nblk * ms * ms * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (hcp && nq1) HIPCHK(hipMemcpyAsync(hcp, W.hcp.p, nq1 * ms * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (hp) std::copy(W.hp.begin(), W.hp.end(), hp);
  if (acf) std::copy(W.acf.begin(), W.acf.end(), acf);
  if (htarg) std::copy(W.htarg.begin(), W.htarg.end(), htarg);
  return 0;
}

int bsrg_dirac_timings(bsrg_ctx* c, double* ms8) {
  if (!c || !ms8) return fail(-2, "null argument");
  std::copy(c->dw.ms, c->dw.ms + 5, ms8);
  std::copy(c->dw.host_ms, c->dw.host_ms + 3, ms8 + 5);
  return 0;
}

int bsrg_dirac_engine(bsrg_ctx* c) { return c ? c->dw.engine : -1; }

int bsrg_dirac_diff(bsrg_ctx* a, bsrg_ctx* ref, double floor_frac, double* rel) {
  if (!a || !ref || !rel) return fail(-2, "null argument");
  const auto& A = a->dw;
  const auto& B = ref->dw;
  if (!A.ok || !B.ok) return fail(-6, "no assembled Dirac wave (bsrg_dirac_assemble)");
  if (a->device != ref->device) return fail(-5, "bsrg_dirac_diff: contexts on different devices");
  if (a->dbs.T.ns != ref->dbs.T.ns || A.nch != B.nch || A.npert != B.npert || A.mk != B.mk || A.ntarg != B.ntarg)
    return fail(-1, "bsrg_dirac_diff: different shapes");
  HIPCHK(hipSetDevice(ref->device));
  HIPCHK(hipStreamSynchronize(a->stream));
  hipStream_t s = ref->stream;
  const size_t ms = 2 * static_cast<size_t>(ref->dbs.T.ns);
  const size_t nh = B.nblk * ms * ms, nv = static_cast<size_t>(B.nch) * B.npert * ms;
  HIPCHK(ref->dtmp.resize(2 * kReduceBlocks));
  std::vector<double> part(2 * kReduceBlocks, 0.0);
  // scale of the reference: largest element of its channel blocks, vectors and scalars
  launch_absmax(nh, B.hch.p, ref->dtmp.p, s);
  if (nv) launch_absmax(nv, B.hcp.p, ref->dtmp.p + kReduceBlocks, s);
  else launch_fill(kReduceBlocks, 0.0, ref->dtmp.p + kReduceBlocks, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(part.data(), ref->dtmp.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  // one scale for the matrix parts (hch, hcp, hp); acf and htarg are other
  // quantities with scales of their own
  double mx = 0.0;
  for (double v : part) mx = std::max(mx, v);
  for (double v : B.hp) mx = std::max(mx, std::fabs(v));
  const double floor = std::max(floor_frac * mx, 1e-300);
  launch_reldiff(nh, A.hch.p, B.hch.p, floor, ref->dtmp.p, s);
  if (nv) launch_reldiff(nv, A.hcp.p, B.hcp.p, floor, ref->dtmp.p + kReduceBlocks, s);
  else launch_fill(kReduceBlocks, 0.0, ref->dtmp.p + kReduceBlocks, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(part.data(), ref->dtmp.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  double r = 0.0;
  for (double v : part) r = std::max(r, v);
  auto host_rel = [floor_frac](const std::vector<double>& a, const std::vector<double>& b, double fl) {
    double m = 0.0, e = 0.0;
    if (fl <= 0.0) {
      for (double v : b) m = std::max(m, std::fabs(v));
      fl = std::max(floor_frac * m, 1e-300);
    }
    for (size_t i = 0; i < b.size(); ++i) e = std::max(e, std::fabs(a[i] - b[i]) / std::max(std::fabs(b[i]), fl));
    return e;
  };
  r = std::max(r, host_rel(A.hp, B.hp, floor));
  r = std::max(r, host_rel(A.acf, B.acf, 0.0));
  r = std::max(r, host_rel(A.htarg, B.htarg, 0.0));
  *rel = r;
  return 0;
}
