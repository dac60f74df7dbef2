// Dirac B-spline R^k bands on gfx950: the four integral families
// PPPP, QQQQ, PQPQ, QPQP of the P (order ksp) and Q (order ksq) bases.
//
// A family f has basis b1 on the r1 pair and b2 on the r2 pair:
//   0 PPPP (P,P)  1 QQQQ (Q,Q)  2 PQPQ (P,Q)  3 QPQP (Q,P)
// Band storage as the reference (rkb(ns,ns,ks,ks), i fastest):
//   rkb_f(i,j,ip,jp) = R^k(B1_i B1_i+ip ; B2_j B2_j+jp)  (0-based offsets)
// with ip < kb1, jp < kb2 (other entries zero).  Every cell is integrated
// (no scaling law), with the nodes and weights of the DBS Gauss rule.
#include <algorithm>

#include "bsrg_internal.h"

namespace bsrg {

namespace {

__device__ __forceinline__ double ipow(double a, int n) {
  bool recip = n < 0;
  unsigned b = recip ? -n : n;
  double r = 1.0;
  for (;;) {
    if (b & 1u) r *= a;
    b >>= 1;
    if (!b) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}

__device__ __forceinline__ int pk2(int a, int b, int kb) { return a * kb - a * (a - 1) / 2 + (b - a); }

__device__ __forceinline__ void upk2(int ii, int kb, int& a, int& b) {
  a = 0;
  while (ii >= kb - a) { ii -= kb - a; ++a; }
  b = a + ii;
}

__device__ void bvals(const double* t, int kb, int left, double x, double* b) {
  double dr[kMaxKs], dl[kMaxKs];
  b[0] = 1.0;
  for (int j = 1; j < kb; ++j) {
    dr[j - 1] = t[left + j - 1] - x;
    dl[j - 1] = x - t[left - j];
    double saved = 0.0;
    for (int r = 0; r < j; ++r) {
      const double term = b[r] / (dr[r] + dl[j - 1 - r]);
      b[r] = saved + dr[r] * term;
      saved = dl[j - 1 - r] * term;
    }
    b[j] = saved;
  }
}

__device__ __forceinline__ void family_bases(int f, int& q1, int& q2) {
  q1 = (f == 1 || f == 3) ? 1 : 0;
  q2 = (f == 1 || f == 2) ? 1 : 0;
}

// moments of basis q: mom[((kk*2+s)*nv+iv)*ikb+ii] = sum_m B_i B_j r^p w,
// p = kk (s=0) or -(kk+1) (s=1) (moment_pp / moment_qq).
// grid (ceil(nv*ikb/256), nk, 2)
__global__ void k_dbs_moments(DevDbs g, int q, double* mom) {
  const int kb = g.kb[q], ikb = kb * (kb + 1) / 2, ks = g.ks, nv = g.nv;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int kk = blockIdx.y, s = blockIdx.z;
  if (gid >= nv * ikb) return;
  const int iv = gid / ikb, ii = gid % ikb;
  int i, j;
  upk2(ii, kb, i, j);
  const int p = s == 0 ? kk : -(kk + 1);
  const double* v = g.val[q];
  double sum = 0.0;
  for (int m = 0; m < ks; ++m) {
    const int c = iv * ks + m;
    double w = g.grw[c];
    if (p > 0) w = g.grw[c] * ipow(g.gr[c], p);
    if (p < 0) w = g.grw[c] * ipow(g.grm[c], -p);
    sum += v[c * kb + i] * (v[c * kb + j] * w);
  }
  mom[(static_cast<size_t>(kk * 2 + s) * nv + iv) * ikb + ii] = sum;
}

// Diagonal cell iv of family f for multipole kk: rkd[((kk*nv+iv)*ik2+jj)*ik1+ii]
// = [r1 pair outer, r2 pair inner] + [r2 pair outer, r1 pair inner]
// (QPQP adds the r2-outer term first, as its reference routine does).
// Inner B-spline values, inner integrals and outer pair values in LDS.
// grid (nv, nk), block 128
__global__ void k_dbs_diag(DevDbs g, int f, double* rkd) {
  const int ks = g.ks, nv = g.nv, iv = blockIdx.x, kk = blockIdx.y;
  int q1, q2;
  family_bases(f, q1, q2);
  const int kb1 = g.kb[q1], kb2 = g.kb[q2];
  const int ik1 = kb1 * (kb1 + 1) / 2, ik2 = kb2 * (kb2 + 1) / 2;
  extern __shared__ double sh[];
  double* bt1 = sh;                          // [m][i][kb1]
  double* bt2 = bt1 + ks * ks * kb1;         // [m][i][kb2]
  double* wk = bt2 + ks * ks * kb2;          // [m][i]
  double* in1 = wk + ks * ks;                // [m][ik1]
  double* in2 = in1 + ks * ik1;              // [m][ik2]
  double* out1 = in2 + ks * ik2;             // [ik1][m]
  double* out2 = out1 + ik1 * ks;            // [ik2][m]
  const double xb = g.t[iv + ks - 1];
  for (int qq = threadIdx.x; qq < ks * ks; qq += blockDim.x) {
    const int m = qq / ks, i = qq % ks;
    const double r = g.gr[iv * ks + m];
    const double x = (r - xb) * g.gx[i] + xb;
    bvals(g.tb[q1], kb1, iv + kb1, x, bt1 + qq * kb1);
    bvals(g.tb[q2], kb2, iv + kb2, x, bt2 + qq * kb2);
    wk[qq] = ((r - xb) * g.gw[i]) * ipow(x, kk);
  }
  __syncthreads();
  for (int qq = threadIdx.x; qq < ks * (ik1 + ik2); qq += blockDim.x) {
    const bool first = qq < ks * ik1;
    const int e = first ? qq : qq - ks * ik1;
    const int ik = first ? ik1 : ik2, kb = first ? kb1 : kb2;
    const double* bt = first ? bt1 : bt2;
    const int m = e / ik, jj = e % ik;
    int j, jp;
    upk2(jj, kb, j, jp);
    double s = 0.0;
    for (int i = 0; i < ks; ++i) s += (wk[m * ks + i] * bt[(m * ks + i) * kb + j]) * bt[(m * ks + i) * kb + jp];
    (first ? in1 : in2)[m * ik + jj] = s;
  }
  for (int qq = threadIdx.x; qq < ks * (ik1 + ik2); qq += blockDim.x) {
    const bool first = qq < ks * ik1;
    const int e = first ? qq : qq - ks * ik1;
    const int kb = first ? kb1 : kb2, q = first ? q1 : q2;
    const int ii = e / ks, m = e % ks;
    int i, ip;
    upk2(ii, kb, i, ip);
    const int c = iv * ks + m;
    const double gg = g.grw[c] * ipow(g.grm[c], kk + 1);
    (first ? out1 : out2)[e] = (g.val[q][c * kb + i] * g.val[q][c * kb + ip]) * gg;
  }
  __syncthreads();
  double* out = rkd + (static_cast<size_t>(kk) * nv + iv) * ik1 * ik2;
  for (int qq = threadIdx.x; qq < ik1 * ik2; qq += blockDim.x) {
    const int jj = qq / ik1, ii = qq % ik1;
    double s1 = 0.0, s2 = 0.0;
    for (int m = 0; m < ks; ++m) s1 += out1[ii * ks + m] * in2[m * ik2 + jj];
    for (int m = 0; m < ks; ++m) s2 += out2[jj * ks + m] * in1[m * ik1 + ii];
    out[jj * ik1 + ii] = f == 3 ? (0.0 + s2) + s1 : (0.0 + s1) + s2;
  }
}

// per-pair sums of the moments of basis q over the cells of the pair:
// msum[((kk*2+s)*kb+ip)*ns+i], pair (i, i+ip) of basis q
// grid (ceil(ns/256), kb, nk*2)
__global__ void k_dbs_pair_sums(DevDbs g, int q, const double* __restrict__ mom, double* __restrict__ msum) {
  const int ns = g.ns, nv = g.nv, kb = g.kb[q], ikb = kb * (kb + 1) / 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int ip = blockIdx.y, kk = blockIdx.z >> 1, s = blockIdx.z & 1;
  if (i >= ns) return;
  const double* m = mom + static_cast<size_t>(kk * 2 + s) * nv * ikb;
  double acc = 0.0;
  if (i + ip < g.nsb[q]) {
    const int ivlo = max(0, i + ip - (kb - 1)), ivhi = min(i, nv - 1);
    for (int iv = ivlo; iv <= ivhi; ++iv) acc += m[iv * ikb + pk2(i - iv, i + ip - iv, kb)];
  }
  msum[(static_cast<size_t>(kk * 2 + s) * kb + ip) * ns + i] = acc;
}

// Far elements: r1 cells all below (or all above) the r2 cells; the cell
// sum factorises into per-pair moment sums.  Near elements are written too
// and overwritten by k_dbs_band_near.  grid (ceil(ns*ns/256), ks*ks, nk)
__global__ void k_dbs_band_far(DevDbs g, int f, const double* __restrict__ msum1, const double* __restrict__ msum2,
                               double* __restrict__ rkb) {
  const int ns = g.ns, ks = g.ks;
  const int gid = blockIdx.x * 256 + threadIdx.x;
  if (gid >= ns * ns) return;
  int q1, q2;
  family_bases(f, q1, q2);
  const int kb1 = g.kb[q1], kb2 = g.kb[q2];
  const int i = gid % ns, j = gid / ns;
  const int r = blockIdx.y, ip = r % ks, jp = r / ks, kk = blockIdx.z;
  double v = 0.0;
  if (ip < kb1 && jp < kb2 && i + ip < g.nsb[q1] && j + jp < g.nsb[q2]) {
    const double* a = msum1 + static_cast<size_t>(kk * 2) * kb1 * ns;  // r1 basis: r^k, then r^-(k+1)
    const double* b = msum2 + static_cast<size_t>(kk * 2) * kb2 * ns;
    const double a_k = a[ip * ns + i], a_m = a[static_cast<size_t>(kb1) * ns + ip * ns + i];
    const double b_k = b[jp * ns + j], b_m = b[static_cast<size_t>(kb2) * ns + jp * ns + j];
    v = j > i ? a_k * b_m : b_k * a_m;
  }
  rkb[(static_cast<size_t>(kk) * ks * ks + r) * ns * ns + gid] = v;
}

// Near elements: d = j - i in [ip-kb1+1, kb2-1-jp]; one thread per i for a
// fixed (ip, jp, d); cells iv = i - o, jv = j - q visited jv ascending then
// iv ascending (the reference's order); cell-major moment copies make the
// loads of a wavefront consecutive.
// grid (ceil(ns/64), 2*ks-1, ks*ks*nk), block 64; blockIdx.y = d + ks - 1
__global__ void k_dbs_band_near(DevDbs g, int f, const double* __restrict__ momT1,
                                const double* __restrict__ momT2, const double* __restrict__ rdT,
                                double* __restrict__ rkb) {
  const int ns = g.ns, ks = g.ks, nv = g.nv;
  int q1, q2;
  family_bases(f, q1, q2);
  const int kb1 = g.kb[q1], kb2 = g.kb[q2];
  const int ik1 = kb1 * (kb1 + 1) / 2, ik2 = kb2 * (kb2 + 1) / 2;
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int d = static_cast<int>(blockIdx.y) - (ks - 1);
  const int kk = blockIdx.z / (ks * ks), r = blockIdx.z % (ks * ks);
  const int ip = r % ks, jp = r / ks;
  const int j = i + d;
  if (ip >= kb1 || jp >= kb2) return;
  if (d < ip - kb1 + 1 || d > kb2 - 1 - jp) return;
  if (i >= ns || j < 0 || i + ip >= g.nsb[q1] || j + jp >= g.nsb[q2]) return;
  const double* m1k = momT1 + static_cast<size_t>(kk * 2) * ik1 * nv;  // [ii][iv]
  const double* m1m = m1k + static_cast<size_t>(ik1) * nv;
  const double* m2k = momT2 + static_cast<size_t>(kk * 2) * ik2 * nv;
  const double* m2m = m2k + static_cast<size_t>(ik2) * nv;
  const double* rd = rdT + static_cast<size_t>(kk) * ik1 * ik2 * nv;    // [jj][ii][iv]
  const int olo = max(0, i - (nv - 1)), ohi = min(kb1 - 1 - ip, i);
  const int qlo = max(0, j - (nv - 1)), qhi = min(kb2 - 1 - jp, j);
  double acc = 0.0;
  for (int q = kb2 - 1 - jp; q >= 0; --q) {
    if (q < qlo || q > qhi) continue;
    const int jv = j - q, jj = pk2(q, q + jp, kb2);
    for (int o = kb1 - 1 - ip; o >= 0; --o) {
      if (o < olo || o > ohi) continue;
      const int iv = i - o, ii = pk2(o, o + ip, kb1);
      double c;
      if (q - o < d) c = m1k[static_cast<size_t>(ii) * nv + iv] * m2m[static_cast<size_t>(jj) * nv + jv];
      else if (q - o > d) c = m2k[static_cast<size_t>(jj) * nv + jv] * m1m[static_cast<size_t>(ii) * nv + iv];
      else c = rd[(static_cast<size_t>(jj) * ik1 + ii) * nv + iv];
      acc += c;
    }
  }
  rkb[(static_cast<size_t>(kk) * ks * ks + r) * ns * ns + i + static_cast<size_t>(ns) * j] = acc;
}

unsigned nblk(size_t n, unsigned b) { return static_cast<unsigned>((n + b - 1) / b); }

}  // namespace

void launch_dbs_moments(const DevDbs& g, int q, int nk, double* mom, double* momT, hipStream_t s) {
  const int ikb = g.kb[q] * (g.kb[q] + 1) / 2;
  hipLaunchKernelGGL(k_dbs_moments, dim3(nblk(static_cast<size_t>(g.nv) * ikb, 256), nk, 2), dim3(256), 0, s, g, q,
                     mom);
  launch_cells_transpose(g.nv, ikb, nk * 2, mom, momT, s);
}

size_t dbs_diag_lds(int ks, int kb1, int kb2) {
  const int ik1 = kb1 * (kb1 + 1) / 2, ik2 = kb2 * (kb2 + 1) / 2;
  return sizeof(double) * (static_cast<size_t>(ks) * ks * (kb1 + kb2 + 1) + 2 * static_cast<size_t>(ks) * (ik1 + ik2));
}

void launch_dbs_band(const DevDbs& g, int f, int nk, const double* mom1, const double* mom2, const double* momT1,
                     const double* momT2, double* msum1, double* msum2, double* rkd, double* rdT, double* rkb,
                     hipStream_t s) {
  const int q1 = (f == 1 || f == 3) ? 1 : 0, q2 = (f == 1 || f == 2) ? 1 : 0;
  const int kb1 = g.kb[q1], kb2 = g.kb[q2], ik1 = kb1 * (kb1 + 1) / 2, ik2 = kb2 * (kb2 + 1) / 2;
  hipLaunchKernelGGL(k_dbs_diag, dim3(g.nv, nk), dim3(128), dbs_diag_lds(g.ks, kb1, kb2), s, g, f, rkd);
  launch_cells_transpose(g.nv, ik1 * ik2, nk, rkd, rdT, s);
  hipLaunchKernelGGL(k_dbs_pair_sums, dim3(nblk(g.ns, 256), kb1, nk * 2), dim3(256), 0, s, g, q1, mom1, msum1);
  hipLaunchKernelGGL(k_dbs_pair_sums, dim3(nblk(g.ns, 256), kb2, nk * 2), dim3(256), 0, s, g, q2, mom2, msum2);
  hipLaunchKernelGGL(k_dbs_band_far, dim3(nblk(static_cast<size_t>(g.ns) * g.ns, 256), g.ks * g.ks, nk), dim3(256), 0,
                     s, g, f, msum1, msum2, rkb);
  hipLaunchKernelGGL(k_dbs_band_near, dim3((g.ns + 63) / 64, 2 * g.ks - 1, g.ks * g.ks * nk), dim3(64), 0, s, g, f,
                     momT1, momT2, rdT, rkb);
}

}  // namespace bsrg
