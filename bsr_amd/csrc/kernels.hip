// FP64 kernels of the R^k radial core for gfx950 (64-wide wavefronts).
//
// Layouts (0-based, column-major as the Fortran interface):
//   rkb_k(i,j,ip,jp) at i + ns*(j + ns*(ip + ks*jp))   pair (i,i+ip) on r1,
//                                                       pair (j,j+jp) on r2
//   upper band d(i,ith) at i + ns*ith                   pair (i,i+ith)
//   hcc block b (x,y)  at b*ns*ns + x + ns*y
#include <algorithm>

#include "bsrg_internal.h"

namespace bsrg {

namespace {

__device__ __forceinline__ double dpowi(double a, int n) {
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

// packed index of (a,b), a<=b<ks, row-major over a: the reference's ii/jj
__device__ __forceinline__ int pack2(int a, int b, int ks) { return a * ks - a * (a - 1) / 2 + (b - a); }

__device__ __forceinline__ void unpack2(int ii, int ks, int& a, int& b) {
  a = 0;
  while (ii >= ks - a) { ii -= ks - a; ++a; }
  b = a + ii;
}

__device__ void dev_bvalues(const double* t, int ks, int left, double x, double* b) {
  double dr[kMaxKs], dl[kMaxKs];
  b[0] = 1.0;
  for (int j = 1; j < ks; ++j) {
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

// ---- a5: one-interval moments <B_i|r^p|B_j>, packed i<=j ---------------
// grid (ceil(nv*ik/256), nk, 2)
__global__ void k_moments(DevGrid g, int k0, double* mom) {
  const int ks = g.ks, nv = g.nv, ik = ks * (ks + 1) / 2;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int kk = blockIdx.y, s = blockIdx.z;
  if (gid >= nv * ik) return;
  const int iv = gid / ik, ii = gid % ik;
  int i, j;
  unpack2(ii, ks, i, j);
  const int k = k0 + kk, p = s == 0 ? k : -(k + 1);
  double sum = 0.0;
  for (int m = 0; m < ks; ++m) {
    const int q = iv * ks + m;
    double w = g.grw[q];
    if (p > 0) w = g.grw[q] * dpowi(g.gr[q], p);
    if (p < 0) w = g.grw[q] * dpowi(g.grm[q], -p);
    sum += g.bsp[q * ks + i] * (g.bsp[q * ks + j] * w);
  }
  mom[(static_cast<size_t>(kk * 2 + s) * nv + iv) * ik + ii] = sum;
}

// scaling law in the geometric region (moments.f90:54-60): serial over iv
// grid (nk*2), block ik
__global__ void k_moments_scale(DevGrid g, int k0, double* mom) {
  const int ks = g.ks, nv = g.nv, ik = ks * (ks + 1) / 2;
  const int ii = threadIdx.x, kk = blockIdx.x / 2, s = blockIdx.x % 2;
  if (ii >= ik) return;
  const int k = k0 + kk, p = s == 0 ? k : -(k + 1);
  const double hp = dpowi(1.0 + g.h, p) * (1.0 + g.h);
  double* m = mom + static_cast<size_t>(kk * 2 + s) * nv * ik;
  const int lo = g.ml + ks - 1, hi = g.ml + g.me - ks + 1;  // 0-based, inclusive
  if (lo > hi || lo >= nv) return;
  double v = m[(lo - 1) * ik + ii];  // carried in a register: same products, no reload
  for (int iv = lo; iv <= hi && iv < nv; ++iv) {
    v *= hp;
    m[iv * ik + ii] = v;
  }
}

// ---- a6: diagonal-cell quadrature (rk_triang) ---------------------------
// one workgroup per (iv, kk); B-splines at the ks*ks inner Gauss points and
// the per-cell integrals are staged in LDS.
__global__ void k_diag_cells(DevGrid g, int k0, double* rkd) {
  const int ks = g.ks, nv = g.nv, ik = ks * (ks + 1) / 2;
  const int iv = blockIdx.x, kk = blockIdx.y, k = k0 + kk;
  extern __shared__ double sh[];
  double* bt = sh;                        // [m][i][j]   ks^3
  double* wk = bt + ks * ks * ks;         // [m][i]      ks^2
  double* in = wk + ks * ks;              // [m][jj]     ks*ik
  double* bi = in + ks * ik;              // [ii][m]     ik*ks
  double* a = bi + ik * ks;               // [jj][ii]    ik*ik
  const int left = iv + ks;               // 1-based knot index of the cell's left end
  const double xb = g.t[left - 1];
  for (int q = threadIdx.x; q < ks * ks; q += blockDim.x) {
    const int m = q / ks, i = q % ks;
    const double r1 = g.gr[iv * ks + m];
    const double x = (r1 - xb) * g.gx[i] + xb;
    dev_bvalues(g.t, ks, left, x, bt + (m * ks + i) * ks);
    const double w = (r1 - xb) * g.gw[i];
    wk[q] = k > 1 ? w * dpowi(x, k) : (k == 1 ? w * x : w);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < ks * ik; q += blockDim.x) {
    const int m = q / ik, jj = q % ik;
    int j, jp;
    unpack2(jj, ks, j, jp);
    double s = 0.0;
    for (int i = 0; i < ks; ++i)
      s += (wk[m * ks + i] * bt[(m * ks + i) * ks + j]) * bt[(m * ks + i) * ks + jp];
    in[m * ik + jj] = s;
  }
  for (int q = threadIdx.x; q < ik * ks; q += blockDim.x) {
    const int ii = q / ks, m = q % ks;
    int i, ip;
    unpack2(ii, ks, i, ip);
    const int c = iv * ks + m;
    const double gx2 = k != 0 ? g.grw[c] * dpowi(g.grm[c], k + 1) : g.grw[c] * g.grm[c];
    bi[q] = g.bsp[c * ks + i] * g.bsp[c * ks + ip] * gx2;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < ik * ik; q += blockDim.x) {
    const int jj = q / ik, ii = q % ik;
    double s = 0.0;
    for (int m = 0; m < ks; ++m) s += bi[ii * ks + m] * in[m * ik + jj];
    a[jj * ik + ii] = s;
  }
  __syncthreads();
  double* out = rkd + (static_cast<size_t>(kk) * nv + iv) * ik * ik;
  for (int q = threadIdx.x; q < ik * ik; q += blockDim.x) {
    const int jj = q / ik, ii = q % ik;
    out[jj * ik + ii] = a[jj * ik + ii] + a[ii * ik + jj];
  }
}

// scaling law for diagonal cells, rkd(iv)=rkd(iv-1)*(1+h) (rk_moments.f90:87-93)
// grid (nk), block 256
__global__ void k_diag_scale(DevGrid g, double* rkd) {
  const int ks = g.ks, nv = g.nv, ik = ks * (ks + 1) / 2, kk = blockIdx.x;
  const double hp = g.h + 1.0;
  const int lo = g.ml + ks - 1, hi = g.ml + g.me - ks + 1;
  double* base = rkd + static_cast<size_t>(kk) * nv * ik * ik;
  if (lo > hi || lo >= nv) return;
  for (int q = threadIdx.x; q < ik * ik; q += blockDim.x) {
    double v = base[static_cast<size_t>(lo - 1) * ik * ik + q];
    for (int iv = lo; iv <= hi && iv < nv; ++iv) {
      v *= hp;
      base[static_cast<size_t>(iv) * ik * ik + q] = v;
    }
  }
}

// ---- a7: band assembly (mrk_cell) ----------------------------------------
// Sum over the cells of one spline pair of each one-cell moment:
// msum[((kk*2+s)*ks+ip)*ns+i] = sum_{iv in cells(i,i+ip)} mom_s(iv; i,i+ip).
__global__ void k_pair_sums(int ns, int ks, int nv, const double* __restrict__ mom, double* __restrict__ msum) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int ip = blockIdx.y, kk = blockIdx.z >> 1, s = blockIdx.z & 1;
  if (i >= ns) return;
  const int ik = ks * (ks + 1) / 2;
  const double* m = mom + static_cast<size_t>(kk * 2 + s) * nv * ik;
  double acc = 0.0;
  if (i + ip < ns) {
    const int ivlo = max(0, i + ip - (ks - 1)), ivhi = min(i, nv - 1);
    for (int iv = ivlo; iv <= ivhi; ++iv) acc += m[iv * ik + pack2(i - iv, i + ip - iv, ks)];
  }
  msum[(static_cast<size_t>(kk * 2 + s) * ks + ip) * ns + i] = acc;
}

// The band splits into two element sets.  For a pair (i,i+ip) on r1 and
// (j,j+jp) on r2, the cells of the two pairs are disjoint and ordered unless
// d = j - i lies in [ip-ks+1, ks-1-jp] ("near"); the clamping of the cell
// ranges at 0 and nv-1 does not change this for elements inside the band.
// Far elements factorise: rkb = (sum of r^k moments) * (sum of r^-(k+1)
// moments); near elements sum the cell pairs in the reference's order, cell
// jv ascending, then cell iv.

// Far elements and the zero padding outside the band: one thread per element
// over the flattened (i,j) plane, fully coalesced writes.  Near elements are
// written too (with the product value) and overwritten by k_band_near.
// grid (ceil(ns*ns/256), ks*ks, nk), block 256
__global__ void k_band_far(int ns, int ks, const double* __restrict__ msum, double* __restrict__ rkb) {
  const int gid = blockIdx.x * 256 + threadIdx.x;
  if (gid >= ns * ns) return;
  const int i = gid % ns, j = gid / ns;
  const int r = blockIdx.y, ip = r % ks, jp = r / ks, kk = blockIdx.z;
  double v = 0.0;
  if (i + ip < ns && j + jp < ns) {
    const double* s1 = msum + static_cast<size_t>(kk * 2) * ks * ns;
    const double* s2 = s1 + static_cast<size_t>(ks) * ns;
    v = j > i ? s1[ip * ns + i] * s2[jp * ns + j] : s1[jp * ns + j] * s2[ip * ns + i];
  }
  rkb[(static_cast<size_t>(kk) * ks * ks + r) * ns * ns + gid] = v;
}

// Cell-major copies of the moments and diagonal-cell integrals for the near
// elements: in[(b*nv + iv)*w + e] -> out[(b*w + e)*nv + iv].
// grid (ceil(w/32), ceil(nv/32), nb), block 256 (32 x 8)
__global__ void k_cells_transpose(int nv, int w, const double* __restrict__ in, double* __restrict__ out) {
  __shared__ double tile[32][33];
  const int e0 = blockIdx.x * 32, v0 = blockIdx.y * 32, b = blockIdx.z;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const double* src = in + static_cast<size_t>(b) * nv * w;
  double* dst = out + static_cast<size_t>(b) * nv * w;
  for (int r = ty; r < 32; r += 8) {
    const int iv = v0 + r, e = e0 + tx;
    if (iv < nv && e < w) tile[r][tx] = src[static_cast<size_t>(iv) * w + e];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int e = e0 + r, iv = v0 + tx;
    if (iv < nv && e < w) dst[static_cast<size_t>(e) * nv + iv] = tile[tx][r];
  }
}

// Near elements: one thread per i for a fixed (ip, jp, d = j - i).  The cells
// are visited as iv = i - o, jv = j - q; for fixed (o, q) the lanes read
// consecutive cells of the cell-major copies, and whether iv < jv depends on
// (o, q, d) only, so the wavefront does not diverge.  Order as the reference:
// cell jv ascending, then cell iv ascending.
// grid (ceil(ns/64), 2*ks-1, ks*ks*nk), block 64; blockIdx.y = d + ks - 1
__global__ void k_band_near(int ns, int ks, int nv, const double* __restrict__ momT, const double* __restrict__ rdT,
                            double* __restrict__ rkb) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int d = static_cast<int>(blockIdx.y) - (ks - 1);
  const int kk = blockIdx.z / (ks * ks), r = blockIdx.z % (ks * ks);
  const int ip = r % ks, jp = r / ks;
  const int j = i + d;
  if (d < ip - ks + 1 || d > ks - 1 - jp) return;
  if (i >= ns || j < 0 || i + ip >= ns || j + jp >= ns) return;
  const int ik = ks * (ks + 1) / 2;
  const double* m1 = momT + static_cast<size_t>(kk * 2) * ik * nv;  // [ii][iv]
  const double* m2 = m1 + static_cast<size_t>(ik) * nv;
  const double* rd = rdT + static_cast<size_t>(kk) * ik * ik * nv;   // [jj][ii][iv]
  const int olo = max(0, i - (nv - 1)), ohi = min(ks - 1 - ip, i);
  const int qlo = max(0, j - (nv - 1)), qhi = min(ks - 1 - jp, j);
  double acc = 0.0;
  for (int q = ks - 1 - jp; q >= 0; --q) {
    if (q < qlo || q > qhi) continue;
    const int jv = j - q, jj = pack2(q, q + jp, ks);
    for (int o = ks - 1 - ip; o >= 0; --o) {
      if (o < olo || o > ohi) continue;
      const int iv = i - o, ii = pack2(o, o + ip, ks);
      double c;
      if (q - o < d) c = m1[static_cast<size_t>(ii) * nv + iv] * m2[static_cast<size_t>(jj) * nv + jv];
      else if (q - o > d) c = m1[static_cast<size_t>(jj) * nv + jv] * m2[static_cast<size_t>(ii) * nv + iv];
      else c = rd[(static_cast<size_t>(jj) * ik + ii) * nv + iv];
      acc += c;
    }
  }
  rkb[(static_cast<size_t>(kk) * ks * ks + r) * ns * ns + i + static_cast<size_t>(ns) * j] = acc;
}

// ---- a8: pair densities ('s' upper band) ----------------------------------
__global__ void k_pair_dens(int ns, int ks, int nb, const double* P, double* dens) {
  const int npair = nb * (nb + 1) / 2;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= npair * ns * ks) return;
  const int p = gid / (ns * ks), r = gid % (ns * ks), i = r % ns, ith = r / ns;
  int a = 0, q = p;
  while (q >= nb - a) { q -= nb - a; ++a; }
  const int b = a + q;
  const double* pa = P + static_cast<size_t>(a) * ns;
  const double* pb = P + static_cast<size_t>(b) * ns;
  double v = 0.0;
  if (ith == 0) v = pa[i] * pb[i];
  else if (i + ith < ns) v = pa[i] * pb[i + ith] + pa[i + ith] * pb[i];
  dens[gid] = v;
}

// ---- a9 direct: D[kk][p](i,ip) = sum_{j,jp} d_p(j,jp) rkb_k(i,j,ip,jp) -----
// Split over the r2 offset jp: one wavefront per (64 rows i, ip, jp, kk,
// chunk of 16 pairs) sums over j; the partials part[kk][p][jp](i,ip) are
// then added in jp order by k_direct_reduce (deterministic).  The pair
// densities of the current jp are staged in LDS.
// Per (ip, jp) slab of the band and up to 16 pair densities:
//   part[kk][p][jp](i, ip) = sum_j dens[p](j, jp) * rkb[kk](i, j, ip, jp).
// 512 threads: wave w covers rows 64 (w & 3) + 256 blockIdx.z + lane and the
// half w >> 2 of the j range; the densities of the slab's jp sit in LDS
// (broadcast reads), the band streams through coalesced row reads, and the
// two j halves are added in a fixed order through LDS.
__global__ __launch_bounds__(512) void k_direct_partial(int ns, int ks, int npair, const double* __restrict__ rkb,
                                                        const double* __restrict__ dens, double* __restrict__ part) {
  extern __shared__ double dl[];  // [16][ns], then reused as [4][64][16]
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.z * 256 + (wave & 3) * 64 + lane, jh = wave >> 2;
  const int ip = blockIdx.x % ks, jp = blockIdx.x / ks;
  const int nchunk = (npair + 15) / 16;
  const int kk = blockIdx.y / nchunk, p0 = (blockIdx.y % nchunk) * 16;
  const int np = min(16, npair - p0);
  for (int q = threadIdx.x; q < 16 * ns; q += 512) {
    const int p = q / ns, j = q % ns;
    dl[q] = p < np ? dens[static_cast<size_t>(p0 + p) * ns * ks + j + static_cast<size_t>(ns) * jp] : 0.0;
  }
  __syncthreads();
  double acc[16];
#pragma unroll
  for (int p = 0; p < 16; ++p) acc[p] = 0.0;
  const int nj = ns - jp, half = (nj + 1) / 2, j0 = jh * half, j1 = min(nj, j0 + half);
  if (i < ns - ip) {
    const double* Rc = rkb + static_cast<size_t>(kk) * ns * ns * ks * ks + i +
                       static_cast<size_t>(ns) * ns * (ip + static_cast<size_t>(ks) * jp);
    int j = j0;
    // eight band reads in flight per lane
    for (; j + 8 <= j1; j += 8) {
      double r[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) r[u] = Rc[static_cast<size_t>(ns) * (j + u)];
#pragma unroll
      for (int p = 0; p < 16; ++p) {
        const double* d = dl + p * ns + j;
#pragma unroll
        for (int u = 0; u < 8; ++u) acc[p] += d[u] * r[u];
      }
    }
    for (; j < j1; ++j) {
      const double r = Rc[static_cast<size_t>(ns) * j];
#pragma unroll
      for (int p = 0; p < 16; ++p) acc[p] += dl[p * ns + j] * r;
    }
  }
  __syncthreads();  // the densities are no longer read: the space takes the second halves
  double* red = dl + ((wave & 3) * 64 + lane) * 16;
  if (jh == 1) {
#pragma unroll
    for (int p = 0; p < 16; ++p) red[p] = acc[p];
  }
  __syncthreads();
  if (jh == 0 && i < ns) {
    for (int p = 0; p < np; ++p)
      part[((static_cast<size_t>(kk) * npair + p0 + p) * ks + jp) * ns * ks + i + static_cast<size_t>(ns) * ip] =
          acc[p] + red[p];
  }
}

__global__ void k_direct_reduce(int ns, int ks, int npair, int nk, const double* __restrict__ part,
                                double* __restrict__ D) {
  const size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t nsk = static_cast<size_t>(ns) * ks;
  if (q >= nsk * npair * nk) return;
  const size_t kp = q / nsk, e = q % nsk;  // kp = kk*npair + p
  const int i = static_cast<int>(e % ns), ip = static_cast<int>(e / ns);
  double v = 0.0;
  if (i < ns - ip)
    for (int jp = 0; jp < ks; ++jp) v += part[(kp * ks + jp) * nsk + e];
  D[kp * nsk + e] = v;
}

// ---- a9 exchange basis ---------------------------------------------------
// E[kk][a][b](x,y) = sum_{z,w} K(x,z,y,w) p_a(z) q_b(w),
// K(x,z,y,w) = rkb(min(x,z), min(y,w), |x-z|, |y-w|): p on the r1 pair (P1),
// q on the r2 pair (P2; the same orbitals as P1 for the BS bands).  The
// nbt x nbt pairs are covered in tiles of at most NB orbitals on each side:
// this launch makes a in [a0, a0+na), b in [b0, b0+nbc) (na, nbc <= NB;
// orbitals past the counts get weight 0 and are not written).  One
// workgroup per output column y; threads over rows x (coalesced band reads
// and writes).
template <int NB, int KS>
__global__ __launch_bounds__(256) void k_exchange_basis(int ns, int ks_rt, int nbt, int a0, int na, int b0, int nbc,
                                                        const double* __restrict__ rkb,
                                                        const double* __restrict__ P, const double* __restrict__ P2,
                                                        double* __restrict__ E) {
  // KS > 0: the order at compile time, so that the 2 KS - 1 band reads of
  // one dz are issued together (out-of-range neighbours read a clamped
  // element with weight 0)
  const int ks = KS > 0 ? KS : ks_rt;
  // 1-D grid over (kk, y).  Neighbouring columns read the same band slabs;
  // workgroups are dealt to the 8 XCDs round robin, so renumber them to put
  // runs of consecutive columns on one XCD (its L2).
  const int nwg = static_cast<int>(gridDim.x), full = nwg - nwg % 8, hw = static_cast<int>(blockIdx.x);
  const int id = hw < full ? (hw % 8) * (full / 8) + hw / 8 : hw;
  const int y = id % ns, kk = id / ns;
  const size_t band = static_cast<size_t>(ns) * ns * ks * ks;
  const double* R = rkb + kk * band;
  __shared__ double pw[NB][2 * kMaxKs];
  for (int q = threadIdx.x; q < NB * (2 * ks - 1); q += blockDim.x) {
    const int b = q / (2 * ks - 1), dw = q % (2 * ks - 1) - (ks - 1), w = y + dw;
    pw[b][dw + ks - 1] = (b < nbc && w >= 0 && w < ns) ? P2[static_cast<size_t>(b0 + b) * ns + w] : 0.0;
  }
  __syncthreads();
  const size_t s_ip = static_cast<size_t>(ns) * ns, s_jp = s_ip * ks;
  constexpr int ND = KS > 0 ? 2 * KS - 1 : 2 * kMaxKs - 1;
  for (int x = threadIdx.x; x < ns; x += blockDim.x) {
    double e[NB * NB];
#pragma unroll
    for (int q = 0; q < NB * NB; ++q) e[q] = 0.0;
    for (int dz = -(ks - 1); dz <= ks - 1; ++dz) {
      const int z = x + dz;
      if (z < 0 || z >= ns) continue;
      const int i0 = min(x, z), ip0 = dz < 0 ? -dz : dz;
      const double* Ri = R + i0 + s_ip * ip0;
      double tb[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) tb[b] = 0.0;
      if (KS > 0) {
        double r[ND];
#pragma unroll
        for (int t = 0; t < ND; ++t) {
          const int dw = t - (ks - 1), w = min(max(y + dw, 0), ns - 1);
          r[t] = Ri[static_cast<size_t>(ns) * min(y, w) + s_jp * (dw < 0 ? -dw : dw)];
        }
#pragma unroll
        for (int t = 0; t < ND; ++t)
#pragma unroll
          for (int b = 0; b < NB; ++b) tb[b] += r[t] * pw[b][t];
      } else {
        for (int dw = -(ks - 1); dw <= ks - 1; ++dw) {
          const int w = y + dw;
          if (w < 0 || w >= ns) continue;
          const double rr = Ri[static_cast<size_t>(ns) * min(y, w) + s_jp * (dw < 0 ? -dw : dw)];
#pragma unroll
          for (int b = 0; b < NB; ++b) tb[b] += rr * pw[b][dw + ks - 1];
        }
      }
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        const double pa = a < na ? P[static_cast<size_t>(a0 + a) * ns + z] : 0.0;
#pragma unroll
        for (int b = 0; b < NB; ++b) e[a * NB + b] += pa * tb[b];
      }
    }
    const size_t pos = x + static_cast<size_t>(ns) * y;
#pragma unroll
    for (int q = 0; q < NB * NB; ++q) {
      const int a = q / NB, b = q % NB;
      if (a < na && b < nbc)
        E[((static_cast<size_t>(kk) * nbt + a0 + a) * nbt + b0 + b) * ns * ns + pos] = e[q];
    }
  }
}

// Exchange bases with the contraction over w done first, per band row:
//   T_b(i, ip) = sum_dw rkb(i, min(y, y+dw), ip, |dw|) q_b(y + dw)
//   E_ab(x, y) = sum_dz p_a(x + dz) T_b(min(x, x+dz), |dz|)
// One workgroup per (64-row block of x, column y, multipole) and tile of NB
// r2 orbitals b in [b0, b0+nbc): phase 1 has one thread per (row i, ip) of
// the block and its halo (2 KS - 1 band reads in flight per thread), T
// lands in LDS; phase 2 has one thread per (x, a) and runs over all nbt r1
// orbitals in chunks of NB, so the band is read once per b tile.
// Same sums as k_exchange_basis, grouped differently (FP64 rounding aside).
template <int NB, int KS>
__global__ __launch_bounds__(((64 + KS - 1) * KS + 63) / 64 * 64) void k_exchange_basis_t(
    int ns, int nbt, int b0, int nbc, const double* __restrict__ rkb, const double* __restrict__ P,
    const double* __restrict__ P2, double* __restrict__ E) {
  constexpr int H = KS - 1, NR = 64 + H, ND = 2 * KS - 1;
  __shared__ double sT[KS][NB][NR];
  __shared__ double pw[NB][ND];
  __shared__ double sp[NB][64 + 2 * H];
  const int nxb = (ns + 63) / 64;
  const int nwg = static_cast<int>(gridDim.x), full = nwg - nwg % 8, hw = static_cast<int>(blockIdx.x);
  const int id = hw < full ? (hw % 8) * (full / 8) + hw / 8 : hw;
  const int xb = id % nxb, y = (id / nxb) % ns, kk = id / (nxb * ns);
  const int x0 = xb * 64;
  const size_t s_ip = static_cast<size_t>(ns) * ns, s_jp = s_ip * KS;
  const double* R = rkb + static_cast<size_t>(kk) * s_jp * KS;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int q = tid; q < NB * ND; q += nt) {
    const int b = q / ND, t = q % ND, w = y + t - H;
    pw[b][t] = (b < nbc && w >= 0 && w < ns) ? P2[static_cast<size_t>(b0 + b) * ns + w] : 0.0;
  }
  __syncthreads();
  // phase 1: thread (i, ip); rows outside [0, ns) only ever meet weight 0
  if (tid < NR * KS) {
    const int ri = tid % NR, ip = tid / NR, i = min(max(x0 - H + ri, 0), ns - 1);
    const double* Ri = R + i + s_ip * ip;
    double r[ND];
#pragma unroll
    for (int t = 0; t < ND; ++t) {
      const int dw = t - H, w = min(max(y + dw, 0), ns - 1);
      r[t] = Ri[static_cast<size_t>(ns) * min(y, w) + s_jp * (dw < 0 ? -dw : dw)];
    }
#pragma unroll 1
    for (int b = 0; b < NB; ++b) {
      double v = 0.0;
#pragma unroll
      for (int t = 0; t < ND; ++t) v += r[t] * pw[b][t];
      sT[ip][b][ri] = v;
    }
  }
  // phase 2: thread (x, a), r1 orbitals in chunks of NB
  for (int a0 = 0; a0 < nbt; a0 += NB) {
    __syncthreads();  // sT complete; the previous chunk no longer reads sp
    for (int q = tid; q < NB * (64 + 2 * H); q += nt) {
      const int a = q / (64 + 2 * H), z = x0 - H + q % (64 + 2 * H);
      sp[a][q % (64 + 2 * H)] = (a0 + a < nbt && z >= 0 && z < ns) ? P[static_cast<size_t>(a0 + a) * ns + z] : 0.0;
    }
    __syncthreads();
    for (int q = tid; q < 64 * NB; q += nt) {
      const int xl = q % 64, a = q / 64, x = x0 + xl;
      if (x >= ns || a0 + a >= nbt) continue;
      double e[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) e[b] = 0.0;
#pragma unroll 1
      for (int t = 0; t < ND; ++t) {
        const int dz = t - H;
        const double pa = sp[a][xl + t];  // p_a(x + dz), 0 outside [0, ns)
        const int ri = xl + H + (dz < 0 ? dz : 0), ip = dz < 0 ? -dz : dz;
#pragma unroll
        for (int b = 0; b < NB; ++b) e[b] += pa * sT[ip][b][ri];
      }
      const size_t pos = x + static_cast<size_t>(ns) * y;
#pragma unroll
      for (int b = 0; b < NB; ++b)
        if (b < nbc) E[((static_cast<size_t>(kk) * nbt + a0 + a) * nbt + b0 + b) * ns * ns + pos] = e[b];
    }
  }
}

// ---- scalar / vector tables ------------------------------------------------
// one wavefront per (p1, p2, kk): strided partial sums, then a fixed-order
// shuffle-tree reduction (deterministic)
__global__ void k_stab(int ns, int ks, int npair, const double* D, const double* dens, double* St) {
  const int p1 = blockIdx.x / npair, p2 = blockIdx.x % npair, kk = blockIdx.y;
  const double* a = D + (static_cast<size_t>(kk) * npair + p1) * ns * ks;
  const double* b = dens + static_cast<size_t>(p2) * ns * ks;
  double x = 0.0;
  for (int q = threadIdx.x; q < ns * ks; q += 64) {
    const int i = q % ns, ith = q / ns;
    if (i < ns - ith) x += a[q] * b[q];
  }
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  if (threadIdx.x == 0) St[(static_cast<size_t>(kk) * npair + p1) * npair + p2] = x;
}

// y = B v, B symmetric upper band; contributions in the reference's order
// (bav.f90:44-56): diagonal, then per off-diagonal ith the lower then the
// upper neighbour.
__global__ void k_vtab(int ns, int ks, int npair, int nb, const double* D, const double* P, double* V) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int pb = blockIdx.y, kk = blockIdx.z;  // pb = p*nb + b
  if (x >= ns) return;
  const int p = pb / nb, b = pb % nb;
  const double* Bd = D + (static_cast<size_t>(kk) * npair + p) * ns * ks;
  const double* v = P + static_cast<size_t>(b) * ns;
  double y = Bd[x] * v[x];
  for (int ith = 1; ith < ks; ++ith) {
    if (x - ith >= 0) y += Bd[(x - ith) + ns * ith] * v[x - ith];
    if (x + ith < ns) y += Bd[x + ns * ith] * v[x + ith];
  }
  V[(static_cast<size_t>(kk) * npair * nb + pb) * ns + x] = y;
}

__global__ void k_sb_upper(int ns, int ks, const double* sb, double* sbu) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= ns * ks) return;
  const int i = gid % ns, ith = gid / ns;
  // sb(row,col) holds <B_row|B_{row+col-(ks-1)}>; pair (i, i+ith) sits at
  // row i+ith, col ks-1-ith.
  sbu[gid] = (i + ith < ns) ? sb[(i + ith) + static_cast<size_t>(ns) * (ks - 1 - ith)] : 0.0;
}

// ---- accumulation of term lists into the channel blocks ------------------
// One pass over a channel block: out = (base + band terms) + rank-1 terms,
// the rank-1 terms in list order, the band terms summed per block first
// (k_term_prep).  The order matches a product kernel that adds the band term
// in its epilogue and leaves the rank-1 terms to this pass (bands off), so
// both routes give the same bits.  base is what hcc holds (the exchange GEMM result,
// or earlier contributions) when has_base, else 0 and every block is
// written.  Term lists are CSR over all blocks (ptr has nblk+1 entries).
// grid nblk * ceil(ns/kAsmCols) * ceil(ns/(64A)) (block, column tile, row
// tile), block 256: lane -> rows x0+lane+64a (a<A), wave -> kAsmCpw columns
// yt+kAsmCpw*wave+b.
// The band entries (i, ith) the workgroup's columns touch come from the
// block's summed band (k_term_prep) through LDS; rank-1 terms go kAsmR1 at
// a time through LDS (u over the rows, f*w over the columns), a small batch
// that keeps the LDS per workgroup low enough for 4-5 waves per SIMD.
// Per-block preparation of the term pass: the summed band of every block
// with band terms, bsum[blk](i, ith) = sum of c_t * bands[bid_t](i, ith) over
// the block's terms in list order, and the factor of every rank-1 term,
// rf[t] = c_t * (stab[sidx_t] if sidx_t >= 0 else 1).
// grid (nblk, ceil(ns*ks/256)), block 256.
__global__ void k_term_prep(int ns, int ks, const int* __restrict__ bptr, const int* __restrict__ bid,
                            const double* __restrict__ bc, const double* __restrict__ bands, double* __restrict__ bsum,
                            const int* __restrict__ rptr, const int* __restrict__ rs, const double* __restrict__ rc,
                            const double* __restrict__ stab, double* __restrict__ rf) {
  const int blk = blockIdx.x;
  if (blockIdx.y == 0)
    for (int t = rptr[blk] + threadIdx.x; t < rptr[blk + 1]; t += 256) rf[t] = rc[t] * (rs[t] >= 0 ? stab[rs[t]] : 1.0);
  const int b0 = bptr[blk], b1 = bptr[blk + 1];
  const int nsk = ns * ks, e = blockIdx.y * 256 + threadIdx.x;
  if (b0 == b1 || e >= nsk) return;
  double s = 0.0;
  for (int t = b0; t < b1; ++t) s += bc[t] * bands[static_cast<size_t>(bid[t]) * nsk + e];
  bsum[static_cast<size_t>(blk) * nsk + e] = s;
}

// Output map: with qnblk = 0 block blk is hcc[blk] (ns x ns, in place);
// with qnblk > 0 the blocks are the quadrants q = 2a+b of qnblk channel
// blocks of order 2ns (Dirac): base[blk] is read from the quadrant-major
// input and the sum goes to rows a*ns.., columns b*ns.. of output block
// blk % qnblk (leading dimension 2ns); base bit q of qmask says whether
// quadrant q has an input.
constexpr int kAsmCpw = 8;                // columns per wavefront
constexpr int kAsmCols = 4 * kAsmCpw;     // columns per workgroup
constexpr int kAsmR1 = 8;                 // rank-1 terms staged in LDS at a time
// A = row groups of 64 per workgroup (rows per tile 64*A, A = ceil(ns/64)
// up to 4), so that short blocks do not idle lanes.
template <int A>
__global__ __launch_bounds__(256) void k_assemble(int ns, int ks, int has_base, const int* __restrict__ bptr,
                                                  const int* __restrict__ bid, const double* __restrict__ bc,
                                                  const double* __restrict__ bands, const int* __restrict__ rptr,
                                                  const int* __restrict__ ru, const int* __restrict__ rw,
                                                  const double* __restrict__ rf, const double* __restrict__ V,
                                                  double* __restrict__ hcc, const double* __restrict__ base,
                                                  int qnblk, int qmask, const double* __restrict__ bsum,
                                                  const unsigned char* __restrict__ live) {
  // one-dimensional grid over (block, column tile, row tile), block-major;
  // workgroups go to the eight XCDs in turn (id mod 8), so they are numbered
  // XCD by XCD: the tiles of a block and neighbouring blocks share an L2
  const int ny = (ns + kAsmCols - 1) / kAsmCols, nz = (ns + 64 * A - 1) / (64 * A);
  unsigned g = blockIdx.x;
  const unsigned per8 = gridDim.x / 8;
  if (g < per8 * 8) g = (g % 8) * per8 + g / 8;
  const int blk = static_cast<int>(g / static_cast<unsigned>(ny * nz));
  const int tyz = static_cast<int>(g % static_cast<unsigned>(ny * nz)), ty = tyz % ny, tz = tyz / ny;
  const int b0 = bptr[blk], b1 = bptr[blk + 1], r0 = rptr[blk], r1 = rptr[blk + 1];
  // a block without terms keeps its input (has_base), or, overwritten, is
  // skipped when it already holds zeros (live[blk] == 0)
  if (qnblk == 0 && b0 == b1 && r0 == r1 && (has_base || (live && !live[blk]))) return;
  // a quadrant updated in place that has no terms keeps its values
  if (qnblk > 0 && base == hcc && ((qmask >> (blk / qnblk)) & 1) && b0 == b1 && r0 == r1) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int R = 64 * A;  // rows per tile
  const int x0 = tz * R, yt = ty * kAsmCols, y0 = yt + wv * kAsmCpw;
  const double* hin = nullptr;
  double* h;
  size_t ldo = ns, ldb = ns;
  if (qnblk > 0) {
    const int q = blk / qnblk, cb = blk % qnblk;
    ldo = 2 * static_cast<size_t>(ns);
    h = hcc + static_cast<size_t>(cb) * ldo * ldo + (q >> 1) * ns + static_cast<size_t>(q & 1) * ns * ldo;
    if ((qmask >> q) & 1) {
      if (base == hcc) {  // in place: the quadrant already holds its input
        hin = h;
        ldb = ldo;
      } else {
        hin = base + static_cast<size_t>(blk) * ns * ns;
      }
    }
  } else {
    h = hcc + static_cast<size_t>(blk) * ns * ns;
    if (has_base) hin = h;
  }
  __shared__ double su[kAsmR1][R];
  __shared__ double sw[kAsmR1][kAsmCols];
  __shared__ double sband[(kAsmCols + kMaxKs - 1) * kMaxKs];
  __shared__ int tu[64], tw[64];
  __shared__ double tf[64];
  // band entries (i, ith) for i in [yt-ks+1, yt+kAsmCols): sband[(i-ilo)*ks + ith],
  // from the block's summed band (k_term_prep)
  const int ilo = yt - (ks - 1), nent = (kAsmCols + ks - 1) * ks;
  const bool has_band = b1 > b0;
  if (has_band) {
    const double* bs = bsum + static_cast<size_t>(blk) * ns * ks;
    for (int q = threadIdx.x; q < nent; q += 256) {
      const int i = ilo + q / ks, ith = q % ks;
      sband[q] = (i >= 0 && i + ith < ns) ? bs[i + ns * ith] : 0.0;
    }
  }
  // the sum starts from the base values, loaded before the rank-1 vectors
  // are staged so their latency overlaps that: out = (base + band terms) +
  // rank-1 terms
  double acc[A][kAsmCpw];
#pragma unroll
  for (int b = 0; b < kAsmCpw; ++b) {
    const int y = min(y0 + b, ns - 1);
#pragma unroll
    for (int a = 0; a < A; ++a) {
      const int x = min(x0 + lane + 64 * a, ns - 1);
      acc[a][b] = hin ? hin[x + ldb * y] : 0.0;
    }
  }
  // a wave's 64 rows of group a meet the band of its columns only when they
  // lie within ks-1 of them (a wave-uniform test); applied once sband is
  // complete (after a barrier), before the first rank-1 term
  bool banded = false;
  const auto add_band = [&]() {
#pragma unroll
    for (int a = 0; a < A; ++a) {
      if (!(has_band && x0 + 64 * a + 63 > y0 - ks && x0 + 64 * a < y0 + kAsmCpw + ks - 1)) continue;
#pragma unroll
      for (int b = 0; b < kAsmCpw; ++b) {
        const int x = x0 + lane + 64 * a, y = y0 + b, d = x - y;
        if (x < ns && y < ns && d > -ks && d < ks) acc[a][b] += sband[(min(x, y) - ilo) * ks + (d < 0 ? -d : d)];
      }
    }
    banded = true;
  };
  // rank-1 terms: metadata and factors of 64 terms at a time into LDS (one
  // round of loads), then their vectors 16 at a time
  const int nrow = min(R, ns - x0);
  for (int tc = r0; tc < r1; tc += 64) {
    const int nc = min(64, r1 - tc);
    __syncthreads();
    if (threadIdx.x < nc) {
      const int t = tc + threadIdx.x;
      tf[threadIdx.x] = rf[t];
      tu[threadIdx.x] = ru[t];
      tw[threadIdx.x] = rw[t];
    }
    for (int tb = 0; tb < nc; tb += kAsmR1) {
      const int nt = min(kAsmR1, nc - tb);
      __syncthreads();
      for (int q = threadIdx.x; q < nt * R; q += 256) {
        const int r = q / R, x = q % R;
        if (x < nrow) su[r][x] = V[static_cast<size_t>(tu[tb + r]) * ns + x0 + x];
      }
      for (int q = threadIdx.x; q < nt * kAsmCols; q += 256) {
        const int r = q / kAsmCols, col = q % kAsmCols, y = yt + col;
        sw[r][col] = y < ns ? tf[tb + r] * V[static_cast<size_t>(tw[tb + r]) * ns + y] : 0.0;
      }
      __syncthreads();
      if (!banded) add_band();
      for (int r = 0; r < nt; ++r) {
        double u[A];
#pragma unroll
        for (int a = 0; a < A; ++a) u[a] = su[r][lane + 64 * a];
#pragma unroll
        for (int b = 0; b < kAsmCpw; ++b) {
          const double w = sw[r][wv * kAsmCpw + b];
#pragma unroll
          for (int a = 0; a < A; ++a) acc[a][b] += u[a] * w;
        }
      }
    }
  }
  if (!banded) {
    __syncthreads();  // sband complete
    add_band();
  }
#pragma unroll
  for (int b = 0; b < kAsmCpw; ++b) {
    const int y = y0 + b;
    if (y >= ns) continue;
    double* hy = h + ldo * y;
#pragma unroll
    for (int a = 0; a < A; ++a) {
      const int x = x0 + lane + 64 * a;
      if (x < ns) hy[x] = acc[a][b];
    }
  }
}

// Channel-perturber vectors: hcb(:, q) = base + sum of c * f * V(:, id), list
// order; CSR over all nq = nch*ncp positions.  One wavefront per vector,
// four vectors per workgroup: the lanes load the factors and vector ids of
// up to 64 terms at once and broadcast them; each lane keeps the elements
// x0 + lane + 64j (j < 4) of a 256-element chunk in registers.
// grid (ceil(nq/4)), block 256.
__global__ __launch_bounds__(256) void k_vec_assemble(int ns, int nq, int has_base, const int* __restrict__ ptr,
                                                      const int* __restrict__ vid, const int* __restrict__ sidx,
                                                      const double* __restrict__ c, const double* __restrict__ stab,
                                                      const double* __restrict__ V, double* __restrict__ hcb) {
  const int q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (q >= nq) return;
  const int t0 = ptr[q], t1 = ptr[q + 1];
  if (has_base && t0 == t1) return;
  double* p = hcb + static_cast<size_t>(q) * ns;
  for (int x0 = 0; x0 < ns; x0 += 256) {
    double acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int x = x0 + lane + 64 * j;
      acc[j] = (has_base && x < ns) ? p[x] : 0.0;
    }
    for (int tc = t0; tc < t1; tc += 64) {
      const int n = min(64, t1 - tc);
      double fl = 0.0;
      int vl = 0;
      if (lane < n) {
        const int t = tc + lane;
        fl = c[t] * (sidx[t] >= 0 ? stab[sidx[t]] : 1.0);
        vl = vid[t];
      }
      for (int tt = 0; tt < n; ++tt) {
        const double f = __shfl(fl, tt, 64);
        const double* v = V + static_cast<size_t>(__shfl(vl, tt, 64)) * ns;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int x = x0 + lane + 64 * j;
          if (x < ns) acc[j] += f * v[x];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int x = x0 + lane + 64 * j;
      if (x < ns) p[x] = acc[j];
    }
  }
}

// Diagonal blocks -> (C + C^T)/2 through 32x32 LDS tiles: the workgroup of
// tile pair (ti >= tj) reads both tiles, writes both.
// grid (nt*(nt+1)/2, nch), block 256 (32 x 8)
__global__ void k_symmetrize_diag(int ns, int nch, double* hcc) {
  const int ich = blockIdx.y;  // 0-based
  const size_t blk = static_cast<size_t>(ich + 1) * (ich + 2) / 2 - 1;
  double* h = hcc + blk * ns * ns;
  int p = blockIdx.x, ti = 0;
  while (p > ti) { p -= ti + 1; ++ti; }
  const int tj = p;  // tj <= ti
  __shared__ double A[32][33], B[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int xa = ti * 32 + tx, ya = tj * 32 + r;  // tile (ti,tj): rows ti, cols tj
    const int xb = tj * 32 + tx, yb = ti * 32 + r;  // tile (tj,ti)
    A[r][tx] = (xa < ns && ya < ns) ? h[xa + static_cast<size_t>(ns) * ya] : 0.0;
    B[r][tx] = (xb < ns && yb < ns) ? h[xb + static_cast<size_t>(ns) * yb] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    // element (x = ti*32+tx, y = tj*32+r) pairs with (y, x) = B[tx][r]
    const double s = (A[r][tx] + B[tx][r]) / 2.0;
    const double sb = (B[r][tx] + A[tx][r]) / 2.0;
    const int xa = ti * 32 + tx, ya = tj * 32 + r;
    const int xb = tj * 32 + tx, yb = ti * 32 + r;
    if (xa < ns && ya < ns) h[xa + static_cast<size_t>(ns) * ya] = s;
    if (ti != tj && xb < ns && yb < ns) h[xb + static_cast<size_t>(ns) * yb] = sb;
  }
}

// ---- comparison of two results (per-job error statement) ------------------
// partial[blockIdx.x] = max |x| over a grid-stride slice; block 256
__global__ void k_absmax(size_t n, const double* __restrict__ x, double* __restrict__ partial) {
  __shared__ double red[4];
  double m = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256)
    m = fmax(m, fabs(x[i]));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// partial[blockIdx.x] = max |a - b| / max(|b|, floor) over a slice
__global__ void k_reldiff(size_t n, const double* __restrict__ a, const double* __restrict__ b, double floor,
                          double* __restrict__ partial) {
  __shared__ double red[4];
  double m = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256)
    m = fmax(m, fabs(a[i] - b[i]) / fmax(fabs(b[i]), floor));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

__global__ void k_scale(size_t n, double a, double* x) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= a;
}

__global__ void k_fill(size_t n, double v, double* x) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ void k_fill_int(size_t n, int v, int* x) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// leaf convol icase 1 (jcase 1) / 2 (jcase 2): a(i,ip) = sum d(j,jp) * R
__global__ void k_convol_direct(int ns, int ks, int jcase, const double* rkb, const double* d, double* a) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= ns * ks) return;
  const int i = gid % ns, ip = gid / ns;
  if (i >= ns - ip) { a[gid] = 0.0; return; }
  double c = 0.0;
  for (int jp = 0; jp < ks; ++jp)
    for (int j = 0; j < ns - jp; ++j) {
      const size_t idx = jcase == 1
          ? i + static_cast<size_t>(ns) * (j + static_cast<size_t>(ns) * (ip + static_cast<size_t>(ks) * jp))
          : j + static_cast<size_t>(ns) * (i + static_cast<size_t>(ns) * (jp + static_cast<size_t>(ks) * ip));
      c += d[j + static_cast<size_t>(ns) * jp] * rkb[idx];
    }
  a[gid] = c;
}

// leaf convol icase 3/4: a(x,y) = sum_{z,w} K(x,z,y,w) d(z,w)
__global__ void k_convol_exchange(int ns, int ks, const double* rkb, const double* d, double* a) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= ns * ns) return;
  const int x = gid % ns, y = gid / ns;
  const size_t s_ip = static_cast<size_t>(ns) * ns, s_jp = s_ip * ks;
  double acc = 0.0;
  for (int dz = -(ks - 1); dz <= ks - 1; ++dz) {
    const int z = x + dz;
    if (z < 0 || z >= ns) continue;
    for (int dw = -(ks - 1); dw <= ks - 1; ++dw) {
      const int w = y + dw;
      if (w < 0 || w >= ns) continue;
      const double r = rkb[min(x, z) + static_cast<size_t>(ns) * min(y, w) + s_ip * (dz < 0 ? -dz : dz) +
                           s_jp * (dw < 0 ? -dw : dw)];
      acc += r * d[z + static_cast<size_t>(ns) * w];
    }
  }
  a[gid] = acc;
}

unsigned nblocks(size_t n, unsigned b) { return static_cast<unsigned>((n + b - 1) / b); }

}  // namespace

void launch_moments(const DevGrid& g, int k0, int nk, double* mom, double* rkd, hipStream_t s) {
  const int ik = g.ks * (g.ks + 1) / 2;
  hipLaunchKernelGGL(k_moments, dim3(nblocks(static_cast<size_t>(g.nv) * ik, 256), nk, 2), dim3(256), 0, s, g,
                     k0, mom);
  const size_t shm = sizeof(double) * (g.ks * g.ks * g.ks + g.ks * g.ks + 2 * g.ks * ik + ik * ik);
  hipLaunchKernelGGL(k_diag_cells, dim3(g.nv, nk), dim3(256), shm, s, g, k0, rkd);
  if (g.me > 0) {
    hipLaunchKernelGGL(k_moments_scale, dim3(nk * 2), dim3(ik), 0, s, g, k0, mom);
    hipLaunchKernelGGL(k_diag_scale, dim3(nk), dim3(256), 0, s, g, rkd);
  }
}

void launch_cells_transpose(int nv, int w, int nb, const double* in, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_cells_transpose, dim3((w + 31) / 32, (nv + 31) / 32, nb), dim3(256), 0, s, nv, w, in, out);
}

void launch_band(const DevGrid& g, int nk, const double* mom, const double* rkd, double* msum, double* momT,
                 double* rdT, double* rkb, hipStream_t s) {
  const int ik = g.ks * (g.ks + 1) / 2;
  hipLaunchKernelGGL(k_pair_sums, dim3(nblocks(g.ns, 256), g.ks, nk * 2), dim3(256), 0, s, g.ns, g.ks, g.nv, mom,
                     msum);
  hipLaunchKernelGGL(k_cells_transpose, dim3((ik + 31) / 32, (g.nv + 31) / 32, nk * 2), dim3(256), 0, s, g.nv, ik,
                     mom, momT);
  hipLaunchKernelGGL(k_cells_transpose, dim3((ik * ik + 31) / 32, (g.nv + 31) / 32, nk), dim3(256), 0, s, g.nv,
                     ik * ik, rkd, rdT);
  hipLaunchKernelGGL(k_band_far, dim3(nblocks(static_cast<size_t>(g.ns) * g.ns, 256), g.ks * g.ks, nk), dim3(256), 0, s,
                     g.ns, g.ks, msum, rkb);
  hipLaunchKernelGGL(k_band_near, dim3((g.ns + 63) / 64, 2 * g.ks - 1, g.ks * g.ks * nk), dim3(64), 0, s, g.ns, g.ks,
                     g.nv, momT, rdT, rkb);
}

void launch_pair_dens(int ns, int ks, int nb, const double* P, double* dens, hipStream_t s) {
  const size_t n = static_cast<size_t>(nb) * (nb + 1) / 2 * ns * ks;
  hipLaunchKernelGGL(k_pair_dens, dim3(nblocks(n, 256)), dim3(256), 0, s, ns, ks, nb, P, dens);
}

void launch_direct_basis(int ns, int ks, int nk, int npair, const double* rkb, const double* dens, double* part,
                         double* D, hipStream_t s) {
  const int nchunk = (npair + 15) / 16;
  const size_t lds = sizeof(double) * std::max(16 * ns, 4 * 64 * 16);
  hipLaunchKernelGGL(k_direct_partial, dim3(ks * ks, nk * nchunk, (ns + 255) / 256), dim3(512), lds, s, ns, ks,
                     npair, rkb, dens, part);
  const size_t n = static_cast<size_t>(ns) * ks * npair * nk;
  hipLaunchKernelGGL(k_direct_reduce, dim3(nblocks(n, 256)), dim3(256), 0, s, ns, ks, npair, nk, part, D);
}

void launch_exchange_basis(int ns, int ks, int nk, int nb, const double* rkb, const double* P, const double* P2,
                           double* E, hipStream_t s) {
  if (nb <= 0) return;
  const dim3 grid(ns * nk), block(ns <= 64 ? 64 : (ns <= 128 ? 128 : 256));
  const dim3 gridt(static_cast<unsigned>((ns + 63) / 64) * ns * nk);
  // up to 8 orbitals a tile exactly; more in tiles of 8
  const int tile = nb <= 8 ? nb : 8;
  for (int b0 = 0; b0 < nb; b0 += tile) {
    const int nbc = std::min(tile, nb - b0);
    switch (tile) {
#define NB_CASE(N)                                                                                        \
  case N:                                                                                                 \
    if (ks == 8)                                                                                          \
      hipLaunchKernelGGL((k_exchange_basis_t<N, 8>), gridt, dim3(((64 + 7) * 8 + 63) / 64 * 64), 0, s, ns, \
                         nb, b0, nbc, rkb, P, P2, E);                                                     \
    else if (ks == 9)                                                                                     \
      hipLaunchKernelGGL((k_exchange_basis_t<N, 9>), gridt, dim3(((64 + 8) * 9 + 63) / 64 * 64), 0, s, ns, \
                         nb, b0, nbc, rkb, P, P2, E);                                                     \
    else                                                                                                  \
      for (int a0 = 0; a0 < nb; a0 += tile)                                                               \
        hipLaunchKernelGGL((k_exchange_basis<N, 0>), grid, block, 0, s, ns, ks, nb, a0,                   \
                           std::min(tile, nb - a0), b0, nbc, rkb, P, P2, E);                              \
    break;
      NB_CASE(1) NB_CASE(2) NB_CASE(3) NB_CASE(4) NB_CASE(5) NB_CASE(6) NB_CASE(7) NB_CASE(8)
#undef NB_CASE
      default: break;
    }
  }
}

void launch_stab(int ns, int ks, int nk, int npair, const double* D, const double* dens, double* St, hipStream_t s) {
  hipLaunchKernelGGL(k_stab, dim3(npair * npair, nk), dim3(64), 0, s, ns, ks, npair, D, dens, St);
}

void launch_vtab(int ns, int ks, int nk, int npair, int nb, const double* D, const double* P, double* V,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_vtab, dim3(nblocks(ns, 256), npair * nb, nk), dim3(256), 0, s, ns, ks, npair, nb, D, P, V);
}

void launch_sb_upper(int ns, int ks, const double* sb, double* sbu, hipStream_t s) {
  hipLaunchKernelGGL(k_sb_upper, dim3(nblocks(static_cast<size_t>(ns) * ks, 256)), dim3(256), 0, s, ns, ks, sb, sbu);
}

void launch_term_prep(int ns, int ks, int nblk, const int* bptr, const int* bid, const double* bc,
                      const double* bands, double* bsum, const int* rptr, const int* rs, const double* rc,
                      const double* stab, double* rf, hipStream_t s) {
  if (nblk <= 0) return;
  hipLaunchKernelGGL(k_term_prep, dim3(nblk, (ns * ks + 255) / 256), dim3(256), 0, s, ns, ks, bptr, bid, bc, bands,
                     bsum, rptr, rs, rc, stab, rf);
}

void launch_assemble(int ns, int ks, int nblk, int has_base, const int* bptr, const int* bid, const double* bc,
                     const double* bands, const int* rptr, const int* ru, const int* rw, const int* rs,
                     const double* rc, const double* stab, const double* V, double* hcc, double* bsum,
                     double* rf, hipStream_t s, const double* base, int qnblk, int qmask, const unsigned char* live,
                     bool prep) {
  if (nblk <= 0) return;
  if (prep)
    hipLaunchKernelGGL(k_term_prep, dim3(nblk, (ns * ks + 255) / 256), dim3(256), 0, s, ns, ks, bptr, bid, bc, bands,
                       bsum, rptr, rs, rc, stab, rf);
  const int A = ns <= 64 ? 1 : (ns <= 128 ? 2 : (ns <= 192 ? 3 : 4));
  const dim3 grid(static_cast<unsigned>(nblk) * ((ns + kAsmCols - 1) / kAsmCols) * ((ns + 64 * A - 1) / (64 * A)));
  switch (A) {
#define ASM_CASE(N)                                                                                             \
  case N:                                                                                                       \
    hipLaunchKernelGGL(k_assemble<N>, grid, dim3(256), 0, s, ns, ks, has_base, bptr, bid, bc, bands, rptr, ru, rw, \
                       rf, V, hcc, base, qnblk, qmask, bsum, live);                                             \
    break;
    ASM_CASE(1) ASM_CASE(2) ASM_CASE(3) ASM_CASE(4)
#undef ASM_CASE
  }
}

void launch_vec_assemble(int ns, int nq, int has_base, const int* ptr, const int* vid, const int* sidx,
                         const double* c, const double* stab, const double* V, double* hcb, hipStream_t s) {
  if (nq <= 0) return;
  hipLaunchKernelGGL(k_vec_assemble, dim3((nq + 3) / 4), dim3(256), 0, s, ns, nq, has_base, ptr, vid, sidx, c, stab,
                     V, hcb);
}

// ---- one-electron operator with the core (Gen_Lval, L_data) ------------
// INT_core (src/bsr/BSR_MAT3/l_core.f90:173-257) in full storage: the core
// exchange sum (k outer, core orbital inner, the order the reference adds
// them), then in the band region the L band and the summed direct core
// potential, both read at the lower element (max, min) and written to both
// halves.  grid (ceil(ns*ns/256)), block 256.
__global__ __launch_bounds__(256) void k_lcore(int ns, int ks, int nt, const int* __restrict__ eid,
                                               const double* __restrict__ ec, const double* __restrict__ E, int nd,
                                               const int* __restrict__ did, const double* __restrict__ dc,
                                               const double* __restrict__ bands, const double* __restrict__ hl,
                                               double* __restrict__ out) {
  const size_t nss = static_cast<size_t>(ns) * ns, e = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= nss) return;
  const int x = static_cast<int>(e % ns), y = static_cast<int>(e / ns);
  const int i = max(x, y), j = min(x, y), d = i - j;
  const bool band = d < ks;
  const size_t src = band ? i + static_cast<size_t>(ns) * j : e;
  double v = 0.0;
  for (int t = 0; t < nt; ++t) v += ec[t] * E[static_cast<size_t>(eid[t]) * nss + src];
  if (band) {
    v += hl[i + static_cast<size_t>(ns) * (ks - 1 - d)];
    if (nd > 0) {
      double h = 0.0;
      for (int u = 0; u < nd; ++u) h += dc[u] * bands[static_cast<size_t>(did[u]) * ns * ks + j + static_cast<size_t>(ns) * d];
      v += h;
    }
  }
  out[e] = v;
}

// L_vec(m, o) = sum_k hl(k, m) P(k, pcol[o]), k ascending (l_core.f90:143).
// grid (ceil(ns/256), n), block 256.
__global__ __launch_bounds__(256) void k_lvec(int ns, const int* __restrict__ pcol, const double* __restrict__ P,
                                              const double* __restrict__ hl, double* __restrict__ out) {
  const int o = blockIdx.y, m = blockIdx.x * 256 + threadIdx.x;
  if (m >= ns) return;
  const double* p = P + static_cast<size_t>(ns) * pcol[o];
  const double* h = hl + static_cast<size_t>(ns) * m;
  double s = 0.0;
  for (int k = 0; k < ns; ++k) s += h[k] * p[k];
  out[m + static_cast<size_t>(ns) * o] = s;
}

// Dense L terms (UPDATE_HX 'x' of C * hl_core, mod_intmat.f90:69-127): each
// listed block adds its terms in list order, c * L(l) or its transpose.
// grid (ceil(ns*ns/256), nlisted), block 256.
__global__ __launch_bounds__(256) void k_dense_terms(int ns, const int* __restrict__ ptr, const int* __restrict__ blk,
                                                     const int* __restrict__ lid, const int* __restrict__ tr,
                                                     const double* __restrict__ c, const double* __restrict__ hlfull,
                                                     double* __restrict__ hcc) {
  const size_t nss = static_cast<size_t>(ns) * ns, e = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= nss) return;
  const int b = blockIdx.y;
  const size_t et = (e / ns) + static_cast<size_t>(ns) * (e % ns);
  double* h = hcc + static_cast<size_t>(blk[b]) * nss;
  double v = h[e];
  for (int t = ptr[b]; t < ptr[b + 1]; ++t) {
    const double* L = hlfull + static_cast<size_t>(lid[t]) * nss;
    v += c[t] * L[tr[t] ? et : e];
  }
  h[e] = v;
}

// ---- BS_ORTH (src/bsr/BSR_MAT3/bs_orth.f90:2-92) -------------------------
// Channel block (i, j), i >= j: H <- (1 - B Z_i Z_i^T) H (1 - Z_j Z_j^T B)
// = H - BZ_i (T - U) - V BZ_j^T, with T = Z_i^T H, V = H Z_j,
// U = (Z_i^T V) BZ_j^T; Z_c holds the nz[c] bound orbitals channel c must be
// orthogonal to (columns zoff[c].. of Z, and BZ = B Z).  T and V are linear
// in the conditions: both are formed from the unmodified block in chunks of
// kOrthMax conditions, so any number of conditions is handled.
// One workgroup of 4 waves per block.  Pass 1 reads the block twice (the
// second time mostly from cache): V with lanes over rows, T with lanes over
// columns, each sum in index order in one register; no cross-lane sums.
// Pass 2 writes the block element by element.
// Work arrays (doubles): V[nz][ns], TU[nz][ns], W[nz][nz], in LDS when they
// fit (GLB = false), else in a global scratch slice per workgroup (blocks
// blk0 .. blk0 + gridDim.x - 1 in one launch).
constexpr int kOrthMax = 16;

__device__ inline void blk_pair(int blk, int& i, int& j) {
  int r = static_cast<int>((sqrt(8.0 * blk + 1.0) + 1.0) * 0.5);
  while (r * (r - 1) / 2 > blk) --r;
  while ((r + 1) * r / 2 <= blk) ++r;
  i = r;
  j = blk - r * (r - 1) / 2 + 1;
}

template <bool GLB>
__global__ __launch_bounds__(256) void k_bs_orth(int ns, int nzmax, int blk0, const int* __restrict__ zoff,
                                                 const int* __restrict__ nz, const double* __restrict__ Z,
                                                 const double* __restrict__ BZ, double* __restrict__ hcc,
                                                 double* __restrict__ scratch) {
  extern __shared__ double lds[];
  const int blk = blk0 + static_cast<int>(blockIdx.x);
  int ci, cj;
  blk_pair(blk, ci, cj);  // 1-based channels, ci >= cj
  const int ni = nz[ci - 1], nj = nz[cj - 1];
  if (ni == 0 && nj == 0) return;
  const double* Zi = Z + static_cast<size_t>(zoff[ci - 1]) * ns;
  const double* Zj = Z + static_cast<size_t>(zoff[cj - 1]) * ns;
  const double* BZi = BZ + static_cast<size_t>(zoff[ci - 1]) * ns;
  const double* BZj = BZ + static_cast<size_t>(zoff[cj - 1]) * ns;
  const size_t per = 2 * static_cast<size_t>(nzmax) * ns + static_cast<size_t>(nzmax) * nzmax;
  double* V = GLB ? scratch + per * blockIdx.x : lds;    // [nzmax][ns]
  double* TU = V + static_cast<size_t>(nzmax) * ns;      // [nzmax][ns]
  double* W = TU + static_cast<size_t>(nzmax) * ns;      // [nzmax][nzmax]
  double* h = hcc + static_cast<size_t>(blk) * ns * ns;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // V = H Z_j: lanes over rows (coalesced), eight columns per round
  for (int q0 = 0; q0 < nj; q0 += kOrthMax) {
    const int nq = min(kOrthMax, nj - q0);
    for (int x0 = 64 * wv; x0 < ns; x0 += 256) {
      const int x = x0 + lane;
      const bool in = x < ns;
      double vacc[kOrthMax];
#pragma unroll
      for (int q = 0; q < kOrthMax; ++q) vacc[q] = 0.0;
      for (int y0 = 0; y0 < ns; y0 += 8) {
        double hv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) hv[u] = (in && y0 + u < ns) ? h[x + static_cast<size_t>(ns) * (y0 + u)] : 0.0;
        const int nu = min(8, ns - y0);
        for (int u = 0; u < nu; ++u)
#pragma unroll
          for (int q = 0; q < kOrthMax; ++q)
            if (q < nq) vacc[q] += hv[u] * Zj[static_cast<size_t>(q0 + q) * ns + y0 + u];
      }
      if (in)
        for (int q = 0; q < nq; ++q) V[static_cast<size_t>(q0 + q) * ns + x] = vacc[q];
    }
  }
  // T = Z_i^T H: lanes over columns (each lane walks its column; a cache
  // line serves eight rows), eight rows per round
  for (int q0 = 0; q0 < ni; q0 += kOrthMax) {
    const int nq = min(kOrthMax, ni - q0);
    for (int yc = 64 * wv; yc < ns; yc += 256) {
      const int y = yc + lane;
      const bool in = y < ns;
      double tacc[kOrthMax];
#pragma unroll
      for (int q = 0; q < kOrthMax; ++q) tacc[q] = 0.0;
      for (int x0 = 0; x0 < ns; x0 += 8) {
        double hv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) hv[u] = (in && x0 + u < ns) ? h[x0 + u + static_cast<size_t>(ns) * y] : 0.0;
        const int nu = min(8, ns - x0);
        for (int u = 0; u < nu; ++u)
#pragma unroll
          for (int q = 0; q < kOrthMax; ++q)
            if (q < nq) tacc[q] += Zi[static_cast<size_t>(q0 + q) * ns + x0 + u] * hv[u];
      }
      if (in)
        for (int q = 0; q < nq; ++q) TU[static_cast<size_t>(q0 + q) * ns + y] = tacc[q];
    }
  }
  __syncthreads();
  // W = Z_i^T V
  for (int t = threadIdx.x; t < ni * nj; t += 256) {
    const int a = t / nj, b = t % nj;
    double s = 0.0;
    for (int x = 0; x < ns; ++x) s += Zi[static_cast<size_t>(a) * ns + x] * V[static_cast<size_t>(b) * ns + x];
    W[a * nzmax + b] = s;
  }
  __syncthreads();
  // TU = T - W BZ_j^T
  for (int q = threadIdx.x; q < ni * ns; q += 256) {
    const int a = q / ns, y = q % ns;
    double u = 0.0;
    for (int b = 0; b < nj; ++b) u += W[a * nzmax + b] * BZj[static_cast<size_t>(b) * ns + y];
    TU[static_cast<size_t>(a) * ns + y] -= u;
  }
  __syncthreads();
  const size_t nss = static_cast<size_t>(ns) * ns;
  for (size_t e0 = threadIdx.x; e0 < nss; e0 += 4 * 256) {
    double hv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) hv[u] = e0 + 256 * u < nss ? h[e0 + 256 * u] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t e = e0 + 256 * u;
      if (e >= nss) break;
      const int x = static_cast<int>(e % ns), y = static_cast<int>(e / ns);
      double v = hv[u];
      for (int a = 0; a < ni; ++a) v -= BZi[static_cast<size_t>(a) * ns + x] * TU[static_cast<size_t>(a) * ns + y];
      for (int b = 0; b < nj; ++b) v -= V[static_cast<size_t>(b) * ns + x] * BZj[static_cast<size_t>(b) * ns + y];
      h[e] = v;
    }
  }
}

// Channel-perturber vectors: hcb(:, ich, ip) <- (1 - B Z Z^T) v, one wave
// per vector; the projections s_a = z_a . v of a wave sit in LDS
// (4 x nzmax doubles).  grid (ceil(nch*ncp/4)), block 256.
__global__ __launch_bounds__(256) void k_bs_orth_vec(int ns, int nch, int nq, int nzmax, const int* __restrict__ zoff,
                                                     const int* __restrict__ nz, const double* __restrict__ Z,
                                                     const double* __restrict__ BZ, double* __restrict__ hcb) {
  extern __shared__ double sl[];
  const int q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (q >= nq) return;
  const int ich = q % nch, n = nz[ich];
  if (n == 0) return;
  double* v = hcb + static_cast<size_t>(q) * ns;
  const double* Zc = Z + static_cast<size_t>(zoff[ich]) * ns;
  const double* BZc = BZ + static_cast<size_t>(zoff[ich]) * ns;
  double* s = sl + static_cast<size_t>(threadIdx.x >> 6) * nzmax;
  for (int a = 0; a < n; ++a) {
    double t = 0.0;
    for (int x = lane; x < ns; x += 64) t += Zc[static_cast<size_t>(a) * ns + x] * v[x];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (lane == 0) s[a] = t;
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  for (int x = lane; x < ns; x += 64) {
    double u = v[x];
    for (int a = 0; a < n; ++a) u -= BZc[static_cast<size_t>(a) * ns + x] * s[a];
    v[x] = u;
  }
}

size_t bs_orth_lds_bytes(int ns, int nzmax) {
  return (2 * static_cast<size_t>(nzmax) * ns + static_cast<size_t>(nzmax) * nzmax) * sizeof(double);
}

size_t bs_orth_scratch_doubles(int ns, int nblk, int nzmax) {
  if (bs_orth_lds_bytes(ns, nzmax) <= 64 * 1024) return 0;
  const size_t per = bs_orth_lds_bytes(ns, nzmax) / sizeof(double);
  const size_t batch = std::max<size_t>(1, std::min<size_t>(nblk, (size_t(32) << 20) / per));
  return batch * per;
}

int launch_bs_orth(int ns, int nblk, int nch, int ncp, int nzmax, const int* zoff, const int* nz, const double* Z,
                   const double* BZ, double* hcc, double* hcb, double* scratch, size_t nscratch, hipStream_t s) {
  if (nzmax < 1) return -1;
  const size_t lds = bs_orth_lds_bytes(ns, nzmax);
  if (nblk > 0) {
    if (lds <= 64 * 1024) {
      hipLaunchKernelGGL(k_bs_orth<false>, dim3(nblk), dim3(256), lds, s, ns, nzmax, 0, zoff, nz, Z, BZ, hcc,
                         static_cast<double*>(nullptr));
    } else {
      // work arrays in global scratch, blocks in batches that fit it
      const size_t per = lds / sizeof(double);
      const int batch = static_cast<int>(std::min<size_t>(nblk, nscratch / per));
      if (batch < 1 || !scratch) return -2;
      for (int b0 = 0; b0 < nblk; b0 += batch)
        hipLaunchKernelGGL(k_bs_orth<true>, dim3(std::min(batch, nblk - b0)), dim3(256), 0, s, ns, nzmax, b0, zoff,
                           nz, Z, BZ, hcc, scratch);
    }
  }
  const int nq = nch * ncp;
  if (nq > 0)
    hipLaunchKernelGGL(k_bs_orth_vec, dim3((nq + 3) / 4), dim3(256), 4 * nzmax * sizeof(double), s, ns, nch, nq,
                       nzmax, zoff, nz, Z, BZ, hcb);
  return 0;
}

void launch_lcore(int ns, int ks, int nt, const int* eid, const double* ec, const double* E, int nd,
                  const int* did, const double* dc, const double* bands, const double* hl, double* out,
                  hipStream_t s) {
  const size_t nss = static_cast<size_t>(ns) * ns;
  hipLaunchKernelGGL(k_lcore, dim3(nblocks(nss, 256)), dim3(256), 0, s, ns, ks, nt, eid, ec, E, nd, did, dc, bands,
                     hl, out);
}

void launch_lvec(int ns, int n, const int* pcol, const double* P, const double* hl, double* out, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lvec, dim3((ns + 255) / 256, n), dim3(256), 0, s, ns, pcol, P, hl, out);
}

void launch_dense_terms(int ns, int nlisted, const int* ptr, const int* blk, const int* lid, const int* tr,
                        const double* c, const double* hlfull, double* hcc, hipStream_t s) {
  if (nlisted <= 0) return;
  const size_t nss = static_cast<size_t>(ns) * ns;
  hipLaunchKernelGGL(k_dense_terms, dim3(nblocks(nss, 256), nlisted), dim3(256), 0, s, ns, ptr, blk, lid, tr, c,
                     hlfull, hcc);
}

void launch_symmetrize_diag(int ns, int nch, double* hcc, hipStream_t s) {
  if (nch <= 0) return;
  const int nt = (ns + 31) / 32;
  hipLaunchKernelGGL(k_symmetrize_diag, dim3(nt * (nt + 1) / 2, nch), dim3(256), 0, s, ns, nch, hcc);
}

void launch_absmax(size_t n, const double* x, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_absmax, dim3(kReduceBlocks), dim3(256), 0, s, n, x, partial);
}

void launch_reldiff(size_t n, const double* a, const double* b, double floor, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_reldiff, dim3(kReduceBlocks), dim3(256), 0, s, n, a, b, floor, partial);
}

void launch_scale(size_t n, double a, double* x, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_scale, dim3(nblocks(n, 256)), dim3(256), 0, s, n, a, x);
}

void launch_fill(size_t n, double v, double* x, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill, dim3(nblocks(n, 256)), dim3(256), 0, s, n, v, x);
}

void launch_fill_int(size_t n, int v, int* x, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_int, dim3(nblocks(n, 256)), dim3(256), 0, s, n, v, x);
}

void launch_convol_direct(int ns, int ks, int jcase, const double* rkb, const double* d, double* a, hipStream_t s) {
  hipLaunchKernelGGL(k_convol_direct, dim3(nblocks(static_cast<size_t>(ns) * ks, 256)), dim3(256), 0, s, ns, ks,
                     jcase, rkb, d, a);
}

void launch_convol_exchange(int ns, int ks, const double* rkb, const double* d, double* a, hipStream_t s) {
  hipLaunchKernelGGL(k_convol_exchange, dim3(nblocks(static_cast<size_t>(ns) * ns, 256)), dim3(256), 0, s, ns, ks,
                     rkb, d, a);
}

}  // namespace bsrg
