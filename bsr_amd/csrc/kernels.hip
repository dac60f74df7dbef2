// FP64 kernels of the R^k radial core for gfx950 (64-wide wavefronts).
//
// Layouts (0-based, column-major as the Fortran interface):
//   rkb_k(i,j,ip,jp) at i + ns*(j + ns*(ip + ks*jp))   pair (i,i+ip) on r1,
//                                                       pair (j,j+jp) on r2
//   upper band d(i,ith) at i + ns*ith                   pair (i,i+ith)
//   hcc block b (x,y)  at b*ns*ns + x + ns*y
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

// One thread per element, i fastest (coalesced writes).  When every cell of
// the r1 pair lies below every cell of the r2 pair (or above), the cell sum
// factorises: rkb = (sum of r^k moments) * (sum of r^-(k+1) moments).
// Otherwise the cells are summed in the reference's order: cell jv
// ascending, then cell iv.
__global__ void k_band(int ns, int ks, int nv, int nk, const double* __restrict__ mom,
                       const double* __restrict__ rkd, const double* __restrict__ msum,
                       double* __restrict__ rkb) {
  // grid (ceil(ns/64), ceil(ns/4), ks*ks*nk), block 256 = 64 i x 4 j
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= ns || j >= ns) return;
  const int kk = blockIdx.z / (ks * ks), r = blockIdx.z % (ks * ks);
  const int ip = r % ks, jp = r / ks;
  const size_t per = static_cast<size_t>(ns) * ns * ks * ks;
  const size_t tid = kk * per + i + static_cast<size_t>(ns) * (j + static_cast<size_t>(ns) * r);
  (void)nk;
  double acc = 0.0;
  if (i + ip < ns && j + jp < ns) {
    const int ik = ks * (ks + 1) / 2;
    const double* m1 = mom + static_cast<size_t>(kk * 2) * nv * ik;
    const double* m2 = m1 + static_cast<size_t>(nv) * ik;
    const double* rd = rkd + static_cast<size_t>(kk) * nv * ik * ik;
    const int ivlo = max(0, i + ip - (ks - 1)), ivhi = min(i, nv - 1);
    const int jvlo = max(0, j + jp - (ks - 1)), jvhi = min(j, nv - 1);
    const double* s1 = msum + static_cast<size_t>(kk * 2) * ks * ns;
    const double* s2 = s1 + static_cast<size_t>(ks) * ns;
    if (ivhi < jvlo) {
      rkb[tid] = s1[ip * ns + i] * s2[jp * ns + j];
      return;
    }
    if (ivlo > jvhi) {
      rkb[tid] = s1[jp * ns + j] * s2[ip * ns + i];
      return;
    }
    for (int jv = jvlo; jv <= jvhi; ++jv) {
      const int jj = pack2(j - jv, j + jp - jv, ks);
      for (int iv = ivlo; iv <= ivhi; ++iv) {
        const int ii = pack2(i - iv, i + ip - iv, ks);
        double c;
        if (iv < jv) c = m1[iv * ik + ii] * m2[jv * ik + jj];
        else if (iv > jv) c = m1[jv * ik + jj] * m2[iv * ik + ii];
        else c = rd[(static_cast<size_t>(iv) * ik + jj) * ik + ii];
        acc += c;
      }
    }
  }
  rkb[tid] = acc;
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
__global__ void k_direct_partial(int ns, int ks, int npair, const double* __restrict__ rkb,
                                 const double* __restrict__ dens, double* __restrict__ part) {
  extern __shared__ double dl[];  // [16][ns]
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int ip = blockIdx.y % ks, jp = blockIdx.y / ks;
  const int nchunk = (npair + 15) / 16;
  const int kk = blockIdx.z / nchunk, p0 = (blockIdx.z % nchunk) * 16;
  const int np = min(16, npair - p0);
  for (int q = threadIdx.x; q < 16 * ns; q += 64) {
    const int p = q / ns, j = q % ns;
    dl[q] = p < np ? dens[static_cast<size_t>(p0 + p) * ns * ks + j + static_cast<size_t>(ns) * jp] : 0.0;
  }
  __syncthreads();
  double acc[16];
#pragma unroll
  for (int p = 0; p < 16; ++p) acc[p] = 0.0;
  if (i < ns - ip) {
    const double* Rc = rkb + static_cast<size_t>(kk) * ns * ns * ks * ks + i +
                       static_cast<size_t>(ns) * ns * (ip + static_cast<size_t>(ks) * jp);
    const int nj = ns - jp;
    int j = 0;
    for (; j + 4 <= nj; j += 4) {
      const double r0 = Rc[static_cast<size_t>(ns) * j], r1 = Rc[static_cast<size_t>(ns) * (j + 1)];
      const double r2 = Rc[static_cast<size_t>(ns) * (j + 2)], r3 = Rc[static_cast<size_t>(ns) * (j + 3)];
#pragma unroll
      for (int p = 0; p < 16; ++p) {
        const double* d = dl + p * ns + j;
        acc[p] += d[0] * r0;
        acc[p] += d[1] * r1;
        acc[p] += d[2] * r2;
        acc[p] += d[3] * r3;
      }
    }
    for (; j < nj; ++j) {
      const double r = Rc[static_cast<size_t>(ns) * j];
#pragma unroll
      for (int p = 0; p < 16; ++p) acc[p] += dl[p * ns + j] * r;
    }
  }
  if (i < ns) {
    for (int p = 0; p < np; ++p)
      part[((static_cast<size_t>(kk) * npair + p0 + p) * ks + jp) * ns * ks + i + static_cast<size_t>(ns) * ip] =
          acc[p];
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
// E[kk][a][b](x,y) = sum_{z,w} K(x,z,y,w) p_a(z) p_b(w),
// K(x,z,y,w) = rkb(min(x,z), min(y,w), |x-z|, |y-w|).  One workgroup per
// output column y; threads over rows x (coalesced band reads and writes).
template <int NB>
__global__ void k_exchange_basis(int ns, int ks, const double* __restrict__ rkb,
                                 const double* __restrict__ P, double* __restrict__ E) {
  const int y = blockIdx.x, kk = blockIdx.y;
  const size_t band = static_cast<size_t>(ns) * ns * ks * ks;
  const double* R = rkb + kk * band;
  __shared__ double pw[NB][2 * kMaxKs];
  for (int q = threadIdx.x; q < NB * (2 * ks - 1); q += blockDim.x) {
    const int b = q / (2 * ks - 1), dw = q % (2 * ks - 1) - (ks - 1), w = y + dw;
    pw[b][dw + ks - 1] = (w >= 0 && w < ns) ? P[static_cast<size_t>(b) * ns + w] : 0.0;
  }
  __syncthreads();
  const size_t s_ip = static_cast<size_t>(ns) * ns, s_jp = s_ip * ks;
  for (int x = threadIdx.x; x < ns; x += blockDim.x) {
    double e[NB * NB];
#pragma unroll
    for (int q = 0; q < NB * NB; ++q) e[q] = 0.0;
    for (int dz = -(ks - 1); dz <= ks - 1; ++dz) {
      const int z = x + dz;
      if (z < 0 || z >= ns) continue;
      const int i0 = min(x, z), ip0 = dz < 0 ? -dz : dz;
      double tb[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) tb[b] = 0.0;
      for (int dw = -(ks - 1); dw <= ks - 1; ++dw) {
        const int w = y + dw;
        if (w < 0 || w >= ns) continue;
        const int j0 = min(y, w), jp0 = dw < 0 ? -dw : dw;
        const double r = R[i0 + static_cast<size_t>(ns) * j0 + s_ip * ip0 + s_jp * jp0];
#pragma unroll
        for (int b = 0; b < NB; ++b) tb[b] += r * pw[b][dw + ks - 1];
      }
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        const double pa = P[static_cast<size_t>(a) * ns + z];
#pragma unroll
        for (int b = 0; b < NB; ++b) e[a * NB + b] += pa * tb[b];
      }
    }
    const size_t pos = x + static_cast<size_t>(ns) * y;
#pragma unroll
    for (int q = 0; q < NB * NB; ++q)
      E[(static_cast<size_t>(kk) * NB * NB + q) * ns * ns + pos] = e[q];
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
// band terms: positions |x-y|<ks; threads walk columns, 2ks-1 rows each.
__global__ void k_band_terms(int ns, int ks, const int* blks, const int* ptr, const int* bid,
                             const double* bc, const double* bands, double* hcc) {
  const int e = blockIdx.x, blk = blks[e];
  const int t0 = ptr[e], t1 = ptr[e + 1];
  const int wid = 2 * ks - 1;
  double* h = hcc + static_cast<size_t>(blk) * ns * ns;
  for (int q = threadIdx.x; q < ns * wid; q += blockDim.x) {
    const int y = q / wid, x = y - (ks - 1) + q % wid;
    if (x < 0 || x >= ns) continue;
    const int i = min(x, y), ith = x > y ? x - y : y - x;
    if (i + ith >= ns) continue;
    double acc = 0.0;
    for (int t = t0; t < t1; ++t) acc += bc[t] * bands[static_cast<size_t>(bid[t]) * ns * ks + i + ns * ith];
    h[x + static_cast<size_t>(ns) * y] += acc;
  }
}

// rank-1 terms: workgroup per (block, 32-column tile); 16 terms per LDS pass
__global__ void k_rank1_terms(int ns, const int* blks, const int* ptr, const int* uid,
                              const int* wid, const int* sidx, const double* c, const double* stab,
                              const double* V, double* hcc) {
  const int e = blockIdx.x, blk = blks[e], y0 = blockIdx.y * 32;
  const int t0 = ptr[e], t1 = ptr[e + 1];
  __shared__ double cw[16][32];
  __shared__ int uu[16];
  double* h = hcc + static_cast<size_t>(blk) * ns * ns;
  const int ncol = min(32, ns - y0);
  for (int xb = 0; xb < ns; xb += blockDim.x) {
    const int x = xb + threadIdx.x;
    double acc[32];
#pragma unroll
    for (int q = 0; q < 32; ++q) acc[q] = 0.0;
    for (int tb = t0; tb < t1; tb += 16) {
      const int nt = min(16, t1 - tb);
      __syncthreads();
      for (int q = threadIdx.x; q < nt * 32; q += blockDim.x) {
        const int r = q / 32, col = q % 32, t = tb + r;
        const double f = c[t] * (sidx[t] >= 0 ? stab[sidx[t]] : 1.0);
        cw[r][col] = col < ncol ? f * V[static_cast<size_t>(wid[t]) * ns + y0 + col] : 0.0;
        if (col == 0) uu[r] = uid[t];
      }
      __syncthreads();
      if (x < ns) {
        for (int r = 0; r < nt; ++r) {
          const double u = V[static_cast<size_t>(uu[r]) * ns + x];
#pragma unroll
          for (int q = 0; q < 32; ++q) acc[q] += u * cw[r][q];
        }
      }
    }
    if (x < ns)
      for (int q = 0; q < ncol; ++q) h[x + static_cast<size_t>(ns) * (y0 + q)] += acc[q];
  }
}

__global__ void k_vec_terms(int ns, const int* qs, const int* ptr, const int* vid, const int* sidx,
                            const double* c, const double* stab, const double* V, double* hcb) {
  const int e = blockIdx.x, q = qs[e];
  const int t0 = ptr[e], t1 = ptr[e + 1];
  for (int x = threadIdx.x; x < ns; x += blockDim.x) {
    double acc = 0.0;
    for (int t = t0; t < t1; ++t) {
      const double f = c[t] * (sidx[t] >= 0 ? stab[sidx[t]] : 1.0);
      acc += f * V[static_cast<size_t>(vid[t]) * ns + x];
    }
    hcb[static_cast<size_t>(q) * ns + x] += acc;
  }
}

__global__ void k_symmetrize_diag(int ns, int nch, double* hcc) {
  const int ich = blockIdx.y;  // 0-based
  const size_t blk = static_cast<size_t>(ich + 1) * (ich + 2) / 2 - 1;
  double* h = hcc + blk * ns * ns;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ns * ns) return;
  const int i = q % ns, j = q / ns;
  if (j > i) return;
  const double s = h[i + static_cast<size_t>(ns) * j] + h[j + static_cast<size_t>(ns) * i];
  h[i + static_cast<size_t>(ns) * j] = s / 2.0;
  h[j + static_cast<size_t>(ns) * i] = s / 2.0;
}

__global__ void k_scale(size_t n, double a, double* x) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= a;
}

__global__ void k_fill(size_t n, double v, double* x) {
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

void launch_band(const DevGrid& g, int nk, const double* mom, const double* rkd, double* msum, double* rkb,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_pair_sums, dim3(nblocks(g.ns, 256), g.ks, nk * 2), dim3(256), 0, s, g.ns, g.ks, g.nv, mom,
                     msum);
  hipLaunchKernelGGL(k_band, dim3((g.ns + 63) / 64, (g.ns + 3) / 4, g.ks * g.ks * nk), dim3(256), 0, s, g.ns, g.ks,
                     g.nv, nk, mom, rkd, msum, rkb);
}

void launch_pair_dens(int ns, int ks, int nb, const double* P, double* dens, hipStream_t s) {
  const size_t n = static_cast<size_t>(nb) * (nb + 1) / 2 * ns * ks;
  hipLaunchKernelGGL(k_pair_dens, dim3(nblocks(n, 256)), dim3(256), 0, s, ns, ks, nb, P, dens);
}

void launch_direct_basis(int ns, int ks, int nk, int npair, const double* rkb, const double* dens, double* part,
                         double* D, hipStream_t s) {
  const int nchunk = (npair + 15) / 16;
  hipLaunchKernelGGL(k_direct_partial, dim3((ns + 63) / 64, ks * ks, nk * nchunk), dim3(64),
                     sizeof(double) * 16 * ns, s, ns, ks, npair, rkb, dens, part);
  const size_t n = static_cast<size_t>(ns) * ks * npair * nk;
  hipLaunchKernelGGL(k_direct_reduce, dim3(nblocks(n, 256)), dim3(256), 0, s, ns, ks, npair, nk, part, D);
}

void launch_exchange_basis(int ns, int ks, int nk, int nb, const double* rkb, const double* P, double* E,
                           hipStream_t s) {
  const dim3 grid(ns, nk), block(ns <= 64 ? 64 : (ns <= 128 ? 128 : 256));
  switch (nb) {
#define NB_CASE(N) \
  case N: hipLaunchKernelGGL(k_exchange_basis<N>, grid, block, 0, s, ns, ks, rkb, P, E); break;
    NB_CASE(1) NB_CASE(2) NB_CASE(3) NB_CASE(4) NB_CASE(5) NB_CASE(6) NB_CASE(7) NB_CASE(8)
#undef NB_CASE
    default: break;
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

void launch_band_terms(int ns, int ks, int nblk, const int* blks, const int* ptr, const int* bid, const double* bc,
                       const double* bands, double* hcc, hipStream_t s) {
  if (nblk <= 0) return;
  hipLaunchKernelGGL(k_band_terms, dim3(nblk), dim3(256), 0, s, ns, ks, blks, ptr, bid, bc, bands, hcc);
}

void launch_rank1_terms(int ns, int nblk, const int* blks, const int* ptr, const int* uid, const int* wid,
                        const int* sidx, const double* c, const double* stab, const double* V, double* hcc,
                        hipStream_t s) {
  if (nblk <= 0) return;
  hipLaunchKernelGGL(k_rank1_terms, dim3(nblk, (ns + 31) / 32), dim3(256), 0, s, ns, blks, ptr, uid, wid, sidx, c,
                     stab, V, hcc);
}

void launch_vec_terms(int ns, int nq, const int* qs, const int* ptr, const int* vid, const int* sidx, const double* c,
                      const double* stab, const double* V, double* hcb, hipStream_t s) {
  if (nq <= 0) return;
  hipLaunchKernelGGL(k_vec_terms, dim3(nq), dim3(256), 0, s, ns, qs, ptr, vid, sidx, c, stab, V, hcb);
}

void launch_symmetrize_diag(int ns, int nch, double* hcc, hipStream_t s) {
  if (nch <= 0) return;
  hipLaunchKernelGGL(k_symmetrize_diag, dim3(nblocks(static_cast<size_t>(ns) * ns, 256), nch), dim3(256), 0, s, ns,
                     nch, hcc);
}

void launch_scale(size_t n, double a, double* x, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_scale, dim3(nblocks(n, 256)), dim3(256), 0, s, n, a, x);
}

void launch_fill(size_t n, double v, double* x, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill, dim3(nblocks(n, 256)), dim3(256), 0, s, n, v, x);
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
