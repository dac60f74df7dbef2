// Lowest eigenvalues of the assembled partial wave, H x = E S x, in FP64 on
// the GPU: the final steps of BSR_HD3's diag_hd (DPOTRF of S, DSYGST, DSYEV,
// src/bsr/BSR_HD3/diag_hd.f90:271-312) on the full basis of the wave --
// without the per-channel truncation diag_hd makes first (:78-190), so the
// problem is the one of the matrices bsr_mat.nnn holds.  It checks what the
// assembly produces (split precision against FP64) where it matters for the
// physics: the bound-state energies.
//
// The channel blocks, channel-perturber vectors and perturber scalars are
// expanded on the device into the lower triangle of dense nhm x nhm
// matrices (nhm = nch ns + ncp, channel i at rows (i-1) ns.., perturbers
// last), then block Lanczos on the standard form (solve_lanczos) or
// rocSOLVER's sygvdx returns eigenvalues il = 1 .. iu = nev.
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bsr_gpu.h"
#include "ctx.h"

namespace {

// Lower triangle of the dense matrix, column C = blockIdx.y, rows R >= C from
// blockIdx.x * 256 + threadIdx.x: channel block (i, j) i >= j is
// hcc[i(i+1)/2 + j] (0-based), row ip of the perturbers against channel ich
// is hcb(:, ich, ip), perturber pair (ip, jp) jp <= ip is hbb[ip(ip+1)/2 + jp].
// Reads run along x, i.e. along the column of a block: coalesced.
__global__ __launch_bounds__(256) void k_densify_lower(int ns, int nch, int ncp, const double* __restrict__ hcc,
                                                        const double* __restrict__ hcb,
                                                        const double* __restrict__ hbb, double* __restrict__ A,
                                                        size_t lda) {
  const size_t n = static_cast<size_t>(nch) * ns + ncp, nchs = static_cast<size_t>(nch) * ns;
  const size_t C = blockIdx.y, R = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (R >= n || R < C) return;
  double v;
  if (R < nchs) {
    const size_t i = R / ns, x = R % ns, j = C / ns, y = C % ns;
    const size_t blk = i * (i + 1) / 2 + j;
    v = hcc[blk * ns * ns + x + ns * y];
  } else if (C < nchs) {
    const size_t ip = R - nchs, ich = C / ns, y = C % ns;
    v = hcb[y + ns * (ich + static_cast<size_t>(nch) * ip)];
  } else {
    const size_t ip = R - nchs, jp = C - nchs;
    v = hbb[ip * (ip + 1) / 2 + jp];
  }
  A[R + lda * C] = v;
}

int densify(const Matrices& M, double* A, size_t lda, DevVec<double>& hbb, hipStream_t s) {
  const size_t n = static_cast<size_t>(M.nch) * M.ns + M.ncp;
  HIPCHK(hbb.resize(std::max<size_t>(M.hbb.size(), 1)));
  if (!M.hbb.empty()) HIPCHK(hipMemcpyAsync(hbb.p, M.hbb.data(), M.hbb.size() * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_densify_lower, dim3(static_cast<unsigned>((n + 255) / 256), static_cast<unsigned>(n)),
                     dim3(256), 0, s, M.ns, M.nch, M.ncp, M.hcc.p, M.hcb.p, hbb.p, A, lda);
  HIPCHK(hipGetLastError());
  return 0;
}

// Upper triangle of a symmetric matrix from its lower one, through 32 x 32
// LDS tiles: tile (ti, tj), ti >= tj, reads the lower tile (rows ti,
// columns tj) and writes its transpose to (rows tj, columns ti); on the
// diagonal tiles only the strictly upper part is written.
// grid nt*(nt+1)/2, block 256 (32 x 8).
__global__ void k_sym_upper(int n, double* A, size_t lda) {
  __shared__ double t[32][33];
  int b = blockIdx.x, ti = 0;
  while (b > ti) { b -= ti + 1; ++ti; }
  const int tj = b;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int row = ti * 32 + tx, col = tj * 32 + r;
    t[r][tx] = (row < n && col < n) ? A[row + lda * col] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int row = tj * 32 + tx, col = ti * 32 + r;  // element (row, col) = A(col, row)
    if (row < n && col < n && row < col) A[row + lda * col] = t[tx][r];
  }
}

// Chunk sums C[q] = sum_b P[q + b mp], q < mp, added in chunk order.
__global__ void k_sum_chunks(size_t mp, int nb, const double* __restrict__ P, double* __restrict__ C) {
  const size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= mp) return;
  double v = 0.0;
  for (int b = 0; b < nb; ++b) v += P[q + static_cast<size_t>(b) * mp];
  C[q] = v;
}

void sym_upper(int n, double* A, hipStream_t s) {
  const int nt = (n + 31) / 32;
  hipLaunchKernelGGL(k_sym_upper, dim3(static_cast<unsigned>(nt) * (nt + 1) / 2), dim3(256), 0, s, n, A,
                     static_cast<size_t>(n));
}

// A deterministic start block: V(i, j) uniform in [-1, 1) from a hash of (i, j).
__global__ void k_start_block(int n, int p, double* V) {
  const size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= static_cast<size_t>(n) * p) return;
  unsigned long long z = q * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  V[q] = static_cast<double>(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// Cholesky factor R (upper, R^T R = G) of a small symmetric p x p matrix in
// place (column-major); false when G is not numerically positive definite.
bool chol_upper(int p, double* G) {
  for (int j = 0; j < p; ++j) {
    double d = G[j + p * j];
    for (int k = 0; k < j; ++k) d -= G[k + p * j] * G[k + p * j];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    G[j + p * j] = d;
    for (int i = j + 1; i < p; ++i) {
      double v = G[j + p * i];
      for (int k = 0; k < j; ++k) v -= G[k + p * j] * G[k + p * i];
      G[j + p * i] = v / d;
    }
    for (int i = j + 1; i < p; ++i) G[i + p * j] = 0.0;
  }
  return true;
}

#define RBCHK(x)                                                                        \
  do {                                                                                  \
    rocblas_status rs_ = (x);                                                           \
    if (rs_ != rocblas_status_success)                                                  \
      return fail(-100, std::string(#x " failed: ") + rocblas_status_to_string(rs_)); \
  } while (0)

// C (m x p, leading dimension m) = A^T B, A n x m and B n x p with leading
// dimensions lda and ldb.  rocBLAS gives a transposed product with a long
// inner dimension and a small result a handful of workgroups (2 ms for
// 16 x 16 x 12400); here the inner dimension is cut into nb chunks, one
// strided-batched product makes the nb partial results and a fixed-order
// sum adds them: enough workgroups, and the same result on every run.
int gemm_tn(rocblas_handle h, hipStream_t s, int m, int p, int n, const double* A, int lda, const double* B, int ldb,
            double* C, DevVec<double>& part) {
  const int tiles = ((m + 63) / 64) * ((p + 63) / 64);
  const int nb = std::max(1, std::min(512 / tiles, n / 64));
  const int kc = n / nb, r = n - kc * nb, nparts = nb + (r > 0 ? 1 : 0);
  const size_t mp = static_cast<size_t>(m) * p;
  const double one = 1.0, zero = 0.0;
  HIPCHK(part.resize(mp * nparts));
  RBCHK(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, m, p, kc, &one, A, lda,
                                      kc, B, ldb, kc, &zero, part.p, m, static_cast<rocblas_stride>(mp), nb));
  if (r > 0)
    RBCHK(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, m, p, r, &one,
                        A + static_cast<size_t>(kc) * nb, lda, B + static_cast<size_t>(kc) * nb, ldb, &zero,
                        part.p + mp * nb, m));
  hipLaunchKernelGGL(k_sum_chunks, dim3(static_cast<unsigned>((mp + 255) / 256)), dim3(256), 0, s, mp, nparts,
                     part.p, C);
  HIPCHK(hipGetLastError());
  return 0;
}

// The nev lowest eigenvalues of H x = E S x by block Lanczos.  S = L L^T
// (potrf), C = L^-1 H L^-T by two triangular solves on H in full storage
// (level-3 work, where sygst's reduction runs thousands of level-2 steps),
// made exactly symmetric from its lower triangle, then block Lanczos on C
// with blocks of p vectors and full reorthogonalisation (two classical
// Gram-Schmidt passes against every earlier block, Cholesky QR twice for
// the new block): T = Q^T C Q is block tridiagonal with the symmetrised
// diagonal blocks and the QR factors below and above the diagonal.  The
// eigenvalues of T (rocSOLVER syevd on the device) are checked at growing
// Krylov dimensions; Ritz pair (theta, Q s) has residual ||B s_last|| (B the
// coupling to the next block, s_last the last block of s), and an
// eigenvalue of C lies within that residual of theta, so the nev lowest are
// returned once each residual is below tol max(1, |theta|).  The start
// block is fixed, so results are reproducible.  Returns 1 when the Krylov
// space reaches its limit without convergence, or a new block is
// numerically rank-deficient, for the dense solver to take over.
int solve_lanczos(bsrg_ctx* c, const Matrices& H, const Matrices& S, int nev, double* e, double tol) {
  const int n = static_cast<int>(static_cast<size_t>(H.nch) * H.ns + H.ncp);
  const int p = std::min(64, std::max(16, nev + 6));
  const int mmax = std::min(n - p, std::max(64 * p, 1024)) / p * p;
  if (mmax < 4 * p) return 1;
  hipStream_t s = c->stream;
  rocblas_handle h = c->blas;
  DevVec<double> A, B, Q, Hh, Tm, Tev, Twk, hbb, hbbS, part;
  DevVec<int> info;
  HIPCHK(A.resize(static_cast<size_t>(n) * n));
  HIPCHK(B.resize(static_cast<size_t>(n) * n));
  HIPCHK(Q.resize(static_cast<size_t>(n) * (mmax + p)));
  HIPCHK(Hh.resize(static_cast<size_t>(mmax + p) * p));
  HIPCHK(info.resize(2));
  if (int rc = densify(H, A.p, n, hbb, s)) return rc;
  if (int rc = densify(S, B.p, n, hbbS, s)) return rc;
  RBCHK(rocsolver_dpotrf(h, rocblas_fill_lower, n, B.p, n, info.p));
  int hinfo = 0;
  HIPCHK(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (hinfo != 0) return fail(-40, "S is not positive definite (DPOTRF: leading minor " + std::to_string(hinfo) + ")");
  const double one = 1.0, mone = -1.0;
  sym_upper(n, A.p, s);
  RBCHK(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit, n,
                      n, &one, B.p, n, A.p, n));
  RBCHK(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                      rocblas_diagonal_non_unit, n, n, &one, B.p, n, A.p, n));
  HIPCHK(hipStreamSynchronize(s));
  B.release();
  sym_upper(n, A.p, s);
  const size_t np = static_cast<size_t>(n) * p;
  hipLaunchKernelGGL(k_start_block, dim3(static_cast<unsigned>((np + 255) / 256)), dim3(256), 0, s, n, p, Q.p);
  HIPCHK(hipGetLastError());
  std::vector<double> G(static_cast<size_t>(p) * p), Rb(static_cast<size_t>(p) * p), R1(static_cast<size_t>(p) * p);
  DevVec<double> Gd;
  HIPCHK(Gd.resize(static_cast<size_t>(p) * p));
  // W <- W R^-1 with R^T R = W^T W, twice; R = R2 R1 on return
  auto chol_qr = [&](double* W, std::vector<double>& R) -> int {
    for (int pass = 0; pass < 2; ++pass) {
      if (int rc = gemm_tn(h, s, p, p, n, W, n, W, n, Gd.p, part)) return rc;
      HIPCHK(hipMemcpyAsync(G.data(), Gd.p, G.size() * sizeof(double), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (!chol_upper(p, G.data())) return 1;
      HIPCHK(hipMemcpyAsync(Gd.p, G.data(), G.size() * sizeof(double), hipMemcpyHostToDevice, s));
      RBCHK(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                          rocblas_diagonal_non_unit, n, p, &one, Gd.p, p, W, n));
      if (pass == 0) {
        R1 = G;
      } else {  // R = G R1 (both upper)
        for (int j = 0; j < p; ++j)
          for (int i = 0; i < p; ++i) {
            double v = 0.0;
            for (int k = i; k <= j; ++k) v += G[i + p * k] * R1[k + p * j];
            R[i + p * j] = v;
          }
      }
    }
    return 0;
  };
  if (int rc = chol_qr(Q.p, Rb)) return rc > 0 ? 1 : rc;
  std::vector<double> T(static_cast<size_t>(mmax) * mmax, 0.0), Aj(static_cast<size_t>(p) * p);
  HIPCHK(Tm.resize(static_cast<size_t>(mmax) * mmax));
  HIPCHK(Tev.resize(mmax));
  HIPCHK(Twk.resize(mmax));
  std::vector<double> ev(mmax), Vlast(static_cast<size_t>(p) * nev);
  int next_check = 4 * p;
  for (int m = p; m <= mmax; m += p) {
    // block j = m/p - 1 at columns [m-p, m); the next one goes to [m, m+p)
    const double* Vj = Q.p + static_cast<size_t>(n) * (m - p);
    double* W = Q.p + static_cast<size_t>(n) * m;
    // W = C Vj as C^T Vj (C is symmetric in full storage)
    if (int rc = gemm_tn(h, s, n, p, n, A.p, n, Vj, n, W, part)) return rc;
    // the diagonal block, then W orthogonalised against every block so far
    if (int rc = gemm_tn(h, s, p, p, n, Vj, n, W, n, Gd.p, part)) return rc;
    HIPCHK(hipMemcpyAsync(Aj.data(), Gd.p, Aj.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    for (int pass = 0; pass < 2; ++pass) {
      if (int rc = gemm_tn(h, s, m, p, n, Q.p, n, W, n, Hh.p, part)) return rc;
      RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, p, m, &mone, Q.p, n, Hh.p, m, &one, W,
                          n));
    }
    HIPCHK(hipStreamSynchronize(s));
    for (int j = 0; j < p; ++j)
      for (int i = 0; i < p; ++i) T[(m - p + i) + static_cast<size_t>(mmax) * (m - p + j)] = 0.5 * (Aj[i + p * j] + Aj[j + p * i]);
    // a new block that is numerically rank-deficient ends the iteration
    // (the dense solver takes over)
    if (chol_qr(W, Rb) != 0) return 1;
    if (m < mmax)
      for (int j = 0; j < p; ++j)
        for (int i = 0; i < p; ++i) {
          T[(m + i) + static_cast<size_t>(mmax) * (m - p + j)] = Rb[i + p * j];
          T[(m - p + j) + static_cast<size_t>(mmax) * (m + i)] = Rb[i + p * j];
        }
    if (m < next_check && m < mmax) continue;
    next_check = m + std::max(4 * p, m / 2);
    // eigenpairs of T(0:m, 0:m) on the device
    HIPCHK(hipMemcpy2DAsync(Tm.p, static_cast<size_t>(m) * sizeof(double), T.data(),
                            static_cast<size_t>(mmax) * sizeof(double), static_cast<size_t>(m) * sizeof(double), m,
                            hipMemcpyHostToDevice, s));
    RBCHK(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, m, Tm.p, m, Tev.p, Twk.p, info.p));
    HIPCHK(hipMemcpyAsync(ev.data(), Tev.p, static_cast<size_t>(m) * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, s));
    // the last block of the nev lowest eigenvectors: rows m-p.. of columns 0..nev
    HIPCHK(hipMemcpy2DAsync(Vlast.data(), static_cast<size_t>(p) * sizeof(double), Tm.p + (m - p),
                            static_cast<size_t>(m) * sizeof(double), static_cast<size_t>(p) * sizeof(double), nev,
                            hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hinfo != 0) return 1;
    double worst = 0.0;
    bool conv = m >= 2 * nev;
    for (int i = 0; i < nev; ++i) {
      double r = 0.0;
      for (int a = 0; a < p; ++a) {  // (B s_last)_a, B = Rb upper
          double v = 0.0;
          for (int b = a; b < p; ++b) v += Rb[a + p * b] * Vlast[b + static_cast<size_t>(p) * i];
          r += v * v;
        }
      r = std::sqrt(r);
      worst = std::max(worst, r);
      conv = conv && r <= tol * std::max(1.0, std::fabs(ev[i]));
    }
    c->eigen_dim = m;
    c->eigen_resid = worst;
    if (conv) {
      std::copy(ev.begin(), ev.begin() + nev, e);
      return 0;
    }
    if (m + p > mmax) break;
  }
  return 1;
}

// The nev lowest eigenvalues of H x = E S x for matrices in the block form.
int solve_lowest(bsrg_ctx* c, const Matrices& H, const Matrices& S, int nev, double* e) {
  const size_t n = static_cast<size_t>(H.nch) * H.ns + H.ncp;
  if (nev < 1 || static_cast<size_t>(nev) > n) return fail(-1, "eigenvalues: nev must be in 1..nhm");
  if (n > 46000) return fail(-1, "eigenvalues: nhm above 46000 (32-bit sizes of the solver)");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->eigen_dim = 0;
  c->eigen_resid = 0.0;
  const bool lanczos = c->eigen_method == 2 || (c->eigen_method == 0 && n >= 2048);
  if (lanczos && 2 * nev + 64 <= static_cast<int>(n)) {
    const int rc = solve_lanczos(c, H, S, nev, e, 1e-11);
    if (rc <= 0) {
      c->eigen_used = 2;
      return rc;
    }
    // not converged within its Krylov space: the dense solver below
  }
  c->eigen_used = 1;
  c->eigen_dim = static_cast<int>(n);
  // dense work memory lives for this call only: 16 nhm^2 bytes
  DevVec<double> A, B, W, Z, hbb, hbbS;
  DevVec<int> info;
  HIPCHK(A.resize(n * n));
  HIPCHK(B.resize(n * n));
  HIPCHK(W.resize(n));
  HIPCHK(Z.resize(n * static_cast<size_t>(nev)));
  HIPCHK(info.resize(2));
  if (int rc = densify(H, A.p, n, hbb, c->stream)) return rc;
  if (int rc = densify(S, B.p, n, hbbS, c->stream)) return rc;
  const rocblas_int ni = static_cast<rocblas_int>(n);
  rocblas_status st = rocsolver_dsygvdx(c->blas, rocblas_eform_ax, rocblas_evect_none, rocblas_erange_index,
                                        rocblas_fill_lower, ni, A.p, ni, B.p, ni, 0.0, 0.0, 1, nev, info.p, W.p, Z.p,
                                        ni, info.p + 1);
  if (st != rocblas_status_success)
    return fail(-100, std::string("rocsolver_dsygvdx failed: ") + rocblas_status_to_string(st));
  int hinfo[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(hinfo, info.p, sizeof(hinfo), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(e, W.p, static_cast<size_t>(nev) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (hinfo[1] > ni) return fail(-40, "S is not positive definite (DPOTRF: leading minor " +
                                          std::to_string(hinfo[1] - ni) + ")");
  if (hinfo[1] != 0) return fail(-41, "eigenvalues did not converge");
  if (hinfo[0] != nev) return fail(-41, "eigen-solve returned " + std::to_string(hinfo[0]) + " eigenvalues");
  return 0;
}

// Fortran unformatted sequential records (4-byte length markers).
struct RecordReader {
  std::FILE* f = nullptr;
  std::vector<char> buf;
  ~RecordReader() { if (f) std::fclose(f); }
  bool next() {
    std::uint32_t len = 0, len2 = 0;
    if (std::fread(&len, 4, 1, f) != 1) return false;
    buf.resize(len);
    if (len && std::fread(buf.data(), 1, len, f) != len) return false;
    if (std::fread(&len2, 4, 1, f) != 1) return false;
    return len == len2;
  }
};

// One matrix (S or H) of the file as BSR_HD3 reads it (diag_hd.f90:42-63):
// kch diagonal blocks, then (ic, jc) headers each followed by an idim x jdim
// record (ns for a channel index, 1 for a perturber), up to ic <= 0.
int read_matrix(RecordReader& r, int ns, int nch, int ncp, Matrices& M, std::vector<double>& hcc,
                std::vector<double>& hcb) {
  const size_t nss = static_cast<size_t>(ns) * ns;
  M.ns = ns; M.nch = nch; M.ncp = ncp;
  hcc.assign(M.nblk() * nss, 0.0);
  hcb.assign(static_cast<size_t>(ns) * nch * ncp, 0.0);
  M.hbb.assign(static_cast<size_t>(ncp) * (ncp + 1) / 2, 0.0);
  for (int i = 1; i <= nch; ++i) {
    if (!r.next() || r.buf.size() != nss * 8) return fail(-50, "bsr_mat file: bad diagonal block record");
    std::memcpy(&hcc[(static_cast<size_t>(i) * (i + 1) / 2 - 1) * nss], r.buf.data(), nss * 8);
  }
  for (;;) {
    if (!r.next() || r.buf.size() < 8) return fail(-50, "bsr_mat file: bad block header");
    int ic, jc;
    std::memcpy(&ic, r.buf.data(), 4);
    std::memcpy(&jc, r.buf.data() + 4, 4);
    if (ic <= 0) return 0;
    if (ic > nch + ncp || jc < 1 || jc > nch + ncp) return fail(-50, "bsr_mat file: block index out of range");
    const size_t idim = ic <= nch ? ns : 1, jdim = jc <= nch ? ns : 1;
    if (!r.next() || r.buf.size() != idim * jdim * 8) return fail(-50, "bsr_mat file: bad block record");
    const double* v = reinterpret_cast<const double*>(r.buf.data());
    if (ic <= nch && jc <= nch) {  // channel block (ic, jc); the lower one is kept
      const int i = std::max(ic, jc), j = std::min(ic, jc);
      double* b = &hcc[(static_cast<size_t>(i) * (i - 1) / 2 + j - 1) * nss];
      for (int y = 0; y < ns; ++y)
        for (int x = 0; x < ns; ++x) b[x + static_cast<size_t>(ns) * y] = ic >= jc ? v[x + ns * y] : v[y + ns * x];
    } else if (ic <= nch || jc <= nch) {  // channel-perturber vector
      const int ich = std::min(ic, jc), ip = std::max(ic, jc) - nch;
      std::memcpy(&hcb[static_cast<size_t>(ns) * (ich - 1 + static_cast<size_t>(nch) * (ip - 1))], v, ns * 8);
    } else {
      const int ip = std::max(ic, jc) - nch, jp = std::min(ic, jc) - nch;
      M.hbb[static_cast<size_t>(ip) * (ip - 1) / 2 + jp - 1] = v[0];
    }
  }
}

}  // namespace

int bsrg_set_eigen_method(bsrg_ctx* c, int method) {
  if (!c) return fail(-2, "null context");
  if (method < 0 || method > 2) return fail(-1, "bsrg_set_eigen_method: method must be 0, 1 or 2");
  c->eigen_method = method;
  return 0;
}

int bsrg_eigen_info(bsrg_ctx* c, int* method, int* dim, double* resid) {
  if (!c) return fail(-2, "null context");
  if (method) *method = c->eigen_used;
  if (dim) *dim = c->eigen_dim;
  if (resid) *resid = c->eigen_resid;
  return 0;
}

int bsrg_wave_eigen(bsrg_ctx* c, int nev, double* e) {
  if (!c) return fail(-2, "null context");
  if (!c->wave_ok) return fail(-8, "no assembled wave (bsrg_wave_assemble)");
  if (!e) return fail(-2, "null argument");
  return solve_lowest(c, c->H, c->S, nev, e);
}

int bsrg_matfile_eigen(bsrg_ctx* c, const char* path, int nev, double* e) {
  if (!c) return fail(-2, "null context");
  if (!path || !e) return fail(-2, "null argument");
  RecordReader r;
  r.f = std::fopen(path, "rb");
  if (!r.f) return fail(-51, std::string("cannot open ") + path);
  if (!r.next() || r.buf.size() != 12) return fail(-50, "bsr_mat file: bad header record");
  int hdr[3];
  std::memcpy(hdr, r.buf.data(), 12);
  const int ns = hdr[0], nch = hdr[1], ncp = hdr[2];
  if (ns < 1 || nch < 1 || ncp < 0) return fail(-50, "bsr_mat file: bad sizes");
  HIPCHK(hipSetDevice(c->device));
  Matrices S, H;
  std::vector<double> hcc, hcb;
  for (Matrices* M : {&S, &H}) {
    if (int rc = read_matrix(r, ns, nch, ncp, *M, hcc, hcb)) return rc;
    HIPCHK(M->hcc.upload(hcc.data(), hcc.size(), c->stream));
    HIPCHK(M->hcb.resize(std::max<size_t>(hcb.size(), 1)));
    if (!hcb.empty()) HIPCHK(M->hcb.upload(hcb.data(), hcb.size(), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return solve_lowest(c, H, S, nev, e);
}
