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
// last), then rocSOLVER's sygvdx returns eigenvalues il = 1 .. iu = nev.
#include <rocsolver/rocsolver.h>

#include <algorithm>
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

// The nev lowest eigenvalues of H x = E S x for matrices in the block form.
int solve_lowest(bsrg_ctx* c, const Matrices& H, const Matrices& S, int nev, double* e) {
  const size_t n = static_cast<size_t>(H.nch) * H.ns + H.ncp;
  if (nev < 1 || static_cast<size_t>(nev) > n) return fail(-1, "eigenvalues: nev must be in 1..nhm");
  if (n > 46000) return fail(-1, "eigenvalues: nhm above 46000 (32-bit sizes of the solver)");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
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
