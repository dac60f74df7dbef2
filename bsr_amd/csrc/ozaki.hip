// Split-precision (Ozaki-style) FP64 GEMM on bf16 MFMA for gfx950.
//
//   C(m x n) = A(m x K) * B(K x n) [+ C],  FP64 in and out, column-major.
//
// Every row of A and column of B is scaled by a power of two so that its
// largest entry is below 1, then cut into S slices of 7 bits:
//   a = sigma * sum_s x_s 2^-7(s+1),   x_s integers, |x_1| <= 128, |x_s| <= 64,
// exact in bf16.  A slice product is an integer of magnitude <= 2^14.  The
// products with s+r <= S-1 are kept; level L = s+r (L+1 products per K
// index) accumulates in FP32 over 128 values of K, i.e. at most 8*128
// products of magnitude <= 2^14: the sum stays <= 2^24 and the bf16 MFMA
// with FP32 accumulation is exact.  Each level is then converted to FP64,
// scaled by 2^-7(L+2), summed in FP64, and the result is multiplied by
// sigma_row * tau_col.  The truncation error per element is of order
// 2^-7S * max|row| * max|col| * K.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "bsrg_internal.h"

namespace bsrg {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBeta = 7;

// power of two above max |row i| (1 for an all-zero row)
__global__ void k_row_scale(int m, int K, const double* A, size_t lda, double* sig) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double mx = 0.0;
  for (int t = 0; t < K; ++t) mx = fmax(mx, fabs(A[i + lda * t]));
  int e = 0;
  if (mx > 0.0) (void)frexp(mx, &e);
  sig[i] = ldexp(1.0, e);
}

__global__ void k_col_scale(int n, int K, const double* B, size_t ldb, double* tau) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double mx = 0.0;
  for (int t = 0; t < K; ++t) mx = fmax(mx, fabs(B[t + ldb * j]));
  int e = 0;
  if (mx > 0.0) (void)frexp(mx, &e);
  tau[j] = ldexp(1.0, e);
}

__device__ __forceinline__ void cut(double x, int S, __bf16* out, size_t stride) {
  for (int s = 0; s < S; ++s) {
    const double y = x * 128.0;
    const double d = rint(y);
    x = y - d;
    out[s * stride] = static_cast<__bf16>(static_cast<float>(d));
  }
}

// A slices, layout As[s][row][t] (t fastest, Kp padded, rows padded to mp):
// 64 rows x 32 t per workgroup, transposed through LDS.
__global__ void k_slice_rows(int m, int K, int mp, int Kp, int S, const double* A, size_t lda, const double* sig,
                             __bf16* As) {
  __shared__ double tile[32][65];
  const int r0 = blockIdx.x * 64, t0 = blockIdx.y * 32;
  for (int q = threadIdx.x; q < 64 * 32; q += blockDim.x) {
    const int r = q % 64, t = q / 64;
    const int i = r0 + r, tt = t0 + t;
    tile[t][r] = (i < m && tt < K) ? A[i + lda * tt] / sig[i] : 0.0;
  }
  __syncthreads();
  const size_t stride = static_cast<size_t>(mp) * Kp;
  for (int q = threadIdx.x; q < 64 * 32; q += blockDim.x) {
    const int t = q % 32, r = q / 32;
    cut(tile[t][r], S, As + static_cast<size_t>(r0 + r) * Kp + t0 + t, stride);
  }
}

// B slices, layout Bs[r][col][t] (t fastest).
__global__ void k_slice_cols(int n, int K, int np, int Kp, int S, const double* B, size_t ldb, const double* tau,
                             __bf16* Bs) {
  const size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= static_cast<size_t>(np) * Kp) return;
  const int t = static_cast<int>(q % Kp), j = static_cast<int>(q / Kp);
  const double x = (j < n && t < K) ? B[t + ldb * j] / tau[j] : 0.0;
  cut(x, S, Bs + q, static_cast<size_t>(np) * Kp);
}

// 64 x 64 output tile per workgroup, 32 x 32 per wave (2 x 2 MFMA 16x16x32).
template <int S>
__global__ __launch_bounds__(256) void k_ozaki(int m, int n, int mp, int np, int Kp, const __bf16* __restrict__ As,
                                                const __bf16* __restrict__ Bs, const double* __restrict__ sig,
                                                const double* __restrict__ tau, double* __restrict__ C, size_t ldc,
                                                int accumulate) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row0 = blockIdx.x * 64 + (wave & 1) * 32, col0 = blockIdx.y * 64 + (wave >> 1) * 32;
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  f32x4 acc[S][2][2];
  double acc64[2][2][4];
#pragma unroll
  for (int L = 0; L < S; ++L)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[L][a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) acc64[a][b][e] = 0.0;
  const size_t sa = static_cast<size_t>(mp) * Kp, sb = static_cast<size_t>(np) * Kp;
  const __bf16* pa = As + static_cast<size_t>(row0 + lr) * Kp + lk;
  const __bf16* pb = Bs + static_cast<size_t>(col0 + lr) * Kp + lk;
  for (int kc = 0; kc < Kp; kc += 32) {
    bf16x8 a[S][2], b[S][2];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        a[s][t] = *reinterpret_cast<const bf16x8*>(pa + s * sa + static_cast<size_t>(t * 16) * Kp + kc);
        b[s][t] = *reinterpret_cast<const bf16x8*>(pb + s * sb + static_cast<size_t>(t * 16) * Kp + kc);
      }
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int r = 0; r < S - s; ++r)
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
          for (int tj = 0; tj < 2; ++tj)
            acc[s + r][ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[s][ti], b[r][tj], acc[s + r][ti][tj], 0, 0, 0);
    if (((kc >> 5) & 3) == 3 || kc + 32 >= Kp) {
#pragma unroll
      for (int L = 0; L < S; ++L) {
        const double f = ldexp(1.0, -kBeta * (L + 2));
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
          for (int tj = 0; tj < 2; ++tj) {
#pragma unroll
            for (int e = 0; e < 4; ++e) acc64[ti][tj][e] += static_cast<double>(acc[L][ti][tj][e]) * f;
            acc[L][ti][tj] = f32x4{0.f, 0.f, 0.f, 0.f};
          }
      }
    }
  }
  // D[i][j] sits in lane j + 16*(i/4), register i%4 of each 16x16 tile
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
      const int col = col0 + tj * 16 + lr;
      if (col >= n) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = row0 + ti * 16 + (lane >> 4) * 4 + e;
        if (row >= m) continue;
        const double v = acc64[ti][tj][e] * sig[row] * tau[col];
        double* c = C + row + ldc * col;
        *c = accumulate ? *c + v : v;
      }
    }
}

// Plain FP64 GEMM on FP64 MFMA (16x16x4): workgroup tile (TM*32) x 64,
// 4 waves of (TM*16) x 32; A column-major (rows fastest), B column-major
// (K fastest).  K is zero-padded to a multiple of 4 by predicated loads.
typedef double f64x4 __attribute__((ext_vector_type(4)));
template <int TM>
__global__ __launch_bounds__(256) void k_dgemm(int m, int n, int K, const double* __restrict__ A, size_t lda,
                                                const double* __restrict__ B, size_t ldb, double* __restrict__ C,
                                                size_t ldc, int accumulate) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row0 = blockIdx.x * (TM * 32) + (wave & 1) * (TM * 16), col0 = blockIdx.y * 64 + (wave >> 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  f64x4 acc[TM][2];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
  int rows[TM], cols[2];
#pragma unroll
  for (int a = 0; a < TM; ++a) rows[a] = min(row0 + a * 16 + lr, m - 1);
#pragma unroll
  for (int b = 0; b < 2; ++b) cols[b] = min(col0 + b * 16 + lr, n - 1);
  for (int k0 = 0; k0 < K; k0 += 4) {
    const int k = k0 + lk;
    const bool kin = k < K;
    double av[TM], bv[2];
#pragma unroll
    for (int a = 0; a < TM; ++a) av[a] = kin ? A[rows[a] + lda * k] : 0.0;
#pragma unroll
    for (int b = 0; b < 2; ++b) bv[b] = kin ? B[k + ldb * cols[b]] : 0.0;
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int col = col0 + b * 16 + lr;
      if (col >= n) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        // FP64 16x16x4 output: D[i][j] in lane j + 16*(i%4), register i/4
        const int row = row0 + a * 16 + lk + 4 * e;
        if (row >= m) continue;
        double* c = C + row + ldc * col;
        *c = accumulate ? *c + acc[a][b][e] : acc[a][b][e];
      }
    }
}

template <int S>
void launch_k(dim3 g, hipStream_t s, int m, int n, int mp, int np, int Kp, const __bf16* As, const __bf16* Bs,
              const double* sig, const double* tau, double* C, size_t ldc, int acc) {
  hipLaunchKernelGGL(k_ozaki<S>, g, dim3(256), 0, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc);
}

}  // namespace

int mfma_dgemm(int m, int n, int K, const double* A, size_t lda, const double* B, size_t ldb, double* C, size_t ldc,
               bool accumulate, int tm, hipStream_t s) {
  if (m <= 0 || n <= 0 || K <= 0) return 0;
  const int acc = accumulate ? 1 : 0;
  if (tm == 4) {
    hipLaunchKernelGGL(k_dgemm<4>, dim3((m + 127) / 128, (n + 63) / 64), dim3(256), 0, s, m, n, K, A, lda, B, ldb, C,
                       ldc, acc);
  } else if (tm == 2) {
    hipLaunchKernelGGL(k_dgemm<2>, dim3((m + 63) / 64, (n + 63) / 64), dim3(256), 0, s, m, n, K, A, lda, B, ldb, C,
                       ldc, acc);
  } else {
    hipLaunchKernelGGL(k_dgemm<1>, dim3((m + 31) / 32, (n + 63) / 64), dim3(256), 0, s, m, n, K, A, lda, B, ldb, C,
                       ldc, acc);
  }
  return hipGetLastError() == hipSuccess ? 0 : fail(-100, "FP64 MFMA GEMM launch failed");
}

size_t ozaki_workspace_bytes(int m, int n, int K, int S) {
  const size_t mp = (static_cast<size_t>(m) + 63) / 64 * 64, np = (static_cast<size_t>(n) + 63) / 64 * 64;
  const size_t Kp = (static_cast<size_t>(K) + 31) / 32 * 32;
  return S * (mp + np) * Kp * sizeof(__bf16) + (mp + np) * sizeof(double) + 256;
}

int ozaki_dgemm(int m, int n, int K, const double* A, size_t lda, const double* B, size_t ldb, double* C, size_t ldc,
                bool accumulate, int S, void* work, hipStream_t s) {
  if (S < 2 || S > 8) return fail(-9, "split precision: slices must be 2..8");
  if (m <= 0 || n <= 0 || K <= 0) return 0;
  const int mp = (m + 63) / 64 * 64, np = (n + 63) / 64 * 64, Kp = (K + 31) / 32 * 32;
  __bf16* As = static_cast<__bf16*>(work);
  __bf16* Bs = As + static_cast<size_t>(S) * mp * Kp;
  double* sig = reinterpret_cast<double*>(
      (reinterpret_cast<uintptr_t>(Bs + static_cast<size_t>(S) * np * Kp) + 15) & ~static_cast<uintptr_t>(15));
  double* tau = sig + mp;
  hipLaunchKernelGGL(k_row_scale, dim3((m + 255) / 256), dim3(256), 0, s, m, K, A, lda, sig);
  hipLaunchKernelGGL(k_col_scale, dim3((n + 255) / 256), dim3(256), 0, s, n, K, B, ldb, tau);
  hipLaunchKernelGGL(k_slice_rows, dim3(mp / 64, Kp / 32), dim3(256), 0, s, m, K, mp, Kp, S, A, lda, sig, As);
  const size_t nb = static_cast<size_t>(np) * Kp;
  hipLaunchKernelGGL(k_slice_cols, dim3(static_cast<unsigned>((nb + 255) / 256)), dim3(256), 0, s, n, K, np, Kp, S,
                     B, ldb, tau, Bs);
  const dim3 g(mp / 64, np / 64);
  const int acc = accumulate ? 1 : 0;
  switch (S) {
    case 2: launch_k<2>(g, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc); break;
    case 3: launch_k<3>(g, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc); break;
    case 4: launch_k<4>(g, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc); break;
    case 5: launch_k<5>(g, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc); break;
    case 6: launch_k<6>(g, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc); break;
    case 7: launch_k<7>(g, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc); break;
    default: launch_k<8>(g, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc); break;
  }
  return hipGetLastError() == hipSuccess ? 0 : fail(-100, "split-precision GEMM launch failed");
}

}  // namespace bsrg
