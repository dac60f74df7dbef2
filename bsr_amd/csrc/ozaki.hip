// Split-precision (Ozaki-style) FP64 GEMM on int8 or bf16 MFMA for gfx950,
// and the plain FP64 MFMA GEMM.
//
//   C(m x n) = A(m x K) * B(K x n) [+ C],  FP64 in and out, column-major.
//
// Every row of A and column of B is scaled by a power of two, 2^(e+1) for a
// largest entry f 2^e (f in [1/2, 1)), which leaves every entry below 1/2,
// then cut into S slices of 7 bits:
//   a = sigma * sum_s x_s 2^-7s,   x_s = rint(y 2^7) of the running
// remainder y, integers in [-64, 64]: exact in int8 and in bf16.  A slice
// product is an integer of magnitude <= 2^12.  The products with
// s + r <= S-1 are kept; level L = s + r accumulates its L+1 products per
// inner index over the whole K:
//   int8 slices, int32 MFMA accumulation: exact while S 2^12 Kp < 2^31;
//   bf16 slices, FP32 MFMA accumulation: every partial sum is an integer of
//     magnitude <= S 2^12 Kp, exact while that is <= 2^24.
// Each level is converted to FP64, scaled by 2^-7(L+2) and summed in FP64,
// and the result is multiplied by sigma_row * tau_col.  The truncation
// error per element is of order 2^-7S * max|row| * max|col| * K.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "bsrg_internal.h"

namespace bsrg {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBeta = 7;

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
  // FP64 16x16x4 output: D[i][j] in lane j + 16*(i%4), register i/4; all
  // loads of the old values before any store
  if (accumulate) {
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int col = min(col0 + b * 16 + lr, n - 1);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = min(row0 + a * 16 + lk + 4 * e, m - 1);
          acc[a][b][e] += C[row + ldc * col];
        }
      }
  }
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int col = col0 + b * 16 + lr;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = row0 + a * 16 + lk + 4 * e;
        if (row < m && col < n) C[row + ldc * col] = acc[a][b][e];
      }
    }
}

// ---- slices on int8 or bf16 MFMA ---------------------------------------------
typedef int i32x4 __attribute__((ext_vector_type(4)));

__global__ void k_row_scale8(int m, int K, const double* A, size_t lda, double* sig) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double mx = 0.0;
  for (int t = 0; t < K; ++t) mx = fmax(mx, fabs(A[i + lda * t]));
  int e = 0;
  if (mx > 0.0) (void)frexp(mx, &e);
  sig[i] = ldexp(1.0, e + 1);
}

// one wavefront per column: coalesced reads along K, shuffle reduction
__global__ void k_col_scale8(int n, int K, const double* B, size_t ldb, double* tau) {
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (j >= n) return;
  double mx = 0.0;
  for (int t = lane; t < K; t += 64) mx = fmax(mx, fabs(B[t + ldb * j]));
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  if (lane == 0) {
    int e = 0;
    if (mx > 0.0) (void)frexp(mx, &e);
    tau[j] = ldexp(1.0, e + 1);
  }
}

template <typename T>
__device__ __forceinline__ void cut8(double x, int S, T* out, size_t stride) {
  for (int s = 0; s < S; ++s) {
    const double y = x * 128.0;
    const double d = rint(y);
    x = y - d;
    out[s * stride] = static_cast<T>(static_cast<float>(d));
  }
}

// K elements in a 64-byte chunk row: 64 int8 or 32 bf16
template <typename T>
constexpr int chunk_k() { return 64 / static_cast<int>(sizeof(T)); }

// Slices are stored chunk-major: X[s][c][row][CK] for K chunk c = t/CK
// (CK = 64 bytes of elements), so the 64 rows x 64 bytes of one chunk that a
// workgroup stages are one contiguous 4 KB block.
template <typename T>
__device__ __forceinline__ size_t at8(int c, int row, int t, int rows_p) {
  return (static_cast<size_t>(c) * rows_p + row) * chunk_k<T>() + t;
}

// A slices (rows padded to mp, K to Kp): 64 rows x CK t per workgroup,
// transposed through LDS.
template <typename T>
__global__ void k_slice8_rows(int m, int K, int mp, int Kp, int S, const double* A, size_t lda, const double* sig,
                              T* As) {
  constexpr int CK = chunk_k<T>();
  __shared__ double tile[CK][65];
  const int r0 = blockIdx.x * 64, t0 = blockIdx.y * CK;
  for (int q = threadIdx.x; q < 64 * CK; q += blockDim.x) {
    const int r = q % 64, t = q / 64;
    const int i = r0 + r, tt = t0 + t;
    tile[t][r] = (i < m && tt < K) ? A[i + lda * tt] / sig[i] : 0.0;
  }
  __syncthreads();
  const size_t stride = static_cast<size_t>(mp) * Kp;
  for (int q = threadIdx.x; q < 64 * CK; q += blockDim.x) {
    const int t = q % CK, r = q / CK;
    cut8(tile[t][r], S, As + at8<T>(blockIdx.y, r0 + r, t, mp), stride);
  }
}

// B slices (columns padded to np, K to Kp), same chunk-major layout
template <typename T>
__global__ void k_slice8_cols(int n, int K, int np, int Kp, int S, const double* B, size_t ldb, const double* tau,
                              T* Bs) {
  constexpr int CK = chunk_k<T>();
  const size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= static_cast<size_t>(np) * Kp) return;
  const int t = static_cast<int>(q % Kp), j = static_cast<int>(q / Kp);
  const double x = (j < n && t < K) ? B[t + ldb * j] / tau[j] : 0.0;
  cut8(x, S, Bs + at8<T>(t / CK, j, t % CK, np), static_cast<size_t>(np) * Kp);
}

// C(m x n) (+)= A B computed as its transpose: MFMA rows are the n columns
// of C (B slices), MFMA columns the m rows of C (A slices), so a wave's
// output lanes run along the rows of C and the FP64 stores are coalesced.
// Workgroup tile 64 (C columns) x 64 (C rows), 4 waves of 32 x 32
// (2 x 2 MFMA 16x16x64 i8 or 16x16x32 bf16), K in chunks of 64 bytes
// staged in LDS (swizzled 64-byte rows, sw8); the fragments of A and B are
// the same 16 bytes of a chunk row per lane, so the two operands' K
// elements correspond whatever order the instruction takes them in.  Every
// level L = s + r <= S-1 has its own int32 or FP32 accumulators; the
// epilogue recombines them in FP64.  Two
// waves per SIMD are asked for up to S = 7 (the accumulators then live in
// VGPRs; at S = 7 the default allocation would leave one wave per SIMD);
// S = 8 keeps one (its LDS tiles allow one workgroup per CU anyway).
// LDS tiles: 64 rows of 64 bytes; the 16-byte piece q of row r sits at
// piece q ^ ((r >> 2) & 3), so that the staging writes (4 lanes a row) and
// the fragment reads (16 rows, one piece) both spread over all banks.
__device__ __forceinline__ int sw8(int r, int byte) { return r * 64 + ((((byte >> 4) ^ (r >> 2)) & 3) << 4); }
// Place of C(row, col): column-major with leading dimension ldc, or, when
// qns > 0, inside block col of order ldo (blocks ldo^2 apart) as the qns x qns
// submatrix whose origin C points at: row = x + qns y -> x + ldo y.
__device__ __forceinline__ size_t out_at(int row, int col, size_t ldc, int qns, int ldo) {
  if (qns <= 0) return row + ldc * col;
  const int y = row / qns, x = row - y * qns;
  return static_cast<size_t>(col) * ldo * ldo + x + static_cast<size_t>(ldo) * y;
}
template <typename T, int S>
__global__ __launch_bounds__(256, S <= 7 ? 2 : 1) void k_ozaki8(int m, int n, int mp, int np, int Kp, const T* __restrict__ As,
                                                 const T* __restrict__ Bs, const double* __restrict__ sig,
                                                 const double* __restrict__ tau, double* __restrict__ C, size_t ldc,
                                                 int accumulate, int sc, int qns, int ldo) {
  constexpr bool kI8 = sizeof(T) == 1;
  constexpr int CK = chunk_k<T>();
  using Acc = typename std::conditional<kI8, i32x4, f32x4>::type;
  __shared__ __attribute__((aligned(16))) int8_t sB[S][64 * 64];  // C columns (blocks)
  __shared__ __attribute__((aligned(16))) int8_t sA[S][64 * 64];  // C rows (positions)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // 1-D grid over (row tile, column tile).  Workgroups are dealt to the 8
  // XCDs round robin; renumber them so that each XCD gets a contiguous run.
  // The run is ordered in column supertiles of sc column tiles (their B
  // slices stay in the XCD's L2) and, within one, row tile by row tile, so
  // the sc workgroups of a row tile share its A slices too.
  const int nct = np / 64, nrt = mp / 64, nwg = static_cast<int>(gridDim.x), full = nwg - nwg % 8;
  const int hw = static_cast<int>(blockIdx.x);
  int id = hw < full ? (hw % 8) * (full / 8) + hw / 8 : hw;
  const int nfull = nct / sc, per = nrt * sc;
  int cs0, cw;
  if (id < nfull * per) {
    cs0 = (id / per) * sc;
    id %= per;
    cw = sc;
  } else {
    id -= nfull * per;
    cs0 = nfull * sc;
    cw = nct - cs0;
  }
  const int c0 = (cs0 + id % cw) * 64, r0 = (id / cw) * 64;  // C column / row tile origins
  const int wc = (wave & 1) * 32, wr = (wave >> 1) * 32;  // wave tile within: columns, rows
  const int li = lane & 15, lk = (lane >> 4) * 16;
  // up to S = 6 the old values of C are loaded before the K loop (their
  // latency hides behind it; 32 more VGPRs still leave two waves per SIMD)
  constexpr bool kPre = S <= 6;
  double old[2][2][4];
  if (kPre && accumulate) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = min(r0 + wr + b * 16 + li, m - 1);
          const int col = min(c0 + wc + a * 16 + (lane >> 4) * 4 + e, n - 1);
          old[a][b][e] = C[out_at(row, col, ldc, qns, ldo)];
        }
  }
  Acc acc[S][2][2];
#pragma unroll
  for (int L = 0; L < S; ++L)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[L][a][b] = Acc{0, 0, 0, 0};
  const size_t sa = static_cast<size_t>(mp) * Kp, sb = static_cast<size_t>(np) * Kp;
  const int pr = threadIdx.x >> 2, pk = (threadIdx.x & 3) * 16;  // staging: row, 16-byte piece
  // software pipeline: the next chunk's global loads are in flight while
  // the current chunk is multiplied out of LDS; the 64 rows of a chunk are
  // one contiguous 4 KB block per slice
  constexpr int kPiece = 16 / static_cast<int>(sizeof(T));  // elements in a thread's 16 bytes
  i32x4 gb[S], ga[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    gb[s] = *reinterpret_cast<const i32x4*>(Bs + s * sb + at8<T>(0, c0, 0, np) + threadIdx.x * kPiece);
    ga[s] = *reinterpret_cast<const i32x4*>(As + s * sa + at8<T>(0, r0, 0, mp) + threadIdx.x * kPiece);
  }
  for (int kc = 0; kc < Kp; kc += CK) {
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S; ++s) {
      *reinterpret_cast<i32x4*>(&sB[s][sw8(pr, pk)]) = gb[s];
      *reinterpret_cast<i32x4*>(&sA[s][sw8(pr, pk)]) = ga[s];
    }
    __syncthreads();
    if (kc + CK < Kp) {
      const int c = kc / CK + 1;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        gb[s] = *reinterpret_cast<const i32x4*>(Bs + s * sb + at8<T>(c, c0, 0, np) + threadIdx.x * kPiece);
        ga[s] = *reinterpret_cast<const i32x4*>(As + s * sa + at8<T>(c, r0, 0, mp) + threadIdx.x * kPiece);
      }
    }
    i32x4 av[S][2];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int t = 0; t < 2; ++t) av[s][t] = *reinterpret_cast<const i32x4*>(&sA[s][sw8(wr + t * 16 + li, lk)]);
#pragma unroll
    for (int r = 0; r < S; ++r) {
      i32x4 bv[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) bv[t] = *reinterpret_cast<const i32x4*>(&sB[r][sw8(wc + t * 16 + li, lk)]);
#pragma unroll
      for (int s = 0; s < S - r; ++s)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b)
            if constexpr (kI8)
              acc[s + r][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(bv[a], av[s][b], acc[s + r][a][b], 0, 0, 0);
            else
              acc[s + r][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, bv[a]),
                                                                        __builtin_bit_cast(bf16x8, av[s][b]),
                                                                        acc[s + r][a][b], 0, 0, 0);
    }
  }
  // tile (a, b): MFMA row i = C column c0+wc+16a+i, MFMA column j = C row
  // r0+wr+16b+j; D[i][j] in lane j + 16*(i/4), register i%4.  All loads of
  // the old values are issued before any store.
  double v[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int row = min(r0 + wr + b * 16 + li, m - 1);
      const double sr = sig[row];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int col = min(c0 + wc + a * 16 + (lane >> 4) * 4 + e, n - 1);
        double x = 0.0;
#pragma unroll
        for (int L = S - 1; L >= 0; --L) x += ldexp(static_cast<double>(acc[L][a][b][e]), -kBeta * (L + 2));
        v[a][b][e] = x * (sr * tau[col]);
      }
    }
  if (accumulate) {
    if (!kPre) {
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int row = min(r0 + wr + b * 16 + li, m - 1);
            const int col = min(c0 + wc + a * 16 + (lane >> 4) * 4 + e, n - 1);
            old[a][b][e] = C[out_at(row, col, ldc, qns, ldo)];
          }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int e = 0; e < 4; ++e) v[a][b][e] += old[a][b][e];
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int row = r0 + wr + b * 16 + li;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int col = c0 + wc + a * 16 + (lane >> 4) * 4 + e;
        if (row < m && col < n) C[out_at(row, col, ldc, qns, ldo)] = v[a][b][e];
      }
    }
}

template <typename T, int S>
void launch_k8(dim3 g, hipStream_t s, int m, int n, int mp, int np, int Kp, const T* As, const T* Bs,
               const double* sig, const double* tau, double* C, size_t ldc, int acc, int qns, int ldo) {
  // column supertile: the B slices of sc column tiles in about 1.5 MB
  const size_t tile = static_cast<size_t>(S) * 64 * Kp * sizeof(T);
  const int sc = std::max(1, std::min(np / 64, static_cast<int>((3u << 19) / tile)));
  hipLaunchKernelGGL((k_ozaki8<T, S>), g, dim3(256), 0, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc, sc, qns,
                     ldo);
}

template <typename T>
void launch_slices(dim3 g, hipStream_t s, int S, int m, int n, int mp, int np, int Kp, const T* As, const T* Bs,
                   const double* sig, const double* tau, double* C, size_t ldc, int acc, int qns, int ldo) {
  switch (S) {
    case 2: launch_k8<T, 2>(g, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc, qns, ldo); break;
    case 3: launch_k8<T, 3>(g, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc, qns, ldo); break;
    case 4: launch_k8<T, 4>(g, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc, qns, ldo); break;
    case 5: launch_k8<T, 5>(g, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc, qns, ldo); break;
    case 6: launch_k8<T, 6>(g, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc, qns, ldo); break;
    case 7: launch_k8<T, 7>(g, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc, qns, ldo); break;
    default: launch_k8<T, 8>(g, s, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc, acc, qns, ldo); break;
  }
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

// bf16 slices: A and B sliced per call into work (A slices, B slices, then
// sig and tau, each 256-byte aligned)
size_t ozaki_workspace_bytes(int m, int n, int K, int S) {
  const size_t mp = (static_cast<size_t>(m) + 63) / 64 * 64, np = (static_cast<size_t>(n) + 63) / 64 * 64;
  const size_t Kp = (static_cast<size_t>(K) + 31) / 32 * 32;
  const auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  return al(S * mp * Kp * sizeof(__bf16)) + al(S * np * Kp * sizeof(__bf16)) + al(mp * sizeof(double)) +
         np * sizeof(double);
}

int ozaki_dgemm(int m, int n, int K, const double* A, size_t lda, const double* B, size_t ldb, double* C, size_t ldc,
                bool accumulate, int S, void* work, hipStream_t s) {
  if (S < 2 || S > 8) return fail(-9, "split precision: slices must be 2..8");
  if (m <= 0 || n <= 0 || K <= 0) return 0;
  constexpr int CK = chunk_k<__bf16>();
  const int mp = (m + 63) / 64 * 64, np = (n + 63) / 64 * 64, Kp = (K + CK - 1) / CK * CK;
  // every partial sum of a level is an integer of magnitude <= S 2^12 Kp:
  // exact in FP32 while that stays <= 2^24
  if (static_cast<long long>(S) * 4096 * Kp > (1LL << 24))
    return fail(-9, "split precision: inner size too large for exact FP32 levels (bf16 slices)");
  const auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  char* w = static_cast<char*>(work);
  __bf16* As = reinterpret_cast<__bf16*>(w);
  w += al(static_cast<size_t>(S) * mp * Kp * sizeof(__bf16));
  __bf16* Bs = reinterpret_cast<__bf16*>(w);
  w += al(static_cast<size_t>(S) * np * Kp * sizeof(__bf16));
  double* sig = reinterpret_cast<double*>(w);
  w += al(static_cast<size_t>(mp) * sizeof(double));
  double* tau = reinterpret_cast<double*>(w);
  hipLaunchKernelGGL(k_row_scale8, dim3((m + 255) / 256), dim3(256), 0, s, m, K, A, lda, sig);
  hipLaunchKernelGGL(k_slice8_rows<__bf16>, dim3(mp / 64, Kp / CK), dim3(256), 0, s, m, K, mp, Kp, S, A, lda, sig, As);
  hipLaunchKernelGGL(k_col_scale8, dim3((n + 3) / 4), dim3(256), 0, s, n, K, B, ldb, tau);
  const size_t nb = static_cast<size_t>(np) * Kp;
  hipLaunchKernelGGL(k_slice8_cols<__bf16>, dim3(static_cast<unsigned>((nb + 255) / 256)), dim3(256), 0, s, n, K, np,
                     Kp, S, B, ldb, tau, Bs);
  launch_slices<__bf16>(dim3((np / 64) * (mp / 64)), s, S, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc,
                        accumulate ? 1 : 0, 0, 0);
  return hipGetLastError() == hipSuccess ? 0 : fail(-100, "split-precision GEMM launch failed");
}

// A side of the int8 split GEMM: slices and row scales, kept while A does
// not change.  Layout: As[S][mp][Kp] int8, then sig[mp] (256-byte aligned).
size_t ozaki8_a_bytes(int m, int K, int S) {
  const size_t mp = (static_cast<size_t>(m) + 63) / 64 * 64, Kp = (static_cast<size_t>(K) + 63) / 64 * 64;
  return (S * mp * Kp + 255) / 256 * 256 + mp * sizeof(double);
}

size_t ozaki8_b_bytes(int n, int K, int S) { return ozaki8_a_bytes(n, K, S); }

int ozaki8_prepare_a(int m, int K, const double* A, size_t lda, int S, void* abuf, hipStream_t s) {
  if (S < 2 || S > 8) return fail(-9, "split precision: slices must be 2..8");
  if (m <= 0 || K <= 0) return 0;
  const int mp = (m + 63) / 64 * 64, Kp = (K + 63) / 64 * 64;
  int8_t* As = static_cast<int8_t*>(abuf);
  double* sig = reinterpret_cast<double*>(As + (static_cast<size_t>(S) * mp * Kp + 255) / 256 * 256);
  hipLaunchKernelGGL(k_row_scale8, dim3((m + 255) / 256), dim3(256), 0, s, m, K, A, lda, sig);
  hipLaunchKernelGGL(k_slice8_rows<int8_t>, dim3(mp / 64, Kp / 64), dim3(256), 0, s, m, K, mp, Kp, S, A, lda, sig,
                     As);
  return hipGetLastError() == hipSuccess ? 0 : fail(-100, "split-precision slicing launch failed");
}

int ozaki8_dgemm(int m, int n, int K, const void* abuf, const double* B, size_t ldb, double* C, size_t ldc,
                 bool accumulate, int S, void* bbuf, hipStream_t s, int qns, int ldo) {
  if (qns > 0 && (static_cast<long long>(qns) * qns != m || ldo < qns)) return fail(-1, "split-precision GEMM: bad block layout");
  if (S < 2 || S > 8) return fail(-9, "split precision: slices must be 2..8");
  if (m <= 0 || n <= 0 || K <= 0) return 0;
  const int mp = (m + 63) / 64 * 64, np = (n + 63) / 64 * 64, Kp = (K + 63) / 64 * 64;
  // a level holds at most S slice products of magnitude <= 2^12 per inner
  // index: its int32 sum is exact while S 2^12 Kp < 2^31
  if (static_cast<long long>(S) * 4096 * Kp >= (1LL << 31))
    return fail(-9, "split precision: inner size too large for exact int32 levels");
  const int8_t* As = static_cast<const int8_t*>(abuf);
  const double* sig = reinterpret_cast<const double*>(As + (static_cast<size_t>(S) * mp * Kp + 255) / 256 * 256);
  int8_t* Bs = static_cast<int8_t*>(bbuf);
  double* tau = reinterpret_cast<double*>(Bs + (static_cast<size_t>(S) * np * Kp + 255) / 256 * 256);
  hipLaunchKernelGGL(k_col_scale8, dim3((n + 3) / 4), dim3(256), 0, s, n, K, B, ldb, tau);
  const size_t nb = static_cast<size_t>(np) * Kp;
  hipLaunchKernelGGL(k_slice8_cols<int8_t>, dim3(static_cast<unsigned>((nb + 255) / 256)), dim3(256), 0, s, n, K, np,
                     Kp, S, B, ldb, tau, Bs);
  launch_slices<int8_t>(dim3((np / 64) * (mp / 64)), s, S, m, n, mp, np, Kp, As, Bs, sig, tau, C, ldc,
                        accumulate ? 1 : 0, qns, ldo);
  return hipGetLastError() == hipSuccess ? 0 : fail(-100, "split-precision GEMM launch failed");
}

}  // namespace bsrg
