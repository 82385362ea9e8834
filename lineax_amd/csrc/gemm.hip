// C <- beta C + sign A B (beta in {0, 1}, sign = +-1) on the gfx950 matrix
// cores, float32 storage, with the products taken in bfloat16, float16 or
// float32 (FP32 accumulation in all), or in split bfloat16 ("bf16x3": each
// operand x = hi + lo with hi = bf16(x), lo = bf16(x - hi), and A B taken as
// A_hi B_hi + A_hi B_lo + A_lo B_hi, about 16 significant bits per operand
// for three bf16 MFMAs).  The Schur-complement updates of
// the blocked LU and its triangular block products run here.  With beta = 0
// the update may be done in place when one operand's tile is exactly the C
// tile it produces (M <= 128 with B aliasing C, or N <= 128 with A aliasing
// C): a workgroup reads all of its operand tiles before it writes C.
//
// Tiling: a 256-thread workgroup owns a 128 x 128 tile of C; each of its four
// waves a 64 x 64 quarter as 2 x 2 blocks of 32 x 32 MFMA accumulators
// (v_mfma_f32_32x32x16_{bf16,f16}, or v_mfma_f32_32x32x2_f32).  K advances 32
// at a time through double-buffered LDS tiles: A as [row][k] and B
// transposed as [col][k], so each lane's MFMA operand (8 consecutive k of one
// row or column) is one 16-byte LDS read; rows are padded by 8 elements.
// A is negated while it is converted when sign < 0; the MFMAs accumulate onto
// the C tile loaded into the accumulators (or onto zero when beta = 0).
//
// 32 x 32 MFMA operand and result layout (lane l, result register r):
//   A operand: row l % 32, k = kper * (l / 32) + [0, kper)
//   B operand: col l % 32, same k
//   D[r]:      row 8 (r / 4) + 4 (l / 32) + r % 4, col l % 32
// with kper = 8 for the 16-bit forms and 1 for float32.

#include "lxa_common.h"

namespace lxa {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

constexpr int kBM = 128, kBN = 128, kBK = 32;
constexpr int kGemmThreads = 256;

template <int MODE>
struct GemmMode;
template <>
struct GemmMode<LXA_GEMM_BF16> {
  using E = __bf16;
  static constexpr int kPad = 8;
};
template <>
struct GemmMode<LXA_GEMM_F16> {
  using E = _Float16;
  static constexpr int kPad = 8;
};
template <>
struct GemmMode<LXA_GEMM_F32> {
  using E = float;
  static constexpr int kPad = 1;
};
template <>
struct GemmMode<LXA_GEMM_BF16X3> {
  using E = __bf16;
  static constexpr int kPad = 8;
};

template <int MODE>
__device__ __forceinline__ void mfma_block(floatx16& acc, const typename GemmMode<MODE>::E* a_row,
                                           const typename GemmMode<MODE>::E* b_col, int lane) {
  if constexpr (MODE == LXA_GEMM_F32) {
    // 16 MFMAs of K = 2 cover the 32-wide k tile.
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 2) {
      const float av = a_row[kk + (lane >> 5)];
      const float bv = b_col[kk + (lane >> 5)];
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
    }
  } else {
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 16) {
      const int ko = kk + 8 * (lane >> 5);
      if constexpr (MODE == LXA_GEMM_BF16 || MODE == LXA_GEMM_BF16X3) {
        const bf16x8 av = *reinterpret_cast<const bf16x8*>(a_row + ko);
        const bf16x8 bv = *reinterpret_cast<const bf16x8*>(b_col + ko);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
      } else {
        const halfx8 av = *reinterpret_cast<const halfx8*>(a_row + ko);
        const halfx8 bv = *reinterpret_cast<const halfx8*>(b_col + ko);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, acc, 0, 0, 0);
      }
    }
  }
}

// TA: A is stored transposed (K x M, element (i, k) at A[k * lda + i]);
// TB: B is stored transposed (N x K, element (k, j) at B[j * ldb + k]).
template <int MODE, bool TA = false, bool TB = false>
__global__ __launch_bounds__(kGemmThreads) void gemm_kernel(int M, int N, int K, const float* A, int64_t lda,
                                                            const float* B, int64_t ldb, float* C, int64_t ldc,
                                                            int beta, float sign, int64_t sA, int64_t sB,
                                                            int64_t sC) {
  // system blockIdx.z of a batch (strides in elements; one system: z = 0)
  A += sA * blockIdx.z;
  B += sB * blockIdx.z;
  C += sC * blockIdx.z;
  using E = typename GemmMode<MODE>::E;
  constexpr int LDK = kBK + GemmMode<MODE>::kPad;
  constexpr bool kSplit = MODE == LXA_GEMM_BF16X3;
  constexpr int kLo = kSplit ? 2 : 1;  // [0] hi (or the only) part, [1] lo part
  __shared__ __align__(16) E sa[kLo][2][kBM][LDK];
  __shared__ __align__(16) E sb[kLo][2][kBN][LDK];

  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  const int wm = (wv >> 1) * 64, wn = (wv & 1) * 64;

  // accumulators <- C tile
  floatx16 acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
      const int col = n0 + wn + 32 * bj + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm + 32 * bi + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        acc[bi][bj][r] = (beta && row < M && col < N) ? C[(int64_t)row * ldc + col] : 0.f;
      }
    }

  // global -> register staging: A tile 128 x 32 and B tile 32 x 128, 4 float4 each per thread.
  // An operand whose k runs along its rows in memory is read 4 consecutive k
  // of one row per float4 (row idx / 8, k 4 (idx % 8)); one whose k runs down
  // its columns is read 4 consecutive rows at one k (k idx / 32, row 4 (idx % 32)).
  float4 ra[4], rb[4];
  auto load_k_along = [&](const float* P, int64_t ld, int r0, int rows, int k0, int idx) {
    const int row = r0 + (idx >> 3), k = k0 + 4 * (idx & 7);
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (row < rows) {
      const float* p = P + (int64_t)row * ld + k;
      if (k + 3 < K && ((reinterpret_cast<uintptr_t>(p) & 15) == 0)) {
        v = *reinterpret_cast<const float4*>(p);
      } else {
        v.x = k < K ? p[0] : 0.f;
        v.y = k + 1 < K ? p[1] : 0.f;
        v.z = k + 2 < K ? p[2] : 0.f;
        v.w = k + 3 < K ? p[3] : 0.f;
      }
    }
    return v;
  };
  auto load_k_down = [&](const float* P, int64_t ld, int r0, int rows, int k0, int idx) {
    const int k = k0 + (idx >> 5), row = r0 + 4 * (idx & 31);
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (k < K) {
      const float* p = P + (int64_t)k * ld + row;
      if (row + 3 < rows && ((reinterpret_cast<uintptr_t>(p) & 15) == 0)) {
        v = *reinterpret_cast<const float4*>(p);
      } else {
        v.x = row < rows ? p[0] : 0.f;
        v.y = row + 1 < rows ? p[1] : 0.f;
        v.z = row + 2 < rows ? p[2] : 0.f;
        v.w = row + 3 < rows ? p[3] : 0.f;
      }
    }
    return v;
  };
  auto load_tiles = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = tid + kGemmThreads * q;
      ra[q] = TA ? load_k_down(A, lda, m0, M, k0, idx) : load_k_along(A, lda, m0, M, k0, idx);
      rb[q] = TB ? load_k_along(B, ldb, n0, N, k0, idx) : load_k_down(B, ldb, n0, N, k0, idx);
    }
  };
  auto put = [&](E (*dst)[2][kBM][LDK], int buf, int r, int k, float v) {
    const E hi = (E)v;
    dst[0][buf][r][k] = hi;
    if constexpr (kSplit) dst[1][buf][r][k] = (E)(v - (float)hi);
  };
  // both tiles live in LDS as [row or column][k]
  auto put4 = [&](E (*dst)[2][kBM][LDK], int buf, bool k_along, int idx, float4 v, float sc) {
    if (k_along) {
      const int r = idx >> 3, k = 4 * (idx & 7);
      put(dst, buf, r, k + 0, sc * v.x);
      put(dst, buf, r, k + 1, sc * v.y);
      put(dst, buf, r, k + 2, sc * v.z);
      put(dst, buf, r, k + 3, sc * v.w);
    } else {
      const int k = idx >> 5, r = 4 * (idx & 31);
      put(dst, buf, r + 0, k, sc * v.x);
      put(dst, buf, r + 1, k, sc * v.y);
      put(dst, buf, r + 2, k, sc * v.z);
      put(dst, buf, r + 3, k, sc * v.w);
    }
  };
  auto store_tiles = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = tid + kGemmThreads * q;
      put4(sa, buf, !TA, idx, ra[q], sign);
      put4(sb, buf, TB, idx, rb[q], 1.f);
    }
  };

  const int ktiles = (K + kBK - 1) / kBK;
  load_tiles(0);
  store_tiles(0);
  __syncthreads();
  for (int kt = 0; kt < ktiles; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < ktiles) load_tiles((kt + 1) * kBK);
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) {
        const E* ah = &sa[0][buf][wm + 32 * bi + (lane & 31)][0];
        const E* bh = &sb[0][buf][wn + 32 * bj + (lane & 31)][0];
        if constexpr (kSplit) {
          // small terms first: lo parts, then hi x hi
          mfma_block<MODE>(acc[bi][bj], ah, &sb[1][buf][wn + 32 * bj + (lane & 31)][0], lane);
          mfma_block<MODE>(acc[bi][bj], &sa[1][buf][wm + 32 * bi + (lane & 31)][0], bh, lane);
        }
        mfma_block<MODE>(acc[bi][bj], ah, bh, lane);
      }
    if (kt + 1 < ktiles) store_tiles(buf ^ 1);
    __syncthreads();
  }

#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
      const int col = n0 + wn + 32 * bj + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm + 32 * bi + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (row < M && col < N) C[(int64_t)row * ldc + col] = acc[bi][bj][r];
      }
    }
}

}  // namespace

int gemm_update_batched(hipStream_t st, int mode, int batch, int M, int N, int K, const float* A, int64_t lda,
                        int64_t sA, const float* B, int64_t ldb, int64_t sB, float* C, int64_t ldc, int64_t sC,
                        int beta, float sign) {
  if (M <= 0 || N <= 0 || batch <= 0) return LXA_OK;
  if (K <= 0 && beta) return LXA_OK;
  const dim3 grid((unsigned)((N + kBN - 1) / kBN), (unsigned)((M + kBM - 1) / kBM), (unsigned)batch);
  switch (mode) {
    case LXA_GEMM_BF16:
      gemm_kernel<LXA_GEMM_BF16><<<grid, kGemmThreads, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, beta, sign, sA, sB,
                                                               sC);
      break;
    case LXA_GEMM_F16:
      gemm_kernel<LXA_GEMM_F16><<<grid, kGemmThreads, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, beta, sign, sA, sB,
                                                              sC);
      break;
    case LXA_GEMM_F32:
      gemm_kernel<LXA_GEMM_F32><<<grid, kGemmThreads, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, beta, sign, sA, sB,
                                                              sC);
      break;
    case LXA_GEMM_F32 | LXA_GEMM_TRANS_A:
      gemm_kernel<LXA_GEMM_F32, true, false><<<grid, kGemmThreads, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, beta,
                                                                           sign, sA, sB, sC);
      break;
    case LXA_GEMM_F32 | LXA_GEMM_TRANS_B:
      gemm_kernel<LXA_GEMM_F32, false, true><<<grid, kGemmThreads, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, beta,
                                                                           sign, sA, sB, sC);
      break;
    case LXA_GEMM_F32 | LXA_GEMM_TRANS_A | LXA_GEMM_TRANS_B:
      gemm_kernel<LXA_GEMM_F32, true, true><<<grid, kGemmThreads, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, beta,
                                                                          sign, sA, sB, sC);
      break;
    case LXA_GEMM_BF16X3:
      gemm_kernel<LXA_GEMM_BF16X3><<<grid, kGemmThreads, 0, st>>>(M, N, K, A, lda, B, ldb, C, ldc, beta, sign, sA,
                                                                 sB, sC);
      break;
    default:
      return set_error(LXA_EINVAL, "unknown gemm mode %d", mode);
  }
  LXA_LAUNCH_CHECK("gemm_kernel");
  return LXA_OK;
}

int gemm_update(hipStream_t st, int mode, int M, int N, int K, const float* A, int64_t lda, const float* B,
                int64_t ldb, float* C, int64_t ldc, int beta, float sign) {
  return gemm_update_batched(st, mode, 1, M, N, K, A, lda, 0, B, ldb, 0, C, ldc, 0, beta, sign);
}

}  // namespace lxa

extern "C" int lxa_gemm_sub(void* stream, int mode, int64_t m, int64_t n, int64_t k, const float* a, int64_t lda,
                            const float* b, int64_t ldb, float* c, int64_t ldc) {
  LXA_REQUIRE(m >= 0 && n >= 0 && k >= 0 && m < (1 << 30) && n < (1 << 30) && k < (1 << 30),
              "invalid sizes m=%lld n=%lld k=%lld", (long long)m, (long long)n, (long long)k);
  if (m == 0 || n == 0 || k == 0) return LXA_OK;
  LXA_REQUIRE(a && b && c, "a, b and c must be non-NULL");
  LXA_REQUIRE(mode >= LXA_GEMM_F32 && mode <= LXA_GEMM_BF16X3, "lxa_gemm_sub: unknown mode %d", mode);
  LXA_REQUIRE(lda >= k && ldb >= n && ldc >= n, "leading dimensions too small (lda=%lld ldb=%lld ldc=%lld)",
              (long long)lda, (long long)ldb, (long long)ldc);
  return lxa::gemm_update(static_cast<hipStream_t>(stream), mode, (int)m, (int)n, (int)k, a, lda, b, ldb, c, ldc, 1,
                          -1.f);
}

extern "C" int lxa_gemm_batched(void* stream, int mode, int64_t batch, int64_t m, int64_t n, int64_t k,
                                const float* a, int64_t lda, int64_t a_stride, const float* b, int64_t ldb,
                                int64_t b_stride, float* c, int64_t ldc, int64_t c_stride) {
  LXA_REQUIRE(batch >= 0 && batch < 65536 && m >= 0 && n >= 0 && k >= 0 && m < (1 << 30) && n < (1 << 30) &&
                  k < (1 << 30),
              "invalid sizes batch=%lld m=%lld n=%lld k=%lld", (long long)batch, (long long)m, (long long)n,
              (long long)k);
  if (batch == 0 || m == 0 || n == 0) return LXA_OK;
  LXA_REQUIRE(a && b && c, "a, b and c must be non-NULL");
  LXA_REQUIRE((mode & ~(LXA_GEMM_TRANS_A | LXA_GEMM_TRANS_B)) == LXA_GEMM_F32 ||
                  (mode & (LXA_GEMM_TRANS_A | LXA_GEMM_TRANS_B)) == 0,
              "transposed operands are supported in float32 mode only (mode %d)", mode);
  const int64_t a_ld_min = (mode & LXA_GEMM_TRANS_A) ? m : k, b_ld_min = (mode & LXA_GEMM_TRANS_B) ? k : n;
  LXA_REQUIRE(lda >= a_ld_min && ldb >= b_ld_min && ldc >= n,
              "leading dimensions too small (lda=%lld ldb=%lld ldc=%lld)", (long long)lda, (long long)ldb,
              (long long)ldc);
  LXA_REQUIRE(a_stride >= 0 && b_stride >= 0 && c_stride >= m * ldc, "bad batch strides");
  return lxa::gemm_update_batched(static_cast<hipStream_t>(stream), mode, (int)batch, (int)m, (int)n, (int)k, a, lda,
                                  a_stride, b, ldb, b_stride, c, ldc, c_stride, 0, 1.f);
}
