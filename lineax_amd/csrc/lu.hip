// Batched LU factorisation (getrf, partial pivoting) and solve (getrs).
//
// Small orders (n <= 64): one lane per matrix row, held in registers; a wave
// holds 64/LPS systems (LPS = lanes per system, the next power of two >= n).
// The pivot search is a shuffle reduction over the system's lanes and the
// pivot row is broadcast through a double-buffered LDS row.  Row
// interchanges are implicit: every lane tracks the LAPACK position of its
// row, and rows are scattered to their final positions on store.  The
// result (LU in place plus 0-based interchange sequence) is exactly
// LAPACK's getrf output format, so jax.scipy's lu_solve semantics apply.
//
// Larger orders use a one-workgroup-per-system right-looking kernel working
// in global memory (correctness path for moderate n; single large float32
// systems take the blocked path of lu_large.hip and gemm.hip instead).
#include <climits>

#include "lxa_common.h"

namespace lxa {
namespace {

template <int N>
struct SmallLanes {
  static constexpr int value = N <= 4 ? 4 : N <= 8 ? 8 : N <= 16 ? 16 : N <= 32 ? 32 : 64;
};

// Total order on |v| with NaN largest, so every lane of a reduction agrees.
__device__ __forceinline__ uint32_t abs_key(float v) { return __float_as_uint(v) & 0x7fffffffu; }
__device__ __forceinline__ uint64_t abs_key(double v) {
  return static_cast<uint64_t>(__double_as_longlong(v)) & 0x7fffffffffffffffull;
}

template <typename T>
struct KeyOf;
template <>
struct KeyOf<float> {
  using type = uint32_t;
};
template <>
struct KeyOf<double> {
  using type = uint64_t;
};

// Orders n <= N are embedded as diag(A, I): the identity rows never win a
// pivot search over a real column (they hold 0 there and sit at higher
// positions), so the real block's factors and interchanges are exactly
// those of the unpadded matrix.
template <typename T, int N>
__global__ __launch_bounds__(64) void lu_factor_small_kernel(int64_t batch, int n, T* __restrict__ a,
                                                             int64_t a_stride, int32_t* __restrict__ piv,
                                                             int32_t* __restrict__ info) {
  constexpr int LPS = SmallLanes<N>::value;
  constexpr int SPW = 64 / LPS;
  using Key = typename KeyOf<T>::type;
  __shared__ T prow[2][SPW][N];

  const int lane = threadIdx.x;
  const int g = lane / LPS;
  const int r = lane % LPS;
  const int64_t sys = static_cast<int64_t>(blockIdx.x) * SPW + g;
  const bool valid = sys < batch;
  const bool owns_row = valid && r < N;  // participates (real or padding row)
  const bool real_row = valid && r < n;   // row of the caller's matrix

  T row[N];
  if (real_row) {
    const T* src = a + sys * a_stride + static_cast<int64_t>(r) * n;
#pragma unroll
    for (int j = 0; j < N; ++j) row[j] = (j < n) ? src[j] : T(0);
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) row[j] = (j == r) ? T(1) : T(0);
  }

  int pos = r;             // LAPACK row position currently holding this lane's row
  bool done = !owns_row;   // rows already used as pivots (and padding lanes)
  int mypiv = r;           // piv[r], recorded by lane r at step r
  int zero_at = 0;         // LAPACK info: first exactly-zero pivot (1-based)

#pragma unroll
  for (int k = 0; k < N; ++k) {
    // Pivot: largest |a(i,k)| over remaining rows, ties to the smallest
    // LAPACK position (idamax returns the first maximum).
    Key ck = done ? Key(0) : abs_key(row[k]);
    int cp = done ? INT_MAX : pos;
#pragma unroll
    for (int off = LPS / 2; off > 0; off >>= 1) {
      const Key ok = __shfl_xor(ck, off, LPS);
      const int op = __shfl_xor(cp, off, LPS);
      if (ok > ck || (ok == ck && op < cp)) {
        ck = ok;
        cp = op;
      }
    }
    const bool is_piv = !done && pos == cp;
    if (is_piv) {
#pragma unroll
      for (int j = 0; j < N; ++j) prow[k & 1][g][j] = row[j];
    }
    __syncthreads();
    const T pval = prow[k & 1][g][k];
    if (r == k) mypiv = cp;
    if (k < n && pval == T(0) && zero_at == 0) zero_at = k + 1;
    if (is_piv) {
      pos = k;
      done = true;
    } else if (!done && pos == k) {
      pos = cp;  // the row at position k moves to the pivot's old position
    }
    if (!done) {
      const T l = (pval != T(0)) ? row[k] / pval : row[k];
      row[k] = l;
#pragma unroll
      for (int j = k + 1; j < N; ++j) row[j] -= l * prow[k & 1][g][j];
    }
  }

  if (real_row) {
    // Real rows end at real positions: padding rows are never pivots for
    // k < n, and the lane at position k < n only moves to a pivot's position.
    T* dst = a + sys * a_stride + static_cast<int64_t>(pos) * n;
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (j < n) dst[j] = row[j];
    piv[sys * n + r] = mypiv;
  }
  if (valid && r == 0) info[sys] = zero_at;
}

template <typename T, int N>
__global__ __launch_bounds__(64) void lu_solve_small_kernel(int64_t batch, int n, const T* __restrict__ lu,
                                                            int64_t lu_stride, const int32_t* __restrict__ piv,
                                                            int64_t piv_stride, T* __restrict__ b,
                                                            int64_t b_stride, int trans) {
  constexpr int LPS = SmallLanes<N>::value;
  constexpr int SPW = 64 / LPS;
  __shared__ T buf[SPW][N];
  __shared__ int pv[SPW][N];

  const int lane = threadIdx.x;
  const int g = lane / LPS;
  const int r = lane % LPS;
  const int64_t sys = static_cast<int64_t>(blockIdx.x) * SPW + g;
  const bool owns_row = sys < batch && r < N;
  const bool real_row = sys < batch && r < n;

  if (owns_row) {
    int q = r;  // padding steps interchange nothing
    if (real_row) {
      q = piv[sys * piv_stride + r];
      q = q < 0 ? 0 : (q >= n ? n - 1 : q);
    }
    pv[g][r] = q;
  }
  __syncthreads();

  T m[N];
  const T* M = lu + sys * lu_stride;
  T* B = b + sys * b_stride;

  if (trans == 0) {
    // Lane r holds row r of LU (identity rows for padding).
    if (real_row) {
#pragma unroll
      for (int j = 0; j < N; ++j) m[j] = (j < n) ? M[r * n + j] : T(0);
    } else {
#pragma unroll
      for (int j = 0; j < N; ++j) m[j] = (j == r) ? T(1) : T(0);
    }
    // Apply the interchanges to b: track where b[r] ends up.
    if (owns_row) {
      int p = r;
      for (int k = 0; k < N; ++k) {
        const int q = pv[g][k];
        p = (p == k) ? q : ((p == q) ? k : p);
      }
      buf[g][p] = real_row ? B[r] : T(0);
    }
    __syncthreads();
    T y = owns_row ? buf[g][r] : T(0);
    // L y = P b (unit lower), column oriented.
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const T yj = __shfl(y, j, LPS);
      if (r > j) y -= m[j] * yj;
    }
    // U x = y.
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
      const T t = y / m[j];
      const T xj = __shfl(t, j, LPS);
      if (r < j) {
        y -= m[j] * xj;
      } else if (r == j) {
        y = xj;
      }
    }
    if (real_row) B[r] = y;
  } else {
    // Lane c holds column c of LU.  A^T = U^T L^T P.
    if (real_row) {
#pragma unroll
      for (int i = 0; i < N; ++i) m[i] = (i < n) ? M[i * n + r] : T(0);
    } else {
#pragma unroll
      for (int i = 0; i < N; ++i) m[i] = (i == r) ? T(1) : T(0);
    }
    T y = real_row ? B[r] : T(0);
    // U^T z = b.
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const T t = y / m[j];
      const T zj = __shfl(t, j, LPS);
      if (r > j) {
        y -= m[j] * zj;
      } else if (r == j) {
        y = zj;
      }
    }
    // L^T w = z (unit upper).
#pragma unroll
    for (int j = N - 1; j >= 0; --j) {
      const T wj = __shfl(y, j, LPS);
      if (r < j) y -= m[j] * wj;
    }
    // x = P^T w: undo the interchanges in reverse order.  Every read of B
    // happened before the first shuffle above consumed it.
    int p = r;
    if (owns_row) {
      for (int k = N - 1; k >= 0; --k) {
        const int q = pv[g][k];
        p = (p == k) ? q : ((p == q) ? k : p);
      }
    }
    if (real_row) B[p] = y;
  }
}

// ---------------------------------------------------------------------------
// Generic orders: one 256-thread workgroup per system, in global memory.
// ---------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void lu_factor_generic_kernel(int n, T* __restrict__ a, int64_t a_stride,
                                                                int32_t* __restrict__ piv,
                                                                int32_t* __restrict__ info) {
  using Key = typename KeyOf<T>::type;
  const int64_t sys = blockIdx.x;
  T* A = a + sys * a_stride;
  int32_t* P = piv + sys * n;
  __shared__ Key s_key[4];
  __shared__ int s_pos[4];
  __shared__ int s_p;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t ld = n;
  int zero_at = 0;
  for (int k = 0; k < n; ++k) {
    Key best = 0;
    int bpos = INT_MAX;
    for (int i = k + tid; i < n; i += 256) {
      const Key kk = abs_key(A[i * ld + k]);
      if (kk > best || (kk == best && i < bpos)) {
        best = kk;
        bpos = i;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const Key ok = __shfl_xor(best, off, 64);
      const int op = __shfl_xor(bpos, off, 64);
      if (ok > best || (ok == best && op < bpos)) {
        best = ok;
        bpos = op;
      }
    }
    if (lane == 0) {
      s_key[w] = best;
      s_pos[w] = bpos;
    }
    __syncthreads();
    if (tid == 0) {
      Key bk = s_key[0];
      int bp = s_pos[0];
      for (int q = 1; q < 4; ++q) {
        if (s_key[q] > bk || (s_key[q] == bk && s_pos[q] < bp)) {
          bk = s_key[q];
          bp = s_pos[q];
        }
      }
      s_p = bp;
      P[k] = bp;
    }
    __syncthreads();
    const int p = s_p;
    if (p != k) {
      for (int j = tid; j < n; j += 256) {
        const T t = A[k * ld + j];
        A[k * ld + j] = A[p * ld + j];
        A[p * ld + j] = t;
      }
    }
    __syncthreads();
    const T pval = A[k * ld + k];
    if (pval == T(0) && zero_at == 0) zero_at = k + 1;
    if (pval != T(0)) {
      for (int i = k + 1 + tid; i < n; i += 256) A[i * ld + k] /= pval;
    }
    __syncthreads();
    const int m = n - k - 1;
    const int64_t mm = static_cast<int64_t>(m) * m;
    for (int64_t t = tid; t < mm; t += 256) {
      const int64_t i = k + 1 + t / m, j = k + 1 + t % m;
      A[i * ld + j] -= A[i * ld + k] * A[k * ld + j];
    }
    __syncthreads();
  }
  if (tid == 0) info[sys] = zero_at;
}

template <typename T>
__global__ __launch_bounds__(256) void lu_solve_generic_kernel(int n, const T* __restrict__ lu,
                                                               int64_t lu_stride, const int32_t* __restrict__ piv,
                                                               int64_t piv_stride, T* __restrict__ b,
                                                               int64_t b_stride, int trans) {
  extern __shared__ unsigned char lxa_smem[];
  T* v = reinterpret_cast<T*>(lxa_smem);
  const int64_t sys = blockIdx.x;
  const T* A = lu + sys * lu_stride;
  const int32_t* P = piv + sys * piv_stride;
  T* B = b + sys * b_stride;
  const int tid = threadIdx.x;
  const int64_t ld = n;
  for (int i = tid; i < n; i += 256) v[i] = B[i];
  __syncthreads();
  if (trans == 0) {
    if (tid == 0) {
      for (int k = 0; k < n; ++k) {
        int q = P[k];
        q = q < 0 ? 0 : (q >= n ? n - 1 : q);
        const T t = v[k];
        v[k] = v[q];
        v[q] = t;
      }
    }
    __syncthreads();
    for (int j = 0; j < n; ++j) {
      const T yj = v[j];
      for (int i = j + 1 + tid; i < n; i += 256) v[i] -= A[i * ld + j] * yj;
      __syncthreads();
    }
    for (int j = n - 1; j >= 0; --j) {
      const T xj = v[j] / A[j * ld + j];
      for (int i = tid; i < j; i += 256) v[i] -= A[i * ld + j] * xj;
      if (tid == 0) B[j] = xj;
      __syncthreads();
    }
  } else {
    // U^T z = b; z_j is final once read at step j and is written to B[j].
    for (int j = 0; j < n; ++j) {
      const T zj = v[j] / A[j * ld + j];
      for (int i = j + 1 + tid; i < n; i += 256) v[i] -= A[j * ld + i] * zj;
      if (tid == 0) B[j] = zj;
      __syncthreads();
    }
    for (int i = tid; i < n; i += 256) v[i] = B[i];
    __syncthreads();
    // L^T w = z (unit upper); w_j is final when read at step j.
    for (int j = n - 1; j >= 0; --j) {
      const T wj = v[j];
      for (int i = tid; i < j; i += 256) v[i] -= A[j * ld + i] * wj;
      __syncthreads();
    }
    if (tid == 0) {
      for (int k = n - 1; k >= 0; --k) {
        int q = P[k];
        q = q < 0 ? 0 : (q >= n ? n - 1 : q);
        const T t = v[k];
        v[k] = v[q];
        v[q] = t;
      }
    }
    __syncthreads();
    for (int i = tid; i < n; i += 256) B[i] = v[i];
  }
}

template <typename T, int N>
void launch_factor_small(hipStream_t s, int64_t batch, int n, T* a, int64_t stride, int32_t* piv, int32_t* info) {
  constexpr int SPW = 64 / SmallLanes<N>::value;
  const int64_t blocks = (batch + SPW - 1) / SPW;
  hipLaunchKernelGGL((lu_factor_small_kernel<T, N>), dim3(static_cast<unsigned>(blocks)), dim3(64), 0, s,
                     batch, n, a, stride, piv, info);
}

template <typename T, int N>
void launch_solve_small(hipStream_t s, int64_t batch, int n, const T* lu, int64_t lus, const int32_t* piv,
                        int64_t ps, T* b, int64_t bs, int trans) {
  constexpr int SPW = 64 / SmallLanes<N>::value;
  const int64_t blocks = (batch + SPW - 1) / SPW;
  hipLaunchKernelGGL((lu_solve_small_kernel<T, N>), dim3(static_cast<unsigned>(blocks)), dim3(64), 0, s,
                     batch, n, lu, lus, piv, ps, b, bs, trans);
}

// Small orders are padded to the next of 8, 16, 32, 64.
template <typename T>
void dispatch_factor_small(int n, hipStream_t s, int64_t batch, T* a, int64_t stride, int32_t* piv,
                           int32_t* info) {
  if (n <= 8) launch_factor_small<T, 8>(s, batch, n, a, stride, piv, info);
  else if (n <= 16) launch_factor_small<T, 16>(s, batch, n, a, stride, piv, info);
  else if (n <= 32) launch_factor_small<T, 32>(s, batch, n, a, stride, piv, info);
  else launch_factor_small<T, 64>(s, batch, n, a, stride, piv, info);
}

template <typename T>
void dispatch_solve_small(int n, hipStream_t s, int64_t batch, const T* lu, int64_t lus, const int32_t* piv,
                          int64_t ps, T* b, int64_t bs, int trans) {
  if (n <= 8) launch_solve_small<T, 8>(s, batch, n, lu, lus, piv, ps, b, bs, trans);
  else if (n <= 16) launch_solve_small<T, 16>(s, batch, n, lu, lus, piv, ps, b, bs, trans);
  else if (n <= 32) launch_solve_small<T, 32>(s, batch, n, lu, lus, piv, ps, b, bs, trans);
  else launch_solve_small<T, 64>(s, batch, n, lu, lus, piv, ps, b, bs, trans);
}

constexpr int kSmallMax = 64;
constexpr int64_t kMaxGrid = 2147483647;

template <typename T>
int lu_factor_t(hipStream_t s, int64_t batch, int64_t n, T* a, int64_t stride, int32_t* piv, int32_t* info) {
  if (n <= kSmallMax) {
    dispatch_factor_small<T>(static_cast<int>(n), s, batch, a, stride, piv, info);
  } else {
    hipLaunchKernelGGL((lu_factor_generic_kernel<T>), dim3(static_cast<unsigned>(batch)), dim3(256), 0, s,
                       static_cast<int>(n), a, stride, piv, info);
  }
  LXA_LAUNCH_CHECK("lxa_lu_factor launch");
  return LXA_OK;
}

template <typename T>
int lu_solve_t(hipStream_t s, int64_t batch, int64_t n, const T* lu, int64_t lus, const int32_t* piv,
               int64_t ps, T* b, int64_t bs, int trans) {
  if (n <= kSmallMax) {
    dispatch_solve_small<T>(static_cast<int>(n), s, batch, lu, lus, piv, ps, b, bs, trans);
  } else {
    const size_t shmem = static_cast<size_t>(n) * sizeof(T);
    if (shmem > 160 * 1024) return set_error(LXA_EUNSUPPORTED, "lxa_lu_solve: n=%lld too large", (long long)n);
    if (shmem > 64 * 1024) {
      hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(lu_solve_generic_kernel<T>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(shmem));
      if (e != hipSuccess) return hip_error(e, "hipFuncSetAttribute(MaxDynamicSharedMemorySize)");
    }
    hipLaunchKernelGGL((lu_solve_generic_kernel<T>), dim3(static_cast<unsigned>(batch)), dim3(256), shmem, s,
                       static_cast<int>(n), lu, lus, piv, ps, b, bs, trans);
  }
  LXA_LAUNCH_CHECK("lxa_lu_solve launch");
  return LXA_OK;
}

}  // namespace
}  // namespace lxa

extern "C" int lxa_lu_factor(void* stream, int dtype, int64_t batch, int64_t n, void* a, int64_t a_batch_stride,
                             int32_t* piv, int32_t* info) {
  LXA_REQUIRE(batch >= 0 && batch <= lxa::kMaxGrid, "lxa_lu_factor: bad batch %lld", (long long)batch);
  LXA_REQUIRE(n >= 1 && n <= 65536, "lxa_lu_factor: bad n %lld", (long long)n);
  LXA_REQUIRE(a && piv && info, "lxa_lu_factor: null pointer");
  LXA_REQUIRE(a_batch_stride >= n * n || batch <= 1, "lxa_lu_factor: a_batch_stride %lld < n*n",
              (long long)a_batch_stride);
  if (batch == 0) return LXA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == LXA_F32) return lxa::lu_factor_t<float>(s, batch, n, static_cast<float*>(a), a_batch_stride, piv, info);
  if (dtype == LXA_F64)
    return lxa::lu_factor_t<double>(s, batch, n, static_cast<double*>(a), a_batch_stride, piv, info);
  return lxa::set_error(LXA_EINVAL, "lxa_lu_factor: bad dtype %d", dtype);
}

extern "C" int lxa_lu_solve(void* stream, int dtype, int64_t batch, int64_t n, const void* lu,
                            int64_t lu_batch_stride, const int32_t* piv, int64_t piv_batch_stride, void* b,
                            int64_t b_batch_stride, int trans) {
  LXA_REQUIRE(batch >= 0 && batch <= lxa::kMaxGrid, "lxa_lu_solve: bad batch %lld", (long long)batch);
  LXA_REQUIRE(n >= 1 && n <= 65536, "lxa_lu_solve: bad n %lld", (long long)n);
  LXA_REQUIRE(lu && piv && b, "lxa_lu_solve: null pointer");
  LXA_REQUIRE(trans == 0 || trans == 1, "lxa_lu_solve: trans must be 0 or 1");
  LXA_REQUIRE(lu_batch_stride >= 0 && piv_batch_stride >= 0 && b_batch_stride >= 0, "lxa_lu_solve: negative stride");
  LXA_REQUIRE(b_batch_stride >= n || batch <= 1, "lxa_lu_solve: b_batch_stride < n");
  if (batch == 0) return LXA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == LXA_F32)
    return lxa::lu_solve_t<float>(s, batch, n, static_cast<const float*>(lu), lu_batch_stride, piv, piv_batch_stride,
                                  static_cast<float*>(b), b_batch_stride, trans);
  if (dtype == LXA_F64)
    return lxa::lu_solve_t<double>(s, batch, n, static_cast<const double*>(lu), lu_batch_stride, piv,
                                   piv_batch_stride, static_cast<double*>(b), b_batch_stride, trans);
  return lxa::set_error(LXA_EINVAL, "lxa_lu_solve: bad dtype %d", dtype);
}
