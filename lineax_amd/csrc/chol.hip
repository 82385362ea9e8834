// Cholesky factorisation (potrf, upper: A = U^T U) and solve (potrs).
//
// Three tiers, chosen by order n:
//  * n <= 64: one lane per row in registers, 64/LPS systems per wave; the
//    scaled pivot row is broadcast through a double-buffered LDS row.
//  * n <= 256 (float) / 181 (double): one 256-thread workgroup per system.
//    Float factors run the register-tiled 16-column panel factorisation of
//    chol_uneven.hip (chol_factor_regs_f32); double factors, and all solves,
//    keep the upper triangle resident in LDS in row-packed form (131.6 KB
//    at n = 256): right-looking factorisation, triangular solves by one
//    wave with the vector in registers.
//  * larger n: one workgroup per system in global memory (correctness path).
//
// Failure semantics follow jax.scipy.linalg.cho_factor as lineax uses it
// (lineax/_solver/cholesky.py:60): a non-positive or NaN pivot makes the
// whole factor NaN, so the later solve yields NaN and lineax reports
// RESULTS.singular (lineax/_solve.py:104-113).
#include "lxa_common.h"

namespace lxa {
namespace {

template <int N>
struct CholLanes {
  static constexpr int value = N <= 8 ? 8 : N <= 16 ? 16 : N <= 32 ? 32 : 64;
};

// ---------------------------------------------------------------------------
// Tier 1: n <= 64, lane per row.  Orders n < N are padded with an identity
// block, which factors to itself and leaves the real block unchanged.
// ---------------------------------------------------------------------------
template <typename T, int N>
__global__ __launch_bounds__(64) void chol_factor_small_kernel(int64_t batch, int n, T* __restrict__ a,
                                                               int64_t a_stride, int negate,
                                                               int32_t* __restrict__ info) {
  constexpr int LPS = CholLanes<N>::value;
  constexpr int SPW = 64 / LPS;
  __shared__ T prow[2][SPW][N];
  const int lane = threadIdx.x;
  const int g = lane / LPS;
  const int r = lane % LPS;
  const int64_t sys = static_cast<int64_t>(blockIdx.x) * SPW + g;
  const bool valid = sys < batch;
  const bool real_row = valid && r < n;
  T* A = a + sys * a_stride;

  T row[N];
  if (real_row) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
      T v = T(0);
      if (j >= r && j < n) v = A[static_cast<int64_t>(r) * n + j];  // upper triangle only
      row[j] = negate ? -v : v;
    }
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) row[j] = (j == r) ? T(1) : T(0);
  }

  int fail = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    if (r == k) {
      const T s = sqrt(row[k]);  // NaN for a negative pivot
#pragma unroll
      for (int j = k; j < N; ++j) prow[k & 1][g][j] = (j == k) ? s : row[j] / s;
    }
    __syncthreads();
    const T ukk = prow[k & 1][g][k];
    if (k < n && !(ukk > T(0)) && fail == 0) fail = k + 1;
    if (r == k) {
#pragma unroll
      for (int j = k; j < N; ++j) row[j] = prow[k & 1][g][j];
    } else if (r > k && r < N) {
      const T ukr = prow[k & 1][g][r];
#pragma unroll
      for (int j = k + 1; j < N; ++j) row[j] -= ukr * prow[k & 1][g][j];
    }
  }

  if (real_row) {
    const T nan = quiet_nan<T>();
#pragma unroll
    for (int j = 0; j < N; ++j) {
      if (j < n) A[static_cast<int64_t>(r) * n + j] = fail ? nan : (j < r ? T(0) : row[j]);
    }
  }
  if (valid && r == 0) info[sys] = fail;
}

template <typename T, int N>
__global__ __launch_bounds__(64) void chol_solve_small_kernel(int64_t batch, int n, const T* __restrict__ u,
                                                              int64_t u_stride, T* __restrict__ b,
                                                              int64_t b_stride, int negate) {
  constexpr int LPS = CholLanes<N>::value;
  const int lane = threadIdx.x;
  const int g = lane / LPS;
  const int r = lane % LPS;
  const int64_t sys = static_cast<int64_t>(blockIdx.x) * (64 / LPS) + g;
  const bool real_row = sys < batch && r < n;
  const T* U = u + sys * u_stride;
  T* B = b + sys * b_stride;

  T m[N];
  // Forward U^T y = b with lane c holding column c of U.
  if (real_row) {
#pragma unroll
    for (int i = 0; i < N; ++i) m[i] = (i <= r) ? U[static_cast<int64_t>(i) * n + r] : T(0);
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) m[i] = (i == r) ? T(1) : T(0);
  }
  T y = real_row ? B[r] : T(0);
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const T t = y / m[j];
    const T yj = __shfl(t, j, LPS);
    if (r > j) {
      y -= m[j] * yj;
    } else if (r == j) {
      y = yj;
    }
  }
  // Backward U x = y with lane r holding row r of U.
  if (real_row) {
#pragma unroll
    for (int j = 0; j < N; ++j) m[j] = (j >= r && j < n) ? U[static_cast<int64_t>(r) * n + j] : T(0);
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) m[j] = (j == r) ? T(1) : T(0);
  }
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
  if (real_row) B[r] = negate ? -y : y;
}

// ---------------------------------------------------------------------------
// Tier 2: one 256-thread workgroup per system, upper triangle in LDS.
// ---------------------------------------------------------------------------
constexpr int kWG = 256;

__device__ __forceinline__ float bcast_lane(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double bcast_lane(double v, int lane) {
  const int64_t bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits & 0xffffffff), lane);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), lane);
  return __longlong_as_double((static_cast<int64_t>(hi) << 32) | static_cast<uint32_t>(lo));
}

// Loads the upper triangle of A (row-major, leading dimension lda) into the
// row-packed LDS triangle L.  Wave w takes rows w, w+4, ...; lanes cover the
// columns, and four rows are in flight per wave.
template <typename T, typename TA>
__device__ void lds_load_upper(int n, const TA* __restrict__ A, int64_t lda, T* L, bool negate) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (int i0 = w; i0 < n; i0 += 16) {
    TA v[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + 4 * u;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int j = l + 64 * c;
        v[u][c] = (i < n && j >= i && j < n) ? A[static_cast<int64_t>(i) * lda + j] : TA(0);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + 4 * u;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int j = l + 64 * c;
        if (i < n && j >= i && j < n) {
          const T x = static_cast<T>(v[u][c]);
          L[packed_upper(n, i, j)] = negate ? -x : x;
        }
      }
    }
  }
}

// Right-looking factorisation of the packed triangle in place.  Returns 0 or
// the 1-based index of the first non-positive pivot (uniform over the block).
template <typename T>
__device__ int lds_chol_factor(int n, T* L) {
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  for (int k = 0; k < n; ++k) {
    const int rk = packed_upper(n, k, k);
    const T akk = L[rk];
    if (!(akk > T(0))) return k + 1;
    const T ukk = sqrt(akk);
    for (int j = k + 1 + tid; j < n; j += kWG) L[rk + (j - k)] /= ukk;
    __syncthreads();
    if (tid == 0) L[rk] = ukk;  // nobody reads L[k][k] until the next barrier
    const int m = n - k - 1;
    for (int ii = ty; ii < m; ii += 16) {
      const int i = k + 1 + ii;
      const T uki = L[rk + (i - k)];
      const int ri = packed_upper(n, i, i);
      int jj = ii + ((tx - ii) & 15);  // first jj >= ii with jj == tx (mod 16)
      for (; jj < m; jj += 16) {
        const int j = k + 1 + jj;
        L[ri + (j - i)] -= uki * L[rk + (j - k)];
      }
    }
    __syncthreads();
  }
  return 0;
}

// Solves U^T U x = v in place, v distributed over one wave: lane l holds
// v[l + 64 c].  Must be called by a whole wave with uniform control flow.
template <typename T>
__device__ void wave_chol_solve(int n, const T* L, T (&v)[4]) {
  const int l = threadIdx.x & 63;
  // U^T y = v, column oriented: y_k = v_k / U_kk, then v_j -= U_kj y_k.
#pragma unroll
  for (int c0 = 0; c0 < 4; ++c0) {
    const int kend = min(64, n - 64 * c0);
    for (int kk = 0; kk < kend; ++kk) {
      const int k = 64 * c0 + kk;
      const int rk = packed_upper(n, k, k);
      const T yk = bcast_lane(v[c0], kk) / L[rk];
      if (l == kk) v[c0] = yk;
#pragma unroll
      for (int c = c0; c < 4; ++c) {
        const int j = l + 64 * c;
        if (j > k && j < n) v[c] -= L[rk + (j - k)] * yk;
      }
    }
  }
  // U x = y, column oriented: x_k = y_k / U_kk, then y_i -= U_ik x_k.
#pragma unroll
  for (int c0 = 3; c0 >= 0; --c0) {
    const int kend = min(64, n - 64 * c0);
    for (int kk = kend - 1; kk >= 0; --kk) {
      const int k = 64 * c0 + kk;
      const T xk = bcast_lane(v[c0], kk) / L[packed_upper(n, k, k)];
      if (l == kk) v[c0] = xk;
#pragma unroll
      for (int c = 0; c <= c0; ++c) {
        const int i = l + 64 * c;
        if (i < k) v[c] -= L[packed_upper(n, i, k)] * xk;
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kWG) void chol_factor_lds_kernel(int n, T* __restrict__ a, int64_t a_stride,
                                                              int negate, int32_t* __restrict__ info) {
  extern __shared__ unsigned char lxa_smem[];
  T* L = reinterpret_cast<T*>(lxa_smem);
  T* A = a + static_cast<int64_t>(blockIdx.x) * a_stride;
  lds_load_upper<T, T>(n, A, n, L, negate != 0);
  __syncthreads();
  const int fail = lds_chol_factor<T>(n, L);
  __syncthreads();
  const T nan = quiet_nan<T>();
  const int64_t nn = static_cast<int64_t>(n) * n;
  for (int64_t t = threadIdx.x; t < nn; t += kWG) {
    const int i = static_cast<int>(t / n), j = static_cast<int>(t % n);
    A[t] = fail ? nan : (j < i ? T(0) : L[packed_upper(n, i, j)]);
  }
  if (threadIdx.x == 0) info[blockIdx.x] = fail;
}

template <typename T>
__global__ __launch_bounds__(kWG) void chol_solve_lds_kernel(int n, const T* __restrict__ u, int64_t u_stride,
                                                             T* __restrict__ b, int64_t b_stride, int negate) {
  extern __shared__ unsigned char lxa_smem[];
  T* L = reinterpret_cast<T*>(lxa_smem);
  const T* U = u + static_cast<int64_t>(blockIdx.x) * u_stride;
  T* B = b + static_cast<int64_t>(blockIdx.x) * b_stride;
  lds_load_upper<T, T>(n, U, n, L, false);
  __syncthreads();
  if ((threadIdx.x >> 6) == 0) {
    const int l = threadIdx.x;
    T v[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = (l + 64 * c < n) ? B[l + 64 * c] : T(0);
    wave_chol_solve<T>(n, L, v);
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (l + 64 * c < n) B[l + 64 * c] = negate ? -v[c] : v[c];
  }
}

// ---------------------------------------------------------------------------
// Tier 3: any n, one workgroup per system in global memory.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kWG) void chol_factor_generic_kernel(int n, T* __restrict__ a, int64_t a_stride,
                                                                  int negate, int32_t* __restrict__ info) {
  T* A = a + static_cast<int64_t>(blockIdx.x) * a_stride;
  const int tid = threadIdx.x;
  const int64_t ld = n;
  if (negate) {
    for (int64_t t = tid; t < ld * ld; t += kWG) {
      const int64_t i = t / ld, j = t % ld;
      if (j >= i) A[t] = -A[t];
    }
    __syncthreads();
  }
  int fail = 0;
  for (int k = 0; k < n; ++k) {
    const T akk = A[k * ld + k];
    if (!(akk > T(0))) {
      fail = k + 1;
      break;
    }
    const T ukk = sqrt(akk);
    for (int j = k + 1 + tid; j < n; j += kWG) A[k * ld + j] /= ukk;
    __syncthreads();
    if (tid == 0) A[k * ld + k] = ukk;
    const int64_t m = n - k - 1;
    for (int64_t t = tid; t < m * m; t += kWG) {
      const int64_t i = k + 1 + t / m, j = k + 1 + t % m;
      if (j >= i) A[i * ld + j] -= A[k * ld + i] * A[k * ld + j];
    }
    __syncthreads();
  }
  __syncthreads();
  const T nan = quiet_nan<T>();
  for (int64_t t = tid; t < ld * ld; t += kWG) {
    const int64_t i = t / ld, j = t % ld;
    if (fail) {
      A[t] = nan;
    } else if (j < i) {
      A[t] = T(0);
    }
  }
  if (tid == 0) info[blockIdx.x] = fail;
}

template <typename T>
__global__ __launch_bounds__(kWG) void chol_solve_generic_kernel(int n, const T* __restrict__ u, int64_t u_stride,
                                                                 T* __restrict__ b, int64_t b_stride, int negate) {
  extern __shared__ unsigned char lxa_smem[];
  T* v = reinterpret_cast<T*>(lxa_smem);
  const T* U = u + static_cast<int64_t>(blockIdx.x) * u_stride;
  T* B = b + static_cast<int64_t>(blockIdx.x) * b_stride;
  const int tid = threadIdx.x;
  const int64_t ld = n;
  for (int i = tid; i < n; i += kWG) v[i] = B[i];
  __syncthreads();
  // U^T y = b: y_j final when read at step j; kept in B.
  for (int j = 0; j < n; ++j) {
    const T yj = v[j] / U[j * ld + j];
    for (int i = j + 1 + tid; i < n; i += kWG) v[i] -= U[j * ld + i] * yj;
    if (tid == 0) B[j] = yj;
    __syncthreads();
  }
  for (int i = tid; i < n; i += kWG) v[i] = B[i];
  __syncthreads();
  for (int j = n - 1; j >= 0; --j) {
    const T xj = v[j] / U[j * ld + j];
    for (int i = tid; i < j; i += kWG) v[i] -= U[i * ld + j] * xj;
    if (tid == 0) B[j] = negate ? -xj : xj;
    __syncthreads();
  }
}

// Kernels using more than 64 KiB of dynamic LDS must say so before launch.
template <typename K>
int allow_lds(K kernel, size_t bytes) {
  if (bytes <= 64 * 1024) return LXA_OK;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
  return e == hipSuccess ? LXA_OK : hip_error(e, "hipFuncSetAttribute(MaxDynamicSharedMemorySize)");
}

template <typename T, int N>
void launch_chol_small(hipStream_t s, int64_t batch, int n, T* a, int64_t stride, int negate, int32_t* info) {
  constexpr int SPW = 64 / CholLanes<N>::value;
  const int64_t blocks = (batch + SPW - 1) / SPW;
  hipLaunchKernelGGL((chol_factor_small_kernel<T, N>), dim3(static_cast<unsigned>(blocks)), dim3(64), 0, s, batch,
                     n, a, stride, negate, info);
}

template <typename T, int N>
void launch_chol_solve_small(hipStream_t s, int64_t batch, int n, const T* u, int64_t us, T* b, int64_t bs,
                             int negate) {
  constexpr int SPW = 64 / CholLanes<N>::value;
  const int64_t blocks = (batch + SPW - 1) / SPW;
  hipLaunchKernelGGL((chol_solve_small_kernel<T, N>), dim3(static_cast<unsigned>(blocks)), dim3(64), 0, s, batch,
                     n, u, us, b, bs, negate);
}

template <typename T>
int lds_max_n() {
  return sizeof(T) == 4 ? 256 : 181;  // packed triangle <= 131.6 KB
}

template <typename T>
int chol_factor_t(hipStream_t s, int64_t batch, int64_t n64, T* a, int64_t stride, int negate, int32_t* info) {
  const int n = static_cast<int>(n64);
  if (n <= 8) launch_chol_small<T, 8>(s, batch, n, a, stride, negate, info);
  else if (n <= 16) launch_chol_small<T, 16>(s, batch, n, a, stride, negate, info);
  else if (n <= 32) launch_chol_small<T, 32>(s, batch, n, a, stride, negate, info);
  else if (n <= 64) launch_chol_small<T, 64>(s, batch, n, a, stride, negate, info);
  else if (sizeof(T) == 4 && n <= 256) {
    // register-tiled 16-column panels (chol_uneven.hip)
    return chol_factor_regs_f32(s, batch, n, reinterpret_cast<float*>(a), stride, negate, info);
  } else if (n <= lds_max_n<T>()) {
    const size_t shmem = static_cast<size_t>(n) * (n + 1) / 2 * sizeof(T);
    if (int rc = allow_lds(chol_factor_lds_kernel<T>, shmem)) return rc;
    hipLaunchKernelGGL((chol_factor_lds_kernel<T>), dim3(static_cast<unsigned>(batch)), dim3(kWG), shmem, s, n, a,
                       stride, negate, info);
  } else {
    hipLaunchKernelGGL((chol_factor_generic_kernel<T>), dim3(static_cast<unsigned>(batch)), dim3(kWG), 0, s, n, a,
                       stride, negate, info);
  }
  LXA_LAUNCH_CHECK("lxa_chol_factor launch");
  return LXA_OK;
}

template <typename T>
int chol_solve_t(hipStream_t s, int64_t batch, int64_t n64, const T* u, int64_t us, T* b, int64_t bs, int negate) {
  const int n = static_cast<int>(n64);
  if (n <= 8) launch_chol_solve_small<T, 8>(s, batch, n, u, us, b, bs, negate);
  else if (n <= 16) launch_chol_solve_small<T, 16>(s, batch, n, u, us, b, bs, negate);
  else if (n <= 32) launch_chol_solve_small<T, 32>(s, batch, n, u, us, b, bs, negate);
  else if (n <= 64) launch_chol_solve_small<T, 64>(s, batch, n, u, us, b, bs, negate);
  else if (n <= lds_max_n<T>()) {
    const size_t shmem = static_cast<size_t>(n) * (n + 1) / 2 * sizeof(T);
    if (int rc = allow_lds(chol_solve_lds_kernel<T>, shmem)) return rc;
    hipLaunchKernelGGL((chol_solve_lds_kernel<T>), dim3(static_cast<unsigned>(batch)), dim3(kWG), shmem, s, n, u, us,
                       b, bs, negate);
  } else {
    const size_t shmem = static_cast<size_t>(n) * sizeof(T);
    if (shmem > 160 * 1024) return set_error(LXA_EUNSUPPORTED, "lxa_chol_solve: n=%d too large", n);
    if (int rc = allow_lds(chol_solve_generic_kernel<T>, shmem)) return rc;
    hipLaunchKernelGGL((chol_solve_generic_kernel<T>), dim3(static_cast<unsigned>(batch)), dim3(kWG), shmem, s, n, u,
                       us, b, bs, negate);
  }
  LXA_LAUNCH_CHECK("lxa_chol_solve launch");
  return LXA_OK;
}

constexpr int64_t kMaxGrid = 2147483647;

}  // namespace
}  // namespace lxa

extern "C" int lxa_chol_factor(void* stream, int dtype, int64_t batch, int64_t n, void* a, int64_t a_batch_stride,
                               int negate, int32_t* info) {
  LXA_REQUIRE(batch >= 0 && batch <= lxa::kMaxGrid, "lxa_chol_factor: bad batch %lld", (long long)batch);
  LXA_REQUIRE(n >= 1 && n <= 65536, "lxa_chol_factor: bad n %lld", (long long)n);
  LXA_REQUIRE(a && info, "lxa_chol_factor: null pointer");
  LXA_REQUIRE(a_batch_stride >= n * n || batch <= 1, "lxa_chol_factor: a_batch_stride < n*n");
  if (batch == 0) return LXA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == LXA_F32)
    return lxa::chol_factor_t<float>(s, batch, n, static_cast<float*>(a), a_batch_stride, negate, info);
  if (dtype == LXA_F64)
    return lxa::chol_factor_t<double>(s, batch, n, static_cast<double*>(a), a_batch_stride, negate, info);
  return lxa::set_error(LXA_EINVAL, "lxa_chol_factor: bad dtype %d", dtype);
}

extern "C" int lxa_chol_solve(void* stream, int dtype, int64_t batch, int64_t n, const void* u,
                              int64_t u_batch_stride, void* b, int64_t b_batch_stride, int negate) {
  LXA_REQUIRE(batch >= 0 && batch <= lxa::kMaxGrid, "lxa_chol_solve: bad batch %lld", (long long)batch);
  LXA_REQUIRE(n >= 1 && n <= 65536, "lxa_chol_solve: bad n %lld", (long long)n);
  LXA_REQUIRE(u && b, "lxa_chol_solve: null pointer");
  LXA_REQUIRE(u_batch_stride >= 0 && b_batch_stride >= 0, "lxa_chol_solve: negative stride");
  LXA_REQUIRE(b_batch_stride >= n || batch <= 1, "lxa_chol_solve: b_batch_stride < n");
  if (batch == 0) return LXA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == LXA_F32)
    return lxa::chol_solve_t<float>(s, batch, n, static_cast<const float*>(u), u_batch_stride,
                                    static_cast<float*>(b), b_batch_stride, negate);
  if (dtype == LXA_F64)
    return lxa::chol_solve_t<double>(s, batch, n, static_cast<const double*>(u), u_batch_stride,
                                     static_cast<double*>(b), b_batch_stride, negate);
  return lxa::set_error(LXA_EINVAL, "lxa_chol_solve: bad dtype %d", dtype);
}
