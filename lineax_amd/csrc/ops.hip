// Dense matrix-vector products and the RESULTS post-processing of a solve.
#include "lxa_common.h"

namespace lxa {
namespace {

// y = A x: one wave per output row, lanes stride the row (coalesced), FP64
// accumulation.  Rows of all batch members are distributed grid-stride.
template <typename T>
__global__ __launch_bounds__(256) void gemv_n_kernel(int64_t batch, int64_t m, int64_t n, const T* __restrict__ a,
                                                     int64_t as, const T* __restrict__ x, int64_t xs,
                                                     T* __restrict__ y, int64_t ys) {
  const int l = threadIdx.x & 63;
  const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = (static_cast<int64_t>(gridDim.x) * blockDim.x) >> 6;
  for (int64_t row = wave; row < batch * m; row += nwaves) {
    const int64_t bi = row / m, i = row % m;
    const T* Ai = a + bi * as + i * n;
    const T* xb = x + bi * xs;
    double acc = 0.0;
    for (int64_t j = l; j < n; j += 64) acc = fma(static_cast<double>(Ai[j]), static_cast<double>(xb[j]), acc);
    acc = wave_sum(acc);
    if (l == 0) y[bi * ys + i] = static_cast<T>(acc);
  }
}

// y = A^T x: thread per output column j, loop over rows (reads of a row are
// coalesced across the block).
template <typename T>
__global__ __launch_bounds__(256) void gemv_t_kernel(int64_t m, int64_t n, const T* __restrict__ a, int64_t as,
                                                     const T* __restrict__ x, int64_t xs, T* __restrict__ y,
                                                     int64_t ys) {
  const int64_t bi = blockIdx.y;
  const int64_t j = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (j >= n) return;
  const T* A = a + bi * as;
  const T* xb = x + bi * xs;
  double acc = 0.0;
  for (int64_t i = 0; i < m; ++i) acc = fma(static_cast<double>(A[i * n + j]), static_cast<double>(xb[i]), acc);
  y[bi * ys + j] = static_cast<T>(acc);
}

template <typename T>
__device__ __forceinline__ bool nonfinite(T v) {
  return !isfinite(v);
}

constexpr int kSolutionBad = 0x100;
constexpr int kVectorBad = 0x200;

// Marks rows holding a non-finite entry: bit 0x100 for the solution, 0x200
// for the vector, OR-ed into result[row].
// Block b scans chunk (b % chunks) of row (b / chunks).
template <typename T>
__global__ __launch_bounds__(256) void nonfinite_scan_kernel(int64_t len, int64_t chunks, const T* __restrict__ sol,
                                                             int64_t ss, const T* __restrict__ vec, int64_t vs,
                                                             int32_t* __restrict__ result) {
  const int64_t row = blockIdx.x / chunks;
  const int64_t chunk = blockIdx.x % chunks;
  bool bad_s = false, bad_v = false;
  for (int64_t t = chunk * 256 + threadIdx.x; t < len; t += chunks * 256) {
    bad_s |= nonfinite(sol[row * ss + t]);
    bad_v |= nonfinite(vec[row * vs + t]);
  }
  const unsigned long long ms = __ballot(bad_s);
  const unsigned long long mv = __ballot(bad_v);
  if ((threadIdx.x & 63) == 0 && (ms | mv)) atomicOr(&result[row], (ms ? kSolutionBad : 0) | (mv ? kVectorBad : 0));
}

// lineax/_solve.py:109-123: successful & bad solution -> singular; then
// singular & bad vector -> nonfinite_input.  Clears the scan bits.
__global__ __launch_bounds__(256) void results_map_kernel(int64_t rows, int32_t* __restrict__ result) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= rows) return;
  const int v = result[i];
  int code = v & 0xff;
  if (code == LXA_RESULT_SUCCESSFUL && (v & kSolutionBad)) code = LXA_RESULT_SINGULAR;
  if (code == LXA_RESULT_SINGULAR && (v & kVectorBad)) code = LXA_RESULT_NONFINITE_INPUT;
  result[i] = code;
}

template <typename T>
int gemv_t(hipStream_t s, int64_t batch, int64_t m, int64_t n, const T* a, int64_t as, const T* x, int64_t xs, T* y,
           int64_t ys, int trans) {
  if (trans == 0) {
    const int64_t waves = batch * m;
    int64_t blocks = (waves + 3) / 4;
    const int64_t cap = static_cast<int64_t>(num_cus()) * 16;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL((gemv_n_kernel<T>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, batch, m, n, a, as,
                       x, xs, y, ys);
  } else {
    LXA_REQUIRE(batch <= 65535, "lxa_gemv: trans=1 batch must be <= 65535");
    hipLaunchKernelGGL((gemv_t_kernel<T>), dim3(static_cast<unsigned>((n + 255) / 256), static_cast<unsigned>(batch)),
                       dim3(256), 0, s, m, n, a, as, x, xs, y, ys);
  }
  LXA_LAUNCH_CHECK("lxa_gemv launch");
  return LXA_OK;
}

template <typename T>
int results_t(hipStream_t s, int64_t rows, int64_t len, const T* sol, int64_t ss, const T* vec, int64_t vs,
              int32_t* result) {
  if (len > 0) {
    int64_t chunks = (len + 256 * 8 - 1) / (256 * 8);
    if (chunks > 1024) chunks = 1024;
    LXA_REQUIRE(chunks * rows <= 2147483647, "lxa_solve_results: too many rows");
    hipLaunchKernelGGL((nonfinite_scan_kernel<T>), dim3(static_cast<unsigned>(chunks * rows)), dim3(256), 0, s, len,
                       chunks, sol, ss, vec, vs, result);
    LXA_LAUNCH_CHECK("lxa_solve_results scan");
  }
  hipLaunchKernelGGL(results_map_kernel, dim3(static_cast<unsigned>((rows + 255) / 256)), dim3(256), 0, s, rows,
                     result);
  LXA_LAUNCH_CHECK("lxa_solve_results map");
  return LXA_OK;
}

}  // namespace
}  // namespace lxa

extern "C" int lxa_gemv(void* stream, int dtype, int64_t batch, int64_t m, int64_t n, const void* a,
                        int64_t a_batch_stride, const void* x, int64_t x_batch_stride, void* y,
                        int64_t y_batch_stride, int trans) {
  LXA_REQUIRE(batch >= 0 && m >= 0 && n >= 0, "lxa_gemv: negative size");
  LXA_REQUIRE(trans == 0 || trans == 1, "lxa_gemv: trans must be 0 or 1");
  LXA_REQUIRE(a && x && y, "lxa_gemv: null pointer");
  if (batch == 0 || (trans == 0 ? m : n) == 0) return LXA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == LXA_F32)
    return lxa::gemv_t<float>(s, batch, m, n, static_cast<const float*>(a), a_batch_stride,
                              static_cast<const float*>(x), x_batch_stride, static_cast<float*>(y), y_batch_stride,
                              trans);
  if (dtype == LXA_F64)
    return lxa::gemv_t<double>(s, batch, m, n, static_cast<const double*>(a), a_batch_stride,
                               static_cast<const double*>(x), x_batch_stride, static_cast<double*>(y),
                               y_batch_stride, trans);
  return lxa::set_error(LXA_EINVAL, "lxa_gemv: bad dtype %d", dtype);
}

extern "C" int lxa_solve_results(void* stream, int dtype, int64_t rows, int64_t len, const void* solution,
                                 int64_t solution_row_stride, const void* vector, int64_t vector_row_stride,
                                 int32_t* result) {
  LXA_REQUIRE(rows >= 0 && len >= 0, "lxa_solve_results: bad sizes");
  LXA_REQUIRE(result && (len == 0 || (solution && vector)), "lxa_solve_results: null pointer");
  if (rows == 0) return LXA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == LXA_F32)
    return lxa::results_t<float>(s, rows, len, static_cast<const float*>(solution), solution_row_stride,
                                 static_cast<const float*>(vector), vector_row_stride, result);
  if (dtype == LXA_F64)
    return lxa::results_t<double>(s, rows, len, static_cast<const double*>(solution), solution_row_stride,
                                  static_cast<const double*>(vector), vector_row_stride, result);
  return lxa::set_error(LXA_EINVAL, "lxa_solve_results: bad dtype %d", dtype);
}
