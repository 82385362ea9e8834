// Dense matrix-vector products and the RESULTS post-processing of a solve.
#include "lxa_common.h"

namespace lxa {
namespace {

// y = A x: one wave per output row, lanes stride the row (coalesced), FP64
// accumulation.  Rows of all batch members are distributed grid-stride.
// Used when rows are not 16-byte aligned multiples of four elements.
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

// Four consecutive elements through 16-byte loads (one for float, two for
// double); the address is 16-byte aligned.
template <typename T>
struct Four {
  double v[4];
};
__device__ __forceinline__ Four<float> load4(const float* p) {
  const float4 q = *reinterpret_cast<const float4*>(p);
  return {{q.x, q.y, q.z, q.w}};
}
__device__ __forceinline__ Four<double> load4(const double* p) {
  const double2 q0 = *reinterpret_cast<const double2*>(p);
  const double2 q1 = *reinterpret_cast<const double2*>(p + 2);
  return {{q0.x, q0.y, q1.x, q1.y}};
}

// y = A x for rows of a multiple of four elements, 16-byte aligned: a wave
// per row, each lane four consecutive columns per load and four loads in
// flight (16 KB per wave for float), four FP64 accumulators.
template <typename T>
__global__ __launch_bounds__(256) void gemv_n4_kernel(int64_t batch, int64_t m, int64_t n, const T* __restrict__ a,
                                                      int64_t as, const T* __restrict__ x, int64_t xs,
                                                      T* __restrict__ y, int64_t ys) {
  const int l = threadIdx.x & 63;
  const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = (static_cast<int64_t>(gridDim.x) * blockDim.x) >> 6;
  for (int64_t row = wave; row < batch * m; row += nwaves) {
    const int64_t bi = row / m, i = row % m;
    const T* Ai = a + bi * as + i * n;
    const T* xb = x + bi * xs;
    double s4[4] = {0.0, 0.0, 0.0, 0.0};
    int64_t j = 4 * l;
    for (; j + 3 * 256 < n; j += 4 * 256) {
      Four<T> av[4], xv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        av[u] = load4(Ai + j + 256 * u);
        xv[u] = load4(xb + j + 256 * u);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int e = 0; e < 4; ++e) s4[u] = fma(av[u].v[e], xv[u].v[e], s4[u]);
    }
    for (; j < n; j += 256) {
      const Four<T> av = load4(Ai + j), xv = load4(xb + j);
#pragma unroll
      for (int e = 0; e < 4; ++e) s4[0] = fma(av.v[e], xv.v[e], s4[0]);
    }
    const double s = wave_sum((s4[0] + s4[1]) + (s4[2] + s4[3]));
    if (l == 0) y[bi * ys + i] = static_cast<T>(s);
  }
}

// y = A^T x: blockIdx.x a strip of 1024 columns (four per thread), blockIdx.y
// a chunk of rows, blockIdx.z the batch member.  Each thread sums its
// columns over the chunk's rows in order (FP64); with one chunk the sums are
// y, otherwise they go to part[bi][chunk][j] and gemv_t_reduce_kernel adds
// the chunks in order, so results do not depend on the launch shape's timing.
template <typename T, bool kVec, bool kDirect>
__global__ __launch_bounds__(256) void gemv_t_part_kernel(int64_t m, int64_t n, int64_t rows_per_chunk,
                                                          const T* __restrict__ a, int64_t as,
                                                          const T* __restrict__ x, int64_t xs, T* __restrict__ y,
                                                          int64_t ys, double* __restrict__ part) {
  const int64_t bi = blockIdx.z, chunk = blockIdx.y;
  const int64_t j0 = (static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x) * 4;
  if (j0 >= n) return;
  const int64_t r0 = chunk * rows_per_chunk;
  const int64_t r1 = r0 + rows_per_chunk < m ? r0 + rows_per_chunk : m;
  const T* A = a + bi * as;
  const T* xb = x + bi * xs;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (kVec) {
    int64_t i = r0;
    for (; i + 3 < r1; i += 4) {
      Four<T> av[4];
      double xv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        av[u] = load4(A + (i + u) * n + j0);
        xv[u] = static_cast<double>(xb[i + u]);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[e] = fma(av[u].v[e], xv[u], acc[e]);
    }
    for (; i < r1; ++i) {
      const Four<T> av = load4(A + i * n + j0);
      const double xv = static_cast<double>(xb[i]);
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[e] = fma(av.v[e], xv, acc[e]);
    }
  } else {
    for (int64_t i = r0; i < r1; ++i) {
      const double xv = static_cast<double>(xb[i]);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (j0 + e < n) acc[e] = fma(static_cast<double>(A[i * n + j0 + e]), xv, acc[e]);
    }
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (j0 + e >= n) break;
    if (kDirect)
      y[bi * ys + j0 + e] = static_cast<T>(acc[e]);
    else
      part[(bi * gridDim.y + chunk) * n + j0 + e] = acc[e];
  }
}

// y[j] = sum of the chunk partials of column j: a workgroup of 16 waves
// takes 64 columns, wave w adds its contiguous range of chunks in order
// (loads eight at a time), then the 16 wave sums are added in order.
template <typename T>
__global__ __launch_bounds__(1024) void gemv_t_reduce_kernel(int64_t n, int64_t chunks, const double* __restrict__ part,
                                                             T* __restrict__ y, int64_t ys) {
  __shared__ double red[16][64];
  const int64_t bi = blockIdx.y;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t j = static_cast<int64_t>(blockIdx.x) * 64 + l;
  const int64_t per = (chunks + 15) / 16;
  const int64_t c0 = w * per;
  const int64_t c1 = c0 + per < chunks ? c0 + per : chunks;
  double s = 0.0;
  if (j < n) {
    const double* p = part + bi * chunks * n + j;
    int64_t c = c0;
    for (; c + 8 <= c1; c += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = p[(c + u) * n];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; c < c1; ++c) s += p[c * n];
  }
  red[w][l] = s;
  __syncthreads();
  if (w == 0 && j < n) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) t += red[k][l];
    y[bi * ys + j] = static_cast<T>(t);
  }
}

// Row chunks of the transposed product: enough workgroups to cover about
// four per compute unit, chunks of at least 16 rows.
struct TChunks {
  int64_t strips, chunks, rows_per_chunk;
};
TChunks t_chunks(int64_t batch, int64_t m, int64_t n) {
  TChunks c;
  c.strips = (n + 1023) / 1024;
  const int64_t want = 4 * static_cast<int64_t>(num_cus());
  int64_t chunks = (want + c.strips * batch - 1) / (c.strips * batch);
  const int64_t most = (m + 15) / 16;
  if (chunks > most) chunks = most;
  if (chunks < 1) chunks = 1;
  c.rows_per_chunk = (m + chunks - 1) / chunks;
  c.chunks = m > 0 ? (m + c.rows_per_chunk - 1) / c.rows_per_chunk : 1;
  return c;
}

size_t t_work_bytes(int64_t batch, int64_t m, int64_t n) {
  const TChunks c = t_chunks(batch, m, n);
  return c.chunks > 1 ? static_cast<size_t>(batch) * c.chunks * n * sizeof(double) : 0;
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

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

template <typename T>
int gemv_t(hipStream_t s, int64_t batch, int64_t m, int64_t n, const T* a, int64_t as, const T* x, int64_t xs, T* y,
           int64_t ys, int trans, void* work, size_t work_bytes) {
  // 16-byte loads need rows (and batch strides) of a multiple of four
  // elements starting on 16-byte boundaries
  const bool vec = n % 4 == 0 && as % 4 == 0 && aligned16(a) && (trans == 1 || (xs % 4 == 0 && aligned16(x)));
  if (trans == 0) {
    const int64_t waves = batch * m;
    int64_t blocks = (waves + 3) / 4;
    const int64_t cap = static_cast<int64_t>(num_cus()) * 16;
    if (blocks > cap) blocks = cap;
    if (vec)
      hipLaunchKernelGGL((gemv_n4_kernel<T>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, batch, m, n, a,
                         as, x, xs, y, ys);
    else
      hipLaunchKernelGGL((gemv_n_kernel<T>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, batch, m, n, a,
                         as, x, xs, y, ys);
  } else {
    LXA_REQUIRE(batch <= 65535, "lxa_gemv: trans=1 batch must be <= 65535");
    const TChunks c = t_chunks(batch, m, n);
    const size_t need = t_work_bytes(batch, m, n);
    LXA_REQUIRE(need == 0 || (work && work_bytes >= need),
                "lxa_gemv: trans=1 needs %zu workspace bytes (lxa_gemv_workspace_size), got %zu", need, work_bytes);
    const dim3 grid(static_cast<unsigned>(c.strips), static_cast<unsigned>(c.chunks), static_cast<unsigned>(batch));
    double* part = static_cast<double*>(work);
    if (c.chunks == 1) {
      if (vec)
        hipLaunchKernelGGL((gemv_t_part_kernel<T, true, true>), grid, dim3(256), 0, s, m, n, c.rows_per_chunk, a, as,
                           x, xs, y, ys, part);
      else
        hipLaunchKernelGGL((gemv_t_part_kernel<T, false, true>), grid, dim3(256), 0, s, m, n, c.rows_per_chunk, a,
                           as, x, xs, y, ys, part);
    } else {
      if (vec)
        hipLaunchKernelGGL((gemv_t_part_kernel<T, true, false>), grid, dim3(256), 0, s, m, n, c.rows_per_chunk, a,
                           as, x, xs, y, ys, part);
      else
        hipLaunchKernelGGL((gemv_t_part_kernel<T, false, false>), grid, dim3(256), 0, s, m, n, c.rows_per_chunk, a,
                           as, x, xs, y, ys, part);
      hipLaunchKernelGGL((gemv_t_reduce_kernel<T>),
                         dim3(static_cast<unsigned>((n + 63) / 64), static_cast<unsigned>(batch)), dim3(1024), 0, s,
                         n, c.chunks, part, y, ys);
    }
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

extern "C" int lxa_gemv_workspace_size(int dtype, int64_t batch, int64_t m, int64_t n, int trans, size_t* bytes) {
  LXA_REQUIRE(bytes, "lxa_gemv_workspace_size: null output");
  LXA_REQUIRE(dtype == LXA_F32 || dtype == LXA_F64, "lxa_gemv_workspace_size: bad dtype %d", dtype);
  LXA_REQUIRE(batch >= 0 && m >= 0 && n >= 0, "lxa_gemv_workspace_size: negative size");
  *bytes = (trans == 1 && batch > 0 && n > 0) ? lxa::t_work_bytes(batch, m, n) : 0;
  return LXA_OK;
}

extern "C" int lxa_gemv(void* stream, int dtype, int64_t batch, int64_t m, int64_t n, const void* a,
                        int64_t a_batch_stride, const void* x, int64_t x_batch_stride, void* y,
                        int64_t y_batch_stride, int trans, void* work, size_t work_bytes) {
  LXA_REQUIRE(batch >= 0 && m >= 0 && n >= 0, "lxa_gemv: negative size");
  LXA_REQUIRE(trans == 0 || trans == 1, "lxa_gemv: trans must be 0 or 1");
  LXA_REQUIRE(a && x && y, "lxa_gemv: null pointer");
  if (batch == 0 || (trans == 0 ? m : n) == 0) return LXA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == LXA_F32)
    return lxa::gemv_t<float>(s, batch, m, n, static_cast<const float*>(a), a_batch_stride,
                              static_cast<const float*>(x), x_batch_stride, static_cast<float*>(y), y_batch_stride,
                              trans, work, work_bytes);
  if (dtype == LXA_F64)
    return lxa::gemv_t<double>(s, batch, m, n, static_cast<const double*>(a), a_batch_stride,
                               static_cast<const double*>(x), x_batch_stride, static_cast<double*>(y),
                               y_batch_stride, trans, work, work_bytes);
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
