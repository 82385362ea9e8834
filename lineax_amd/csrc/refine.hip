// Batched FP64 products with float32 operators of uneven shapes, for the
// FP64 refinement residuals of the request loop: per system
//   trans = 0:  y = b - A x   (or y = A x when b is null),  length m;
//   trans = 1:  y = A^T x,                                  length n;
// A float32 row-major (leading dimension lda), x, y float64, b float32.
// Each system is described by device addresses, so operators can live
// anywhere in device memory (the request's own operator tensor).
//
// trans = 0: a wave per row (lanes stride the columns, wave reduction);
// trans = 1: a lane per column, the block's waves splitting the rows.
// blockIdx.y is the system, blockIdx.x a chunk of its rows / columns.  Sums
// are in a fixed order: the results do not depend on the batch.
#include "lxa_common.h"

namespace lxa {
namespace {

constexpr int kT = 256;

__global__ __launch_bounds__(kT) void gemv_uneven_n_kernel(const lxa_gemv_desc* __restrict__ descs) {
  const lxa_gemv_desc d = descs[blockIdx.y];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const float* A = reinterpret_cast<const float*>(d.a);
  const double* x = reinterpret_cast<const double*>(d.x);
  const float* b = reinterpret_cast<const float*>(d.b);
  double* y = reinterpret_cast<double*>(d.y);
  const int waves = gridDim.x * (kT / 64);
  for (int i = blockIdx.x * (kT / 64) + w; i < d.m; i += waves) {
    const float* Ai = A + static_cast<int64_t>(i) * d.lda;
    double s4[4] = {0.0, 0.0, 0.0, 0.0};
    int j = l;
    for (; j + 192 < d.n; j += 256) {
#pragma unroll
      for (int u = 0; u < 4; ++u) s4[u] = fma(static_cast<double>(Ai[j + 64 * u]), x[j + 64 * u], s4[u]);
    }
    for (; j < d.n; j += 64) s4[0] = fma(static_cast<double>(Ai[j]), x[j], s4[0]);
    double s = wave_sum((s4[0] + s4[1]) + (s4[2] + s4[3]));
    if (l == 0) y[i] = b ? static_cast<double>(b[i]) - s : s;
  }
}

// 64 columns per block (a lane per column); the four waves take every
// fourth row and are summed in a fixed order.
__global__ __launch_bounds__(kT) void gemv_uneven_t_kernel(const lxa_gemv_desc* __restrict__ descs) {
  __shared__ double part[kT / 64][64];
  const lxa_gemv_desc d = descs[blockIdx.y];
  const float* A = reinterpret_cast<const float*>(d.a);
  const double* x = reinterpret_cast<const double*>(d.x);
  double* y = reinterpret_cast<double*>(d.y);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (int j0 = blockIdx.x * 64; j0 < d.n; j0 += gridDim.x * 64) {
    const int j = j0 + l;
    double s4[4] = {0.0, 0.0, 0.0, 0.0};
    if (j < d.n) {
      int i = w;
      for (; i + 12 < d.m; i += 16) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
          s4[u] = fma(static_cast<double>(A[static_cast<int64_t>(i + 4 * u) * d.lda + j]), x[i + 4 * u], s4[u]);
      }
      for (; i < d.m; i += kT / 64) s4[0] = fma(static_cast<double>(A[static_cast<int64_t>(i) * d.lda + j]), x[i], s4[0]);
    }
    part[w][l] = (s4[0] + s4[1]) + (s4[2] + s4[3]);
    __syncthreads();
    if (w == 0 && j < d.n) y[j] = (part[0][l] + part[1][l]) + (part[2][l] + part[3][l]);
    __syncthreads();
  }
}

}  // namespace
}  // namespace lxa

extern "C" int lxa_gemv_uneven(void* stream, int64_t count, const lxa_gemv_desc* descs, int trans, int max_m,
                               int max_n) {
  LXA_REQUIRE(count >= 0 && count <= 65535, "lxa_gemv_uneven: count must be in [0, 65535]");
  LXA_REQUIRE(trans == 0 || trans == 1, "lxa_gemv_uneven: trans must be 0 or 1");
  LXA_REQUIRE(max_m >= 0 && max_n >= 0, "lxa_gemv_uneven: bad sizes");
  if (count == 0 || max_m == 0 || max_n == 0) return LXA_OK;
  LXA_REQUIRE(descs, "lxa_gemv_uneven: null descriptors");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (trans == 0) {
    // about 16 rows per wave
    const int chunks = (max_m + 16 * (lxa::kT / 64) - 1) / (16 * (lxa::kT / 64));
    hipLaunchKernelGGL(lxa::gemv_uneven_n_kernel, dim3(static_cast<unsigned>(chunks), static_cast<unsigned>(count)),
                       dim3(lxa::kT), 0, s, descs);
  } else {
    const int chunks = (max_n + 63) / 64;
    hipLaunchKernelGGL(lxa::gemv_uneven_t_kernel, dim3(static_cast<unsigned>(chunks), static_cast<unsigned>(count)),
                       dim3(lxa::kT), 0, s, descs);
  }
  LXA_LAUNCH_CHECK("lxa_gemv_uneven");
  return LXA_OK;
}
