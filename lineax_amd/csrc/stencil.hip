// Five-point stencils on a regular nx x ny grid with zero Dirichlet
// boundary values, batched over rows (one grid function per right-hand
// side).  These are the matrix-free operators of the Krylov workloads:
//   y[i][j] = c * x[i][j] + w * x[i][j-1] + e * x[i][j+1]
//                         + s * x[i-1][j] + n * x[i+1][j]
// with out-of-grid neighbours taken as zero.  One thread per point, j the
// fastest index so loads are coalesced; neighbour rows are re-read through
// the caches.  Accumulation is FP64.
#include "lxa_common.h"

namespace lxa {
namespace {

struct Coeffs {
  double c, w, e, s, n;
};

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void stencil5_kernel(int64_t rows, int nx, int ny, Coeffs k,
                                                       const TI* __restrict__ x, int64_t xs, TO* __restrict__ y,
                                                       int64_t ys) {
  const int64_t row = blockIdx.z;
  const int i = blockIdx.y;
  const int j = (blockIdx.x * 256 + threadIdx.x);
  if (j >= ny || row >= rows) return;
  const TI* X = x + row * xs + static_cast<int64_t>(i) * ny;
  const double xc = static_cast<double>(X[j]);
  const double xw = j > 0 ? static_cast<double>(X[j - 1]) : 0.0;
  const double xe = j + 1 < ny ? static_cast<double>(X[j + 1]) : 0.0;
  const double xsn = i > 0 ? static_cast<double>(X[j - ny]) : 0.0;
  const double xn = i + 1 < nx ? static_cast<double>(X[j + ny]) : 0.0;
  const double v = k.c * xc + k.w * xw + k.e * xe + k.s * xsn + k.n * xn;
  y[row * ys + static_cast<int64_t>(i) * ny + j] = static_cast<TO>(v);
}

template <typename TI, typename TO>
int launch(hipStream_t st, int64_t rows, int nx, int ny, Coeffs k, const void* x, int64_t xs, void* y, int64_t ys) {
  dim3 grid(static_cast<unsigned>((ny + 255) / 256), static_cast<unsigned>(nx), static_cast<unsigned>(rows));
  hipLaunchKernelGGL((stencil5_kernel<TI, TO>), grid, dim3(256), 0, st, rows, nx, ny, k, static_cast<const TI*>(x), xs,
                     static_cast<TO*>(y), ys);
  LXA_LAUNCH_CHECK("lxa_stencil5_apply launch");
  return LXA_OK;
}

}  // namespace
}  // namespace lxa

extern "C" int lxa_stencil5_apply(void* stream, int in_dtype, int out_dtype, int64_t rows, int64_t nx, int64_t ny,
                                  const double* coeffs, const void* x, int64_t x_row_stride, void* y,
                                  int64_t y_row_stride) {
  LXA_REQUIRE(rows >= 1 && rows <= 65535, "lxa_stencil5_apply: rows must be in [1, 65535]");
  LXA_REQUIRE(nx >= 1 && ny >= 1 && nx <= 65535 && ny <= (1 << 24), "lxa_stencil5_apply: bad grid %lld x %lld",
              (long long)nx, (long long)ny);
  LXA_REQUIRE(coeffs && x && y, "lxa_stencil5_apply: null pointer");
  LXA_REQUIRE(x_row_stride >= nx * ny && y_row_stride >= nx * ny, "lxa_stencil5_apply: row stride < nx*ny");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const lxa::Coeffs k{coeffs[0], coeffs[1], coeffs[2], coeffs[3], coeffs[4]};
  const int a = static_cast<int>(nx), bdim = static_cast<int>(ny);
  if (in_dtype == LXA_F64 && out_dtype == LXA_F64)
    return lxa::launch<double, double>(st, rows, a, bdim, k, x, x_row_stride, y, y_row_stride);
  if (in_dtype == LXA_F32 && out_dtype == LXA_F32)
    return lxa::launch<float, float>(st, rows, a, bdim, k, x, x_row_stride, y, y_row_stride);
  if (in_dtype == LXA_F32 && out_dtype == LXA_F64)
    return lxa::launch<float, double>(st, rows, a, bdim, k, x, x_row_stride, y, y_row_stride);
  if (in_dtype == LXA_F64 && out_dtype == LXA_F32)
    return lxa::launch<double, float>(st, rows, a, bdim, k, x, x_row_stride, y, y_row_stride);
  return lxa::set_error(LXA_EINVAL, "lxa_stencil5_apply: bad dtypes");
}
