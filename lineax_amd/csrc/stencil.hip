// Five-point stencils on a regular nx x ny grid with zero Dirichlet
// boundary values, batched over rows (one grid function per right-hand
// side).  These are the matrix-free operators of the Krylov workloads:
//   y[i][j] = c * x[i][j] + w * x[i][j-1] + e * x[i][j+1]
//                         + s * x[i-1][j] + n * x[i+1][j]
// with out-of-grid neighbours taken as zero.  Accumulation is FP64.  Grids
// whose width is a multiple of 4 (16-byte aligned rows) run
// stencil5_rows4_kernel (four columns per thread, rows walked with the
// neighbouring rows kept in registers); others one thread per point, j the
// fastest index so loads are coalesced, neighbour rows re-read through the
// caches.
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

// Four consecutive grid points as doubles (16-byte loads: one for float, two
// for double) and their store in the output type.
__device__ __forceinline__ void load4(const float* p, double (&o)[4]) {
  const float4 q = *reinterpret_cast<const float4*>(p);
  o[0] = q.x;
  o[1] = q.y;
  o[2] = q.z;
  o[3] = q.w;
}
__device__ __forceinline__ void load4(const double* p, double (&o)[4]) {
  const double2 q0 = *reinterpret_cast<const double2*>(p);
  const double2 q1 = *reinterpret_cast<const double2*>(p + 2);
  o[0] = q0.x;
  o[1] = q0.y;
  o[2] = q1.x;
  o[3] = q1.y;
}
__device__ __forceinline__ void store4(float* p, const double (&v)[4]) {
  *reinterpret_cast<float4*>(p) = make_float4(static_cast<float>(v[0]), static_cast<float>(v[1]),
                                              static_cast<float>(v[2]), static_cast<float>(v[3]));
}
__device__ __forceinline__ void store4(double* p, const double (&v)[4]) {
  *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
  *reinterpret_cast<double2*>(p + 2) = make_double2(v[2], v[3]);
}

constexpr int kRows4 = 16;  // grid rows per workgroup
constexpr int kAhead = 4;   // rows loaded per step

// Grid widths that are a multiple of 4 with 16-byte aligned rows: a thread
// owns four adjacent columns and walks kRows4 grid rows, keeping the rows
// above and below in registers, so each point is read once from memory
// (the columns either side of the four come from the cache).  kAhead rows
// are loaded per step.  Same expression and order as stencil5_kernel.
template <typename TI, typename TO>
__global__ __launch_bounds__(256) void stencil5_rows4_kernel(int nx, int ny, Coeffs k, const TI* __restrict__ x,
                                                             int64_t xs, TO* __restrict__ y, int64_t ys) {
  const int64_t row = blockIdx.z;
  const int i0 = blockIdx.y * kRows4;
  const int i1 = i0 + kRows4 < nx ? i0 + kRows4 : nx;
  const int j0 = (blockIdx.x * 256 + threadIdx.x) * 4;
  if (j0 >= ny) return;
  const TI* X = x + row * xs;
  TO* Y = y + row * ys;
  // loads are unconditional (clamped rows and columns, then zeroed outside
  // the grid): in branches the compiler waited for each load in turn
  auto ldrow = [&](int i, double (&o)[4]) {
    const bool in = i >= 0 && i < nx;
    load4(X + static_cast<int64_t>(in ? i : 0) * ny + j0, o);
#pragma unroll
    for (int e = 0; e < 4; ++e) o[e] = in ? o[e] : 0.0;
  };
  double up[4], cu[4];
  ldrow(i0 - 1, up);
  ldrow(i0, cu);
  for (int ib = i0; ib < i1; ib += kAhead) {
    double q[kAhead][4], we[kAhead][2];
#pragma unroll
    for (int u = 0; u < kAhead; ++u) {
      ldrow(ib + 1 + u, q[u]);
      const int i = ib + u;
      const TI* Xi = X + static_cast<int64_t>(i < i1 ? i : i1 - 1) * ny;
      const double wl = static_cast<double>(Xi[j0 > 0 ? j0 - 1 : 0]);
      const double wr = static_cast<double>(Xi[j0 + 4 < ny ? j0 + 4 : ny - 1]);
      we[u][0] = j0 > 0 ? wl : 0.0;
      we[u][1] = j0 + 4 < ny ? wr : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kAhead; ++u) {
      const int i = ib + u;
      if (i < i1) {
        const double* dn = q[u];
        double v[4];
        v[0] = k.c * cu[0] + k.w * we[u][0] + k.e * cu[1] + k.s * up[0] + k.n * dn[0];
        v[1] = k.c * cu[1] + k.w * cu[0] + k.e * cu[2] + k.s * up[1] + k.n * dn[1];
        v[2] = k.c * cu[2] + k.w * cu[1] + k.e * cu[3] + k.s * up[2] + k.n * dn[2];
        v[3] = k.c * cu[3] + k.w * cu[2] + k.e * we[u][1] + k.s * up[3] + k.n * dn[3];
        store4(Y + static_cast<int64_t>(i) * ny + j0, v);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        up[e] = cu[e];
        cu[e] = q[u][e];
      }
    }
  }
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

template <typename TI, typename TO>
int launch(hipStream_t st, int64_t rows, int nx, int ny, Coeffs k, const void* x, int64_t xs, void* y, int64_t ys) {
  if (ny % 4 == 0 && xs % 4 == 0 && ys % 4 == 0 && aligned16(x) && aligned16(y)) {
    dim3 grid(static_cast<unsigned>((ny / 4 + 255) / 256), static_cast<unsigned>((nx + kRows4 - 1) / kRows4),
              static_cast<unsigned>(rows));
    hipLaunchKernelGGL((stencil5_rows4_kernel<TI, TO>), grid, dim3(256), 0, st, nx, ny, k,
                       static_cast<const TI*>(x), xs, static_cast<TO*>(y), ys);
  } else {
    dim3 grid(static_cast<unsigned>((ny + 255) / 256), static_cast<unsigned>(nx), static_cast<unsigned>(rows));
    hipLaunchKernelGGL((stencil5_kernel<TI, TO>), grid, dim3(256), 0, st, rows, nx, ny, k, static_cast<const TI*>(x),
                       xs, static_cast<TO*>(y), ys);
  }
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
