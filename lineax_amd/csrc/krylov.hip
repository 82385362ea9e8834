// Building blocks of lineax's CG and GMRES on the device.
//
// A batch of right-hand sides is a set of rows [rows, n].  Each row carries
// its own scalar state (lxa_krylov_row, include/lineax_amd.h) so rows
// converge, break down or stagnate independently, exactly as lineax under
// jax.vmap (finished rows are frozen).  Every reduction accumulates in
// FP64 in two deterministic phases: per-block partials, then a per-row
// finalize in a fixed order.  NaN propagates through the max-norms as it
// does through jnp.max, so a NaN residual ends a solve the way it ends
// lineax's while loop.
//
// Vector storage: TY for the iterate y and right-hand side b, TV for the
// working vectors (CG: r, p, q, z; GMRES: basis, w, r).  <double, double>
// and <float, float> follow lineax in that dtype; <double, float> keeps y
// and b in FP64 with FP32 working vectors.
#include "lxa_common.h"

namespace lxa {
namespace {

constexpr int kB = 256;
constexpr int kMaxRestart = 128;

__device__ __forceinline__ double nanmax(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return a > b ? a : b;
}

__device__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  const double out = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return out;
}

__device__ double wave_nanmax(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = nanmax(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ double block_nanmax(double v, double* red) {
  v = wave_nanmax(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  const double out = nanmax(nanmax(red[0], red[1]), nanmax(red[2], red[3]));
  __syncthreads();
  return out;
}

// Norm of a scaled vector: kind 0 max |.|, 1 two-norm, 2 rms.  Per-element
// contribution and how partials combine.
__device__ __forceinline__ double norm_elem(int kind, double v) { return kind == 0 ? fabs(v) : v * v; }
__device__ __forceinline__ double norm_block(int kind, double v, double* red) {
  return kind == 0 ? block_nanmax(v, red) : block_sum(v, red);
}
__device__ __forceinline__ double norm_accum(int kind, double a, double b) { return kind == 0 ? nanmax(a, b) : a + b; }
__device__ __forceinline__ double norm_final(int kind, double v, int64_t n) {
  if (kind == 0) return v;
  if (kind == 1) return sqrt(v);
  return n > 0 ? sqrt(v / static_cast<double>(n)) : 0.0;
}

// Row-block geometry: blockIdx.y is the row, blockIdx.x the chunk.
struct Chunk {
  int64_t begin, end;
};
__device__ __forceinline__ Chunk chunk_of(int64_t n) {
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  Chunk c;
  c.begin = static_cast<int64_t>(blockIdx.x) * per;
  c.end = c.begin + per < n ? c.begin + per : n;
  return c;
}

__device__ __forceinline__ bool row_active(const lxa_krylov_row* st) { return st[blockIdx.y].active != 0; }

// ---------------------------------------------------------------------------
// Generic reductions
// ---------------------------------------------------------------------------

// part[row * pstride + poff + blk] = sum x*y over the chunk.
template <typename TA, typename TB>
__global__ __launch_bounds__(kB) void dot_partial_kernel(int64_t n, const TA* __restrict__ x, const TB* __restrict__ y,
                                                         const lxa_krylov_row* __restrict__ st, int gate,
                                                         double* __restrict__ part, int64_t pstride, int64_t poff) {
  __shared__ double red[4];
  if (gate && !row_active(st)) return;
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(n);
  const TA* xr = x + row * n;
  const TB* yr = y + row * n;
  double acc = 0.0;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB)
    acc = fma(static_cast<double>(xr[i]), static_cast<double>(yr[i]), acc);
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) part[row * pstride + poff + blockIdx.x] = acc;
}

// Sum / NaN-propagating max of a row's block partials, in index order.  The
// loads go out sixteen at a time (independent of the running sum), so the
// one-thread-per-row finish kernels wait for nblk / 16 memory round trips,
// not nblk; the order of the additions is unchanged.
__device__ double sum_partials(const double* p, int nblk) {
  double s = 0.0;
  int k = 0;
  for (; k + 16 <= nblk; k += 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = p[k + u];
#pragma unroll
    for (int u = 0; u < 16; ++u) s += v[u];
  }
  for (; k < nblk; ++k) s += p[k];
  return s;
}

__device__ double max_partials(const double* p, int nblk) {
  double s = 0.0;
  int k = 0;
  for (; k + 16 <= nblk; k += 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = p[k + u];
#pragma unroll
    for (int u = 0; u < 16; ++u) s = nanmax(s, v[u]);
  }
  for (; k < nblk; ++k) s = nanmax(s, p[k]);
  return s;
}

// Three reductions over a row's [3][nblk] partials, each in index order:
// acc[q] accumulates with norm kind kinds[q] (0 NaN-propagating max, else
// sum).  Loads go out eight entries of each at a time.
__device__ void reduce3_partials(const double* pr, int nblk, const int (&kinds)[3], double (&acc)[3]) {
  int k = 0;
  for (; k + 8 <= nblk; k += 8) {
    double v[3][8];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int u = 0; u < 8; ++u) v[q][u] = pr[q * nblk + k + u];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int q = 0; q < 3; ++q) acc[q] = norm_accum(kinds[q], acc[q], v[q][u]);
  }
  for (; k < nblk; ++k)
#pragma unroll
    for (int q = 0; q < 3; ++q) acc[q] = norm_accum(kinds[q], acc[q], pr[q * nblk + k]);
}

// ---------------------------------------------------------------------------
// CG (lineax/_solver/cg.py:114-227)
// ---------------------------------------------------------------------------

struct CgParams {
  double rtol, atol, rcond;
  int64_t max_steps;
  int has_scale, norm_kind, max_steps_given;
};

// gamma = <z, r>; active = gamma > 0 & max_steps > 0 (the first test of
// cond_fun: diff = inf makes the tolerance test true).  cg.py:128-167.
__global__ void cg_begin_kernel(int64_t rows, int nblk, const double* __restrict__ part, CgParams prm,
                                lxa_krylov_row* __restrict__ st) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  lxa_krylov_row s = {};
  s.gamma = sum_partials(part + row * nblk, nblk);
  s.step = 0;
  s.norm_diff = INFINITY;
  s.active = (s.gamma > 0.0) && (prm.max_steps > 0);
  s.result = 0;
  st[row] = s;
}

// alpha = gamma / <Ap, p>, NaN when |<Ap, p>| <= 100 rcond |gamma| (cg.py:172-178).
__global__ void cg_alpha_kernel(int64_t rows, int nblk, const double* __restrict__ part, CgParams prm,
                                lxa_krylov_row* __restrict__ st) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows || !st[row].active) return;
  const double ip = sum_partials(part + row * nblk, nblk);
  const double gamma = st[row].gamma;
  double alpha = gamma / ip;
  if (!(fabs(ip) > 100.0 * prm.rcond * fabs(gamma))) alpha = __builtin_nan("");
  st[row].alpha = alpha;
  st[row].ip = ip;
}

// y += alpha p; unless `stable`, r -= alpha q.  Partials: [0] |diff / y_scale|
// norm contribution, and when fuse_gamma (identity preconditioner, not a
// stable step) [1] sum r^2 and [2] |r / b_scale| contribution.
template <typename TY, typename TV>
__global__ __launch_bounds__(kB) void cg_update_kernel(int64_t n, const TV* __restrict__ p, const TV* __restrict__ q,
                                                       TY* __restrict__ y, TV* __restrict__ r,
                                                       const TY* __restrict__ b, CgParams prm, int stable,
                                                       int fuse_gamma, const lxa_krylov_row* __restrict__ st,
                                                       double* __restrict__ part) {
  __shared__ double red[4];
  if (!row_active(st)) return;
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(n);
  const double alpha = st[row].alpha;
  const int kind = prm.norm_kind;
  double n2 = 0.0, rr = 0.0, n1 = 0.0;
  const int64_t o = row * n;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    const double d = alpha * static_cast<double>(p[o + i]);
    const double yn = static_cast<double>(y[o + i]) + d;
    y[o + i] = static_cast<TY>(yn);
    if (prm.has_scale) n2 = norm_accum(kind, n2, norm_elem(kind, d / (prm.atol + prm.rtol * fabs(yn))));
    if (!stable) {
      const double rn = static_cast<double>(r[o + i]) - alpha * static_cast<double>(q[o + i]);
      const TV rs = static_cast<TV>(rn);
      r[o + i] = rs;
      if (fuse_gamma) {
        const double rv = static_cast<double>(rs);
        rr = fma(rv, rv, rr);
        if (prm.has_scale)
          n1 = norm_accum(kind, n1,
                          norm_elem(kind, rv / (prm.atol + prm.rtol * fabs(static_cast<double>(b[o + i])))));
      }
    }
  }
  n2 = norm_block(kind, n2, red);
  if (fuse_gamma) {
    rr = block_sum(rr, red);
    n1 = norm_block(kind, n1, red);
  }
  if (threadIdx.x == 0) {
    const int64_t base = (row * 3) * gridDim.x + blockIdx.x;
    part[base] = n2;
    if (fuse_gamma) {
      part[base + gridDim.x] = rr;
      part[base + 2 * gridDim.x] = n1;
    }
  }
}

// r = b - t with t = A y in the vector dtype TY (stable step, cg.py:187-188);
// partials as cg_update [1],[2] when fuse_gamma.
template <typename TY, typename TV>
__global__ __launch_bounds__(kB) void cg_residual_kernel(int64_t n, const TY* __restrict__ b, const TY* __restrict__ t,
                                                         TV* __restrict__ r, CgParams prm, int fuse_gamma,
                                                         const lxa_krylov_row* __restrict__ st,
                                                         double* __restrict__ part) {
  __shared__ double red[4];
  if (!row_active(st)) return;
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(n);
  const int kind = prm.norm_kind;
  double rr = 0.0, n1 = 0.0;
  const int64_t o = row * n;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    const double bv = static_cast<double>(b[o + i]);
    const TV rs = static_cast<TV>(bv - static_cast<double>(t[o + i]));
    r[o + i] = rs;
    if (fuse_gamma) {
      const double rv = static_cast<double>(rs);
      rr = fma(rv, rv, rr);
      if (prm.has_scale) n1 = norm_accum(kind, n1, norm_elem(kind, rv / (prm.atol + prm.rtol * fabs(bv))));
    }
  }
  if (fuse_gamma) {
    rr = block_sum(rr, red);
    n1 = norm_block(kind, n1, red);
    if (threadIdx.x == 0) {
      const int64_t base = (row * 3) * gridDim.x + blockIdx.x;
      part[base + gridDim.x] = rr;
      part[base + 2 * gridDim.x] = n1;
    }
  }
}

// |r / b_scale| partials only (preconditioned solves), into part[1] and [2]:
// [1] receives <z, r> from a separate dot, [2] the norm contribution.
template <typename TY, typename TV>
__global__ __launch_bounds__(kB) void cg_rnorm_kernel(int64_t n, const TY* __restrict__ b, const TV* __restrict__ r,
                                                      CgParams prm, const lxa_krylov_row* __restrict__ st,
                                                      double* __restrict__ part) {
  __shared__ double red[4];
  if (!row_active(st)) return;
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(n);
  const int kind = prm.norm_kind;
  double n1 = 0.0;
  const int64_t o = row * n;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB)
    n1 = norm_accum(kind, n1,
                    norm_elem(kind, static_cast<double>(r[o + i]) /
                                        (prm.atol + prm.rtol * fabs(static_cast<double>(b[o + i])))));
  n1 = norm_block(kind, n1, red);
  if (threadIdx.x == 0) part[(row * 3 + 2) * gridDim.x + blockIdx.x] = n1;
}

// gamma' = <z, r>, beta = gamma'/gamma, step += 1, then cond_fun
// (cg.py:162-167, 202-206).  part layout [row][3][nblk].
__global__ void cg_finish_kernel(int64_t rows, int nblk, int64_t n, const double* __restrict__ part, CgParams prm,
                                 lxa_krylov_row* __restrict__ st) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows || !st[row].active) return;
  const double* pr = part + row * 3 * nblk;
  const int kind = prm.norm_kind;
  double acc[3] = {0.0, 0.0, 0.0};
  reduce3_partials(pr, nblk, {kind, 1, kind}, acc);
  const double n2 = acc[0], rr = acc[1], n1 = acc[2];
  lxa_krylov_row s = st[row];
  const double gamma_prev = s.gamma;
  s.gamma = rr;
  s.beta = rr / gamma_prev;
  s.step += 1;
  s.norm_r = norm_final(kind, n1, n);
  s.norm_diff = norm_final(kind, n2, n);
  bool not_converged = true;
  if (prm.has_scale) not_converged = (s.norm_r > 1.0) || (s.norm_diff > 1.0);
  s.active = (s.gamma > 0.0) && (s.step < prm.max_steps) && not_converged;
  st[row] = s;
}

// p = z + beta p (cg.py:206), rows that are still active.
template <typename TV>
__global__ __launch_bounds__(kB) void cg_direction_kernel(int64_t n, const TV* __restrict__ z, TV* __restrict__ p,
                                                          const lxa_krylov_row* __restrict__ st) {
  if (!row_active(st)) return;
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(n);
  const double beta = st[row].beta;
  const int64_t o = row * n;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB)
    p[o + i] = static_cast<TV>(static_cast<double>(z[o + i]) + beta * static_cast<double>(p[o + i]));
}

// RESULTS (cg.py:213-222): num_steps == max_steps -> singular (max_steps not
// given) or max_steps_reached (given, with tolerances).
__global__ void cg_result_kernel(int64_t rows, CgParams prm, lxa_krylov_row* __restrict__ st) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  int res = LXA_RESULT_SUCCESSFUL;
  const bool capped = st[row].step == prm.max_steps;
  if (!prm.max_steps_given) {
    if (capped) res = LXA_RESULT_SINGULAR;
  } else if (prm.has_scale) {
    if (capped) res = LXA_RESULT_MAX_STEPS_REACHED;
  }
  st[row].result = res;
}

// ---------------------------------------------------------------------------
// BiCGStab (lineax/_solver/bicgstab.py:78-205), right preconditioning.
// Row state: gamma = rho, alpha = alpha, ip = omega (lineax's carry),
// beta = <r0, r> of the current residual (the next iteration's rho_new),
// norm_r / norm_diff the two tolerance norms.  Iteration k:
//   direction: p = r + beta_k (p - omega v), beta_k = (rho_new / rho)(alpha / omega)
//   [caller: x = M p, v = A x]
//   alpha:     alpha = rho_new / <r0, v>, rho = rho_new
//   s:         s = r - alpha v  [caller: z = M s, t = A z]
//   omega:     omega = <s, t> / <t, t>
//   update:    diff = alpha x + omega z, y += diff, r = s - omega t, the
//              next rho_new = <r0, r>, the norms, step += 1 and cond_fun
//              (bicgstab.py:107-168) on the carry (omega, alpha, rho).
// ---------------------------------------------------------------------------

// Breakdown test (bicgstab.py:107-113): exact zeros when threshold == 0 (a
// float64 solve, as with jax_enable_x64), else value < threshold.
__device__ __forceinline__ bool bcg_breakdown(double omega, double alpha, double rho, double threshold) {
  if (threshold == 0.0) return omega == 0.0 || alpha == 0.0 || rho == 0.0;
  return omega < threshold || alpha < threshold || rho < threshold;
}

// r = r0 = b - t; partials [row][nblk] of <r0, r0> (the stored values).
template <typename TY, typename TV>
__global__ __launch_bounds__(kB) void bcg_begin_kernel(int64_t n, const TY* __restrict__ b, const TY* __restrict__ t,
                                                       TV* __restrict__ r, TV* __restrict__ r0,
                                                       double* __restrict__ part) {
  __shared__ double red[4];
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(n);
  const int64_t o = row * n;
  double acc = 0.0;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    const TV rv = static_cast<TV>(static_cast<double>(b[o + i]) - static_cast<double>(t[o + i]));
    r[o + i] = rv;
    r0[o + i] = rv;
    acc = fma(static_cast<double>(rv), static_cast<double>(rv), acc);
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) part[row * gridDim.x + blockIdx.x] = acc;
}

// The initial carry (bicgstab.py:170-183): alpha = omega = rho = 1, diff =
// inf, so the first cond_fun is step < max_steps.
__global__ void bcg_begin_finish_kernel(int64_t rows, int nblk, const double* __restrict__ part, CgParams prm,
                                        lxa_krylov_row* __restrict__ st) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  lxa_krylov_row s{};
  s.gamma = 1.0;
  s.alpha = 1.0;
  s.ip = 1.0;
  s.beta = sum_partials(part + row * nblk, nblk);
  s.norm_r = INFINITY;
  s.norm_diff = INFINITY;
  s.step = 0;
  s.active = prm.max_steps > 0;
  st[row] = s;
}

// p = r (first) or p = r + beta_k (p - omega v).
template <typename TV>
__global__ __launch_bounds__(kB) void bcg_direction_kernel(int64_t n, const TV* __restrict__ r, TV* __restrict__ p,
                                                           const TV* __restrict__ v, int first,
                                                           const lxa_krylov_row* __restrict__ st) {
  if (!row_active(st)) return;
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(n);
  const lxa_krylov_row s = st[row];
  const double beta = (s.beta / s.gamma) * (s.alpha / s.ip);
  const double omega = s.ip;
  const int64_t o = row * n;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    const double rv = static_cast<double>(r[o + i]);
    p[o + i] = first ? static_cast<TV>(rv)
                     : static_cast<TV>(rv + beta * (static_cast<double>(p[o + i]) -
                                                    omega * static_cast<double>(v[o + i])));
  }
}

// alpha = rho_new / <r0, v>; the carry's rho becomes rho_new.
__global__ void bcg_alpha_kernel(int64_t rows, int nblk, const double* __restrict__ part,
                                 lxa_krylov_row* __restrict__ st) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows || !st[row].active) return;
  const double rho_new = st[row].beta;
  st[row].alpha = rho_new / sum_partials(part + row * nblk, nblk);
  st[row].gamma = rho_new;
}

// s = r - alpha v.
template <typename TV>
__global__ __launch_bounds__(kB) void bcg_s_kernel(int64_t n, const TV* __restrict__ r, const TV* __restrict__ v,
                                                   TV* __restrict__ s, const lxa_krylov_row* __restrict__ st) {
  if (!row_active(st)) return;
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(n);
  const double alpha = st[row].alpha;
  const int64_t o = row * n;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB)
    s[o + i] = static_cast<TV>(static_cast<double>(r[o + i]) - alpha * static_cast<double>(v[o + i]));
}

// Partials [row][2][nblk]: <s, t> and <t, t> in one pass.
template <typename TV>
__global__ __launch_bounds__(kB) void bcg_st_kernel(int64_t n, const TV* __restrict__ s, const TV* __restrict__ t,
                                                    const lxa_krylov_row* __restrict__ st,
                                                    double* __restrict__ part) {
  __shared__ double red[4];
  if (!row_active(st)) return;
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(n);
  const int64_t o = row * n;
  double a = 0.0, q = 0.0;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    const double tv = static_cast<double>(t[o + i]);
    a = fma(static_cast<double>(s[o + i]), tv, a);
    q = fma(tv, tv, q);
  }
  a = block_sum(a, red);
  q = block_sum(q, red);
  if (threadIdx.x == 0) {
    part[(row * 2) * gridDim.x + blockIdx.x] = a;
    part[(row * 2 + 1) * gridDim.x + blockIdx.x] = q;
  }
}

__global__ void bcg_omega_kernel(int64_t rows, int nblk, const double* __restrict__ part,
                                 lxa_krylov_row* __restrict__ st) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows || !st[row].active) return;
  const double* pr = part + row * 2 * nblk;
  st[row].ip = sum_partials(pr, nblk) / sum_partials(pr + nblk, nblk);
}

// diff = alpha x + omega z; y += diff; r = s - omega t.  Partials
// [row][3][nblk]: [0] <r0, r>, [1] |r / b_scale|, [2] |diff / y_scale|.
template <typename TY, typename TV>
__global__ __launch_bounds__(kB) void bcg_update_kernel(int64_t n, const TV* __restrict__ x, const TV* __restrict__ z,
                                                        const TV* __restrict__ s, const TV* __restrict__ t,
                                                        const TV* __restrict__ r0, TY* __restrict__ y,
                                                        TV* __restrict__ r, const TY* __restrict__ b, CgParams prm,
                                                        const lxa_krylov_row* __restrict__ st,
                                                        double* __restrict__ part) {
  __shared__ double red[4];
  if (!row_active(st)) return;
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(n);
  const double alpha = st[row].alpha, omega = st[row].ip;
  const int kind = prm.norm_kind;
  const int64_t o = row * n;
  double rr = 0.0, n1 = 0.0, n2 = 0.0;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    const double d = alpha * static_cast<double>(x[o + i]) + omega * static_cast<double>(z[o + i]);
    const double yn = static_cast<double>(y[o + i]) + d;
    y[o + i] = static_cast<TY>(yn);
    const TV rs = static_cast<TV>(static_cast<double>(s[o + i]) - omega * static_cast<double>(t[o + i]));
    r[o + i] = rs;
    const double rv = static_cast<double>(rs);
    rr = fma(static_cast<double>(r0[o + i]), rv, rr);
    if (prm.has_scale) {
      n1 = norm_accum(kind, n1, norm_elem(kind, rv / (prm.atol + prm.rtol * fabs(static_cast<double>(b[o + i])))));
      n2 = norm_accum(kind, n2, norm_elem(kind, d / (prm.atol + prm.rtol * fabs(static_cast<double>(y[o + i])))));
    }
  }
  rr = block_sum(rr, red);
  n1 = norm_block(kind, n1, red);
  n2 = norm_block(kind, n2, red);
  if (threadIdx.x == 0) {
    const int64_t base = (row * 3) * gridDim.x + blockIdx.x;
    part[base] = rr;
    part[base + gridDim.x] = n1;
    part[base + 2 * gridDim.x] = n2;
  }
}

// step += 1, the next rho_new, the norms and cond_fun on the carry.
__global__ void bcg_update_finish_kernel(int64_t rows, int nblk, int64_t n, const double* __restrict__ part,
                                         CgParams prm, double threshold, lxa_krylov_row* __restrict__ st) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows || !st[row].active) return;
  const double* pr = part + row * 3 * nblk;
  const int kind = prm.norm_kind;
  double acc[3] = {0.0, 0.0, 0.0};
  reduce3_partials(pr, nblk, {1, kind, kind}, acc);
  const double rr = acc[0], n1 = acc[1], n2 = acc[2];
  lxa_krylov_row s = st[row];
  s.beta = rr;
  s.norm_r = norm_final(kind, n1, n);
  s.norm_diff = norm_final(kind, n2, n);
  s.step += 1;
  const bool not_converged = prm.has_scale ? (s.norm_r > 1.0 || s.norm_diff > 1.0) : true;
  s.active = !bcg_breakdown(s.ip, s.alpha, s.gamma, threshold) && not_converged && s.step < prm.max_steps;
  st[row] = s;
}

// RESULTS (bicgstab.py:188-203): the step cap as for CG, then breakdown
// when the last carry broke down without converging.
__global__ void bcg_result_kernel(int64_t rows, CgParams prm, double threshold, lxa_krylov_row* __restrict__ st) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  const lxa_krylov_row s = st[row];
  int res = LXA_RESULT_SUCCESSFUL;
  const bool capped = s.step == prm.max_steps;
  if (!prm.max_steps_given) {
    if (capped) res = LXA_RESULT_SINGULAR;
  } else if (prm.has_scale) {
    if (capped) res = LXA_RESULT_MAX_STEPS_REACHED;
  }
  const bool not_converged = prm.has_scale ? (s.norm_r > 1.0 || s.norm_diff > 1.0) : true;
  if (bcg_breakdown(s.ip, s.alpha, s.gamma, threshold) && not_converged) res = LXA_RESULT_BREAKDOWN;
  st[row].result = res;
}

// ---------------------------------------------------------------------------
// LSMR (lineax/_solver/lsmr.py:94-359), damp 0, two_norm.  Vectors of a row
// are stored unscaled; u / sel(beta) and v / sel(alpha) are the unit
// vectors of the bidiagonalisation (sel(z) = z, or 1 when z is 0, as
// lax.select in lsmr.py:142,153,219,223).
// ---------------------------------------------------------------------------

struct LsParams {
  double rtol, atol, conlim;
  int64_t max_steps;
  int has_scale, max_steps_given;
};

__device__ __forceinline__ double sel1(double z) { return z == 0.0 ? 1.0 : z; }
__device__ __forceinline__ double sgn(double z) { return z > 0.0 ? 1.0 : (z < 0.0 ? -1.0 : 0.0); }
__device__ __forceinline__ double nanmin(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return a < b ? a : b;
}

// c, s, r with [c -s; s c] [a; b] = [r; 0] (lsmr.py:361-409)
__device__ void ls_givens(double a, double b, double& c, double& s, double& r) {
  if (a == 0.0 || b == 0.0) {
    if (b == 0.0) {
      c = sgn(a);
      s = 0.0;
      r = fabs(a);
    } else {
      c = 0.0;
      s = sgn(b);
      r = fabs(b);
    }
  } else if (fabs(b) > fabs(a)) {
    const double tau = a / b;
    s = sgn(b) / sqrt(1.0 + tau * tau);
    c = s * tau;
    r = b / s;
  } else {
    const double tau = b / a;
    c = sgn(a) / sqrt(1.0 + tau * tau);
    s = c * tau;
    r = a / c;
  }
}

__device__ __forceinline__ bool ls_active(const lxa_lsmr_row* st) { return st[blockIdx.y].active != 0; }

// u = b - t; partials [row][2][nblk]: sum u^2, sum b^2.
template <typename TY, typename TV>
__global__ __launch_bounds__(kB) void ls_begin_kernel(int64_t m, const TY* __restrict__ b, const TY* __restrict__ t,
                                                      TV* __restrict__ u, double* __restrict__ part) {
  __shared__ double red[4];
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(m);
  const int64_t o = row * m;
  double uu = 0.0, bb = 0.0;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    const double bv = static_cast<double>(b[o + i]);
    const TV uv = static_cast<TV>(bv - static_cast<double>(t[o + i]));
    u[o + i] = uv;
    uu = fma(static_cast<double>(uv), static_cast<double>(uv), uu);
    bb = fma(bv, bv, bb);
  }
  uu = block_sum(uu, red);
  bb = block_sum(bb, red);
  if (threadIdx.x == 0) {
    part[(row * 2) * gridDim.x + blockIdx.x] = uu;
    part[(row * 2 + 1) * gridDim.x + blockIdx.x] = bb;
  }
}

__global__ void ls_begin_finish_kernel(int64_t rows, int nblk, const double* __restrict__ part,
                                       lxa_lsmr_row* __restrict__ st) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  const double* pr = part + row * 2 * nblk;
  lxa_lsmr_row s{};
  s.beta = sqrt(sum_partials(pr, nblk));
  s.normb = sqrt(sum_partials(pr + nblk, nblk));
  s.active = 1;
  st[row] = s;
}

// v = z / beta (beta nonzero) or 0; partials [row][nblk]: sum v^2.
template <typename TV>
__global__ __launch_bounds__(kB) void ls_v0_kernel(int64_t n, const TV* __restrict__ z, TV* __restrict__ v,
                                                   const lxa_lsmr_row* __restrict__ st, double* __restrict__ part) {
  __shared__ double red[4];
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(n);
  const double beta = st[row].beta;
  const int64_t o = row * n;
  double vv = 0.0;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    const TV x = beta != 0.0 ? static_cast<TV>(static_cast<double>(z[o + i]) / beta) : TV(0);
    v[o + i] = x;
    vv = fma(static_cast<double>(x), static_cast<double>(x), vv);
  }
  vv = block_sum(vv, red);
  if (threadIdx.x == 0) part[row * gridDim.x + blockIdx.x] = vv;
}

// alpha and the loop state of lsmr.py:137-200.
__global__ void ls_init_finish_kernel(int64_t rows, int nblk, const double* __restrict__ part,
                                      lxa_lsmr_row* __restrict__ st) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  lxa_lsmr_row s = st[row];
  const double alpha = s.beta != 0.0 ? sqrt(sum_partials(part + row * nblk, nblk)) : 0.0;
  s.alpha = alpha;
  s.zetabar = alpha * s.beta;
  s.alphabar = alpha;
  s.rho = 1.0;
  s.rhobar = 1.0;
  s.cbar = 1.0;
  s.sbar = 0.0;
  s.betadd = s.beta;
  s.betad = 0.0;
  s.rhodold = 1.0;
  s.tautildeold = 0.0;
  s.thetatilde = 0.0;
  s.zeta = 0.0;
  s.delta = 0.0;
  s.normA2 = alpha * alpha;
  s.maxrbar = 0.0;
  s.minrbar = 1.79769313486231570815e308;  // the largest double (jnp.finfo(dtype).max in float64)
  s.condA = 1.0;
  s.normr = s.beta;
  s.normAr = alpha * s.beta;
  s.normx = 0.0;
  s.itn = 0;
  s.istop = 0;
  if (alpha == 0.0) s.istop = 2;  // x0 solves the least-squares problem
  if (s.beta == 0.0) s.istop = 1;  // x0 solves A x = b
  s.active = s.istop == 0;
  st[row] = s;
}

// h = v / alpha, hbar = 0.
template <typename TV>
__global__ __launch_bounds__(kB) void ls_h0_kernel(int64_t n, const TV* __restrict__ v, TV* __restrict__ h,
                                                   TV* __restrict__ hbar, const lxa_lsmr_row* __restrict__ st) {
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(n);
  const double ia = 1.0 / sel1(st[row].alpha);
  const int64_t o = row * n;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    h[o + i] = static_cast<TV>(static_cast<double>(v[o + i]) * ia);
    hbar[o + i] = TV(0);
  }
}

// u = w / alpha - alpha u / beta (w = A v unscaled; lsmr.py:214-219);
// partials [row][nblk]: sum u^2.
template <typename TV>
__global__ __launch_bounds__(kB) void ls_u_kernel(int64_t m, const TV* __restrict__ w, TV* __restrict__ u,
                                                  const lxa_lsmr_row* __restrict__ st, double* __restrict__ part) {
  __shared__ double red[4];
  if (!ls_active(st)) return;
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(m);
  const double alpha = st[row].alpha;
  const double ia = 1.0 / sel1(alpha), ab = alpha / sel1(st[row].beta);
  const int64_t o = row * m;
  double uu = 0.0;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    const TV x = static_cast<TV>(static_cast<double>(w[o + i]) * ia - ab * static_cast<double>(u[o + i]));
    u[o + i] = x;
    uu = fma(static_cast<double>(x), static_cast<double>(x), uu);
  }
  uu = block_sum(uu, red);
  if (threadIdx.x == 0) part[row * gridDim.x + blockIdx.x] = uu;
}

__global__ void ls_u_finish_kernel(int64_t rows, int nblk, const double* __restrict__ part,
                                   lxa_lsmr_row* __restrict__ st) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows || !st[row].active) return;
  st[row].beta = sqrt(sum_partials(part + row * nblk, nblk));
}

// When beta is nonzero: v = z / beta - beta v / alpha (z = A^T u unscaled;
// lsmr.py:220-223); partials [row][nblk]: sum v^2.
template <typename TV>
__global__ __launch_bounds__(kB) void ls_v_kernel(int64_t n, const TV* __restrict__ z, TV* __restrict__ v,
                                                  const lxa_lsmr_row* __restrict__ st, double* __restrict__ part) {
  __shared__ double red[4];
  if (!ls_active(st)) return;
  const int64_t row = blockIdx.y;
  const double beta = st[row].beta;
  if (beta == 0.0) return;  // lsmr.py:226-230: u, v and alpha unchanged (uniform per row)
  const Chunk c = chunk_of(n);
  const double ib = 1.0 / beta, ba = beta / sel1(st[row].alpha);
  const int64_t o = row * n;
  double vv = 0.0;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    const TV x = static_cast<TV>(static_cast<double>(z[o + i]) * ib - ba * static_cast<double>(v[o + i]));
    v[o + i] = x;
    vv = fma(static_cast<double>(x), static_cast<double>(x), vv);
  }
  vv = block_sum(vv, red);
  if (threadIdx.x == 0) part[row * gridDim.x + blockIdx.x] = vv;
}

// The new alpha, then the rotations and estimates of lsmr.py:240-312.
__global__ void ls_v_finish_kernel(int64_t rows, int nblk, const double* __restrict__ part,
                                   lxa_lsmr_row* __restrict__ st) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows || !st[row].active) return;
  lxa_lsmr_row s = st[row];
  s.itn += 1;
  if (s.beta != 0.0) s.alpha = sqrt(sum_partials(part + row * nblk, nblk));
  const double alpha = s.alpha, beta = s.beta;
  // Qhat_{k,2k+1} (damp = 0), then Q_i turning B_i into R_i
  double chat, shat, alphahat;
  ls_givens(s.alphabar, 0.0, chat, shat, alphahat);
  const double rhoold = s.rho;
  double c, sn, rho;
  ls_givens(alphahat, beta, c, sn, rho);
  s.rho = rho;
  const double thetanew = sn * alpha;
  s.alphabar = c * alpha;
  // Qbar_i turning R_i^T into R_i^bar
  const double rhobarold = s.rhobar, zetaold = s.zeta;
  const double thetabar = s.sbar * rho, rhotemp = s.cbar * rho;
  double cbar, sbar, rhobar;
  ls_givens(s.cbar * rho, thetanew, cbar, sbar, rhobar);
  s.cbar = cbar;
  s.sbar = sbar;
  s.rhobar = rhobar;
  s.zeta = cbar * s.zetabar;
  s.zetabar = -sbar * s.zetabar;
  // coefficients of the h, hbar and x updates (lsmr.py:260-272)
  s.c_hbar = -(thetabar * rho / (rhoold * rhobarold));
  s.c_x = s.zeta / (rho * rhobar);
  s.c_h = -(thetanew / rho);
  // estimate of |r| (lsmr.py:274-300)
  const double betaacute = chat * s.betadd, betacheck = -shat * s.betadd;
  const double betahat = c * betaacute;
  s.betadd = -sn * betaacute;
  const double thetatildeold = s.thetatilde;
  double ctildeold, stildeold, rhotildeold;
  ls_givens(s.rhodold, thetabar, ctildeold, stildeold, rhotildeold);
  s.thetatilde = stildeold * rhobar;
  s.rhodold = ctildeold * rhobar;
  s.betad = -stildeold * s.betad + ctildeold * betahat;
  s.tautildeold = (zetaold - thetatildeold * s.tautildeold) / rhotildeold;
  const double taud = (s.zeta - s.thetatilde * s.tautildeold) / s.rhodold;
  s.delta = s.delta + betacheck * betacheck;
  s.normr = sqrt(s.delta + (s.betad - taud) * (s.betad - taud) + s.betadd * s.betadd);
  // |A| and cond(A) (lsmr.py:302-312)
  s.normA2 = s.normA2 + beta * beta;
  s.normA = sqrt(s.normA2);
  s.normA2 = s.normA2 + alpha * alpha;
  s.maxrbar = nanmax(s.maxrbar, rhobarold);
  if (s.itn > 1) s.minrbar = nanmin(s.minrbar, rhobarold);
  s.condA = nanmax(s.maxrbar, rhotemp) / nanmin(s.minrbar, rhotemp);
  s.normAr = fabs(s.zetabar);
  st[row] = s;
}

// hbar = h + c_hbar hbar; x += c_x hbar; h = v / alpha + c_h h; partials
// [row][nblk]: sum x^2.
template <typename TY, typename TV>
__global__ __launch_bounds__(kB) void ls_x_kernel(int64_t n, const TV* __restrict__ v, TV* __restrict__ h,
                                                  TV* __restrict__ hbar, TY* __restrict__ x,
                                                  const lxa_lsmr_row* __restrict__ st, double* __restrict__ part) {
  __shared__ double red[4];
  if (!ls_active(st)) return;
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(n);
  const lxa_lsmr_row& s = st[row];
  const double chb = s.c_hbar, cx = s.c_x, ch = s.c_h, ia = 1.0 / sel1(s.alpha);
  const int64_t o = row * n;
  double xx = 0.0;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    const double hv = static_cast<double>(h[o + i]);
    const TV hb = static_cast<TV>(hv + chb * static_cast<double>(hbar[o + i]));
    hbar[o + i] = hb;
    const TY xn = static_cast<TY>(static_cast<double>(x[o + i]) + cx * static_cast<double>(hb));
    x[o + i] = xn;
    h[o + i] = static_cast<TV>(static_cast<double>(v[o + i]) * ia + ch * hv);
    xx = fma(static_cast<double>(xn), static_cast<double>(xn), xx);
  }
  xx = block_sum(xx, red);
  if (threadIdx.x == 0) part[row * gridDim.x + blockIdx.x] = xx;
}

// |x| and the stopping rules (lsmr.py:313-329).
__global__ void ls_x_finish_kernel(int64_t rows, int nblk, const double* __restrict__ part, LsParams prm,
                                   lxa_lsmr_row* __restrict__ st) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows || !st[row].active) return;
  lxa_lsmr_row s = st[row];
  s.normx = sqrt(sum_partials(part + row * nblk, nblk));
  const double well_posed_tol = prm.atol + prm.rtol * (s.normA * s.normx + s.normb);
  const double least_squares_tol = prm.atol + prm.rtol * (s.normA * s.normr);
  if (s.itn >= prm.max_steps) s.istop = 4;
  if (s.condA > prm.conlim) s.istop = 3;
  if (s.normAr < least_squares_tol) s.istop = 2;
  if (s.normr < well_posed_tol) s.istop = 1;
  s.active = s.istop == 0;
  st[row] = s;
}

__global__ void ls_result_kernel(int64_t rows, LsParams prm, lxa_lsmr_row* __restrict__ st) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  const lxa_lsmr_row s = st[row];
  int res = LXA_RESULT_SUCCESSFUL;
  const bool capped = s.itn == prm.max_steps;
  if (!prm.max_steps_given) {
    if (capped) res = LXA_RESULT_SINGULAR;
  } else if (prm.has_scale) {
    if (capped) res = LXA_RESULT_MAX_STEPS_REACHED;
  }
  if (s.istop < 3) res = LXA_RESULT_SUCCESSFUL;
  if (s.istop == 3) res = LXA_RESULT_CONLIM;
  st[row].result = res;
}

// ---------------------------------------------------------------------------
// GMRES (lineax/_solver/gmres.py:106-413)
// ---------------------------------------------------------------------------

struct GmParams {
  double rtol, atol, eps;
  int64_t max_steps;
  int has_scale, norm_kind, max_steps_given, restart, stagnation_iters;
};

// part[row][blk] = sum x^2 (TV storage).
template <typename TV>
__global__ __launch_bounds__(kB) void sumsq_partial_kernel(int64_t n, const TV* __restrict__ x, int64_t row_stride,
                                                           const lxa_krylov_row* __restrict__ st,
                                                           double* __restrict__ part) {
  __shared__ double red[4];
  if (!row_active(st)) return;
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(n);
  const TV* xr = x + row * row_stride;
  double acc = 0.0;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    const double v = static_cast<double>(xr[i]);
    acc = fma(v, v, acc);
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) part[row * gridDim.x + blockIdx.x] = acc;
}

// Start of a cycle (gmres.py:255-268): r_norm = |r|_2, initial breakdown
// if r_norm < eps; the Hessenberg buffer becomes eye(m, m+1); inner step 0.
__global__ void gm_cycle_begin_kernel(int64_t rows, int nblk, int64_t n, const double* __restrict__ part,
                                      GmParams prm, double* __restrict__ H, lxa_krylov_row* __restrict__ st,
                                      lxa_gmres_row* __restrict__ gs) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows || !st[row].active) return;
  const int m = prm.restart;
  const double ss = sum_partials(part + row * nblk, nblk);
  // two_norm of a single element is |x| (lineax/_norm.py:74-80): sqrt(x^2) equals it.
  const double nrm = sqrt(ss);
  lxa_gmres_row g = gs[row];
  g.r_norm = nrm;
  g.breakdown = nrm < prm.eps ? 1 : 0;
  g.scale = g.breakdown ? 0.0 : 1.0 / nrm;
  g.inner = 0;
  g.columns = 1;
  g.inner_done = g.breakdown;
  gs[row] = g;
  double* h = H + row * static_cast<int64_t>(m) * (m + 1);
  for (int i = 0; i < m; ++i)
    for (int k = 0; k <= m; ++k) h[i * (m + 1) + k] = (i == k) ? 1.0 : 0.0;
}

// dst = src * scale(row): normalisation of the starting vector / a new basis
// vector (gmres.py:401-413).  `which` 0 uses r_norm's scale, 1 the new
// column's scale.
template <typename TS, typename TD>
__global__ __launch_bounds__(kB) void gm_scale_kernel(int64_t n, const TS* __restrict__ src, int64_t src_stride,
                                                      TD* __restrict__ dst, int64_t dst_stride,
                                                      const lxa_krylov_row* __restrict__ st,
                                                      const lxa_gmres_row* __restrict__ gs, int gate_inner) {
  if (!row_active(st)) return;
  const int64_t row = blockIdx.y;
  if (gate_inner && gs[row].inner_done) return;
  const Chunk c = chunk_of(n);
  const double sc = gs[row].scale;
  const TS* s = src + row * src_stride;
  TD* d = dst + row * dst_stride;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) d[i] = static_cast<TD>(static_cast<double>(s[i]) * sc);
}

// h_k = <v_k, w> for k <= j and |w|^2 (gmres.py:355-363).
// part[row][k][blk], k = 0..j for the dots, k = m + 1 for |w|^2.
template <typename TV>
__global__ __launch_bounds__(kB) void gm_project_kernel(int64_t n, const TV* __restrict__ V, const TV* __restrict__ w,
                                                        int j, int m, const lxa_krylov_row* __restrict__ st,
                                                        const lxa_gmres_row* __restrict__ gs, double* __restrict__ part) {
  __shared__ double red[4];
  if (!row_active(st)) return;
  const int64_t row = blockIdx.y;
  if (gs[row].inner_done) return;
  const Chunk c = chunk_of(n);
  const TV* wr = w + row * n;
  const TV* Vr = V + row * static_cast<int64_t>(m + 1) * n;
  const int64_t pbase = row * static_cast<int64_t>(m + 2) * gridDim.x + blockIdx.x;
  double ww = 0.0;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    const double v = static_cast<double>(wr[i]);
    ww = fma(v, v, ww);
  }
  ww = block_sum(ww, red);
  if (threadIdx.x == 0) part[pbase + static_cast<int64_t>(m + 1) * gridDim.x] = ww;
  // eight basis vectors per pass over the chunk: eight independent loads per
  // element and one block reduction per eight coefficients
  __shared__ double red8[kB / 64][8];
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  for (int k0 = 0; k0 <= j; k0 += 8) {
    double acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = 0.0;
    for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
      const double wi = static_cast<double>(wr[i]);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (k0 + u <= j) acc[u] = fma(static_cast<double>(Vr[static_cast<int64_t>(k0 + u) * n + i]), wi, acc[u]);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = wave_sum(acc[u]);
    if (ln == 0) {
#pragma unroll
      for (int u = 0; u < 8; ++u) red8[wv][u] = acc[u];
    }
    __syncthreads();
    if (threadIdx.x < 8 && k0 + static_cast<int>(threadIdx.x) <= j) {
      const int u = threadIdx.x;
      const double v = (red8[0][u] + red8[1][u]) + (red8[2][u] + red8[3][u]);
      part[pbase + static_cast<int64_t>(k0 + u) * gridDim.x] = v;
    }
    __syncthreads();
  }
}

// Finalize the projections: proj[row][k] = h_k, and step_norm = |w|_2.  One
// block per row; each wave sums whole coefficients (lanes strided over the
// block partials, then a fixed butterfly), so the order is fixed.
__global__ __launch_bounds__(kB) void gm_project_finish_kernel(int64_t rows, int nblk, int j, int m,
                                                               const double* __restrict__ part,
                                                               double* __restrict__ proj,
                                                               const lxa_krylov_row* __restrict__ st,
                                                               lxa_gmres_row* __restrict__ gs) {
  const int64_t row = blockIdx.x;
  if (row >= rows || !st[row].active || gs[row].inner_done) return;
  const double* pr = part + row * static_cast<int64_t>(m + 2) * nblk;
  double* h = proj + row * (m + 1);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (int k = w; k <= m + 1; k += kB / 64) {
    double v = 0.0;
    if (k <= j || k == m + 1)
      for (int q = l; q < nblk; q += 64) v += pr[static_cast<int64_t>(k) * nblk + q];
    v = wave_sum(v);
    if (l == 0) {
      if (k == m + 1) gs[row].step_norm = sqrt(v);
      else h[k] = k <= j ? v : 0.0;
    }
  }
}

// w' = w - sum_{k<=j} h_k v_k, written to basis column j+1 (unnormalised);
// partial |w'|^2 (gmres.py:364-366).
template <typename TV>
__global__ __launch_bounds__(kB) void gm_orthogonalise_kernel(int64_t n, TV* __restrict__ V, const TV* __restrict__ w,
                                                              const double* __restrict__ proj, int j, int m,
                                                              const lxa_krylov_row* __restrict__ st,
                                                              const lxa_gmres_row* __restrict__ gs,
                                                              double* __restrict__ part) {
  __shared__ double red[4];
  __shared__ double hs[kMaxRestart + 1];
  if (!row_active(st)) return;
  const int64_t row = blockIdx.y;
  if (gs[row].inner_done) return;
  const Chunk c = chunk_of(n);
  const TV* wr = w + row * n;
  TV* Vr = V + row * static_cast<int64_t>(m + 1) * n;
  for (int k = threadIdx.x; k <= j; k += kB) hs[k] = proj[row * (m + 1) + k];
  __syncthreads();
  TV* out = Vr + static_cast<int64_t>(j + 1) * n;
  double acc = 0.0;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    double v = static_cast<double>(wr[i]);
    for (int k = 0; k <= j; ++k) v -= hs[k] * static_cast<double>(Vr[static_cast<int64_t>(k) * n + i]);
    const TV vs = static_cast<TV>(v);
    out[i] = vs;
    const double vv = static_cast<double>(vs);
    acc = fma(vv, vv, acc);
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) part[row * gridDim.x + blockIdx.x] = acc;
}

// |w'|, breakdown test |w'| < |w| eps, Hessenberg row j (gmres.py:367-398),
// scale for the new column, inner loop bookkeeping (gmres.py:269-271).
__global__ void gm_arnoldi_finish_kernel(int64_t rows, int nblk, int j, int m, const double* __restrict__ part,
                                         double* __restrict__ proj, GmParams prm, double* __restrict__ H,
                                         const lxa_krylov_row* __restrict__ st, lxa_gmres_row* __restrict__ gs) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows || !st[row].active || gs[row].inner_done) return;
  lxa_gmres_row g = gs[row];
  const double nrm = sqrt(sum_partials(part + row * nblk, nblk));
  const double eps = g.step_norm * prm.eps;
  const int bd = nrm < eps;
  g.scale = bd ? 0.0 : 1.0 / nrm;
  double* h = proj + row * (m + 1);
  h[j + 1] = nrm;
  double* Hr = H + row * static_cast<int64_t>(m) * (m + 1);
  for (int k = 0; k <= m; ++k) Hr[j * (m + 1) + k] = h[k];
  g.breakdown = bd;
  g.inner = j + 1;
  g.columns = j + 2;
  g.inner_done = (g.inner >= m) || bd;
  gs[row] = g;
}

// Least squares min |H^T z - r_norm e1| by Householder QR, one workgroup per
// row: the nested linear_solve(MatrixLinearOperator(coeff_mat.T), beta_vec,
// QR()) of gmres.py:295-303 (geqrf, Q^T b, back substitution).
__global__ __launch_bounds__(64) void gm_lstsq_kernel(int64_t rows, int m, const double* __restrict__ H,
                                                      double* __restrict__ z, const lxa_krylov_row* __restrict__ st,
                                                      const lxa_gmres_row* __restrict__ gs) {
  extern __shared__ double sm[];
  const int64_t row = blockIdx.x;
  if (!st[row].active) return;
  const int M = m + 1;  // rows of A = H^T
  double* A = sm;       // [M][m] row-major
  double* c = sm + M * m;
  double* tau = c + M;
  const double* Hr = H + row * static_cast<int64_t>(m) * (m + 1);
  const int l = threadIdx.x;
  for (int t = l; t < M * m; t += 64) {
    const int i = t / m, k = t % m;
    A[i * m + k] = Hr[k * (m + 1) + i];  // A = coeff_mat^T
  }
  for (int i = l; i < M; i += 64) c[i] = (i == 0) ? gs[row].r_norm : 0.0;
  __syncthreads();
  for (int k = 0; k < m; ++k) {
    // dlarfg on column k, rows k..M-1
    double part = 0.0;
    for (int i = k + 1 + l; i < M; i += 64) part = fma(A[i * m + k], A[i * m + k], part);
    part = wave_sum(part);
    const double alpha = A[k * m + k];
    const double xnorm = sqrt(part);
    double tk = 0.0, beta = alpha, scal = 0.0;
    if (xnorm != 0.0) {
      beta = -copysign(hypot(alpha, xnorm), alpha);
      tk = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    __syncthreads();
    for (int i = k + 1 + l; i < M; i += 64) A[i * m + k] *= (xnorm != 0.0 ? scal : 1.0);
    if (l == 0) {
      A[k * m + k] = beta;
      tau[k] = tk;
    }
    __syncthreads();
    // apply H_k to the remaining columns and to c
    for (int col = k + 1 + l; col <= m; col += 64) {
      // col == m means the right-hand side c
      double wsum = (col == m) ? c[k] : A[k * m + col];
      for (int i = k + 1; i < M; ++i) wsum += A[i * m + k] * ((col == m) ? c[i] : A[i * m + col]);
      wsum *= tk;
      if (col == m) {
        c[k] -= wsum;
        for (int i = k + 1; i < M; ++i) c[i] -= A[i * m + k] * wsum;
      } else {
        A[k * m + col] -= wsum;
        for (int i = k + 1; i < M; ++i) A[i * m + col] -= A[i * m + k] * wsum;
      }
    }
    __syncthreads();
  }
  if (l == 0) {
    double* zr = z + row * m;
    for (int i = m - 1; i >= 0; --i) {
      double s = c[i];
      for (int k = i + 1; k < m; ++k) s -= A[i * m + k] * zr[k];
      zr[i] = s / A[i * m + i];
    }
  }
}

// diff = sum_k z_k v_k over the written columns; y += diff; partial
// |diff / y_scale| (gmres.py:304-310, 130-141).
// colscale (may be null): basis column k is stored as v_k / colscale[k]
// (the fused stencil Arnoldi steps), so its coefficient is z_k colscale[k].
template <typename TY, typename TV>
__global__ __launch_bounds__(kB) void gm_update_kernel(int64_t n, const TV* __restrict__ V, const double* __restrict__ z,
                                                       int m, TY* __restrict__ y, GmParams prm,
                                                       const lxa_krylov_row* __restrict__ st,
                                                       const lxa_gmres_row* __restrict__ gs,
                                                       const double* __restrict__ colscale,
                                                       double* __restrict__ part) {
  __shared__ double red[4];
  __shared__ double zs[kMaxRestart];
  if (!row_active(st)) return;
  const int64_t row = blockIdx.y;
  const int cols = gs[row].columns < m ? gs[row].columns : m;
  const Chunk c = chunk_of(n);
  const TV* Vr = V + row * static_cast<int64_t>(m + 1) * n;
  for (int k = threadIdx.x; k < m; k += kB)
    zs[k] = colscale ? z[row * m + k] * colscale[row * (m + 1) + k] : z[row * m + k];
  __syncthreads();
  const int kind = prm.norm_kind;
  double n2 = 0.0;
  auto point = [&](int64_t i, double d) {
    const double yn = static_cast<double>(y[row * n + i]) + d;
    y[row * n + i] = static_cast<TY>(yn);
    if (prm.has_scale) n2 = norm_accum(kind, n2, norm_elem(kind, d / (prm.atol + prm.rtol * fabs(yn))));
  };
  bool done = false;
  if constexpr (sizeof(TV) == 4) {
    if (n % 4 == 0) {
      done = true;
    // four points per thread with 16-byte basis loads (one contiguous 1 KiB
    // request per wave and column instead of 256 bytes)
    const int64_t q0 = c.begin & ~int64_t(3), q1 = blockIdx.x + 1 == gridDim.x ? n : (c.end & ~int64_t(3));
    for (int64_t i = q0 + 4 * static_cast<int64_t>(threadIdx.x); i < q1; i += 4 * kB) {
      double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
      for (int k = 0; k < cols; ++k) {
        const float4 v = *reinterpret_cast<const float4*>(Vr + static_cast<int64_t>(k) * n + i);
        d0 = fma(zs[k], static_cast<double>(v.x), d0);
        d1 = fma(zs[k], static_cast<double>(v.y), d1);
        d2 = fma(zs[k], static_cast<double>(v.z), d2);
        d3 = fma(zs[k], static_cast<double>(v.w), d3);
      }
      point(i, d0);
      point(i + 1, d1);
      point(i + 2, d2);
      point(i + 3, d3);
    }
    }
  }
  if (!done) {
    for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
      double d = 0.0;
      for (int k = 0; k < cols; ++k) d = fma(zs[k], static_cast<double>(Vr[static_cast<int64_t>(k) * n + i]), d);
      point(i, d);
    }
  }
  n2 = norm_block(kind, n2, red);
  if (threadIdx.x == 0) part[(row * 3) * gridDim.x + blockIdx.x] = n2;
}

// r = b - t (then the caller may precondition); partials: [1] norm(r) in
// the solver's norm (stagnation, gmres.py:175), [2] |r / b_scale|.
template <typename TY, typename TV, typename TR>
__global__ __launch_bounds__(kB) void gm_residual_kernel(int64_t n, const TY* __restrict__ b, const TV* __restrict__ t,
                                                         TR* __restrict__ r, GmParams prm, int write,
                                                         const lxa_krylov_row* __restrict__ st,
                                                         double* __restrict__ part) {
  __shared__ double red[4];
  if (!row_active(st)) return;
  const int64_t row = blockIdx.y;
  const Chunk c = chunk_of(n);
  const int kind = prm.norm_kind;
  double nr = 0.0, n1 = 0.0;
  const int64_t o = row * n;
  for (int64_t i = c.begin + threadIdx.x; i < c.end; i += kB) {
    const double bv = static_cast<double>(b[o + i]);
    double rv;
    if (write) {
      rv = static_cast<double>(static_cast<TR>(bv - static_cast<double>(t[o + i])));
      r[o + i] = static_cast<TR>(rv);
    } else {
      rv = static_cast<double>(r[o + i]);  // r already holds M (b - A y)
    }
    nr = norm_accum(kind, nr, norm_elem(kind, rv));
    if (prm.has_scale) n1 = norm_accum(kind, n1, norm_elem(kind, rv / (prm.atol + prm.rtol * fabs(bv))));
  }
  nr = norm_block(kind, nr, red);
  n1 = norm_block(kind, n1, red);
  if (threadIdx.x == 0) {
    const int64_t base = (row * 3) * gridDim.x + blockIdx.x;
    part[base + gridDim.x] = nr;
    part[base + 2 * gridDim.x] = n1;
  }
}

// End of a restart cycle (gmres.py:143-190): stagnation counter, deferred
// breakdown, step += 1, then cond_fun.  first_pass: diff = inf.
__global__ void gm_cycle_end_kernel(int64_t rows, int nblk, int64_t n, const double* __restrict__ part, GmParams prm,
                                    int first_pass, lxa_krylov_row* __restrict__ st, lxa_gmres_row* __restrict__ gs) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows || !st[row].active) return;
  const int kind = prm.norm_kind;
  const double* pr = part + row * 3 * nblk;
  double acc[3] = {0.0, 0.0, 0.0};
  reduce3_partials(pr, nblk, {kind, kind, kind}, acc);
  const double n2 = first_pass ? 0.0 : acc[0], nr = acc[1], n1 = acc[2];
  lxa_krylov_row s = st[row];
  lxa_gmres_row g = gs[row];
  const double r_new_norm = norm_final(kind, nr, n);
  if ((r_new_norm - g.r_min) < 0.0) {
    g.stagnation = 0;
  } else {
    g.stagnation += 1;
  }
  // jnp.minimum propagates NaN
  g.r_min = (r_new_norm != r_new_norm || g.r_min != g.r_min) ? __builtin_nan("") : fmin(r_new_norm, g.r_min);
  const int new_breakdown = first_pass ? 0 : g.breakdown;
  g.deferred = g.carry_breakdown;
  g.carry_breakdown = new_breakdown;
  s.step += 1;
  s.norm_r = norm_final(kind, n1, n);
  s.norm_diff = first_pass ? INFINITY : norm_final(kind, n2, n);
  bool not_converged = true;
  if (prm.has_scale) not_converged = (s.norm_r > 1.0) || (s.norm_diff > 1.0);
  const bool cont = (!g.deferred) && (g.stagnation < prm.stagnation_iters) && not_converged &&
                    (s.step < prm.max_steps);
  s.active = cont ? 1 : 0;
  g.inner_done = 0;
  st[row] = s;
  gs[row] = g;
}

// RESULTS (gmres.py:217-237).
__global__ void gm_result_kernel(int64_t rows, GmParams prm, lxa_krylov_row* __restrict__ st,
                                 const lxa_gmres_row* __restrict__ gs) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  const lxa_krylov_row s = st[row];
  const lxa_gmres_row g = gs[row];
  int res = LXA_RESULT_SUCCESSFUL;
  const bool capped = s.step == prm.max_steps;
  if (!prm.max_steps_given) {
    if (capped) res = LXA_RESULT_SINGULAR;
  } else if (prm.has_scale) {
    if (capped) res = LXA_RESULT_MAX_STEPS_REACHED;
  }
  if (g.stagnation >= prm.stagnation_iters) res = LXA_RESULT_STAGNATION;
  bool not_converged = true;
  if (prm.has_scale) not_converged = (s.norm_r > 1.0) || (s.norm_diff > 1.0);
  if (g.deferred && not_converged) res = LXA_RESULT_BREAKDOWN;
  st[row].result = res;
}

__global__ void gm_init_kernel(int64_t rows, lxa_krylov_row* __restrict__ st, lxa_gmres_row* __restrict__ gs) {
  const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  lxa_krylov_row s = {};
  s.active = 1;  // the dummy first pass always runs (cond_fun's `step == 0`)
  s.norm_diff = INFINITY;
  st[row] = s;
  lxa_gmres_row g = {};
  g.r_min = INFINITY;
  gs[row] = g;
}

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------

int vtype(int code, int* out) {
  if (code != LXA_F32 && code != LXA_F64) return set_error(LXA_EINVAL, "bad dtype %d", code);
  *out = code;
  return LXA_OK;
}

int nblocks_for(int64_t n, int64_t rows) {
  int64_t target = static_cast<int64_t>(num_cus()) * 8;
  int64_t per_row = (target + rows - 1) / (rows > 0 ? rows : 1);
  int64_t byn = (n + kB * 4 - 1) / (kB * 4);
  int64_t nb = per_row < byn ? per_row : byn;
  if (nb < 1) nb = 1;
  if (nb > 1024) nb = 1024;
  return static_cast<int>(nb);
}

dim3 grid2(int nblk, int64_t rows) { return dim3(static_cast<unsigned>(nblk), static_cast<unsigned>(rows)); }
dim3 grid_rows(int64_t rows) { return dim3(static_cast<unsigned>((rows + 63) / 64)); }

#define LXA_DISPATCH_YV(ydt, vdt, ...)                                              \
  do {                                                                              \
    if ((ydt) == LXA_F64 && (vdt) == LXA_F64) {                                     \
      using TY = double;                                                            \
      using TV = double;                                                            \
      __VA_ARGS__;                                                                  \
    } else if ((ydt) == LXA_F64 && (vdt) == LXA_F32) {                              \
      using TY = double;                                                            \
      using TV = float;                                                             \
      __VA_ARGS__;                                                                  \
    } else if ((ydt) == LXA_F32 && (vdt) == LXA_F32) {                              \
      using TY = float;                                                             \
      using TV = float;                                                             \
      __VA_ARGS__;                                                                  \
    } else {                                                                        \
      return ::lxa::set_error(LXA_EINVAL, "unsupported dtype pair (y %d, vectors %d)", ydt, vdt); \
    }                                                                               \
  } while (0)

}  // namespace
}  // namespace lxa

using lxa::CgParams;
using lxa::GmParams;

static CgParams cg_params(const lxa_krylov_params* p) {
  CgParams c;
  c.rtol = p->rtol;
  c.atol = p->atol;
  c.rcond = p->rcond;
  c.max_steps = p->max_steps;
  c.has_scale = !(p->rtol == 0.0 && p->atol == 0.0);
  c.norm_kind = p->norm_kind;
  c.max_steps_given = p->max_steps_given;
  return c;
}

static GmParams gm_params(const lxa_krylov_params* p) {
  GmParams g;
  g.rtol = p->rtol;
  g.atol = p->atol;
  g.eps = p->eps;
  g.max_steps = p->max_steps;
  g.has_scale = !(p->rtol == 0.0 && p->atol == 0.0);
  g.norm_kind = p->norm_kind;
  g.max_steps_given = p->max_steps_given;
  g.restart = p->restart;
  g.stagnation_iters = p->stagnation_iters;
  return g;
}

extern "C" int lxa_krylov_partials_size(int64_t rows, int64_t n, int restart, size_t* bytes) {
  LXA_REQUIRE(rows >= 1 && n >= 1 && restart >= 0 && bytes, "lxa_krylov_partials_size: bad argument");
  const int nb = lxa::nblocks_for(n, rows);
  const int64_t q = restart + 2 > 4 ? restart + 2 : 4;
  *bytes = static_cast<size_t>(rows) * static_cast<size_t>(q) * static_cast<size_t>(nb) * sizeof(double);
  return LXA_OK;
}

static int check_common(int64_t rows, int64_t n, const lxa_krylov_params* prm, const void* st, const void* part) {
  LXA_REQUIRE(rows >= 1 && rows <= 65535, "Krylov: rows must be in [1, 65535]");
  LXA_REQUIRE(n >= 1, "Krylov: n must be positive");
  LXA_REQUIRE(prm && st && part, "Krylov: null pointer");
  LXA_REQUIRE(prm->norm_kind >= 0 && prm->norm_kind <= 2, "Krylov: bad norm_kind");
  return LXA_OK;
}

template <typename TV>
static void launch_dot(hipStream_t s, int64_t rows, int64_t n, int nb, const void* x, const void* y,
                       const lxa_krylov_row* st, int gate, double* part, int64_t pstride, int64_t poff) {
  hipLaunchKernelGGL((lxa::dot_partial_kernel<TV, TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                     static_cast<const TV*>(x), static_cast<const TV*>(y), st, gate, part, pstride, poff);
}

// ----------------------------------------------------------------------------- CG

extern "C" int lxa_cg_begin(void* stream, int vdtype, int64_t rows, int64_t n, const void* z, const void* r,
                            const lxa_krylov_params* prm, lxa_krylov_row* state, double* partials) {
  if (int rc = check_common(rows, n, prm, state, partials)) return rc;
  LXA_REQUIRE(z && r, "lxa_cg_begin: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  if (vdtype == LXA_F64) launch_dot<double>(s, rows, n, nb, z, r, state, 0, partials, nb, 0);
  else if (vdtype == LXA_F32) launch_dot<float>(s, rows, n, nb, z, r, state, 0, partials, nb, 0);
  else return lxa::set_error(LXA_EINVAL, "lxa_cg_begin: bad dtype");
  hipLaunchKernelGGL(lxa::cg_begin_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, partials, cg_params(prm),
                     state);
  LXA_LAUNCH_CHECK("lxa_cg_begin");
  return LXA_OK;
}

extern "C" int lxa_cg_alpha(void* stream, int vdtype, int64_t rows, int64_t n, const void* q, const void* p,
                            const lxa_krylov_params* prm, lxa_krylov_row* state, double* partials) {
  if (int rc = check_common(rows, n, prm, state, partials)) return rc;
  LXA_REQUIRE(q && p, "lxa_cg_alpha: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  if (vdtype == LXA_F64) launch_dot<double>(s, rows, n, nb, q, p, state, 1, partials, nb, 0);
  else if (vdtype == LXA_F32) launch_dot<float>(s, rows, n, nb, q, p, state, 1, partials, nb, 0);
  else return lxa::set_error(LXA_EINVAL, "lxa_cg_alpha: bad dtype");
  hipLaunchKernelGGL(lxa::cg_alpha_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, partials, cg_params(prm),
                     state);
  LXA_LAUNCH_CHECK("lxa_cg_alpha");
  return LXA_OK;
}

extern "C" int lxa_cg_update(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* p,
                             const void* q, void* y, void* r, const void* b, int stable, int fuse_gamma,
                             const lxa_krylov_params* prm, lxa_krylov_row* state, double* partials) {
  if (int rc = check_common(rows, n, prm, state, partials)) return rc;
  LXA_REQUIRE(p && y && r && b && (stable || q), "lxa_cg_update: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  const CgParams c = cg_params(prm);
  LXA_DISPATCH_YV(ydtype, vdtype,
                  hipLaunchKernelGGL((lxa::cg_update_kernel<TY, TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                                     static_cast<const TV*>(p), static_cast<const TV*>(q), static_cast<TY*>(y),
                                     static_cast<TV*>(r), static_cast<const TY*>(b), c, stable, fuse_gamma, state,
                                     partials));
  LXA_LAUNCH_CHECK("lxa_cg_update");
  return LXA_OK;
}

extern "C" int lxa_cg_residual(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* b,
                               const void* t, void* r, int fuse_gamma, const lxa_krylov_params* prm,
                               lxa_krylov_row* state, double* partials) {
  if (int rc = check_common(rows, n, prm, state, partials)) return rc;
  LXA_REQUIRE(b && t && r, "lxa_cg_residual: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  const CgParams c = cg_params(prm);
  LXA_DISPATCH_YV(ydtype, vdtype,
                  hipLaunchKernelGGL((lxa::cg_residual_kernel<TY, TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                                     static_cast<const TY*>(b), static_cast<const TY*>(t), static_cast<TV*>(r), c,
                                     fuse_gamma, state, partials));
  LXA_LAUNCH_CHECK("lxa_cg_residual");
  return LXA_OK;
}

extern "C" int lxa_cg_precond_gamma(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* b,
                                    const void* z, const void* r, const lxa_krylov_params* prm,
                                    lxa_krylov_row* state, double* partials) {
  if (int rc = check_common(rows, n, prm, state, partials)) return rc;
  LXA_REQUIRE(b && z && r, "lxa_cg_precond_gamma: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  const CgParams c = cg_params(prm);
  // <z, r> into slot 1 and |r / b_scale| into slot 2 of the [row][3][nb] layout.
  LXA_DISPATCH_YV(ydtype, vdtype, {
    launch_dot<TV>(s, rows, n, nb, z, r, state, 1, partials, 3 * nb, nb);
    hipLaunchKernelGGL((lxa::cg_rnorm_kernel<TY, TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const TY*>(b), static_cast<const TV*>(r), c, state, partials);
  });
  LXA_LAUNCH_CHECK("lxa_cg_precond_gamma");
  return LXA_OK;
}

extern "C" int lxa_cg_finish(void* stream, int vdtype, int64_t rows, int64_t n, const void* z, void* p,
                             const lxa_krylov_params* prm, lxa_krylov_row* state, double* partials) {
  if (int rc = check_common(rows, n, prm, state, partials)) return rc;
  LXA_REQUIRE(z && p, "lxa_cg_finish: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  const CgParams c = cg_params(prm);
  hipLaunchKernelGGL(lxa::cg_finish_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, n, partials, c, state);
  if (vdtype == LXA_F64)
    hipLaunchKernelGGL((lxa::cg_direction_kernel<double>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const double*>(z), static_cast<double*>(p), state);
  else if (vdtype == LXA_F32)
    hipLaunchKernelGGL((lxa::cg_direction_kernel<float>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const float*>(z), static_cast<float*>(p), state);
  else
    return lxa::set_error(LXA_EINVAL, "lxa_cg_finish: bad dtype");
  LXA_LAUNCH_CHECK("lxa_cg_finish");
  return LXA_OK;
}

extern "C" int lxa_cg_result(void* stream, int64_t rows, const lxa_krylov_params* prm, lxa_krylov_row* state) {
  LXA_REQUIRE(rows >= 1 && prm && state, "lxa_cg_result: bad argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(lxa::cg_result_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, cg_params(prm), state);
  LXA_LAUNCH_CHECK("lxa_cg_result");
  return LXA_OK;
}

// ----------------------------------------------------------------------------- BiCGStab

extern "C" int lxa_bicgstab_begin(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* b,
                                  const void* t, void* r, void* r0, const lxa_krylov_params* prm,
                                  lxa_krylov_row* state, double* partials) {
  if (int rc = check_common(rows, n, prm, state, partials)) return rc;
  LXA_REQUIRE(b && t && r && r0, "lxa_bicgstab_begin: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  LXA_DISPATCH_YV(ydtype, vdtype,
                  hipLaunchKernelGGL((lxa::bcg_begin_kernel<TY, TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                                     static_cast<const TY*>(b), static_cast<const TY*>(t), static_cast<TV*>(r),
                                     static_cast<TV*>(r0), partials));
  hipLaunchKernelGGL(lxa::bcg_begin_finish_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, partials,
                     cg_params(prm), state);
  LXA_LAUNCH_CHECK("lxa_bicgstab_begin");
  return LXA_OK;
}

extern "C" int lxa_bicgstab_direction(void* stream, int vdtype, int64_t rows, int64_t n, const void* r, void* p,
                                      const void* v, int first, lxa_krylov_row* state) {
  LXA_REQUIRE(rows >= 1 && rows <= 65535 && n >= 1 && state && r && p && v, "lxa_bicgstab_direction: bad argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  if (vdtype == LXA_F64)
    hipLaunchKernelGGL((lxa::bcg_direction_kernel<double>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const double*>(r), static_cast<double*>(p), static_cast<const double*>(v), first,
                       state);
  else if (vdtype == LXA_F32)
    hipLaunchKernelGGL((lxa::bcg_direction_kernel<float>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const float*>(r), static_cast<float*>(p), static_cast<const float*>(v), first,
                       state);
  else
    return lxa::set_error(LXA_EINVAL, "lxa_bicgstab_direction: bad dtype");
  LXA_LAUNCH_CHECK("lxa_bicgstab_direction");
  return LXA_OK;
}

extern "C" int lxa_bicgstab_alpha(void* stream, int vdtype, int64_t rows, int64_t n, const void* r0, const void* v,
                                  const lxa_krylov_params* prm, lxa_krylov_row* state, double* partials) {
  if (int rc = check_common(rows, n, prm, state, partials)) return rc;
  LXA_REQUIRE(r0 && v, "lxa_bicgstab_alpha: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  if (vdtype == LXA_F64) launch_dot<double>(s, rows, n, nb, r0, v, state, 1, partials, nb, 0);
  else if (vdtype == LXA_F32) launch_dot<float>(s, rows, n, nb, r0, v, state, 1, partials, nb, 0);
  else return lxa::set_error(LXA_EINVAL, "lxa_bicgstab_alpha: bad dtype");
  hipLaunchKernelGGL(lxa::bcg_alpha_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, partials, state);
  LXA_LAUNCH_CHECK("lxa_bicgstab_alpha");
  return LXA_OK;
}

extern "C" int lxa_bicgstab_s(void* stream, int vdtype, int64_t rows, int64_t n, const void* r, const void* v,
                              void* s_out, lxa_krylov_row* state) {
  LXA_REQUIRE(rows >= 1 && rows <= 65535 && n >= 1 && state && r && v && s_out, "lxa_bicgstab_s: bad argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  if (vdtype == LXA_F64)
    hipLaunchKernelGGL((lxa::bcg_s_kernel<double>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const double*>(r), static_cast<const double*>(v), static_cast<double*>(s_out),
                       state);
  else if (vdtype == LXA_F32)
    hipLaunchKernelGGL((lxa::bcg_s_kernel<float>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const float*>(r), static_cast<const float*>(v), static_cast<float*>(s_out), state);
  else
    return lxa::set_error(LXA_EINVAL, "lxa_bicgstab_s: bad dtype");
  LXA_LAUNCH_CHECK("lxa_bicgstab_s");
  return LXA_OK;
}

extern "C" int lxa_bicgstab_omega(void* stream, int vdtype, int64_t rows, int64_t n, const void* s_in, const void* t,
                                  const lxa_krylov_params* prm, lxa_krylov_row* state, double* partials) {
  if (int rc = check_common(rows, n, prm, state, partials)) return rc;
  LXA_REQUIRE(s_in && t, "lxa_bicgstab_omega: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  if (vdtype == LXA_F64)
    hipLaunchKernelGGL((lxa::bcg_st_kernel<double>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const double*>(s_in), static_cast<const double*>(t), state, partials);
  else if (vdtype == LXA_F32)
    hipLaunchKernelGGL((lxa::bcg_st_kernel<float>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const float*>(s_in), static_cast<const float*>(t), state, partials);
  else
    return lxa::set_error(LXA_EINVAL, "lxa_bicgstab_omega: bad dtype");
  hipLaunchKernelGGL(lxa::bcg_omega_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, partials, state);
  LXA_LAUNCH_CHECK("lxa_bicgstab_omega");
  return LXA_OK;
}

extern "C" int lxa_bicgstab_update(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* x,
                                   const void* z, const void* s_in, const void* t, const void* r0, void* y, void* r,
                                   const void* b, const lxa_krylov_params* prm, lxa_krylov_row* state,
                                   double* partials) {
  if (int rc = check_common(rows, n, prm, state, partials)) return rc;
  LXA_REQUIRE(x && z && s_in && t && r0 && y && r && b, "lxa_bicgstab_update: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  const CgParams c = cg_params(prm);
  LXA_DISPATCH_YV(ydtype, vdtype,
                  hipLaunchKernelGGL((lxa::bcg_update_kernel<TY, TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                                     static_cast<const TV*>(x), static_cast<const TV*>(z),
                                     static_cast<const TV*>(s_in), static_cast<const TV*>(t),
                                     static_cast<const TV*>(r0), static_cast<TY*>(y), static_cast<TV*>(r),
                                     static_cast<const TY*>(b), c, state, partials));
  hipLaunchKernelGGL(lxa::bcg_update_finish_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, n, partials, c,
                     prm->eps, state);
  LXA_LAUNCH_CHECK("lxa_bicgstab_update");
  return LXA_OK;
}

extern "C" int lxa_bicgstab_result(void* stream, int64_t rows, const lxa_krylov_params* prm,
                                   lxa_krylov_row* state) {
  LXA_REQUIRE(rows >= 1 && prm && state, "lxa_bicgstab_result: bad argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(lxa::bcg_result_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, cg_params(prm), prm->eps,
                     state);
  LXA_LAUNCH_CHECK("lxa_bicgstab_result");
  return LXA_OK;
}

// ----------------------------------------------------------------------------- LSMR

static_assert(sizeof(lxa_lsmr_row) == 232, "lxa_lsmr_row layout is part of the C ABI");
static_assert(sizeof(lxa_lsmr_params) == 40, "lxa_lsmr_params layout is part of the C ABI");

static lxa::LsParams ls_params(const lxa_lsmr_params* p) {
  lxa::LsParams l;
  l.rtol = p->rtol;
  l.atol = p->atol;
  l.conlim = p->conlim;
  l.max_steps = p->max_steps;
  l.has_scale = !(p->rtol == 0.0 && p->atol == 0.0);
  l.max_steps_given = p->max_steps_given;
  return l;
}

static int check_lsmr(int64_t rows, int64_t len, const void* prm, const void* st, const void* part) {
  LXA_REQUIRE(rows >= 1 && rows <= 65535, "LSMR: rows must be in [1, 65535]");
  LXA_REQUIRE(len >= 1, "LSMR: vector length must be positive");
  LXA_REQUIRE(prm && st && part, "LSMR: null pointer");
  return LXA_OK;
}

#define LXA_DISPATCH_V(vdt, ...)                                                     \
  do {                                                                               \
    if ((vdt) == LXA_F64) {                                                          \
      using TV = double;                                                             \
      __VA_ARGS__;                                                                   \
    } else if ((vdt) == LXA_F32) {                                                   \
      using TV = float;                                                              \
      __VA_ARGS__;                                                                   \
    } else {                                                                         \
      return ::lxa::set_error(LXA_EINVAL, "unsupported vector dtype %d", vdt);       \
    }                                                                                \
  } while (0)

extern "C" int lxa_lsmr_begin(void* stream, int ydtype, int vdtype, int64_t rows, int64_t m, const void* b,
                              const void* t, void* u, const lxa_lsmr_params* prm, lxa_lsmr_row* state,
                              double* partials) {
  if (int rc = check_lsmr(rows, m, prm, state, partials)) return rc;
  LXA_REQUIRE(b && t && u, "lxa_lsmr_begin: null vector");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(m, rows);
  LXA_DISPATCH_YV(ydtype, vdtype,
                  hipLaunchKernelGGL((lxa::ls_begin_kernel<TY, TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, m,
                                     static_cast<const TY*>(b), static_cast<const TY*>(t), static_cast<TV*>(u),
                                     partials));
  hipLaunchKernelGGL(lxa::ls_begin_finish_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, partials, state);
  LXA_LAUNCH_CHECK("lxa_lsmr_begin");
  return LXA_OK;
}

extern "C" int lxa_lsmr_init(void* stream, int vdtype, int64_t rows, int64_t n, const void* z, void* v, void* h,
                             void* hbar, const lxa_lsmr_params* prm, lxa_lsmr_row* state, double* partials) {
  if (int rc = check_lsmr(rows, n, prm, state, partials)) return rc;
  LXA_REQUIRE(z && v && h && hbar, "lxa_lsmr_init: null vector");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  LXA_DISPATCH_V(vdtype, {
    hipLaunchKernelGGL((lxa::ls_v0_kernel<TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const TV*>(z), static_cast<TV*>(v), state, partials);
    hipLaunchKernelGGL(lxa::ls_init_finish_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, partials, state);
    hipLaunchKernelGGL((lxa::ls_h0_kernel<TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const TV*>(v), static_cast<TV*>(h), static_cast<TV*>(hbar), state);
  });
  LXA_LAUNCH_CHECK("lxa_lsmr_init");
  return LXA_OK;
}

extern "C" int lxa_lsmr_step_u(void* stream, int vdtype, int64_t rows, int64_t m, const void* w, void* u,
                               const lxa_lsmr_params* prm, lxa_lsmr_row* state, double* partials) {
  if (int rc = check_lsmr(rows, m, prm, state, partials)) return rc;
  LXA_REQUIRE(w && u, "lxa_lsmr_step_u: null vector");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(m, rows);
  LXA_DISPATCH_V(vdtype, hipLaunchKernelGGL((lxa::ls_u_kernel<TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, m,
                                            static_cast<const TV*>(w), static_cast<TV*>(u), state, partials));
  hipLaunchKernelGGL(lxa::ls_u_finish_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, partials, state);
  LXA_LAUNCH_CHECK("lxa_lsmr_step_u");
  return LXA_OK;
}

extern "C" int lxa_lsmr_step_v(void* stream, int vdtype, int64_t rows, int64_t n, const void* z, void* v,
                               const lxa_lsmr_params* prm, lxa_lsmr_row* state, double* partials) {
  if (int rc = check_lsmr(rows, n, prm, state, partials)) return rc;
  LXA_REQUIRE(z && v, "lxa_lsmr_step_v: null vector");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  LXA_DISPATCH_V(vdtype, hipLaunchKernelGGL((lxa::ls_v_kernel<TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                                            static_cast<const TV*>(z), static_cast<TV*>(v), state, partials));
  hipLaunchKernelGGL(lxa::ls_v_finish_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, partials, state);
  LXA_LAUNCH_CHECK("lxa_lsmr_step_v");
  return LXA_OK;
}

extern "C" int lxa_lsmr_step_x(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* v, void* h,
                               void* hbar, void* x, const lxa_lsmr_params* prm, lxa_lsmr_row* state,
                               double* partials) {
  if (int rc = check_lsmr(rows, n, prm, state, partials)) return rc;
  LXA_REQUIRE(v && h && hbar && x, "lxa_lsmr_step_x: null vector");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  LXA_DISPATCH_YV(ydtype, vdtype,
                  hipLaunchKernelGGL((lxa::ls_x_kernel<TY, TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                                     static_cast<const TV*>(v), static_cast<TV*>(h), static_cast<TV*>(hbar),
                                     static_cast<TY*>(x), state, partials));
  hipLaunchKernelGGL(lxa::ls_x_finish_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, partials,
                     ls_params(prm), state);
  LXA_LAUNCH_CHECK("lxa_lsmr_step_x");
  return LXA_OK;
}

extern "C" int lxa_lsmr_result(void* stream, int64_t rows, const lxa_lsmr_params* prm, lxa_lsmr_row* state) {
  LXA_REQUIRE(rows >= 1 && prm && state, "lxa_lsmr_result: bad argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(lxa::ls_result_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, ls_params(prm), state);
  LXA_LAUNCH_CHECK("lxa_lsmr_result");
  return LXA_OK;
}

// ----------------------------------------------------------------------------- GMRES

static int check_gmres(int64_t rows, int64_t n, const lxa_krylov_params* prm, const void* st, const void* gs,
                       const void* part) {
  if (int rc = check_common(rows, n, prm, st, part)) return rc;
  LXA_REQUIRE(gs, "GMRES: null row state");
  LXA_REQUIRE(prm->restart >= 1 && prm->restart <= lxa::kMaxRestart, "GMRES: restart must be in [1, %d]",
              lxa::kMaxRestart);
  return LXA_OK;
}

extern "C" int lxa_gmres_init(void* stream, int64_t rows, lxa_krylov_row* state, lxa_gmres_row* gstate) {
  LXA_REQUIRE(rows >= 1 && state && gstate, "lxa_gmres_init: bad argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(lxa::gm_init_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, state, gstate);
  LXA_LAUNCH_CHECK("lxa_gmres_init");
  return LXA_OK;
}

extern "C" int lxa_gmres_cycle_begin(void* stream, int vdtype, int64_t rows, int64_t n, const void* r, void* basis,
                                     const lxa_krylov_params* prm, lxa_krylov_row* state, lxa_gmres_row* gstate,
                                     double* hessenberg, double* partials) {
  if (int rc = check_gmres(rows, n, prm, state, gstate, partials)) return rc;
  LXA_REQUIRE(r && basis && hessenberg, "lxa_gmres_cycle_begin: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  const int m = prm->restart;
  const int64_t vstride = static_cast<int64_t>(m + 1) * n;
  if (vdtype == LXA_F64) {
    hipLaunchKernelGGL((lxa::sumsq_partial_kernel<double>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const double*>(r), n, state, partials);
  } else if (vdtype == LXA_F32) {
    hipLaunchKernelGGL((lxa::sumsq_partial_kernel<float>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const float*>(r), n, state, partials);
  } else {
    return lxa::set_error(LXA_EINVAL, "lxa_gmres_cycle_begin: bad dtype");
  }
  hipLaunchKernelGGL(lxa::gm_cycle_begin_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, n, partials,
                     gm_params(prm), hessenberg, state, gstate);
  if (vdtype == LXA_F64)
    hipLaunchKernelGGL((lxa::gm_scale_kernel<double, double>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const double*>(r), n, static_cast<double*>(basis), vstride, state, gstate, 0);
  else
    hipLaunchKernelGGL((lxa::gm_scale_kernel<float, float>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const float*>(r), n, static_cast<float*>(basis), vstride, state, gstate, 0);
  LXA_LAUNCH_CHECK("lxa_gmres_cycle_begin");
  return LXA_OK;
}

extern "C" int lxa_gmres_arnoldi(void* stream, int vdtype, int64_t rows, int64_t n, int j, void* basis, const void* w,
                                 const lxa_krylov_params* prm, lxa_krylov_row* state, lxa_gmres_row* gstate,
                                 double* hessenberg, double* proj, double* partials) {
  if (int rc = check_gmres(rows, n, prm, state, gstate, partials)) return rc;
  LXA_REQUIRE(basis && w && hessenberg && proj, "lxa_gmres_arnoldi: null pointer");
  const int m = prm->restart;
  LXA_REQUIRE(j >= 0 && j < m, "lxa_gmres_arnoldi: step %d outside [0, %d)", j, m);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  const GmParams g = gm_params(prm);
  const int64_t vstride = static_cast<int64_t>(m + 1) * n;
  if (vdtype == LXA_F64) {
    using TV = double;
    hipLaunchKernelGGL((lxa::gm_project_kernel<TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const TV*>(basis), static_cast<const TV*>(w), j, m, state, gstate, partials);
    hipLaunchKernelGGL(lxa::gm_project_finish_kernel, dim3(static_cast<unsigned>(rows)), dim3(lxa::kB), 0, s, rows,
                       nb, j, m, partials, proj, state, gstate);
    hipLaunchKernelGGL((lxa::gm_orthogonalise_kernel<TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<TV*>(basis), static_cast<const TV*>(w), proj, j, m, state, gstate, partials);
    hipLaunchKernelGGL(lxa::gm_arnoldi_finish_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, j, m, partials,
                       proj, g, hessenberg, state, gstate);
    TV* col = static_cast<TV*>(basis) + static_cast<int64_t>(j + 1) * n;
    hipLaunchKernelGGL((lxa::gm_scale_kernel<TV, TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n, col, vstride,
                       col, vstride, state, gstate, 0);
  } else if (vdtype == LXA_F32) {
    using TV = float;
    hipLaunchKernelGGL((lxa::gm_project_kernel<TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<const TV*>(basis), static_cast<const TV*>(w), j, m, state, gstate, partials);
    hipLaunchKernelGGL(lxa::gm_project_finish_kernel, dim3(static_cast<unsigned>(rows)), dim3(lxa::kB), 0, s, rows,
                       nb, j, m, partials, proj, state, gstate);
    hipLaunchKernelGGL((lxa::gm_orthogonalise_kernel<TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                       static_cast<TV*>(basis), static_cast<const TV*>(w), proj, j, m, state, gstate, partials);
    hipLaunchKernelGGL(lxa::gm_arnoldi_finish_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, j, m, partials,
                       proj, g, hessenberg, state, gstate);
    TV* col = static_cast<TV*>(basis) + static_cast<int64_t>(j + 1) * n;
    hipLaunchKernelGGL((lxa::gm_scale_kernel<TV, TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n, col, vstride,
                       col, vstride, state, gstate, 0);
  } else {
    return lxa::set_error(LXA_EINVAL, "lxa_gmres_arnoldi: bad dtype");
  }
  LXA_LAUNCH_CHECK("lxa_gmres_arnoldi");
  return LXA_OK;
}

static int gmres_update_impl(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* basis,
                             void* y, const lxa_krylov_params* prm, lxa_krylov_row* state, lxa_gmres_row* gstate,
                             const double* hessenberg, double* z, const double* colscale, double* partials);

extern "C" int lxa_gmres_update(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* basis,
                                void* y, const lxa_krylov_params* prm, lxa_krylov_row* state, lxa_gmres_row* gstate,
                                const double* hessenberg, double* z, double* partials) {
  return gmres_update_impl(stream, ydtype, vdtype, rows, n, basis, y, prm, state, gstate, hessenberg, z, nullptr,
                           partials);
}

extern "C" int lxa_gmres_update_scaled(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n,
                                       const void* basis, void* y, const lxa_krylov_params* prm,
                                       lxa_krylov_row* state, lxa_gmres_row* gstate, const double* hessenberg,
                                       double* z, const double* colscale, double* partials) {
  LXA_REQUIRE(colscale, "lxa_gmres_update_scaled: null colscale");
  return gmres_update_impl(stream, ydtype, vdtype, rows, n, basis, y, prm, state, gstate, hessenberg, z, colscale,
                           partials);
}

static int gmres_update_impl(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* basis,
                             void* y, const lxa_krylov_params* prm, lxa_krylov_row* state, lxa_gmres_row* gstate,
                             const double* hessenberg, double* z, const double* colscale, double* partials) {
  if (int rc = check_gmres(rows, n, prm, state, gstate, partials)) return rc;
  LXA_REQUIRE(basis && y && hessenberg && z, "lxa_gmres_update: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  const int m = prm->restart;
  const size_t lds = (static_cast<size_t>(m + 1) * m + (m + 1) + m) * sizeof(double);
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(lxa::gm_lstsq_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return lxa::hip_error(e, "hipFuncSetAttribute(gm_lstsq_kernel)");
  }
  hipLaunchKernelGGL(lxa::gm_lstsq_kernel, dim3(static_cast<unsigned>(rows)), dim3(64), lds, s, rows, m, hessenberg,
                     z, state, gstate);
  const GmParams g = gm_params(prm);
  LXA_DISPATCH_YV(ydtype, vdtype,
                  hipLaunchKernelGGL((lxa::gm_update_kernel<TY, TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0, s, n,
                                     static_cast<const TV*>(basis), z, m, static_cast<TY*>(y), g, state, gstate,
                                     colscale, partials));
  LXA_LAUNCH_CHECK("lxa_gmres_update");
  return LXA_OK;
}

extern "C" int lxa_gmres_residual(void* stream, int ydtype, int vdtype, int64_t rows, int64_t n, const void* b,
                                  const void* t, void* r, int write, const lxa_krylov_params* prm,
                                  lxa_krylov_row* state, double* partials) {
  if (int rc = check_common(rows, n, prm, state, partials)) return rc;
  LXA_REQUIRE(b && r && (t || !write), "lxa_gmres_residual: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  const GmParams g = gm_params(prm);
  LXA_DISPATCH_YV(ydtype, vdtype,
                  hipLaunchKernelGGL((lxa::gm_residual_kernel<TY, TY, TV>), lxa::grid2(nb, rows), dim3(lxa::kB), 0,
                                     s, n, static_cast<const TY*>(b), static_cast<const TY*>(t),
                                     static_cast<TV*>(r), g, write, state, partials));
  LXA_LAUNCH_CHECK("lxa_gmres_residual");
  return LXA_OK;
}

extern "C" int lxa_gmres_cycle_end(void* stream, int64_t rows, int64_t n, int first_pass,
                                   const lxa_krylov_params* prm, lxa_krylov_row* state, lxa_gmres_row* gstate,
                                   double* partials) {
  if (int rc = check_gmres(rows, n, prm, state, gstate, partials)) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = lxa::nblocks_for(n, rows);
  hipLaunchKernelGGL(lxa::gm_cycle_end_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, nb, n, partials,
                     gm_params(prm), first_pass, state, gstate);
  LXA_LAUNCH_CHECK("lxa_gmres_cycle_end");
  return LXA_OK;
}

extern "C" int lxa_gmres_result(void* stream, int64_t rows, const lxa_krylov_params* prm, lxa_krylov_row* state,
                                const lxa_gmres_row* gstate) {
  LXA_REQUIRE(rows >= 1 && prm && state && gstate, "lxa_gmres_result: bad argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(lxa::gm_result_kernel, lxa::grid_rows(rows), dim3(64), 0, s, rows, gm_params(prm), state,
                     gstate);
  LXA_LAUNCH_CHECK("lxa_gmres_result");
  return LXA_OK;
}
