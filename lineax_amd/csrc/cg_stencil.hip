// Fused CG iterations for constant-coefficient five-point stencil operators.
//
// lineax's CG body (lineax/_solver/cg.py:169-206) with the identity
// preconditioner is, per iteration,
//   q = A p;  alpha = gamma / <q, p>;  y += alpha p;
//   r = (stable step) ? b - A y : r - alpha q;
//   gamma' = <r, r>;  beta = gamma' / gamma;  p = r + beta p;
// plus the stopping test on |r / b_scale| and |alpha p / y_scale|.  Run as
// separate kernels it moves 17 vector passes per point (y and b in FP64 make
// it more).  Here one iteration is two sweeps over the grid:
//   direction sweep: p_new = r + beta p_old (the last step's direction
//     update, deferred), store p_new, and <A p_new, p_new> with A p_new
//     formed on the fly from a rolling three-row window;
//   update sweep: recompute q = A p_new from p_new, y += alpha p_new,
//     r -= alpha q (or r = b - A y_new on stabilising steps, with y_new
//     formed on the fly the same way), gamma' and both stopping norms.
// Each sweep reads its vectors once from HBM (grid rows above and below a
// tile, and one column each side, are re-read through the caches).
//
// Scalars: after each sweep a small kernel (one block per right-hand side)
// reduces the row's block partials in a fixed order and applies the lineax
// scalar step (alpha with the rcond guard; gamma, beta, step, norms, the
// loop condition).  No host round trip, and the results do not depend on
// block timing.  (A last-block-done ticket would save the launch but needs
// a device-scope fence in every block, which costs more than the launch.)
#include "lxa_common.h"

namespace lxa {
namespace {

constexpr int kTW = 256;  // grid columns per tile, one per thread
constexpr int kTH = 64;   // grid rows per tile
constexpr int kQ = 8;     // rows of loads in flight per thread
// The row loop is unrolled by kQ so each queue slot is a fixed register:
// slot u is consumed and refilled in place, never moved (moving a register
// that a load is still filling would wait for that load).  Rows alternate
// between two LDS row buffers, one barrier per row.

struct St5 {
  double c, w, e, s, n;
};

struct CgP {
  double rtol, atol, rcond;
  int64_t max_steps;
  int has_scale, norm_kind;
};

__device__ __forceinline__ double nanmax2(double a, double b) {
  if (a != a) return a;
  if (b != b) return b;
  return a > b ? a : b;
}

__device__ __forceinline__ double norm_elem(int kind, double v) { return kind == 0 ? fabs(v) : v * v; }
__device__ __forceinline__ double norm_acc(int kind, double a, double b) { return kind == 0 ? nanmax2(a, b) : a + b; }
__device__ __forceinline__ double norm_final(int kind, double v, int64_t n) {
  if (kind == 0) return v;
  if (kind == 1) return sqrt(v);
  return n > 0 ? sqrt(v / static_cast<double>(n)) : 0.0;
}

// Reduces three values over a block of nt threads (a multiple of 64, at
// most 1024; sum, sum, or norm).  Every thread gets the results; red holds
// 3 * (nt / 64) doubles.
__device__ void block_reduce3_n(int kind, double& a, double& b, double& c, double* red, int nt) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a = norm_acc(kind, a, __shfl_xor(a, off, 64));
    b += __shfl_xor(b, off, 64);
    c = norm_acc(kind, c, __shfl_xor(c, off, 64));
  }
  const int w = threadIdx.x >> 6, nw = nt >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = a;
    red[16 + w] = b;
    red[32 + w] = c;
  }
  __syncthreads();
  a = red[0];
  b = red[16];
  c = red[32];
  for (int q = 1; q < nw; ++q) {
    a = norm_acc(kind, a, red[q]);
    b += red[16 + q];
    c = norm_acc(kind, c, red[32 + q]);
  }
  __syncthreads();
}

// Reduces three values over the 256-thread block (sum, sum, or norm).  Every
// thread gets the results.
__device__ void block_reduce3(int kind, double& a, double& b, double& c, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a = norm_acc(kind, a, __shfl_xor(a, off, 64));
    b += __shfl_xor(b, off, 64);
    c = norm_acc(kind, c, __shfl_xor(c, off, 64));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = a;
    red[4 + w] = b;
    red[8 + w] = c;
  }
  __syncthreads();
  a = norm_acc(kind, norm_acc(kind, red[0], red[1]), norm_acc(kind, red[2], red[3]));
  b = (red[4] + red[5]) + (red[6] + red[7]);
  c = norm_acc(kind, norm_acc(kind, red[8], red[9]), norm_acc(kind, red[10], red[11]));
  __syncthreads();
}

struct TileGeom {
  int i0, i1, j0, j;
};

__device__ __forceinline__ TileGeom tile_geom(int nx, int ny, int tiles_j) {
  const int tile = blockIdx.x;
  const int ti = tile / tiles_j;
  const int tj = tile - ti * tiles_j;
  TileGeom g;
  g.i0 = ti * kTH;
  g.i1 = g.i0 + kTH < nx ? g.i0 + kTH : nx;
  g.j0 = tj * kTW;
  g.j = g.j0 + static_cast<int>(threadIdx.x);
  return g;
}

// Column of this thread's halo load: thread 0 the column left of the tile,
// thread kTW-1 the column right of it, others none (-1).
__device__ __forceinline__ int halo_col(const TileGeom& g) {
  if (threadIdx.x == 0) return g.j0 - 1;
  if (threadIdx.x == kTW - 1) return g.j0 + kTW;
  return -1;
}

// base[i][j], zero outside the grid and past row ilim (rows a tile never
// uses are not fetched).
template <typename T>
__device__ __forceinline__ T ld(const T* base, int nx, int ny, int i, int j, int ilim = 1 << 30) {
  if (i < 0 || i >= nx || i > ilim || j < 0 || j >= ny) return T(0);
  return base[static_cast<int64_t>(i) * ny + j];
}

// Unconditional load at an in-grid index (the caller clamps the index and
// masks the value where it is used): a queue slot is then a plain load into
// a fixed register, with no branch whose join would wait for it.
template <typename T>
__device__ __forceinline__ T ldc(const T* base, int ny, int i, int j) {
  return base[static_cast<int64_t>(i) * ny + j];
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---------------------------------------------------------------------------
// Direction sweep: p_new = r + beta p_old (p_new = r on the first
// iteration), <A p_new, p_new>; the last block sets alpha (cg.py:169-178).
// ---------------------------------------------------------------------------
template <typename TV>
__global__ __launch_bounds__(kTW) void cgs_direction_kernel(int nx, int ny, int tiles_j, St5 k,
                                                            const TV* __restrict__ r, const TV* __restrict__ pold,
                                                            TV* __restrict__ pnew, int first, int64_t vstride,
                                                            const lxa_krylov_row* __restrict__ st,
                                                            double* __restrict__ part) {
  __shared__ double lrow[2][kTW + 2];
  __shared__ double red[12];
  const int row = blockIdx.y;
  if (!st[row].active) return;
  const TileGeom g = tile_geom(nx, ny, tiles_j);
  const int t = threadIdx.x;
  const int hj = halo_col(g);
  const double beta = first ? 0.0 : st[row].beta;
  const TV* R = r + row * vstride;
  const TV* P = pold + row * vstride;
  TV* PN = pnew + row * vstride;

  // p_new rounded to storage, as the update sweep will read it back
  auto pn = [&](double rv, double pv) -> double {
    const double v = first ? rv : fma(beta, pv, rv);
    return static_cast<double>(static_cast<TV>(v));
  };
  // Load queue: slot u holds rows i+1 of (r, p_old) for the own column and
  // the halo column.  Indices are clamped into the tile's rows (up to its
  // lower halo row) and the grid; out-of-range values are masked at use.
  // p_old is read even on the first iteration (a valid buffer; unused).
  const int imax = g.i1 < nx - 1 ? g.i1 : nx - 1;
  const bool jin = g.j < ny;
  const int jc = jin ? g.j : ny - 1;
  const bool hin = hj >= 0 && hj < ny;
  const int hc = hin ? hj : jc;
  TV qr[kQ], qp[kQ], hr[kQ], hp[kQ];
#pragma unroll
  for (int u = 0; u < kQ; ++u) {
    const int i = clampi(g.i0 + 1 + u, 0, imax);
    qr[u] = ldc(R, ny, i, jc);
    qp[u] = ldc(P, ny, i, jc);
    hr[u] = ldc(R, ny, i, hc);
    hp[u] = ldc(P, ny, i, hc);
  }
  double prev = pn(ld(R, nx, ny, g.i0 - 1, g.j), first ? 0.0 : ld(P, nx, ny, g.i0 - 1, g.j));
  double cur = pn(ld(R, nx, ny, g.i0, g.j), first ? 0.0 : ld(P, nx, ny, g.i0, g.j));
  double hcur = hj >= 0 ? pn(ld(R, nx, ny, g.i0, hj), first ? 0.0 : ld(P, nx, ny, g.i0, hj)) : 0.0;
  double acc = 0.0;
  for (int ib = g.i0; ib < g.i1; ib += kQ) {
#pragma unroll
    for (int u = 0; u < kQ; ++u) {
      // rows past the tile in the last group run masked (no branch around
      // the slot, so its register is not merged and never waited on)
      const int i = ib + u;
      const bool live = i < g.i1;
      {
        // slot u holds row i + 1; refill it with row i + 1 + kQ
        const bool nin = i + 1 <= imax;
        const double next = (nin && jin) ? pn(qr[u], qp[u]) : 0.0;
        const double hnext = (nin && hin) ? pn(hr[u], hp[u]) : 0.0;
        const int il = clampi(i + 1 + kQ, 0, imax);
        qr[u] = ldc(R, ny, il, jc);
        qp[u] = ldc(P, ny, il, jc);
        hr[u] = ldc(R, ny, il, hc);
        hp[u] = ldc(P, ny, il, hc);
        double* lr = lrow[u & 1];
        lr[t + 1] = cur;
        if (t == 0) lr[0] = hcur;
        if (t == kTW - 1) lr[kTW + 1] = hcur;
        lds_barrier();
        if (live && g.j < ny) {
          const double ap = k.c * cur + k.w * lr[t] + k.e * lr[t + 2] + k.s * prev + k.n * next;
          acc = fma(ap, cur, acc);
          PN[static_cast<int64_t>(i) * ny + g.j] = static_cast<TV>(cur);
        }
        prev = cur;
        cur = next;
        hcur = hnext;
      }
    }
  }
  double z0 = 0.0, z1 = 0.0;
  block_reduce3(1, z0, acc, z1, red);
  if (t == 0) part[static_cast<int64_t>(row) * 3 * gridDim.x + blockIdx.x] = acc;
}

// alpha = gamma / <A p, p>, NaN when |<A p, p>| <= 100 rcond |gamma|
// (cg.py:172-178).  One block per row.
__global__ __launch_bounds__(kTW) void cgs_alpha_kernel(int nblk, CgP prm, lxa_krylov_row* __restrict__ st,
                                                        const double* __restrict__ part) {
  __shared__ double red[12];
  const int row = blockIdx.x;
  if (!st[row].active) return;
  const double* pr = part + static_cast<int64_t>(row) * 3 * nblk;
  double ip = 0.0, z0 = 0.0, z1 = 0.0;
  for (int q = threadIdx.x; q < nblk; q += kTW) ip += pr[q];
  block_reduce3(1, z0, ip, z1, red);
  if (threadIdx.x == 0) {
    const double gamma = st[row].gamma;
    double alpha = gamma / ip;
    if (!(fabs(ip) > 100.0 * prm.rcond * fabs(gamma))) alpha = __builtin_nan("");
    st[row].alpha = alpha;
    st[row].ip = ip;
  }
}

// ---------------------------------------------------------------------------
// Update sweep: y += alpha p; r -= alpha A p, or r = b - A y on stabilising
// steps; gamma', beta, step, norms and the loop condition in the last block
// (cg.py:179-206 and cond_fun 162-167).
// ---------------------------------------------------------------------------
// Stabilising steps read y at the tile halo, so they write y_new to a second
// buffer (y_out != y_in); plain steps update in place (y_out == y_in).
template <bool kStable, typename TY, typename TV, typename TB>
__global__ __launch_bounds__(kTW) void cgs_update_kernel(int nx, int ny, int tiles_j, St5 k,
                                                         const TV* __restrict__ p, const TY* y_in, TY* y_out,
                                                         TV* __restrict__ r, const TB* __restrict__ b,
                                                         int64_t vstride, CgP prm,
                                                         const lxa_krylov_row* __restrict__ st,
                                                         double* __restrict__ part) {
  __shared__ double lrow[2][kTW + 2];
  __shared__ double red[12];
  const int row = blockIdx.y;
  if (!st[row].active) {
    // a finished row keeps its iterate across the buffer swap
    if (kStable) {
      const TileGeom g = tile_geom(nx, ny, tiles_j);
      if (g.j < ny) {
        const int64_t base = row * vstride;
        for (int i = g.i0; i < g.i1; ++i) {
          const int64_t o = base + static_cast<int64_t>(i) * ny + g.j;
          y_out[o] = y_in[o];
        }
      }
    }
    return;
  }
  const TileGeom g = tile_geom(nx, ny, tiles_j);
  const int t = threadIdx.x;
  const int hj = halo_col(g);
  const double alpha = st[row].alpha;
  const int kind = prm.norm_kind;
  const TV* P = p + row * vstride;
  const TY* Y = y_in + row * vstride;
  TY* YO = y_out + row * vstride;
  TV* Rv = r + row * vstride;
  const TB* B = b + row * vstride;

  // The stencil operand: p (plain step) or y_new = y + alpha p (stable step),
  // in the precision the unfused kernels apply A to (p as stored; y_new
  // rounded to the iterate's storage).
  auto operand = [&](double pv, double yv) -> double {
    if (kStable) return static_cast<double>(static_cast<TY>(yv + alpha * pv));
    return pv;
  };
  // Queues (see the direction sweep): window slot u holds row i+1 of p (and
  // y, stable) for the own and halo columns; own-point slot u holds row i of
  // y, r and b.  Clamped indices, values masked at use.
  const int imax = g.i1 < nx - 1 ? g.i1 : nx - 1;
  const int own = g.i1 - 1;
  const bool jin = g.j < ny;
  const int jc = jin ? g.j : ny - 1;
  const bool hin = hj >= 0 && hj < ny;
  const int hc = hin ? hj : jc;
  TV qp[kQ], hp[kQ], cr[kQ];
  TY qy[kQ], hy[kQ], cy[kQ];
  TB cb[kQ];
#pragma unroll
  for (int u = 0; u < kQ; ++u) {
    const int i = clampi(g.i0 + 1 + u, 0, imax);
    qp[u] = ldc(P, ny, i, jc);
    hp[u] = ldc(P, ny, i, hc);
    if (kStable) {
      qy[u] = ldc(Y, ny, i, jc);
      hy[u] = ldc(Y, ny, i, hc);
    } else {
      qy[u] = hy[u] = TY(0);
    }
    const int ic = clampi(g.i0 + u, 0, own);
    if (!kStable) {
      cy[u] = ldc(Y, ny, ic, jc);
      cr[u] = ldc(Rv, ny, ic, jc);
    } else {
      cy[u] = TY(0);
      cr[u] = TV(0);
    }
    cb[u] = ldc(B, ny, ic, jc);
  }
  const double p_prev = ld(P, nx, ny, g.i0 - 1, g.j);
  const double p_cur = ld(P, nx, ny, g.i0, g.j);
  double prev = operand(p_prev, kStable ? static_cast<double>(ld(Y, nx, ny, g.i0 - 1, g.j)) : 0.0);
  double ycur = kStable ? static_cast<double>(ld(Y, nx, ny, g.i0, g.j)) : 0.0;
  double pcur = p_cur;
  double cur = operand(p_cur, ycur);
  double hcur = hj >= 0 ? operand(ld(P, nx, ny, g.i0, hj), kStable ? static_cast<double>(ld(Y, nx, ny, g.i0, hj))
                                                                    : 0.0)
                        : 0.0;
  double n2 = 0.0, rr = 0.0, n1 = 0.0;
  for (int ib = g.i0; ib < g.i1; ib += kQ) {
#pragma unroll
    for (int u = 0; u < kQ; ++u) {
      const int i = ib + u;
      const bool live = i < g.i1;
      {
        // window slot u holds row i + 1, own-point slot u row i
        const bool nin = i + 1 <= imax;
        const double pnext = (nin && jin) ? static_cast<double>(qp[u]) : 0.0;
        const double ynext = (nin && jin) ? static_cast<double>(qy[u]) : 0.0;
        const double next = operand(pnext, ynext);
        const double hnext = (nin && hin) ? operand(hp[u], hy[u]) : 0.0;
        const double yold = kStable ? ycur : static_cast<double>(cy[u]);
        const double rold = cr[u];
        const double bv = cb[u];
        {
          const int il = clampi(i + 1 + kQ, 0, imax);
          qp[u] = ldc(P, ny, il, jc);
          hp[u] = ldc(P, ny, il, hc);
          if (kStable) {
            qy[u] = ldc(Y, ny, il, jc);
            hy[u] = ldc(Y, ny, il, hc);
          }
          const int ic = clampi(i + kQ, 0, own);
          if (!kStable) {
            cy[u] = ldc(Y, ny, ic, jc);
            cr[u] = ldc(Rv, ny, ic, jc);
          }
          cb[u] = ldc(B, ny, ic, jc);
        }
        double* lr = lrow[u & 1];
        lr[t + 1] = cur;
        if (t == 0) lr[0] = hcur;
        if (t == kTW - 1) lr[kTW + 1] = hcur;
        lds_barrier();
        if (live && g.j < ny) {
          const double av = k.c * cur + k.w * lr[t] + k.e * lr[t + 2] + k.s * prev + k.n * next;
          const int64_t o = static_cast<int64_t>(i) * ny + g.j;
          const double d = alpha * pcur;
          const double yn = yold + d;
          YO[o] = static_cast<TY>(yn);
          if (prm.has_scale) n2 = norm_acc(kind, n2, norm_elem(kind, d / (prm.atol + prm.rtol * fabs(yn))));
          const double rn = kStable ? bv - av : rold - alpha * av;
          const TV rs = static_cast<TV>(rn);
          Rv[o] = rs;
          const double rv = static_cast<double>(rs);
          rr = fma(rv, rv, rr);
          if (prm.has_scale) n1 = norm_acc(kind, n1, norm_elem(kind, rv / (prm.atol + prm.rtol * fabs(bv))));
        }
        prev = cur;
        cur = next;
        hcur = hnext;
        pcur = pnext;
        ycur = ynext;
      }
    }
  }
  block_reduce3(kind, n2, rr, n1, red);
  if (t == 0) {
    double* pr = part + static_cast<int64_t>(row) * 3 * gridDim.x;
    pr[blockIdx.x] = n2;
    pr[gridDim.x + blockIdx.x] = rr;
    pr[2 * gridDim.x + blockIdx.x] = n1;
  }
}

// gamma' = <r, r>, beta = gamma' / gamma, step += 1, the norms and the loop
// condition (cg.py:150-167, 202-206).  One block per row.
__global__ __launch_bounds__(kTW) void cgs_finish_kernel(int nblk, int64_t n, CgP prm,
                                                         lxa_krylov_row* __restrict__ st,
                                                         const double* __restrict__ part) {
  __shared__ double red[12];
  const int row = blockIdx.x;
  if (!st[row].active) return;
  const int kind = prm.norm_kind;
  const double* pr = part + static_cast<int64_t>(row) * 3 * nblk;
  double a2 = 0.0, ar = 0.0, a1 = 0.0;
  for (int q = threadIdx.x; q < nblk; q += kTW) {
    a2 = norm_acc(kind, a2, pr[q]);
    ar += pr[nblk + q];
    a1 = norm_acc(kind, a1, pr[2 * nblk + q]);
  }
  block_reduce3(kind, a2, ar, a1, red);
  if (threadIdx.x == 0) {
    lxa_krylov_row s = st[row];
    const double gamma_prev = s.gamma;
    s.gamma = ar;
    s.beta = ar / gamma_prev;
    s.step += 1;
    s.norm_r = norm_final(kind, a1, n);
    s.norm_diff = norm_final(kind, a2, n);
    bool not_converged = true;
    if (prm.has_scale) not_converged = (s.norm_r > 1.0) || (s.norm_diff > 1.0);
    s.active = (s.gamma > 0.0) && (s.step < prm.max_steps) && not_converged;
    st[row] = s;
  }
}

// ---------------------------------------------------------------------------
// Full-row variants for grid widths that are a multiple of 4 (at most 2048):
// thread t owns the four adjacent columns 4t .. 4t + 3 (16-byte loads of
// float vectors), a workgroup covers kTH4 whole grid rows, so there is no
// column halo; the columns either side of a thread's group come from its
// neighbours through LDS.  Same arithmetic, order and results as the tiled
// kernels above.
// ---------------------------------------------------------------------------
constexpr int kTH4 = 64;     // grid rows per tile, update sweep
constexpr int kTH4Dir = 16;  // grid rows per tile, direction sweep
constexpr int kQ4 = 4;       // rows of loads in flight, update sweep
constexpr int kQ4Dir = 8;    // rows of loads in flight, direction sweep
constexpr int kMaxThreads4 = 512;  // grid widths up to 2048

template <typename T>
struct Quad;
template <>
struct Quad<float> {
  using type = float4;
};
template <>
struct Quad<double> {
  using type = double4;
};

template <typename T>
__device__ __forceinline__ typename Quad<T>::type ld4(const T* base, int ny, int i, int c) {
  return *reinterpret_cast<const typename Quad<T>::type*>(base + static_cast<int64_t>(i) * ny + c);
}

template <typename Q>
__device__ __forceinline__ double q_at(const Q& q, int e) {
  return e == 0 ? static_cast<double>(q.x) : e == 1 ? static_cast<double>(q.y) : e == 2 ? static_cast<double>(q.z)
                                                                                     : static_cast<double>(q.w);
}

// The four stencil values of a thread's columns from the window rows and the
// neighbouring columns (west of column 0, east of column 3).
__device__ __forceinline__ void stencil4(const St5& k, const double (&pr)[4], const double (&cu)[4],
                                         const double (&nx4)[4], double west, double east, double (&out)[4]) {
  out[0] = k.c * cu[0] + k.w * west + k.e * cu[1] + k.s * pr[0] + k.n * nx4[0];
  out[1] = k.c * cu[1] + k.w * cu[0] + k.e * cu[2] + k.s * pr[1] + k.n * nx4[1];
  out[2] = k.c * cu[2] + k.w * cu[1] + k.e * cu[3] + k.s * pr[2] + k.n * nx4[2];
  out[3] = k.c * cu[3] + k.w * cu[2] + k.e * east + k.s * pr[3] + k.n * nx4[3];
}

template <typename TV>
__global__ __launch_bounds__(kMaxThreads4) void cgs_direction4_kernel(int nx, int ny, St5 k,
                                                                      const TV* __restrict__ r,
                                                                      const TV* __restrict__ pold,
                                                                      TV* __restrict__ pnew, int first,
                                                                      int64_t vstride,
                                                                      const lxa_krylov_row* __restrict__ st,
                                                                      double* __restrict__ part) {
  using Q = typename Quad<TV>::type;
  __shared__ double eL[2][kMaxThreads4 + 1], eR[2][kMaxThreads4 + 1];
  __shared__ double red[48];
  const int row = blockIdx.y;
  if (!st[row].active) return;
  const int t = threadIdx.x, nt = blockDim.x;
  const int i0 = blockIdx.x * kTH4Dir, i1 = i0 + kTH4Dir < nx ? i0 + kTH4Dir : nx;
  const int imax = i1 < nx - 1 ? i1 : nx - 1;
  const int c0 = 4 * t;
  const bool cin = c0 < ny;
  const int cc = cin ? c0 : ny - 4;
  const double beta = first ? 0.0 : st[row].beta;
  const TV* R = r + row * vstride;
  const TV* P = pold + row * vstride;
  TV* PN = pnew + row * vstride;
  auto pn = [&](double rv, double pv) -> double {
    const double v = first ? rv : fma(beta, pv, rv);
    return static_cast<double>(static_cast<TV>(v));
  };
  auto row4 = [&](const Q& qr, const Q& qp, bool in, double (&out)[4]) {
#pragma unroll
    for (int e = 0; e < 4; ++e) out[e] = in ? pn(q_at(qr, e), q_at(qp, e)) : 0.0;
  };
  Q qr[kQ4Dir], qp[kQ4Dir];
#pragma unroll
  for (int u = 0; u < kQ4Dir; ++u) {
    const int i = clampi(i0 + 1 + u, 0, imax);
    qr[u] = ld4(R, ny, i, cc);
    qp[u] = ld4(P, ny, i, cc);
  }
  double prev[4], cur[4], next[4];
  {
    const int ip = clampi(i0 - 1, 0, nx - 1);
    row4(ld4(R, ny, ip, cc), ld4(P, ny, ip, cc), cin && i0 >= 1, prev);
    row4(ld4(R, ny, i0, cc), ld4(P, ny, i0, cc), cin, cur);
  }
  double acc = 0.0;
  for (int ib = i0; ib < i1; ib += kQ4Dir) {
#pragma unroll
    for (int u = 0; u < kQ4Dir; ++u) {
      const int i = ib + u;
      const bool live = i < i1;
      row4(qr[u], qp[u], cin && i + 1 <= imax, next);
      const int il = clampi(i + 1 + kQ4Dir, 0, imax);
      qr[u] = ld4(R, ny, il, cc);
      qp[u] = ld4(P, ny, il, cc);
      const int buf = u & 1;
      eL[buf][t] = cur[0];
      eR[buf][t] = cur[3];
      lds_barrier();
      if (live && cin) {
        const double west = t > 0 ? eR[buf][t - 1] : 0.0;
        const double east = t + 1 < nt ? eL[buf][t + 1] : 0.0;
        double ap[4];
        stencil4(k, prev, cur, next, west, east, ap);
#pragma unroll
        for (int e = 0; e < 4; ++e) acc = fma(ap[e], cur[e], acc);
        Q o;
        o.x = static_cast<TV>(cur[0]);
        o.y = static_cast<TV>(cur[1]);
        o.z = static_cast<TV>(cur[2]);
        o.w = static_cast<TV>(cur[3]);
        *reinterpret_cast<Q*>(PN + static_cast<int64_t>(i) * ny + c0) = o;
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        prev[e] = cur[e];
        cur[e] = next[e];
      }
    }
  }
  double z0 = 0.0, z1 = 0.0;
  block_reduce3_n(1, z0, acc, z1, red, nt);
  if (t == 0) part[static_cast<int64_t>(row) * 3 * gridDim.x + blockIdx.x] = acc;
}

template <bool kStable, typename TY, typename TV, typename TB>
__global__ __launch_bounds__(kMaxThreads4) void cgs_update4_kernel(int nx, int ny, St5 k, const TV* __restrict__ p,
                                                                   const TY* y_in, TY* y_out, TV* __restrict__ r,
                                                                   const TB* __restrict__ b, int64_t vstride,
                                                                   CgP prm, const lxa_krylov_row* __restrict__ st,
                                                                   double* __restrict__ part) {
  using QV = typename Quad<TV>::type;
  using QY = typename Quad<TY>::type;
  using QB = typename Quad<TB>::type;
  __shared__ double eL[2][kMaxThreads4 + 1], eR[2][kMaxThreads4 + 1];
  __shared__ double red[48];
  const int row = blockIdx.y;
  const int t = threadIdx.x, nt = blockDim.x;
  const int i0 = blockIdx.x * kTH4, i1 = i0 + kTH4 < nx ? i0 + kTH4 : nx;
  const int c0 = 4 * t;
  const bool cin = c0 < ny;
  if (!st[row].active) {
    // a finished row keeps its iterate across the buffer swap
    if (kStable && cin)
      for (int i = i0; i < i1; ++i) {
        const int64_t o = row * vstride + static_cast<int64_t>(i) * ny + c0;
        *reinterpret_cast<QY*>(y_out + o) = *reinterpret_cast<const QY*>(y_in + o);
      }
    return;
  }
  const int imax = i1 < nx - 1 ? i1 : nx - 1;
  const int own = i1 - 1;
  const int cc = cin ? c0 : ny - 4;
  const double alpha = st[row].alpha;
  const int kind = prm.norm_kind;
  const TV* P = p + row * vstride;
  const TY* Y = y_in + row * vstride;
  TY* YO = y_out + row * vstride;
  TV* Rv = r + row * vstride;
  const TB* B = b + row * vstride;
  // stencil operand: p (plain step) or y_new = y + alpha p rounded to storage
  auto opnd = [&](const QV& qp, const QY& qy, bool in, double (&out)[4]) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double pv = q_at(qp, e);
      out[e] = !in ? 0.0 : (kStable ? static_cast<double>(static_cast<TY>(q_at(qy, e) + alpha * pv)) : pv);
    }
  };
  QV qp[kQ4], cr[kQ4];
  QY qy[kQ4], cy[kQ4];
  QB cb[kQ4];
#pragma unroll
  for (int u = 0; u < kQ4; ++u) {
    const int i = clampi(i0 + 1 + u, 0, imax);
    qp[u] = ld4(P, ny, i, cc);
    if (kStable) qy[u] = ld4(Y, ny, i, cc);
    const int ic = clampi(i0 + u, 0, own);
    if (!kStable) {
      cy[u] = ld4(Y, ny, ic, cc);
      cr[u] = ld4(Rv, ny, ic, cc);
    }
    cb[u] = ld4(B, ny, ic, cc);
  }
  double prev[4], cur[4], next[4], pcur[4], ycur[4];
  {
    const int ip = clampi(i0 - 1, 0, nx - 1);
    QY yz{};
    opnd(ld4(P, ny, ip, cc), kStable ? ld4(Y, ny, ip, cc) : yz, cin && i0 >= 1, prev);
    const QV p0 = ld4(P, ny, i0, cc);
    const QY y0 = kStable ? ld4(Y, ny, i0, cc) : yz;
    opnd(p0, y0, cin, cur);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      pcur[e] = q_at(p0, e);
      ycur[e] = q_at(y0, e);
    }
  }
  double n2 = 0.0, rr = 0.0, n1 = 0.0;
  // lineax's loop condition is (norm1 > 1) | (norm2 > 1) (cg.py:150-158).
  // Once a wave has seen norm2's contribution exceed 1 (max norm, or the
  // two-norm's partial sum), the row cannot stop this step whatever norm1
  // is, so the wave stops reading b for norm1 (plain steps; stabilising steps
  // need b for the residual).  Should norm2 later turn NaN (NaN compares
  // false), norm1 is unknown and is set to +inf: the row runs one more step.
  const bool can_skip = !kStable && prm.has_scale && (kind == 0 || kind == 1);
  bool skipped = false;
  for (int ib = i0; ib < i1; ib += kQ4) {
#pragma unroll
    for (int u = 0; u < kQ4; ++u) {
      const int i = ib + u;
      const bool live = i < i1;
      const bool nin = cin && i + 1 <= imax;
      double pnext[4], ynext[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        pnext[e] = nin ? q_at(qp[u], e) : 0.0;
        ynext[e] = (kStable && nin) ? q_at(qy[u], e) : 0.0;
      }
      opnd(qp[u], qy[u], nin, next);
      double yold[4], rold[4], bv[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        yold[e] = kStable ? ycur[e] : q_at(cy[u], e);
        rold[e] = kStable ? 0.0 : q_at(cr[u], e);
        bv[e] = q_at(cb[u], e);
      }
      {
        const int il = clampi(i + 1 + kQ4, 0, imax);
        qp[u] = ld4(P, ny, il, cc);
        if (kStable) qy[u] = ld4(Y, ny, il, cc);
        const int ic = clampi(i + kQ4, 0, own);
        if (!kStable) {
          cy[u] = ld4(Y, ny, ic, cc);
          cr[u] = ld4(Rv, ny, ic, cc);
        }
        if (!skipped) cb[u] = ld4(B, ny, ic, cc);
      }
      const int buf = u & 1;
      eL[buf][t] = cur[0];
      eR[buf][t] = cur[3];
      lds_barrier();
      if (live && cin) {
        const double west = t > 0 ? eR[buf][t - 1] : 0.0;
        const double east = t + 1 < nt ? eL[buf][t + 1] : 0.0;
        double av[4];
        stencil4(k, prev, cur, next, west, east, av);
        QY yo;
        QV ro;
        double ys[4], rs[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const double d = alpha * pcur[e];
          const double yn = yold[e] + d;
          ys[e] = yn;
          if (prm.has_scale) n2 = norm_acc(kind, n2, norm_elem(kind, d / (prm.atol + prm.rtol * fabs(yn))));
          const double rn = kStable ? bv[e] - av[e] : rold[e] - alpha * av[e];
          rs[e] = static_cast<double>(static_cast<TV>(rn));
          rr = fma(rs[e], rs[e], rr);
          if (prm.has_scale && !skipped)
            n1 = norm_acc(kind, n1, norm_elem(kind, rs[e] / (prm.atol + prm.rtol * fabs(bv[e]))));
        }
        yo.x = static_cast<TY>(ys[0]);
        yo.y = static_cast<TY>(ys[1]);
        yo.z = static_cast<TY>(ys[2]);
        yo.w = static_cast<TY>(ys[3]);
        ro.x = static_cast<TV>(rs[0]);
        ro.y = static_cast<TV>(rs[1]);
        ro.z = static_cast<TV>(rs[2]);
        ro.w = static_cast<TV>(rs[3]);
        const int64_t o = static_cast<int64_t>(i) * ny + c0;
        *reinterpret_cast<QY*>(YO + o) = yo;
        *reinterpret_cast<QV*>(Rv + o) = ro;
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        prev[e] = cur[e];
        cur[e] = next[e];
        pcur[e] = pnext[e];
        ycur[e] = ynext[e];
      }
      if (can_skip && !skipped && __any(n2 > 1.0)) skipped = true;
    }
  }
  if (skipped && n2 != n2) n1 = INFINITY;
  block_reduce3_n(kind, n2, rr, n1, red, nt);
  if (t == 0) {
    double* pr = part + static_cast<int64_t>(row) * 3 * gridDim.x;
    pr[blockIdx.x] = n2;
    pr[gridDim.x + blockIdx.x] = rr;
    pr[2 * gridDim.x + blockIdx.x] = n1;
  }
}

bool use_vec4(int64_t ny) { return ny % 4 == 0 && ny <= 4 * kMaxThreads4; }
int threads4(int64_t ny) { return static_cast<int>(((ny / 4 + 63) / 64) * 64); }
int tiles4(int64_t nx) { return static_cast<int>((nx + kTH4 - 1) / kTH4); }
int tiles4_dir(int64_t nx) { return static_cast<int>((nx + kTH4Dir - 1) / kTH4Dir); }

int tiles_of(int64_t nx, int64_t ny) {
  return static_cast<int>(((nx + kTH - 1) / kTH) * ((ny + kTW - 1) / kTW));
}

CgP cgp_of(const lxa_krylov_params* p) {
  CgP c;
  c.rtol = p->rtol;
  c.atol = p->atol;
  c.rcond = p->rcond;
  c.max_steps = p->max_steps;
  c.has_scale = !(p->rtol == 0.0 && p->atol == 0.0);
  c.norm_kind = p->norm_kind;
  return c;
}

int check_args(int64_t rows, int64_t nx, int64_t ny, const double* coeffs, const lxa_krylov_params* prm,
               const void* st, const void* part) {
  LXA_REQUIRE(rows >= 1 && rows <= 65535, "lxa_cg_stencil: rows must be in [1, 65535]");
  LXA_REQUIRE(nx >= 1 && ny >= 1 && nx <= (1 << 24) && ny <= (1 << 24), "lxa_cg_stencil: bad grid %lld x %lld",
              (long long)nx, (long long)ny);
  LXA_REQUIRE(tiles_of(nx, ny) <= (1 << 30), "lxa_cg_stencil: grid too large");
  LXA_REQUIRE(coeffs && prm && st && part, "lxa_cg_stencil: null pointer");
  LXA_REQUIRE(prm->norm_kind >= 0 && prm->norm_kind <= 2, "lxa_cg_stencil: bad norm_kind");
  return LXA_OK;
}

template <bool S, typename TY, typename TV, typename TB>
void launch_update(hipStream_t s, dim3 grid, int nx, int ny, int tj, St5 k, const void* p, const void* yi, void* yo,
                   void* r, const void* b, int64_t vs, CgP c, const lxa_krylov_row* st, double* part) {
  if (use_vec4(ny))
    hipLaunchKernelGGL((cgs_update4_kernel<S, TY, TV, TB>), grid, dim3(static_cast<unsigned>(threads4(ny))), 0, s, nx,
                       ny, k, static_cast<const TV*>(p), static_cast<const TY*>(yi), static_cast<TY*>(yo),
                       static_cast<TV*>(r), static_cast<const TB*>(b), vs, c, st, part);
  else
    hipLaunchKernelGGL((cgs_update_kernel<S, TY, TV, TB>), grid, dim3(kTW), 0, s, nx, ny, tj, k,
                       static_cast<const TV*>(p), static_cast<const TY*>(yi), static_cast<TY*>(yo),
                       static_cast<TV*>(r), static_cast<const TB*>(b), vs, c, st, part);
}

template <typename TY, typename TV, typename TB>
void launch_update2(int stable, hipStream_t s, dim3 grid, int nx, int ny, int tj, St5 k, const void* p,
                    const void* yi, void* yo, void* r, const void* b, int64_t vs, CgP c, const lxa_krylov_row* st,
                    double* part) {
  if (stable)
    launch_update<true, TY, TV, TB>(s, grid, nx, ny, tj, k, p, yi, yo, r, b, vs, c, st, part);
  else
    launch_update<false, TY, TV, TB>(s, grid, nx, ny, tj, k, p, yi, yo, r, b, vs, c, st, part);
}

}  // namespace
}  // namespace lxa

extern "C" int lxa_cg_stencil_workspace(int64_t rows, int64_t nx, int64_t ny, size_t* partial_bytes) {
  LXA_REQUIRE(rows >= 1 && nx >= 1 && ny >= 1 && partial_bytes, "lxa_cg_stencil_workspace: bad argument");
  int64_t tiles = lxa::tiles_of(nx, ny);
  if (lxa::tiles4(nx) > tiles) tiles = lxa::tiles4(nx);
  if (lxa::tiles4_dir(nx) > tiles) tiles = lxa::tiles4_dir(nx);
  *partial_bytes = static_cast<size_t>(rows) * 3 * static_cast<size_t>(tiles) * sizeof(double);
  return LXA_OK;
}

extern "C" int lxa_cg_stencil_direction(void* stream, int vdtype, int64_t rows, int64_t nx, int64_t ny,
                                        const double* coeffs, const void* r, const void* p_old, void* p_new,
                                        int first, const lxa_krylov_params* prm, lxa_krylov_row* state,
                                        double* partials) {
  if (int rc = lxa::check_args(rows, nx, ny, coeffs, prm, state, partials)) return rc;
  LXA_REQUIRE(r && p_new && p_old, "lxa_cg_stencil_direction: null vector");
  LXA_REQUIRE(p_new != p_old && p_new != r, "lxa_cg_stencil_direction: p_new must not alias r or p_old");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const lxa::St5 k{coeffs[0], coeffs[1], coeffs[2], coeffs[3], coeffs[4]};
  const int tj = static_cast<int>((ny + lxa::kTW - 1) / lxa::kTW);
  const bool v4 = lxa::use_vec4(ny);
  const int nblk = v4 ? lxa::tiles4_dir(nx) : lxa::tiles_of(nx, ny);
  const dim3 grid(static_cast<unsigned>(nblk), static_cast<unsigned>(rows));
  const int64_t vs = nx * ny;
  const lxa::CgP c = lxa::cgp_of(prm);
  if (v4) {
    const dim3 blk(static_cast<unsigned>(lxa::threads4(ny)));
    if (vdtype == LXA_F32)
      hipLaunchKernelGGL(lxa::cgs_direction4_kernel<float>, grid, blk, 0, s, static_cast<int>(nx), static_cast<int>(ny),
                         k, static_cast<const float*>(r), static_cast<const float*>(p_old), static_cast<float*>(p_new),
                         first, vs, state, partials);
    else if (vdtype == LXA_F64)
      hipLaunchKernelGGL(lxa::cgs_direction4_kernel<double>, grid, blk, 0, s, static_cast<int>(nx),
                         static_cast<int>(ny), k, static_cast<const double*>(r), static_cast<const double*>(p_old),
                         static_cast<double*>(p_new), first, vs, state, partials);
    else
      return lxa::set_error(LXA_EINVAL, "lxa_cg_stencil_direction: bad dtype");
  } else if (vdtype == LXA_F32)
    hipLaunchKernelGGL(lxa::cgs_direction_kernel<float>, grid, dim3(lxa::kTW), 0, s, static_cast<int>(nx),
                       static_cast<int>(ny), tj, k, static_cast<const float*>(r), static_cast<const float*>(p_old),
                       static_cast<float*>(p_new), first, vs, state, partials);
  else if (vdtype == LXA_F64)
    hipLaunchKernelGGL(lxa::cgs_direction_kernel<double>, grid, dim3(lxa::kTW), 0, s, static_cast<int>(nx),
                       static_cast<int>(ny), tj, k, static_cast<const double*>(r), static_cast<const double*>(p_old),
                       static_cast<double*>(p_new), first, vs, state, partials);
  else
    return lxa::set_error(LXA_EINVAL, "lxa_cg_stencil_direction: bad dtype");
  hipLaunchKernelGGL(lxa::cgs_alpha_kernel, dim3(static_cast<unsigned>(rows)), dim3(lxa::kTW), 0, s, nblk, c, state,
                     partials);
  LXA_LAUNCH_CHECK("lxa_cg_stencil_direction");
  return LXA_OK;
}

extern "C" int lxa_cg_stencil_update(void* stream, int ydtype, int vdtype, int bdtype, int64_t rows, int64_t nx,
                                     int64_t ny, const double* coeffs, const void* p, const void* y_in,
                                     void* y_out, void* r, const void* b, int stable, const lxa_krylov_params* prm,
                                     lxa_krylov_row* state, double* partials) {
  if (int rc = lxa::check_args(rows, nx, ny, coeffs, prm, state, partials)) return rc;
  LXA_REQUIRE(p && y_in && y_out && r && b, "lxa_cg_stencil_update: null vector");
  LXA_REQUIRE(stable ? y_in != y_out : y_in == y_out,
              "lxa_cg_stencil_update: stabilising steps need y_out != y_in, plain steps y_out == y_in");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const lxa::St5 k{coeffs[0], coeffs[1], coeffs[2], coeffs[3], coeffs[4]};
  const int tj = static_cast<int>((ny + lxa::kTW - 1) / lxa::kTW);
  const int nblk = lxa::use_vec4(ny) ? lxa::tiles4(nx) : lxa::tiles_of(nx, ny);
  const dim3 grid(static_cast<unsigned>(nblk), static_cast<unsigned>(rows));
  const int64_t vs = nx * ny;
  const lxa::CgP c = lxa::cgp_of(prm);
  const int a = static_cast<int>(nx), bn = static_cast<int>(ny);
  using lxa::launch_update2;
  if (ydtype == LXA_F64 && vdtype == LXA_F32 && bdtype == LXA_F32)
    launch_update2<double, float, float>(stable, s, grid, a, bn, tj, k, p, y_in, y_out, r, b, vs, c, state, partials);
  else if (ydtype == LXA_F64 && vdtype == LXA_F32 && bdtype == LXA_F64)
    launch_update2<double, float, double>(stable, s, grid, a, bn, tj, k, p, y_in, y_out, r, b, vs, c, state,
                                          partials);
  else if (ydtype == LXA_F64 && vdtype == LXA_F64 && bdtype == LXA_F64)
    launch_update2<double, double, double>(stable, s, grid, a, bn, tj, k, p, y_in, y_out, r, b, vs, c, state,
                                           partials);
  else if (ydtype == LXA_F32 && vdtype == LXA_F32 && bdtype == LXA_F32)
    launch_update2<float, float, float>(stable, s, grid, a, bn, tj, k, p, y_in, y_out, r, b, vs, c, state, partials);
  else
    return lxa::set_error(LXA_EINVAL, "lxa_cg_stencil_update: unsupported dtypes (y %d, vectors %d, b %d)", ydtype,
                          vdtype, bdtype);
  hipLaunchKernelGGL(lxa::cgs_finish_kernel, dim3(static_cast<unsigned>(rows)), dim3(lxa::kTW), 0, s, nblk, nx * ny, c,
                     state, partials);
  LXA_LAUNCH_CHECK("lxa_cg_stencil_update");
  return LXA_OK;
}
