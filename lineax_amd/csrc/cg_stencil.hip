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
  hipLaunchKernelGGL((cgs_update_kernel<S, TY, TV, TB>), grid, dim3(kTW), 0, s, nx, ny, tj, k,
                     static_cast<const TV*>(p), static_cast<const TY*>(yi), static_cast<TY*>(yo), static_cast<TV*>(r),
                     static_cast<const TB*>(b), vs, c, st, part);
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
  *partial_bytes = static_cast<size_t>(rows) * 3 * static_cast<size_t>(lxa::tiles_of(nx, ny)) * sizeof(double);
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
  const int nblk = lxa::tiles_of(nx, ny);
  const dim3 grid(static_cast<unsigned>(nblk), static_cast<unsigned>(rows));
  const int64_t vs = nx * ny;
  const lxa::CgP c = lxa::cgp_of(prm);
  if (vdtype == LXA_F32)
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
  const int nblk = lxa::tiles_of(nx, ny);
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
