// Fused Arnoldi steps of GMRES for constant-coefficient five-point stencil
// operators (no preconditioner), grids at most 1024 points wide.
//
// lineax's inner step (lineax/_solver/gmres.py:355-398) is classical
// Gram-Schmidt:  w = A v_j;  h = V^T w;  w' = w - V h;  nrm = |w'|;
// breakdown if nrm < |w| eps;  v_{j+1} = w' / nrm.  Run as separate kernels
// it reads the basis twice per step (the projection, then the subtraction).
// Here one sweep does the subtraction of step j-1 and the projection of
// step j together, reading the basis once:
//   u_j = alpha w_{j-1}^raw - sum_{k<j} coef_k u_k    (u_j = w'_{j-1})
//   t_j = A u_j                                        (on the fly, rolling rows)
//   partial sums of u_k . t_j (k < j), u_j . t_j, u_j . u_j, t_j . t_j.
// u_j is stored unnormalised as basis column j with its scale s_j = 1 / |u_j|
// kept per column (v_j = s_j u_j), and t_j as the raw next w.  The per-row
// finish kernel turns the sums into lineax's quantities:
//   nrm_{j-1} = |u_j|, the breakdown test and Hessenberg row j-1;
//   h_{j,k} = s_k s_j (u_k . t_j), h_{j,j} = s_j^2 (u_j . t_j),
//   |w_j| = s_j |t_j|;  and the coefficients of the next sweep.
// A last sweep (mode 2) forms |u_m| to close row m-1.  Column j of the basis
// is stored unnormalised, so the update uses z_k s_k.
//
// One workgroup covers whole grid rows (a thread per grid column) for kTH
// rows; the rows above and below the tile are recomputed (u there depends
// only on data that this sweep does not write).  Reductions are two-phase
// and in a fixed order.
#include "lxa_common.h"

namespace lxa {
namespace {

constexpr int kMaxK = 32;         // largest restart on this path
constexpr int kRowWidth = 1024;  // widest grid row
constexpr int kThreads = 512;    // threads per workgroup
constexpr int kCols = kRowWidth / kThreads;  // grid columns per thread
constexpr int kTH = 16;           // grid rows per tile
constexpr int kSlots = kMaxK + 3;  // partial sums per row: u_k.t (k < j), u.t, u.u, t.t
constexpr int kUT = kMaxK, kUU = kMaxK + 1, kTT = kMaxK + 2;

struct St5 {
  double c, w, e, s, n;
};

struct GmsParams {
  double eps;
  int restart;
};

template <typename K>
int allow_vrow(K kernel, size_t bytes) {
  if (bytes == 0) return LXA_OK;
  const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
  return e == hipSuccess ? LXA_OK : hip_error(e, "hipFuncSetAttribute(gms_pass_kernel)");
}

// One sweep.  MODE 0: u = v_0 (step 0).  MODE 1: step j >= 1 as above
// (mode 2, the closing sweep, is gms_close_kernel).  512 threads, each owning
// grid columns t and t + 512 (handled one after the other, so one column's
// basis values are live at a time).
template <typename T>
struct Pair2;
template <>
struct Pair2<float> {
  using type = float2;
};
template <>
struct Pair2<double> {
  using type = double2;
};

// PAIR: thread t owns the adjacent grid columns 2t, 2t + 1 (8- or 16-byte
// loads; needs an even grid width); otherwise columns t and t + 512.
template <typename TV, int MODE, bool PAIR>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 2))) void gms_pass_kernel(int nx, int ny, St5 k, int j, int m,
                                                            TV* __restrict__ basis, const TV* __restrict__ w_in,
                                                            TV* __restrict__ w_out, const double* __restrict__ coefs,
                                                            const lxa_krylov_row* __restrict__ st,
                                                            const lxa_gmres_row* __restrict__ gs,
                                                            double* __restrict__ part) {
  __shared__ double lrow[2][kRowWidth + 2];
  __shared__ double red[kSlots][kThreads / 64];
  extern __shared__ float vrow[];  // [kMaxK][kRowWidth] float basis values of the row behind
  const int row = blockIdx.y;
  if (!st[row].active || gs[row].inner_done) return;
  const int64_t n = static_cast<int64_t>(nx) * ny;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int i0 = blockIdx.x * kTH, i1 = i0 + kTH < nx ? i0 + kTH : nx;
  TV* V = basis + static_cast<int64_t>(row) * (m + 1) * n;
  const TV* W = w_in + static_cast<int64_t>(row) * n;
  TV* WO = w_out + static_cast<int64_t>(row) * n;
  const int nk = MODE == 0 ? 0 : j;  // basis columns combined into u
  const double* __restrict__ cf = coefs + static_cast<int64_t>(row) * (kMaxK + 1);
  // float basis values of the row behind wait in LDS (each thread's own
  // columns, so no barrier); double ones are read again through the caches
  constexpr bool kPark = MODE == 1 && sizeof(TV) == 4;

  // u at (i, col) (0 outside the grid), rounded to storage; with kPark the
  // row's basis values go to `vals` for parking
  auto u_at = [&](int i, int col, TV (&vals)[kMaxK]) -> double {
    const bool in = i >= 0 && i < nx && col < ny;
    const int64_t o = static_cast<int64_t>(in ? i : 0) * ny + (col < ny ? col : ny - 1);
    if (MODE == 0) return in ? static_cast<double>(V[o]) : 0.0;
    // the coefficients are uniform over the block: read through the scalar
    // cache (SGPRs), not held in vector registers
    double v = cf[0] * static_cast<double>(W[o]);
#pragma unroll
    for (int kk = 0; kk < kMaxK; ++kk) vals[kk] = kk < nk ? V[kk * n + o] : TV(0);
#pragma unroll
    for (int kk = 0; kk < kMaxK; ++kk)
      if (kk < nk) v = fma(-cf[1 + kk], static_cast<double>(vals[kk]), v);
    return in ? static_cast<double>(static_cast<TV>(v)) : 0.0;
  };
  auto park = [&](int col, const TV (&vals)[kMaxK]) {
    if (kPark) {
#pragma unroll
      for (int kk = 0; kk < kMaxK; ++kk)
        if (kk < nk) vrow[kk * kRowWidth + col] = static_cast<float>(vals[kk]);
    }
  };

  double acc[kMaxK];
#pragma unroll
  for (int q = 0; q < kMaxK; ++q) acc[q] = 0.0;
  double ut = 0.0, uu = 0.0, tt = 0.0;
  if constexpr (PAIR) {
    using P = typename Pair2<TV>::type;
    const int c0 = 2 * t;  // columns c0, c0 + 1 (ny even: both in or both out)
    const bool cin = c0 < ny;
    const int cc = cin ? c0 : ny - 2;
    P pv[kMaxK];
    // u at (i, c0..c0+1) and the row's basis values
    auto u2 = [&](int i, double& ua, double& ub) {
      const bool in = i >= 0 && i < nx && cin;
      const int64_t o = static_cast<int64_t>(in ? i : 0) * ny + cc;
      if (MODE == 0) {
        const P v = *reinterpret_cast<const P*>(V + o);
        ua = in ? static_cast<double>(v.x) : 0.0;
        ub = in ? static_cast<double>(v.y) : 0.0;
        return;
      }
      const P wv2 = *reinterpret_cast<const P*>(W + o);
#pragma unroll
      for (int kk = 0; kk < kMaxK; ++kk)
        if (kk < nk) pv[kk] = *reinterpret_cast<const P*>(V + kk * n + o);
      double va = cf[0] * static_cast<double>(wv2.x), vb = cf[0] * static_cast<double>(wv2.y);
#pragma unroll
      for (int kk = 0; kk < kMaxK; ++kk)
        if (kk < nk) {
          va = fma(-cf[1 + kk], static_cast<double>(pv[kk].x), va);
          vb = fma(-cf[1 + kk], static_cast<double>(pv[kk].y), vb);
        }
      ua = in ? static_cast<double>(static_cast<TV>(va)) : 0.0;
      ub = in ? static_cast<double>(static_cast<TV>(vb)) : 0.0;
    };
    auto park2 = [&]() {
      if (kPark) {
#pragma unroll
        for (int kk = 0; kk < kMaxK; ++kk)
          if (kk < nk)
            *reinterpret_cast<float2*>(vrow + kk * kRowWidth + c0) =
                make_float2(static_cast<float>(pv[kk].x), static_cast<float>(pv[kk].y));
      }
    };
    double pa, pb, ca, cb, na, nb;
    u2(i0 - 1, pa, pb);
    u2(i0, ca, cb);
    park2();
    for (int i = i0; i < i1; ++i) {
      u2(i + 1, na, nb);
      double* lr = lrow[i & 1];
      lr[c0 + 1] = ca;
      lr[c0 + 2] = cb;
      if (t == 0) {
        lr[0] = 0.0;
        lr[kRowWidth + 1] = 0.0;
      }
      lds_barrier();
      if (cin) {
        const double ta = k.c * ca + k.w * lr[c0] + k.e * cb + k.s * pa + k.n * na;
        const double tb = k.c * cb + k.w * ca + k.e * lr[c0 + 3] + k.s * pb + k.n * nb;
        const int64_t o = static_cast<int64_t>(i) * ny + c0;
        P uo, to;
        uo.x = static_cast<TV>(ca);
        uo.y = static_cast<TV>(cb);
        to.x = static_cast<TV>(ta);
        to.y = static_cast<TV>(tb);
        if (MODE == 1) *reinterpret_cast<P*>(V + static_cast<int64_t>(j) * n + o) = uo;
        *reinterpret_cast<P*>(WO + o) = to;
        const double tra = static_cast<double>(to.x), trb = static_cast<double>(to.y);
        if (kPark) {
#pragma unroll
          for (int kk = 0; kk < kMaxK; ++kk)
            if (kk < nk) {
              const float2 q = *reinterpret_cast<const float2*>(vrow + kk * kRowWidth + c0);
              acc[kk] = fma(static_cast<double>(q.y), trb, fma(static_cast<double>(q.x), tra, acc[kk]));
            }
        } else if (MODE == 1) {
          for (int kk = 0; kk < nk; ++kk) {
            const P q = *reinterpret_cast<const P*>(V + kk * n + o);
            const double d = fma(static_cast<double>(q.y), trb, static_cast<double>(q.x) * tra);
#pragma unroll
            for (int qq = 0; qq < kMaxK; ++qq)
              if (qq == kk) acc[qq] += d;
          }
        }
        ut = fma(cb, trb, fma(ca, tra, ut));
        uu = fma(cb, cb, fma(ca, ca, uu));
        tt = fma(trb, trb, fma(tra, tra, tt));
      }
      park2();
      pa = ca;
      pb = cb;
      ca = na;
      cb = nb;
    }
  } else {
    TV vals[kMaxK];
    double prev[kCols], cur[kCols], next[kCols];
#pragma unroll
    for (int h = 0; h < kCols; ++h) {
      const int col = t + kThreads * h;
      prev[h] = u_at(i0 - 1, col, vals);
      cur[h] = u_at(i0, col, vals);
      park(col, vals);
    }
    for (int i = i0; i < i1; ++i) {
      double* lr = lrow[i & 1];
#pragma unroll
      for (int h = 0; h < kCols; ++h) lr[t + kThreads * h + 1] = cur[h];
      if (t == 0) {
        lr[0] = 0.0;
        lr[kRowWidth + 1] = 0.0;
      }
      lds_barrier();
#pragma unroll
      for (int h = 0; h < kCols; ++h) {
        const int col = t + kThreads * h;
        next[h] = u_at(i + 1, col, vals);
        if (col < ny) {
          const double tv = k.c * cur[h] + k.w * lr[col] + k.e * lr[col + 2] + k.s * prev[h] + k.n * next[h];
          const int64_t o = static_cast<int64_t>(i) * ny + col;
          if (MODE == 1) V[static_cast<int64_t>(j) * n + o] = static_cast<TV>(cur[h]);
          WO[o] = static_cast<TV>(tv);
          const double tr = static_cast<double>(static_cast<TV>(tv));
          if (kPark) {
#pragma unroll
            for (int kk = 0; kk < kMaxK; ++kk)
              if (kk < nk) acc[kk] = fma(static_cast<double>(vrow[kk * kRowWidth + col]), tr, acc[kk]);
          } else if (MODE == 1) {
            for (int kk = 0; kk < nk; ++kk) {
              const double vk = static_cast<double>(V[kk * n + o]);
#pragma unroll
              for (int q = 0; q < kMaxK; ++q)
                if (q == kk) acc[q] = fma(vk, tr, acc[q]);
            }
          }
          ut = fma(cur[h], tr, ut);
          uu = fma(cur[h], cur[h], uu);
          tt = fma(tr, tr, tt);
        }
        park(col, vals);
        prev[h] = cur[h];
        cur[h] = next[h];
        // one column's loads and values at a time
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  // block partials: per-wave butterflies, then the waves in a fixed order
  auto wave_reduce = [&](double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
  };
#pragma unroll
  for (int kk = 0; kk < kMaxK; ++kk) {
    if (kk < nk) {
      const double v = wave_reduce(acc[kk]);
      if (lane == 0) red[kk][wv] = v;
    }
  }
  {
    const double a = wave_reduce(ut), b = wave_reduce(uu), c = wave_reduce(tt);
    if (lane == 0) {
      red[kUT][wv] = a;
      red[kUU][wv] = b;
      red[kTT][wv] = c;
    }
  }
  __syncthreads();
  if (t < kSlots && (t < nk || t >= kMaxK)) {
    double s = 0.0;
    for (int q = 0; q < kThreads / 64; ++q) s += red[t][q];
    part[(static_cast<int64_t>(row) * kSlots + t) * gridDim.x + blockIdx.x] = s;
  }
}

// The closing sweep: |u_m|^2 with u_m = alpha w - sum_k coef_k u_k, a plain
// streaming reduction (no stencil), written to the same partial slot as
// gms_pass_kernel's u.u (the other slots are zeroed).
constexpr int kCloseThreads = 256;

template <typename T>
struct Quad4;
template <>
struct Quad4<float> {
  using type = float4;
};
template <>
struct Quad4<double> {
  using type = double4;
};

// VEC4: four consecutive points per thread with 16/32-byte loads (n % 4 == 0).
template <typename TV, bool VEC4>
__global__ __launch_bounds__(kCloseThreads) void gms_close_kernel(int64_t n, int m, int nblk_row,
                                                                  const TV* __restrict__ basis,
                                                                  const TV* __restrict__ w_in,
                                                                  const double* __restrict__ coefs,
                                                                  const lxa_krylov_row* __restrict__ st,
                                                                  const lxa_gmres_row* __restrict__ gs,
                                                                  double* __restrict__ part) {
  __shared__ double red[kCloseThreads / 64];
  const int row = blockIdx.y;
  if (!st[row].active || gs[row].inner_done) return;
  const TV* V = basis + static_cast<int64_t>(row) * (m + 1) * n;
  const TV* W = w_in + static_cast<int64_t>(row) * n;
  const double* __restrict__ cf = coefs + static_cast<int64_t>(row) * (kMaxK + 1);
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t b0 = static_cast<int64_t>(blockIdx.x) * per;
  const int64_t b1 = b0 + per < n ? b0 + per : n;
  double uu = 0.0;
  if constexpr (VEC4) {
    using Q = typename Quad4<TV>::type;
    // block ranges rounded to multiples of 4 points
    const int64_t q0 = b0 & ~int64_t(3), q1 = blockIdx.x + 1 == gridDim.x ? n : (b1 & ~int64_t(3));
    for (int64_t o = q0 + 4 * static_cast<int64_t>(threadIdx.x); o < q1; o += 4 * kCloseThreads) {
      const Q wq = *reinterpret_cast<const Q*>(W + o);
      double v0 = cf[0] * static_cast<double>(wq.x), v1 = cf[0] * static_cast<double>(wq.y);
      double v2 = cf[0] * static_cast<double>(wq.z), v3 = cf[0] * static_cast<double>(wq.w);
#pragma unroll 8
      for (int kk = 0; kk < m; ++kk) {
        const Q q = *reinterpret_cast<const Q*>(V + kk * n + o);
        const double c = cf[1 + kk];
        v0 = fma(-c, static_cast<double>(q.x), v0);
        v1 = fma(-c, static_cast<double>(q.y), v1);
        v2 = fma(-c, static_cast<double>(q.z), v2);
        v3 = fma(-c, static_cast<double>(q.w), v3);
      }
      const double u0 = static_cast<TV>(v0), u1 = static_cast<TV>(v1), u2 = static_cast<TV>(v2),
                   u3 = static_cast<TV>(v3);
      uu = fma(u3, u3, fma(u2, u2, fma(u1, u1, fma(u0, u0, uu))));
    }
  } else {
    for (int64_t o = b0 + threadIdx.x; o < b1; o += kCloseThreads) {
      TV vals[kMaxK];
#pragma unroll
      for (int kk = 0; kk < kMaxK; ++kk) vals[kk] = kk < m ? V[kk * n + o] : TV(0);
      double v = cf[0] * static_cast<double>(W[o]);
#pragma unroll
      for (int kk = 0; kk < kMaxK; ++kk)
        if (kk < m) v = fma(-cf[1 + kk], static_cast<double>(vals[kk]), v);
      const double u = static_cast<double>(static_cast<TV>(v));
      uu = fma(u, u, uu);
    }
  }
  uu = wave_sum(uu);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = uu;
  __syncthreads();
  if (threadIdx.x < kSlots) {
    double s = 0.0;
    if (threadIdx.x == kUU) s = (red[0] + red[1]) + (red[2] + red[3]);
    if (blockIdx.x < static_cast<unsigned>(nblk_row))
      part[(static_cast<int64_t>(row) * kSlots + threadIdx.x) * nblk_row + blockIdx.x] = s;
  }
}

// Per-row finish of a sweep (one block per row).  See the file comment.
// proj holds h of the step being closed; colscale[k] = s_k; coefs[0] = alpha,
// coefs[1 + k] = coef_k for the next sweep.
__global__ __launch_bounds__(256) void gms_finish_kernel(int nblk, int j, int mode, GmsParams prm,
                                                         const double* __restrict__ part, double* __restrict__ coefs,
                                                         double* __restrict__ colscale, double* __restrict__ proj,
                                                         double* __restrict__ H, const lxa_krylov_row* __restrict__ st,
                                                         lxa_gmres_row* __restrict__ gs) {
  __shared__ double sums[kSlots];
  const int row = blockIdx.x;
  if (!st[row].active || gs[row].inner_done) return;
  const int m = prm.restart;
  const int nk = mode == 0 ? 0 : j;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (int q = w; q < kSlots; q += 4) {
    double v = 0.0;
    if (q < nk || q >= kMaxK)
      for (int b = l; b < nblk; b += 64) v += part[(static_cast<int64_t>(row) * kSlots + q) * nblk + b];
    v = wave_sum(v);
    if (l == 0) sums[q] = v;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  lxa_gmres_row g = gs[row];
  double* cs = colscale + static_cast<int64_t>(row) * (m + 1);
  double* h = proj + static_cast<int64_t>(row) * (m + 1);
  double* cf = coefs + static_cast<int64_t>(row) * (kMaxK + 1);
  double sj = 1.0;
  if (mode == 0) {
    cs[0] = 1.0;  // v_0 is stored normalised
    for (int q = 1; q <= m; ++q) cs[q] = 0.0;
  } else {
    // close step j - 1: nrm = |u_j|, breakdown test, Hessenberg row j - 1
    const double nrm = sqrt(sums[kUU]);
    const int bd = nrm < g.step_norm * prm.eps;
    sj = bd ? 0.0 : 1.0 / nrm;
    cs[j] = sj;
    h[j] = nrm;
    double* Hr = H + static_cast<int64_t>(row) * m * (m + 1) + static_cast<int64_t>(j - 1) * (m + 1);
    for (int q = 0; q <= m; ++q) Hr[q] = q <= j ? h[q] : 0.0;
    g.scale = sj;
    g.breakdown = bd;
    g.inner = j;
    g.columns = j + 1;
    g.inner_done = (j >= m) || bd;
  }
  if (!g.inner_done && mode != 2) {
    // step j: projections and |w_j| (w_j = s_j t_j)
    for (int q = 0; q < j; ++q) h[q] = cs[q] * sj * sums[q];
    h[j] = sj * sj * sums[kUT];
    for (int q = j + 1; q <= m; ++q) h[q] = 0.0;
    g.step_norm = sj * sqrt(sums[kTT]);
    cf[0] = sj;
    for (int q = 0; q <= j; ++q) cf[1 + q] = h[q] * cs[q];
  }
  gs[row] = g;
}

}  // namespace
}  // namespace lxa

extern "C" int lxa_gmres_stencil_workspace(int64_t rows, int64_t nx, int64_t ny, size_t* partial_bytes,
                                           size_t* coef_bytes) {
  LXA_REQUIRE(rows >= 1 && nx >= 1 && ny >= 1 && partial_bytes && coef_bytes,
              "lxa_gmres_stencil_workspace: bad argument");
  const int64_t tiles = (nx + lxa::kTH - 1) / lxa::kTH;
  *partial_bytes = static_cast<size_t>(rows * lxa::kSlots * tiles) * sizeof(double);
  *coef_bytes = static_cast<size_t>(rows * (lxa::kMaxK + 1)) * sizeof(double);
  return LXA_OK;
}

extern "C" int lxa_gmres_stencil_step(void* stream, int vdtype, int64_t rows, int64_t nx, int64_t ny,
                                      const double* coeffs, int mode, int j, const lxa_krylov_params* prm,
                                      void* basis, const void* w_in, void* w_out, double* coefs, double* colscale,
                                      double* proj, double* hessenberg, lxa_krylov_row* state,
                                      lxa_gmres_row* gstate, double* partials) {
  LXA_REQUIRE(rows >= 1 && rows <= 65535, "lxa_gmres_stencil_step: rows must be in [1, 65535]");
  LXA_REQUIRE(nx >= 1 && ny >= 1 && ny <= lxa::kRowWidth && nx <= (1 << 24),
              "lxa_gmres_stencil_step: grid %lld x %lld (at most %d wide)", (long long)nx, (long long)ny,
              lxa::kRowWidth);
  LXA_REQUIRE(coeffs && prm && basis && coefs && colscale && proj && hessenberg && state && gstate && partials,
              "lxa_gmres_stencil_step: null pointer");
  const int m = prm->restart;
  LXA_REQUIRE(m >= 1 && m <= lxa::kMaxK, "lxa_gmres_stencil_step: restart must be in [1, %d]", lxa::kMaxK);
  LXA_REQUIRE(mode >= 0 && mode <= 2, "lxa_gmres_stencil_step: bad mode");
  LXA_REQUIRE((mode == 0 && j == 0) || (mode == 1 && j >= 1 && j < m) || (mode == 2 && j == m),
              "lxa_gmres_stencil_step: mode %d with step %d (restart %d)", mode, j, m);
  LXA_REQUIRE(mode == 0 || w_in, "lxa_gmres_stencil_step: w_in required");
  LXA_REQUIRE(mode == 2 || (w_out && w_out != w_in), "lxa_gmres_stencil_step: w_out required, not aliasing w_in");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const lxa::St5 k{coeffs[0], coeffs[1], coeffs[2], coeffs[3], coeffs[4]};
  const int tiles = static_cast<int>((nx + lxa::kTH - 1) / lxa::kTH);
  const dim3 grid(static_cast<unsigned>(tiles), static_cast<unsigned>(rows));
  const int a = static_cast<int>(nx), b = static_cast<int>(ny);
  const size_t vbytes = static_cast<size_t>(lxa::kMaxK) * lxa::kRowWidth * sizeof(float);
#define LXA_GMS_LAUNCH1(TV, MODE, PAIR)                                                                             \
  do {                                                                                                             \
    const size_t vb = (MODE == 1 && sizeof(TV) == 4) ? vbytes : 0;                                                 \
    if (int rc = lxa::allow_vrow((lxa::gms_pass_kernel<TV, MODE, PAIR>), vb)) return rc;                            \
    hipLaunchKernelGGL((lxa::gms_pass_kernel<TV, MODE, PAIR>), grid, dim3(lxa::kThreads), vb, s, a, b, k, j, m,   \
                       static_cast<TV*>(basis), static_cast<const TV*>(w_in), static_cast<TV*>(w_out), coefs,       \
                       state, gstate, partials);                                                                    \
  } while (0)
#define LXA_GMS_LAUNCH(TV, MODE)         \
  do {                                   \
    if (ny % 2 == 0)                     \
      LXA_GMS_LAUNCH1(TV, MODE, true);   \
    else                                 \
      LXA_GMS_LAUNCH1(TV, MODE, false);  \
  } while (0)
  const int64_t n = nx * ny;
#define LXA_GMS_CLOSE1(TV, VEC4)                                                                                    \
  hipLaunchKernelGGL((lxa::gms_close_kernel<TV, VEC4>), grid, dim3(lxa::kCloseThreads), 0, s, n, m, tiles,          \
                     static_cast<const TV*>(basis), static_cast<const TV*>(w_in), coefs, state, gstate, partials)
#define LXA_GMS_CLOSE(TV)            \
  do {                               \
    if (n % 4 == 0)                  \
      LXA_GMS_CLOSE1(TV, true);      \
    else                             \
      LXA_GMS_CLOSE1(TV, false);     \
  } while (0)
  if (vdtype == LXA_F32) {
    if (mode == 0) LXA_GMS_LAUNCH(float, 0);
    else if (mode == 1) LXA_GMS_LAUNCH(float, 1);
    else LXA_GMS_CLOSE(float);
  } else if (vdtype == LXA_F64) {
    if (mode == 0) LXA_GMS_LAUNCH(double, 0);
    else if (mode == 1) LXA_GMS_LAUNCH(double, 1);
    else LXA_GMS_CLOSE(double);
  } else {
    return lxa::set_error(LXA_EINVAL, "lxa_gmres_stencil_step: bad dtype");
  }
#undef LXA_GMS_LAUNCH
#undef LXA_GMS_LAUNCH1
#undef LXA_GMS_CLOSE
#undef LXA_GMS_CLOSE1
  const lxa::GmsParams gp{prm->eps, m};
  hipLaunchKernelGGL(lxa::gms_finish_kernel, dim3(static_cast<unsigned>(rows)), dim3(256), 0, s, tiles, j, mode, gp,
                     partials, coefs, colscale, proj, hessenberg, state, gstate);
  LXA_LAUNCH_CHECK("lxa_gmres_stencil_step");
  return LXA_OK;
}
