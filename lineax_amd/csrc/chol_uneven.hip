// Uneven batch of Cholesky solves (orders 1..256) in one launch.
//
// One 256-thread workgroup (a 16 x 16 thread grid, four waves) owns one
// system at a time.  The trailing matrix lives in registers, distributed
// cyclically: thread (ty, tx) holds A(ty + 16a, tx + 16b) for a <= b, i.e.
// the upper triangle in 136 registers.  The factorisation runs in 16-row
// panels:
//   (a) the owners copy panel rows K into an LDS triangle T (row-packed,
//       order N16 = n rounded up to 16; padding rows are the identity);
//   (b) wave 0 factors the 16x16 diagonal block in registers (lane q
//       holds column q; broadcasts by readlane) and stores U_KK in T; then
//       every wave solves U_KK^T U_K,J = A_K,J for its own 64-column chunk
//       of the panel (U_KK read by LDS broadcast) and writes the result to
//       T and to a transposed, bank-conflict-free copy PT[j][p];
//   (c) all threads apply the rank-16 update to their register slots below
//       the panel, reading U values from PT two at a time.
// Three barriers per panel.  T ends up holding U, which one wave uses for the
// triangular solves (vector in registers, reciprocal diagonal in LDS).
// With refinement, residuals r = b - A x are formed in FP64 from A's upper
// triangle in global memory, and x (FP64, LDS) is corrected with FP32
// solves until the normwise backward error meets the tolerance.
//
// Work arrives either from a device queue (persistent launch: workgroups
// pull the next system; callers order the queue largest first) or one
// system per workgroup.
#include <mutex>

#include "lxa_common.h"

namespace lxa {
namespace {

constexpr int kT = 256;
#ifndef LXA_SMALL_WAVES
#define LXA_SMALL_WAVES 3
#endif
constexpr int kMaxN = 256;
constexpr int kPTStride = 18;  // 16 panel values + 2 pad: b64 reads hit distinct banks
constexpr size_t kDoubleBytes = (2048 + 16) * sizeof(double);  // x, r, rowpart, rowabs, acc[4], red
constexpr size_t kRinvBytes = 256 * sizeof(float);
constexpr size_t kItemBytes = 16;

__host__ __device__ inline int round16(int n) { return (n + 15) & ~15; }

// The panel copy PT (rows up to the largest order of the launch) and the
// refinement's doubles share the front of LDS.  Sized by the launch's largest
// order, so that orders up to 176 fit two workgroups in 160 KB.
__host__ __device__ inline size_t union_bytes(int max_n) {
  const size_t pt = static_cast<size_t>(round16(max_n)) * kPTStride * sizeof(float);
  return pt > kDoubleBytes ? pt : kDoubleBytes;
}

__host__ __device__ inline size_t uneven_lds_bytes(int max_n) {
  const size_t n16 = static_cast<size_t>(round16(max_n));
  return union_bytes(max_n) + kRinvBytes + kItemBytes + n16 * (n16 + 1) / 2 * sizeof(float);
}

__device__ __forceinline__ float bcast(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ int pku(int n, int i, int j) { return i * n - ((i * (i - 1)) >> 1) + (j - i); }

struct Lds {
  float* PT;         // [256][18] transposed panel (factorisation phase)
  double* x;         // [256] FP64 solution      } refinement phase,
  double* r;         // [256] residual           } aliasing PT
  double* rowpart;   // [256]                    }
  double* rowabs;    // [256]                    }
  double* acc;       // [4][256]                 }
  double* red;       // [16]                     }
  float* rinv;       // [256] 1 / U_kk
  int* item;         // work-queue slot
  float* T;          // packed U, order N16
};

// lds_n: the largest order the launch's LDS was sized for (uneven_lds_bytes).
__device__ Lds carve(unsigned char* base, int lds_n) {
  const size_t ub = union_bytes(lds_n);
  Lds s;
  s.PT = reinterpret_cast<float*>(base);
  double* d = reinterpret_cast<double*>(base);
  s.x = d;
  s.r = d + 256;
  s.rowpart = d + 512;
  s.rowabs = d + 768;
  s.acc = d + 1024;
  s.red = d + 2048;
  s.rinv = reinterpret_cast<float*>(base + ub);
  s.item = reinterpret_cast<int*>(base + ub + kRinvBytes);
  s.T = reinterpret_cast<float*>(base + ub + kRinvBytes + kItemBytes);
  return s;
}

// Factorises the system held in R (see the file comment).  Returns 0 or the
// 1-based index of the first non-positive pivot; uniform over the block.
template <int KR>
__device__ int factor_registers(float (&R)[KR][KR], int n, const Lds& s) {
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4, w = tid >> 6, l = tid & 63;
  const int N16 = round16(n);
  const int nk = N16 >> 4;
  float* T = s.T;
  float* PT = s.PT;
  for (int K = 0; K < nk; ++K) {
    const int k0 = 16 * K;
    // (a) panel rows to T
#pragma unroll
    for (int a = 0; a < KR; ++a) {
      if (a == K) {
#pragma unroll
        for (int b = a; b < KR; ++b) {
          if (b < nk) {
            const int i = ty + 16 * a, j = tx + 16 * b;
            if (j >= i) T[pku(N16, i, j)] = R[a][b];
          }
        }
      }
    }
    __syncthreads();

    // (b1) wave 0 factors the diagonal block in registers, lane q holding
    // column q (broadcasts by readlane), and stores U_KK and 1/diag to LDS.
    if (w == 0) {
      const int q = l & 15;
      float c[16];
#pragma unroll
      for (int p = 0; p < 16; ++p) c[p] = (p <= q) ? T[pku(N16, k0 + p, k0 + q)] : 0.0f;
      int dfail = 0;
#pragma unroll
      for (int st = 0; st < 16; ++st) {
        const float dss = bcast(c[st], st);
        if (!(dss > 0.0f) && dfail == 0) dfail = st + 1;
        const float d = sqrtf(dss);
        c[st] = (q == st) ? d : (q > st ? c[st] / d : c[st]);
#pragma unroll
        for (int p = st + 1; p < 16; ++p) {
          const float usp = bcast(c[st], p);
          c[p] = (q >= p) ? c[p] - usp * c[st] : c[p];
        }
      }
      if (l < 16) {
#pragma unroll
        for (int p = 0; p < 16; ++p)
          if (p <= q) T[pku(N16, k0 + p, k0 + q)] = c[p];
        s.rinv[k0 + q] = 1.0f / c[q];
      }
      if (l == 0) s.item[1] = dfail ? k0 + dfail : 0;  // padding rows never fail
    }
    __syncthreads();
    if (s.item[1]) return s.item[1];

    // (b2) block row for this wave's 64 columns: U_KK^T u = a, with U_KK
    // read by broadcast from LDS.  Lanes without a column of their own work
    // on the last column, which always exists, and do not store.
    {
      const int j = 64 * w + l;
      const bool jvalid = j >= k0 + 16 && j < N16;
      const int jc = jvalid ? j : N16 - 1;
      float u[16];
#pragma unroll
      for (int p = 0; p < 16; ++p) u[p] = T[pku(N16, k0 + p, jc)];
      // column sweep: once u[st] is final its multiples leave every later
      // entry at once (15 independent FMAs), and the pivot division is a
      // multiplication by the stored reciprocal
#pragma unroll
      for (int st = 0; st < 16; ++st) {
        u[st] *= s.rinv[k0 + st];
#pragma unroll
        for (int p = st + 1; p < 16; ++p) u[p] -= T[pku(N16, k0 + st, k0 + p)] * u[st];
      }
      if (jvalid) {
#pragma unroll
        for (int p = 0; p < 16; ++p) T[pku(N16, k0 + p, j)] = u[p];
        float2* dst = reinterpret_cast<float2*>(PT + j * kPTStride);
#pragma unroll
        for (int p = 0; p < 16; p += 2) dst[p >> 1] = make_float2(u[p], u[p + 1]);
      }
    }
    __syncthreads();

    // (c) rank-16 update of the register slots below the panel.  The p loop
    // stays rolled: unrolled, the per-panel code would outgrow the
    // instruction cache.  K is a runtime value, so a guard per register
    // tile is a branch per tile.  With 16 x 16 tiles (136 of them) the
    // branches cost more than the arithmetic: the panel values of rows and
    // columns outside (K, nk) are selected to zero instead (PT rows hold
    // stale values there, so a select, not a multiplication) and only whole
    // groups of four tile rows are skipped; tiles in a group's inactive
    // rows or columns subtract an exact zero from entries that are never
    // read again (already copied to T, or padding).  The 8 x 8 tiles keep
    // the per-tile guards, which measure faster there.
#pragma unroll 1
    for (int pc = 0; pc < 16; pc += 2) {
      const float2 zero = make_float2(0.0f, 0.0f);
      float2 uj[KR];
      if constexpr (KR > 8) {
#pragma unroll
        for (int b = 0; b < KR; ++b) {
          const float2 t = *reinterpret_cast<const float2*>(PT + (tx + 16 * b) * kPTStride + pc);
          uj[b] = (b > K && b < nk) ? t : zero;
        }
#pragma unroll
        for (int ag = 0; ag < KR; ag += 4) {
          if (ag + 3 > K && ag < nk) {
#pragma unroll
            for (int a = ag; a < ag + 4 && a < KR; ++a) {
              const float2 t = *reinterpret_cast<const float2*>(PT + (ty + 16 * a) * kPTStride + pc);
              const float2 ui = (a > K && a < nk) ? t : zero;
#pragma unroll
              for (int b = a; b < KR; ++b) {
                R[a][b] -= ui.x * uj[b].x;
                R[a][b] -= ui.y * uj[b].y;
              }
            }
          }
        }
      } else {
#pragma unroll
        for (int b = 0; b < KR; ++b) {
          uj[b] = zero;
          if (b > K && b < nk) uj[b] = *reinterpret_cast<const float2*>(PT + (tx + 16 * b) * kPTStride + pc);
        }
#pragma unroll
        for (int a = 0; a < KR; ++a) {
          if (a > K && a < nk) {
            const float2 ui = *reinterpret_cast<const float2*>(PT + (ty + 16 * a) * kPTStride + pc);
#pragma unroll
            for (int b = a; b < KR; ++b) {
              if (b < nk) {
                R[a][b] -= ui.x * uj[b].x;
                R[a][b] -= ui.y * uj[b].y;
              }
            }
          }
        }
      }
    }
  }
  __syncthreads();
  return 0;
}

// U^T U x = v for v held by one wave (lane l: v[l + 64c]); uses k < n only.
// The solves are a chain of n dependent steps (broadcast the finished entry,
// scale by the reciprocal pivot, update the rest).  The LDS operands of the
// next group of steps (a row of U forward, a column backward, and the
// reciprocal pivots) are loaded while the current group runs, so the LDS
// latency stays off the chain.  (Groups of kG steps.)
// The loads are unconditional (an out-of-range entry reads element 0 and is
// selected away): with no branch around them the compiler knows how many
// LDS reads are in flight and waits only for the group it is about to use,
// so the next group's reads stay in flight during this group's steps.
template <int NC>
__device__ __forceinline__ void solve_load_row(const Lds& s, int N16, int n, int l, int k, bool live, float (&t)[NC],
                                               float& r) {
  const float* T = s.T;
  const float rv = s.rinv[live ? k : 0];
  r = live ? rv : 0.0f;
  const int rk = pku(N16, k, k);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int jj = l + 64 * c;
    const bool in = live && jj > k && jj < n;
    const float v = T[in ? rk + (jj - k) : 0];
    t[c] = in ? v : 0.0f;
  }
}

template <int NC>
__device__ __forceinline__ void solve_load_col(const Lds& s, int N16, int l, int k, bool live, float (&t)[NC],
                                               float& r) {
  const float* T = s.T;
  const float rv = s.rinv[live ? k : 0];
  r = live ? rv : 0.0f;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int i = l + 64 * c;
    const bool in = live && i < k;
    const float v = T[in ? pku(N16, i, k) : 0];
    t[c] = in ? v : 0.0f;
  }
}

constexpr int kG = 4;  // steps per prefetch group (4 fits both tile classes without scratch)
#ifndef LXA_RES_ROWS
#define LXA_RES_ROWS 4
#endif
constexpr int kRB = LXA_RES_ROWS;  // rows per residual batch (per wave)

template <int NC>
__device__ void wave_solve(int n, const Lds& s, float (&v)[NC]) {
  const int l = threadIdx.x & 63;
  const int N16 = round16(n);
#pragma unroll
  for (int c0 = 0; c0 < NC; ++c0) {
    const int kend = min(64, n - 64 * c0);
    if (kend <= 0) continue;
    float ta[kG][NC], ra[kG];
#pragma unroll
    for (int u = 0; u < kG; ++u) solve_load_row(s, N16, n, l, 64 * c0 + u, u < kend, ta[u], ra[u]);
    for (int g = 0; g < kend; g += kG) {
      float tb[kG][NC], rb[kG];
#pragma unroll
      for (int u = 0; u < kG; ++u) solve_load_row(s, N16, n, l, 64 * c0 + g + kG + u, g + kG + u < kend, tb[u], rb[u]);
#pragma unroll
      for (int u = 0; u < kG; ++u) {
        const int kk = g + u;
        if (kk < kend) {
          const float yk = bcast(v[c0], kk) * ra[u];
          if (l == kk) v[c0] = yk;
#pragma unroll
          for (int c = c0; c < NC; ++c) v[c] -= ta[u][c] * yk;
        }
      }
#pragma unroll
      for (int u = 0; u < kG; ++u) {
        ra[u] = rb[u];
#pragma unroll
        for (int c = 0; c < NC; ++c) ta[u][c] = tb[u][c];
      }
    }
  }
#pragma unroll
  for (int c0 = NC - 1; c0 >= 0; --c0) {
    const int kend = min(64, n - 64 * c0);
    if (kend <= 0) continue;
    float ta[kG][NC], ra[kG];
#pragma unroll
    for (int u = 0; u < kG; ++u) solve_load_col(s, N16, l, 64 * c0 + kend - 1 - u, kend - 1 - u >= 0, ta[u], ra[u]);
    for (int g = kend - 1; g >= 0; g -= kG) {
      float tb[kG][NC], rb[kG];
#pragma unroll
      for (int u = 0; u < kG; ++u) solve_load_col(s, N16, l, 64 * c0 + g - kG - u, g - kG - u >= 0, tb[u], rb[u]);
#pragma unroll
      for (int u = 0; u < kG; ++u) {
        const int kk = g - u;
        if (kk >= 0) {
          const float xk = bcast(v[c0], kk) * ra[u];
          if (l == kk) v[c0] = xk;
#pragma unroll
          for (int c = 0; c <= c0; ++c) v[c] -= ta[u][c] * xk;
        }
      }
#pragma unroll
      for (int u = 0; u < kG; ++u) {
        ra[u] = rb[u];
#pragma unroll
        for (int c = 0; c < NC; ++c) ta[u][c] = tb[u][c];
      }
    }
  }
}

// Four block-wide maxima in one exchange (red holds 4 doubles per wave).
__device__ void block_max4(double (&v)[4], double* red) {
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = wave_max(v[i]);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) red[4 * w + i] = v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = fmax(fmax(red[i], red[4 + i]), fmax(red[8 + i], red[12 + i]));
  __syncthreads();
}


// r = b - S x, S the symmetric matrix of A's upper triangle; optionally the
// row sums of |S|.  Every upper-triangle element is read once from global
// memory: row j gives sum_{i>=j} A_ji x_i to r_j (wave reduction) and
// A_ji x_j to r_i for i > j (lane accumulators).
template <int NC>
__device__ void residual_sym_upper(int n, const float* __restrict__ A, const float* __restrict__ b, const Lds& s,
                                   bool want_abs) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  double acc[NC], aab[NC], xi[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    acc[c] = aab[c] = 0.0;
    xi[c] = (l + 64 * c < n) ? s.x[l + 64 * c] : 0.0;
  }
  // Rows j0 + 4u (u < kRB) per batch; the next batch's loads are issued before
  // this batch is reduced (the row reductions are independent, so their
  // shuffles interleave).  The loads are unconditional, so the compiler
  // waits only for the batch it uses and the next batch's loads stay in
  // flight: a lane left of the diagonal reads the row's diagonal element
  // (a cache line the row loads anyway), a row past n reads row n - 1, and
  // both are selected away.
  auto load_rows = [&](int j0, float (&av)[kRB][NC]) {
#pragma unroll
    for (int u = 0; u < kRB; ++u) {
      const int j = j0 + 4 * u;
      const int jr = j < n ? j : n - 1;
      const float* Aj = A + static_cast<int64_t>(jr) * n;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int i = l + 64 * c;
        const int ic = i < jr ? jr : (i < n ? i : n - 1);
        const float v = Aj[ic];
        av[u][c] = (j < n && i >= j && i < n) ? v : 0.0f;
      }
    }
  };
  float av[kRB][NC];
  load_rows(w, av);
  for (int j0 = w; j0 < n; j0 += 4 * kRB) {
    float nv[kRB][NC];
    load_rows(j0 + 4 * kRB, nv);  // rows past n load nothing (zero)
    double red[2 * kRB];
#pragma unroll
    for (int u = 0; u < kRB; ++u) {
      const int j = j0 + 4 * u;
      const double xj = j < n ? s.x[j] : 0.0;
      double dot = 0.0, rab = 0.0;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int i = l + 64 * c;
        const double aji = static_cast<double>(av[u][c]);
        dot = fma(aji, xi[c], dot);
        rab += fabs(aji);
        if (i > j) {
          acc[c] = fma(aji, xj, acc[c]);
          aab[c] += fabs(aji);
        }
      }
      red[u] = dot;
      red[kRB + u] = rab;
    }
#pragma unroll
    for (int u = 0; u < kRB; ++u) {
      const int j = j0 + 4 * u;
      const double dot = wave_sum(red[u]);
      const double rab = want_abs ? wave_sum(red[kRB + u]) : 0.0;
      if (l == 0 && j < n) {
        s.rowpart[j] = dot;
        if (want_abs) s.rowabs[j] = rab;
      }
    }
#pragma unroll
    for (int u = 0; u < kRB; ++u)
#pragma unroll
      for (int c = 0; c < NC; ++c) av[u][c] = nv[u][c];
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int i = l + 64 * c;
    if (i < n) s.acc[w * 256 + i] = acc[c];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kT) {
    const double tx = s.rowpart[i] + ((s.acc[i] + s.acc[256 + i]) + (s.acc[512 + i] + s.acc[768 + i]));
    s.r[i] = static_cast<double>(b[i]) - tx;
  }
  if (want_abs) {
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int i = l + 64 * c;
      if (i < n) s.acc[w * 256 + i] = aab[c];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kT)
      s.rowabs[i] += (s.acc[i] + s.acc[256 + i]) + (s.acc[512 + i] + s.acc[768 + i]);
  }
  __syncthreads();
}

// KR: register tile extent; a workgroup holds orders up to 16 KR.  Each
// launch takes the systems with n_lo < n <= n_hi from the shared order and
// skips the rest, so two launches with different tiles split one batch.
// report_invalid: this launch marks orders outside [1, max_n].
template <int KR, bool kRefine, typename TX>
__global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(KR == 8 ? LXA_SMALL_WAVES : (KR < 16 ? 2 : 1))))
void chol_uneven_kernel(int64_t count, int max_n, int n_lo, int n_hi,
                                                         int report_invalid, const int32_t* __restrict__ ns,
                                                         const int64_t* __restrict__ a_off,
                                                         const float* __restrict__ a,
                                                         const int64_t* __restrict__ b_off,
                                                         const float* __restrict__ b, TX* __restrict__ x,
                                                         const int32_t* __restrict__ order, int persistent,
                                                         unsigned int* __restrict__ queue, int max_refine, double tol,
                                                         int32_t* __restrict__ info, int32_t* __restrict__ iters,
                                                         double* __restrict__ eta) {
  extern __shared__ __align__(16) unsigned char lxa_smem[];
  const Lds s = carve(lxa_smem, n_hi < 16 * KR ? n_hi : 16 * KR);  // launch_class's lds_n
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int tx = tid & 15, ty = tid >> 4;
  int64_t next = blockIdx.x;
  for (;;) {
    if (persistent) {
      if (tid == 0) s.item[0] = static_cast<int>(atomicAdd(queue, 1u));
      __syncthreads();
      next = s.item[0];
      __syncthreads();
    }
    if (next >= count) break;
    const int64_t sys = order[next];
    if (!persistent) next = count;
    const int n = ns[sys];
    if (n < 1 || n > max_n) {
      if (tid == 0 && report_invalid) {
        info[sys] = -1;
        if (iters) iters[sys] = 0;
        if (eta) eta[sys] = -1.0;
      }
      __syncthreads();
      continue;
    }
    if (n <= n_lo || n > n_hi) continue;  // the other tile class's system (uniform)
    const float* A = a + a_off[sys];
    const float* B = b + b_off[sys];
    TX* X = x + b_off[sys];
    const int nk = round16(n) >> 4;

    constexpr int NC = (KR + 3) / 4;  // 64-entry chunks of a vector (orders <= 64 NC)
    float R[KR][KR];
#pragma unroll
    for (int aa = 0; aa < KR; ++aa) {
#pragma unroll
      for (int bb = aa; bb < KR; ++bb) {
        R[aa][bb] = 0.0f;
        if (aa < nk && bb < nk) {
          const int i = ty + 16 * aa, j = tx + 16 * bb;
          if (i < n && j < n) {
            if (j >= i) R[aa][bb] = A[static_cast<int64_t>(i) * n + j];
          } else if (i == j) {
            R[aa][bb] = 1.0f;
          }
        }
      }
    }

    const int fail = factor_registers<KR>(R, n, s);
    if (fail) {
      __syncthreads();
      for (int i = tid; i < n; i += kT) X[i] = static_cast<TX>(quiet_nan<double>());
      if (tid == 0) {
        info[sys] = fail;
        if (iters) iters[sys] = 0;
        if (eta) eta[sys] = -1.0;
      }
      __syncthreads();
      continue;
    }
    // Pass -1 solves with b; later passes form the FP64 residual, test the
    // backward error and solve for the correction (one solve call site).
    int steps = 0;
    double last_eta = -1.0;
    double bnorm = 0.0, anorm = 0.0;
    const int passes = kRefine ? max_refine : 0;
    for (int it = -1; it < passes; ++it) {
      if constexpr (kRefine) {
        if (it >= 0) {
          residual_sym_upper<NC>(n, A, B, s, it == 0);
          // |r|_inf, |x|_inf and (first pass) |A|_inf, |b|_inf in one exchange
          double m4[4] = {0.0, 0.0, 0.0, 0.0};
          for (int i = tid; i < n; i += kT) {
            m4[0] = fmax(m4[0], fabs(s.r[i]));
            m4[1] = fmax(m4[1], fabs(s.x[i]));
            if (it == 0) {
              m4[2] = fmax(m4[2], s.rowabs[i]);
              m4[3] = fmax(m4[3], fabs(static_cast<double>(B[i])));
            }
          }
          block_max4(m4, s.red);
          const double rmax = m4[0], xmax = m4[1];
          if (it == 0) {
            anorm = m4[2];
            bnorm = m4[3];
          }
          const double denom = anorm * xmax + bnorm;
          last_eta = denom > 0.0 ? rmax / denom : (rmax > 0.0 ? INFINITY : 0.0);
          ++steps;
          if (!(last_eta > tol)) break;  // also stops on NaN
        }
      }
      if (w == 0) {
        float v[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const int i = l + 64 * c;
          v[c] = i < n ? (it < 0 ? B[i] : static_cast<float>(s.r[i])) : 0.0f;
        }
        wave_solve<NC>(n, s, v);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const int i = l + 64 * c;
          if (i < n) s.x[i] = (it < 0 ? 0.0 : s.x[i]) + static_cast<double>(v[c]);
        }
      }
      __syncthreads();
    }
    for (int i = tid; i < n; i += kT) X[i] = static_cast<TX>(s.x[i]);
    if (tid == 0) {
      info[sys] = 0;
      if (iters) iters[sys] = steps;
      if (eta) eta[sys] = last_eta;
    }
    __syncthreads();
  }
}

// Factor only, for uniform float32 batches (lxa_chol_factor, orders 65..256):
// the register-tiled factorisation above on A's upper triangle (negated when
// `negate`), then U (row-major, lower triangle zero) overwrites A.  A failed
// factorisation leaves the whole matrix NaN and info the 1-based index of the
// first non-positive pivot (chol.hip's convention for every tier).
template <int KR>
__global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(KR == 8 ? LXA_SMALL_WAVES : 1)))
void chol_factor_regs_kernel(int n, float* __restrict__ a, int64_t a_stride, int negate,
                             int32_t* __restrict__ info) {
  extern __shared__ __align__(16) unsigned char lxa_smem[];
  const Lds s = carve(lxa_smem, n);  // launch_factor_regs sizes LDS for n
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  float* A = a + static_cast<int64_t>(blockIdx.x) * a_stride;
  const int N16 = round16(n);
  const int nk = N16 >> 4;
  float R[KR][KR];
#pragma unroll
  for (int aa = 0; aa < KR; ++aa) {
#pragma unroll
    for (int bb = aa; bb < KR; ++bb) {
      R[aa][bb] = 0.0f;
      if (aa < nk && bb < nk) {
        const int i = ty + 16 * aa, j = tx + 16 * bb;
        if (i < n && j < n) {
          if (j >= i) {
            const float v = A[static_cast<int64_t>(i) * n + j];
            R[aa][bb] = negate ? -v : v;
          }
        } else if (i == j) {
          R[aa][bb] = 1.0f;
        }
      }
    }
  }
  const int fail = factor_registers<KR>(R, n, s);
  // rows by wave, columns by lane (no index division)
  const int w = tid >> 6, l = tid & 63;
  for (int i = w; i < n; i += kT / 64) {
    float* Ai = A + static_cast<int64_t>(i) * n;
    const int ri = pku(N16, i, i) - i;
    for (int j = l; j < n; j += 64) Ai[j] = fail ? quiet_nan<float>() : (j < i ? 0.0f : s.T[ri + j]);
  }
  if (tid == 0) info[blockIdx.x] = fail;
}

__global__ void zero_u32_kernel(unsigned int* p) { *p = 0u; }

template <typename K>
int allow_lds(K kernel, size_t bytes) {
  if (bytes <= 64 * 1024) return LXA_OK;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
  return e == hipSuccess ? LXA_OK : hip_error(e, "hipFuncSetAttribute(MaxDynamicSharedMemorySize)");
}

constexpr int64_t kMaxGrid = 2147483647;

// One launch of the KR tile class over `order`; persistent launches are sized
// by the occupancy the kernel's registers and LDS allow.
template <int KR, bool kRefine, typename TX>
int launch_class(hipStream_t s, int64_t count, int max_n, int n_lo, int n_hi, int report, const int32_t* n,
                 const int64_t* a_offset, const float* a, const int64_t* b_offset, const float* b, void* x,
                 const int32_t* order, bool persistent, unsigned int* queue, int max_refine, double tol, int32_t* info,
                 int32_t* iters, double* eta) {
  auto kernel = chol_uneven_kernel<KR, kRefine, TX>;
  const int lds_n = n_hi < 16 * KR ? n_hi : 16 * KR;
  const size_t shmem = uneven_lds_bytes(lds_n);
  if (shmem > 160 * 1024) return set_error(LXA_EINVAL, "lxa_chol_uneven_solve: LDS need %zu exceeds 160 KiB", shmem);
  if (int rc = allow_lds(kernel, shmem)) return rc;
  int64_t grid = count;
  if (persistent) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, queue);
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kernel), kT,
                                                                shmem);
    if (e != hipSuccess) return hip_error(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (per_cu < 1) per_cu = 1;
    grid = static_cast<int64_t>(num_cus()) * per_cu;
    if (grid > count) grid = count;
  }
  hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(grid)), dim3(kT), shmem, s, count, max_n, n_lo, n_hi, report,
                     n, a_offset, a, b_offset, b, static_cast<TX*>(x), order, persistent ? 1 : 0, queue, max_refine,
                     tol, info, iters, eta);
  LXA_LAUNCH_CHECK("lxa_chol_uneven_solve launch");
  return LXA_OK;
}

}  // namespace

template <int KR>
static int launch_factor_regs(hipStream_t s, int64_t batch, int n, float* a, int64_t stride, int negate,
                              int32_t* info) {
  const size_t shmem = uneven_lds_bytes(n);
  if (int rc = allow_lds(chol_factor_regs_kernel<KR>, shmem)) return rc;
  hipLaunchKernelGGL((chol_factor_regs_kernel<KR>), dim3(static_cast<unsigned>(batch)), dim3(kT), shmem, s, n, a,
                     stride, negate, info);
  LXA_LAUNCH_CHECK("lxa_chol_factor (register tiles) launch");
  return LXA_OK;
}

int chol_factor_regs_f32(hipStream_t s, int64_t batch, int n, float* a, int64_t stride, int negate,
                         int32_t* info) {
  if (n < 1 || n > kMaxN) return set_error(LXA_EINVAL, "chol_factor_regs_f32: n=%d outside [1, %d]", n, kMaxN);
  if (batch > kMaxGrid) return set_error(LXA_EINVAL, "chol_factor_regs_f32: batch too large");
  return n <= 128 ? launch_factor_regs<8>(s, batch, n, a, stride, negate, info)
                  : launch_factor_regs<16>(s, batch, n, a, stride, negate, info);
}

}  // namespace lxa

namespace lxa {

namespace {

// Work-queue counters of the persistent launches: one per register-tile class.
constexpr int kQueueCounters = 3;
constexpr size_t kUnevenWorkspaceBytes = 16;  // kQueueCounters 4-byte counters, rounded up
static_assert(kQueueCounters * sizeof(unsigned int) <= kUnevenWorkspaceBytes, "workspace holds the queue counters");

// Device scratch of lxa_chol_uneven_solve_noalloc: the queue counters live in
// this file-scope array instead of a caller's workspace.  Each GPU the
// library runs on holds its own copy until the process ends.  Host code uses
// the address hipGetSymbolAddress returns for the current device, never the
// array's name (a host-side stand-in, not a device address).
constexpr size_t kNoAllocScratchBytes = 256;
static_assert(kUnevenWorkspaceBytes <= kNoAllocScratchBytes, "queue counters exceed the no-alloc scratch");
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept; 0 on success.
int noalloc_scratch(unsigned int** out) {
  static std::mutex mu;
  static void* ptr[kMaxDevices] = {};
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_error(e, "lxa_chol_uneven_solve_noalloc: hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(LXA_EUNSUPPORTED, "lxa_chol_uneven_solve_noalloc: device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_error(e, "lxa_chol_uneven_solve_noalloc: hipGetSymbolAddress(scratch)");
    if (!p) return set_error(LXA_EHIP, "lxa_chol_uneven_solve_noalloc: null scratch address");
    ptr[dev] = p;
  }
  *out = static_cast<unsigned int*>(ptr[dev]);
  return LXA_OK;
}

int uneven_check(int64_t count, int max_n, const int32_t* n, const int64_t* a_offset, const float* a,
                 const int64_t* b_offset, const float* b, void* x, int x_dtype, const int32_t* order,
                 int max_refine, int32_t* info) {
  LXA_REQUIRE(count >= 0 && count <= kMaxGrid, "lxa_chol_uneven_solve: bad count");
  LXA_REQUIRE(max_n >= 1 && max_n <= kMaxN, "lxa_chol_uneven_solve: max_n must be in [1, %d]", kMaxN);
  LXA_REQUIRE(n && a_offset && a && b_offset && b && x && order && info, "lxa_chol_uneven_solve: null pointer");
  LXA_REQUIRE(x_dtype == LXA_F32 || x_dtype == LXA_F64, "lxa_chol_uneven_solve: bad x_dtype");
  LXA_REQUIRE(max_refine >= 0 && max_refine <= 64, "lxa_chol_uneven_solve: max_refine in [0, 64]");
  LXA_REQUIRE(max_refine == 0 || x_dtype == LXA_F64, "lxa_chol_uneven_solve: refinement needs x_dtype F64");
  return LXA_OK;
}

// The launches of one uneven solve; `queue` holds kQueueCounters counters
// (persistent launches only; each launch clears its own with a fill kernel).
int uneven_launch(hipStream_t s, int64_t count, int max_n, const int32_t* n, const int64_t* a_offset,
                  const float* a, const int64_t* b_offset, const float* b, void* x, int x_dtype,
                  const int32_t* order, int flags, int max_refine, double tol, int32_t* info, int32_t* iters,
                  double* eta, unsigned int* queue) {
  const bool persistent = (flags & LXA_UNEVEN_PERSISTENT) != 0;
  const bool refine = max_refine > 0;
  // Three register-tile classes: orders <= 128 run with 8 x 8 tiles (three
  // systems per compute unit), 129..176 with 11 x 11 (two per compute unit:
  // their LDS fits twice in 160 KB), larger ones with 16 x 16 (one).  A
  // persistent solve is one launch per class over the same order (the large
  // classes first, each with its own queue counter); one workgroup per
  // system uses the class that covers max_n.
  const int split = 128, split2 = 176;
  int rc = LXA_OK;
#define LXA_UNEVEN_CLASS(KR, LO, HI, REPORT, Q)                                                                  \
  do {                                                                                                           \
    if (refine)                                                                                                  \
      rc = launch_class<KR, true, double>(s, count, max_n, LO, HI, REPORT, n, a_offset, a, b_offset, b, x,        \
                                          order, persistent, Q, max_refine, tol, info, iters, eta);             \
    else if (x_dtype == LXA_F64)                                                                                 \
      rc = launch_class<KR, false, double>(s, count, max_n, LO, HI, REPORT, n, a_offset, a, b_offset, b, x,       \
                                           order, persistent, Q, 0, tol, info, iters, eta);                     \
    else                                                                                                         \
      rc = launch_class<KR, false, float>(s, count, max_n, LO, HI, REPORT, n, a_offset, a, b_offset, b, x,        \
                                          order, persistent, Q, 0, tol, info, iters, eta);                      \
  } while (0)
  if (!persistent) {
    if (max_n > split2) LXA_UNEVEN_CLASS(16, 0, max_n, 1, queue);
    else if (max_n > split) LXA_UNEVEN_CLASS(11, 0, max_n, 1, queue);
    else LXA_UNEVEN_CLASS(8, 0, max_n, 1, queue);
  } else if (max_n > split2) {
    LXA_UNEVEN_CLASS(16, split2, max_n, 1, queue);
    if (rc) return rc;
    LXA_UNEVEN_CLASS(11, split, split2, 0, queue + 1);
    if (rc) return rc;
    LXA_UNEVEN_CLASS(8, 0, split, 0, queue + 2);
  } else if (max_n > split) {
    LXA_UNEVEN_CLASS(11, split, max_n, 1, queue);
    if (rc) return rc;
    LXA_UNEVEN_CLASS(8, 0, split, 0, queue + 1);
  } else {
    LXA_UNEVEN_CLASS(8, 0, max_n, 1, queue);
  }
#undef LXA_UNEVEN_CLASS
  if (rc) return rc;
  LXA_LAUNCH_CHECK("lxa_chol_uneven_solve launch");
  return LXA_OK;
}

}  // namespace
}  // namespace lxa

extern "C" int lxa_chol_uneven_workspace_size(int64_t count, size_t* bytes) {
  LXA_REQUIRE(count >= 0 && bytes, "lxa_chol_uneven_workspace_size: bad argument");
  *bytes = lxa::kUnevenWorkspaceBytes;  // work-queue counters
  return LXA_OK;
}

extern "C" int lxa_chol_uneven_solve(void* stream, int64_t count, int max_n, const int32_t* n,
                                     const int64_t* a_offset, const float* a, const int64_t* b_offset,
                                     const float* b, void* x, int x_dtype, const int32_t* order, int flags,
                                     int max_refine, double tol, int32_t* info, int32_t* iters, double* eta,
                                     void* workspace, size_t workspace_bytes) {
  if (int rc = lxa::uneven_check(count, max_n, n, a_offset, a, b_offset, b, x, x_dtype, order, max_refine, info))
    return rc;
  const bool persistent = (flags & LXA_UNEVEN_PERSISTENT) != 0;
  LXA_REQUIRE(!persistent || (workspace && workspace_bytes >= lxa::kUnevenWorkspaceBytes),
              "lxa_chol_uneven_solve: the persistent launch needs a workspace of %zu bytes (got %zu)",
              lxa::kUnevenWorkspaceBytes, workspace ? workspace_bytes : static_cast<size_t>(0));
  if (count == 0) return LXA_OK;
  return lxa::uneven_launch(static_cast<hipStream_t>(stream), count, max_n, n, a_offset, a, b_offset, b, x, x_dtype,
                            order, flags, max_refine, tol, info, iters, eta, static_cast<unsigned int*>(workspace));
}

extern "C" int lxa_chol_uneven_solve_noalloc(void* stream, int64_t count, int max_n, const int32_t* n,
                                             const int64_t* a_offset, const float* a, const int64_t* b_offset,
                                             const float* b, void* x, int x_dtype, const int32_t* order, int flags,
                                             int max_refine, double tol, int32_t* info, int32_t* iters,
                                             double* eta) {
  if (int rc = lxa::uneven_check(count, max_n, n, a_offset, a, b_offset, b, x, x_dtype, order, max_refine, info))
    return rc;
  size_t need = 0;
  if (int rc = lxa_chol_uneven_workspace_size(count, &need)) return rc;
  LXA_REQUIRE(need <= lxa::kNoAllocScratchBytes,
              "lxa_chol_uneven_solve_noalloc: needs %zu bytes of scratch, more than its %zu", need,
              lxa::kNoAllocScratchBytes);
  unsigned int* queue = nullptr;
  if (int rc = lxa::noalloc_scratch(&queue)) return rc;
  if (count == 0) return LXA_OK;
  return lxa::uneven_launch(static_cast<hipStream_t>(stream), count, max_n, n, a_offset, a, b_offset, b, x, x_dtype,
                            order, flags, max_refine, tol, info, iters, eta, queue);
}
