// Householder QR over uneven batches (geqrf / ormqr) and triangular solves
// for gfx950.
//
// Factorisation (lxa_qr_factor) is blocked, right-looking, compact WY:
// each panel step factors nb columns of every system and then updates the
// systems' trailing columns:
//
//   panel   one 512-thread workgroup per system.  The panel (rows k..m-1,
//           nb columns) lives in registers, up to eight rows x nb columns
//           per thread, i.e. 4096 rows (a global-memory variant covers
//           taller panels).  Per column one block reduction carries the
//           norm below the diagonal (larfg) and the raw products of that
//           column with the other panel columns, from which v^T A (larf)
//           and v_a^T v_j (for the T factor) follow.
//           The workgroup ends by forming T (larft) and writing tau, T and
//           the panel back.
//   trail   C <- (I - V T V^T)^T C = C - V (T^T (V^T C)) on the columns
//           right of the panel: (1) partial V^T C per 256-row chunk and
//           64-column block, (2) the chunk sum and T^T, (3) C -= V W.  Each
//           pass is a coalesced sweep over C (read, read, read+write).
//
// Block sums are deterministic: a wave reduce-scatter (log2 K halving
// exchanges, then a full reduction of one value per lane), then a fixed
// order over waves.
//
// Reference: lineax/_solver/qr.py:55-92 calls jnp.linalg.qr(mode="raw")
// (geqrf), jax.lax.linalg.ormqr and jax.scipy.linalg.solve_triangular; the
// reflector convention here is LAPACK's (larfg: beta = -sign(alpha) ||x||,
// tau = (beta - alpha) / beta, v = x / (alpha - beta), tau = 0 when x = 0).

#include <type_traits>

#include "lxa_common.h"

namespace lxa {
namespace {

template <typename T>
struct QrCfg;
template <>
struct QrCfg<float> {
  static constexpr int NB = 16;
};
template <>
struct QrCfg<double> {
  static constexpr int NB = 8;
};

constexpr int kPanelThreads = 512;   // 2 waves per SIMD: 256 VGPRs each
constexpr int kPanelWaves = kPanelThreads / kWave;
constexpr int kPanelMaxRpt = 8;      // register panel: rows <= 4096
constexpr int kBigThreads = 1024;    // apply / triangular solve workgroups
constexpr int kBigWaves = kBigThreads / kWave;
constexpr int kTrailCols = 64;
constexpr int kTrailRows = 256;
constexpr int kTrailThreads = 256;

// Block sum of K values per thread over a workgroup of WAVES waves.
// stage: LDS [WAVES][K]; out: LDS [K].  Every thread gets the sums.
template <int K, int WAVES, typename R>
__device__ __forceinline__ void block_sum(R (&v)[K], R* stage, R* out, int tid) {
  const int lane = tid & (kWave - 1);
  const int wv = tid >> 6;
  const R x = wave_reduce_scatter<K>(v, lane);
  constexpr int kShift = 6 - Log2<K>::value;
  if ((lane & ((1 << kShift) - 1)) == 0) stage[wv * K + (lane >> kShift)] = x;
  __syncthreads();
  if (tid < K) {
    R s = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) s += stage[w * K + tid];
    out[tid] = s;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < K; ++i) v[i] = out[i];
}

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Two block sums (K1 values of type R1, K2 of type R2 per thread) behind one
// pair of barriers: the second array's cross-wave pass runs on another wave.
template <int K1, int K2, int WAVES, typename R1, typename R2>
__device__ __forceinline__ void block_sum2(R1 (&u)[K1], R1* stage1, R1* out1, R2 (&v)[K2], R2* stage2, R2* out2,
                                           int tid) {
  static_assert(WAVES >= 2, "the two cross-wave passes run on waves 0 and 1");
  const int lane = tid & (kWave - 1);
  const int wv = tid >> 6;
  const R1 x1 = wave_reduce_scatter<K1>(u, lane);
  const R2 x2 = wave_reduce_scatter<K2>(v, lane);
  constexpr int kShift1 = 6 - Log2<K1>::value, kShift2 = 6 - Log2<K2>::value;
  if ((lane & ((1 << kShift1) - 1)) == 0) stage1[wv * K1 + (lane >> kShift1)] = x1;
  if ((lane & ((1 << kShift2) - 1)) == 0) stage2[wv * K2 + (lane >> kShift2)] = x2;
  __syncthreads();
  if (tid < K1) {
    R1 t = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) t += stage1[w * K1 + tid];
    out1[tid] = t;
  } else if (tid >= kWave && tid < kWave + K2) {
    R2 t = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) t += stage2[w * K2 + tid - kWave];
    out2[tid - kWave] = t;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < K1; ++i) u[i] = out1[i];
#pragma unroll
  for (int i = 0; i < K2; ++i) v[i] = out2[i];
}

// Householder vector entry j of panel-relative row (row - k): 1 on the
// diagonal, the stored value below it, 0 above.
template <typename T>
__device__ __forceinline__ T vpick(int row, int diag, T stored) {
  return row > diag ? stored : (row == diag ? T(1) : T(0));
}

// ---------------------------------------------------------------------------
// Panel factorisation.  RPT > 0: rows in registers (RPT per thread);
// RPT == 0: rows in global memory (any height, slower).
// ---------------------------------------------------------------------------
template <typename T, int RPT>
__global__ __launch_bounds__(kPanelThreads) void qr_panel_kernel(const lxa_qr_desc* __restrict__ descs, int k_all,
                                                                 const int32_t* __restrict__ kcol,
                                                                 T* __restrict__ a, T* __restrict__ tau,
                                                                 T* __restrict__ tblk) {
  constexpr int NB = QrCfg<T>::NB;
  constexpr int KW = NB;  // width of the per-column dot reduction (>= NB - 1)
  __shared__ double s_stage2[kPanelWaves];
  __shared__ double s_out2[1];
  __shared__ T s_stage[kPanelWaves * KW];
  __shared__ T s_out[KW];
  __shared__ T s_diag[2][NB];  // the diagonal row of the current column
  __shared__ T s_rows[RPT > 0 ? kPanelThreads : 1][NB + 1];  // staging of 512 panel rows
  __shared__ T s_g[NB][NB];    // s_g[a][j] = v_a^T v_j, a < j

  const lxa_qr_desc d = descs[blockIdx.x];
  const int m = d.m, n = d.n, lda = d.lda;
  const int kmax = min(m, n);
  const int k = kcol ? kcol[blockIdx.x] : k_all;
  if (k >= kmax) return;
  const int nb = min(NB, kmax - k);
  const int tid = threadIdx.x;
  T* A = a + d.a_offset;
  const int nr = RPT ? RPT : (m - k + kPanelThreads - 1) / kPanelThreads;

  T regs[RPT ? RPT : 1][NB];
  auto P = [&](int r, int c) -> T& {
    if constexpr (RPT > 0) {
      return regs[r][c];
    } else {
      return A[(int64_t)(k + tid + r * kPanelThreads) * lda + k + c];
    }
  };

  if constexpr (RPT > 0) {
    // rows k + r * 512 + [0, 512) pass through LDS: NB consecutive lanes read
    // one row's NB contiguous entries, then each thread takes its own row
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int base = k + r * kPanelThreads;
      if (base < m) {  // uniform
        // all NB loads in flight before any LDS write
        T v[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          const int e = tid + kPanelThreads * u, rr = e / NB, c = e % NB;
          const int row = base + rr;
          v[u] = (row < m && c < nb) ? A[(int64_t)row * lda + k + c] : T(0);
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          const int e = tid + kPanelThreads * u;
          s_rows[e / NB][e % NB] = v[u];
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NB; ++c) regs[r][c] = s_rows[tid][c];
        __syncthreads();
      } else {
#pragma unroll
        for (int c = 0; c < NB; ++c) regs[r][c] = T(0);
      }
    }
  }

  T taus[NB];
  // the column loop at compile time (every register index a constant)
  static_for<0, NB>([&](auto J_) __attribute__((always_inline)) {
    constexpr int j = decltype(J_)::value;
    taus[j] = T(0);
    if (j >= nb) return;  // uniform
    const int diag = k + j;
    // ---- one block reduction per column.  With x = column j below the
    // diagonal, it sums ||x||^2 and, for every other panel column c, the raw
    // product x^T A(:, c) over the same rows; the diagonal row goes to LDS
    // (thread j holds it).  v = [1; x / (alpha - beta)] then gives
    //   v^T A(:, c)  = A(diag, c) + scale * x^T A(:, c)    (c > j, larf)
    //   v_c^T v      = v_c(diag)  + scale * v_c^T x        (c < j, larft)
    // so the norm and the products share one pair of barriers.
    T* sd = s_diag[j & 1];  // double-buffered: readers of column j - 1 may lag
    double nv[1] = {0.0};
    T dots[KW];
#pragma unroll
    for (int c = 0; c < KW; ++c) dots[c] = T(0);
#pragma unroll
    for (int r = 0; r < nr; ++r) {
      const int row = k + tid + r * kPanelThreads;
      if (row < m && row > diag) {
        const T x = P(r, j);
        nv[0] += (double)x * (double)x;
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c != j && c < nb) dots[c] += x * P(r, c);
      }
    }
    if (tid == j) {
#pragma unroll
      for (int c = 0; c < NB; ++c) sd[c] = c < nb ? P(0, c) : T(0);
    }
    block_sum2<1, KW, kPanelWaves>(nv, s_stage2, s_out2, dots, s_stage, s_out, tid);
    const double ss = nv[0], alpha = (double)sd[j];
    T tj = T(0), beta = (T)alpha, scale = T(1);
    if (ss != 0.0) {
      const double bd = -copysign(sqrt(alpha * alpha + ss), alpha);
      beta = (T)bd;
      tj = (T)((bd - alpha) / bd);
      scale = (T)(1.0 / (alpha - bd));
    }
    taus[j] = tj;
    T acc[KW];
#pragma unroll
    for (int c = 0; c < KW; ++c) acc[c] = (c != j && c < nb) ? sd[c] + scale * dots[c] : T(0);
    if (tid == 0) {
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < j) s_g[c][j] = acc[c];
    }
    // ---- scale x into v, put beta on the diagonal, and apply
    // (I - tau v v^T) to the panel columns right of j
#pragma unroll
    for (int r = 0; r < nr; ++r) {
      const int row = k + tid + r * kPanelThreads;
      if (row < m && row >= diag) {
        T vj = T(1);
        if (row > diag) {
          vj = P(r, j) * scale;
          if (ss != 0.0) P(r, j) = vj;
        } else {
          P(r, j) = beta;
        }
        if (tj != T(0)) {
          const T tv = tj * vj;
#pragma unroll
          for (int c = 0; c < NB; ++c)
            if (c > j && c < nb) P(r, c) -= tv * acc[c];
        }
      }
    }
  });
  __syncthreads();  // s_g complete

  // ---- T (larft, forward, columnwise): lane a owns row a of T.
  if (tid < NB) {
    const int ra = tid;
    T trow[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      T v = T(0);
      if (i < nb) {
        if (ra == i) {
          v = taus[i];
        } else if (ra < i) {
          T s = T(0);
#pragma unroll
          for (int b = 0; b < NB; ++b)
            if (b >= ra && b < i) s += trow[b] * s_g[b][i];
          v = -taus[i] * s;
        }
      }
      trow[i] = v;
    }
    T* Tp = tblk + d.t_offset + (int64_t)(k / NB) * NB * NB + ra * NB;
#pragma unroll
    for (int i = 0; i < NB; ++i) Tp[i] = trow[i];
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if (i == ra && i < nb) tau[d.tau_offset + k + i] = taus[i];
  }

  if constexpr (RPT > 0) {
    // the panel back through LDS, in the same coalesced order as the load
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const int base = k + r * kPanelThreads;
      if (base < m) {  // uniform
        __syncthreads();  // the previous chunk's readers are done
#pragma unroll
        for (int c = 0; c < NB; ++c) s_rows[tid][c] = regs[r][c];
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NB; ++u) {
          const int e = tid + kPanelThreads * u, rr = e / NB, c = e % NB;
          const int row = base + rr;
          if (row < m && c < nb) A[(int64_t)row * lda + k + c] = s_rows[rr][c];
        }
      }
    }
  }
}

// Loads the Householder vectors of rows [r0, r0 + kTrailRows) of the panel at
// column k into LDS (with the implied unit diagonal and zeros above).
template <typename T, int NB>
__device__ __forceinline__ void load_v_chunk(T (*sv)[NB], const T* A, int lda, int m, int k, int nb, int r0,
                                             int tid) {
  // NB consecutive lanes read one row's NB entries (contiguous in memory)
  static_assert(kTrailThreads % NB == 0 && kTrailRows % (kTrailThreads / NB) == 0, "tile shape");
  const int j = tid % NB;
  const int diag = k + j;
  constexpr int kPer = kTrailRows / (kTrailThreads / NB);
#pragma unroll
  for (int t = 0; t < kPer; ++t) {
    const int i = tid / NB + t * (kTrailThreads / NB);
    const int row = r0 + i;
    T v = T(0);
    if (row < m && j < nb) v = row > diag ? A[(int64_t)row * lda + k + j] : (row == diag ? T(1) : T(0));
    sv[i][j] = v;
  }
}

// Rows of a trailing chunk handled per wave between waits: each wave issues
// this many independent loads of C before using any of them.
constexpr int kTrailBatch = 8;

// (1) partial W = V^T C over one 256-row chunk and one 64-column block.
template <typename T>
__global__ __launch_bounds__(kTrailThreads) void qr_trail_partial_kernel(const lxa_qr_desc* __restrict__ descs,
                                                                        int k_all, const int32_t* __restrict__ kcol,
                                                                        const T* __restrict__ a, T* __restrict__ work) {
  constexpr int NB = QrCfg<T>::NB;
  __shared__ T sv[kTrailRows][NB];
  __shared__ T sp[kTrailThreads / kWave][NB][kWave];
  const lxa_qr_desc d = descs[blockIdx.z];
  const int m = d.m, n = d.n, lda = d.lda;
  const int kmax = min(m, n);
  const int k = kcol ? kcol[blockIdx.z] : k_all;
  if (k >= kmax) return;
  const int nb = min(NB, kmax - k);
  const int c0 = k + nb + kTrailCols * (int)blockIdx.x;
  const int r0 = k + kTrailRows * (int)blockIdx.y;
  if (c0 >= n || r0 >= m) return;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid >> 6;
  const T* A = a + d.a_offset;
  load_v_chunk<T, NB>(sv, A, lda, m, k, nb, r0, tid);
  __syncthreads();
  const int col = c0 + lane;
  const bool cv = col < n;
  T acc[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) acc[j] = T(0);
  const int rows = min(kTrailRows, m - r0);
  constexpr int kStep = kTrailThreads / kWave;
  for (int i0 = wv; i0 < rows; i0 += kTrailBatch * kStep) {
    T cb[kTrailBatch];
#pragma unroll
    for (int u = 0; u < kTrailBatch; ++u) {
      const int i = i0 + u * kStep;
      cb[u] = (cv && i < rows) ? A[(int64_t)(r0 + i) * lda + col] : T(0);
    }
    // rows in the same order as one at a time: the sums are unchanged
#pragma unroll
    for (int u = 0; u < kTrailBatch; ++u) {
      const int i = i0 + u * kStep;
      if (i < rows) {
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[j] += sv[i][j] * cb[u];
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NB; ++j) sp[wv][j][lane] = acc[j];
  __syncthreads();
  T* W = work + d.w_offset + (int64_t)blockIdx.y * NB * n;
  for (int idx = tid; idx < NB * kWave; idx += kTrailThreads) {
    const int j = idx >> 6, l = idx & (kWave - 1);
    T s = T(0);
#pragma unroll
    for (int w = 0; w < kTrailThreads / kWave; ++w) s += sp[w][j][l];
    if (c0 + l < n) W[(int64_t)j * n + c0 + l] = s;
  }
}

// (2) W = sum over chunks, then W <- T^T W; stored in chunk slot 0.
template <typename T>
__global__ __launch_bounds__(kTrailThreads) void qr_trail_reduce_kernel(const lxa_qr_desc* __restrict__ descs,
                                                                       int k_all, const int32_t* __restrict__ kcol,
                                                                       const T* __restrict__ tblk,
                                                                       T* __restrict__ work) {
  constexpr int NB = QrCfg<T>::NB;
  __shared__ T sw[NB][kWave];
  __shared__ T st[NB][NB];
  const lxa_qr_desc d = descs[blockIdx.y];
  const int m = d.m, n = d.n;
  const int kmax = min(m, n);
  const int k = kcol ? kcol[blockIdx.y] : k_all;
  if (k >= kmax) return;
  const int nb = min(NB, kmax - k);
  const int c0 = k + nb + kTrailCols * (int)blockIdx.x;
  if (c0 >= n) return;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), g = tid >> 6;
  const int chunks = (m - k + kTrailRows - 1) / kTrailRows;
  T* W = work + d.w_offset;
  const T* Tp = tblk + d.t_offset + (int64_t)(k / NB) * NB * NB;
  if (tid < NB * NB) st[tid / NB][tid % NB] = Tp[tid];
  const int col = c0 + lane;
  for (int j = g; j < NB; j += kTrailThreads / kWave) {
    T s = T(0);
    if (col < n)
      for (int ch = 0; ch < chunks; ++ch) s += W[((int64_t)ch * NB + j) * n + col];
    sw[j][lane] = s;
  }
  __syncthreads();
  for (int j = g; j < NB; j += kTrailThreads / kWave) {
    T s = T(0);
#pragma unroll
    for (int i = 0; i < NB; ++i)
      if (i <= j) s += st[i][j] * sw[i][lane];
    if (col < n) W[(int64_t)j * n + col] = s;
  }
}

// (3) C -= V W over one chunk and column block.
template <typename T>
__global__ __launch_bounds__(kTrailThreads) void qr_trail_update_kernel(const lxa_qr_desc* __restrict__ descs,
                                                                       int k_all, const int32_t* __restrict__ kcol,
                                                                       T* __restrict__ a,
                                                                       const T* __restrict__ work) {
  constexpr int NB = QrCfg<T>::NB;
  __shared__ T sv[kTrailRows][NB];
  const lxa_qr_desc d = descs[blockIdx.z];
  const int m = d.m, n = d.n, lda = d.lda;
  const int kmax = min(m, n);
  const int k = kcol ? kcol[blockIdx.z] : k_all;
  if (k >= kmax) return;
  const int nb = min(NB, kmax - k);
  const int c0 = k + nb + kTrailCols * (int)blockIdx.x;
  const int r0 = k + kTrailRows * (int)blockIdx.y;
  if (c0 >= n || r0 >= m) return;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid >> 6;
  T* A = a + d.a_offset;
  load_v_chunk<T, NB>(sv, A, lda, m, k, nb, r0, tid);
  const int col = c0 + lane;
  const bool cv = col < n;
  const T* W = work + d.w_offset;
  T w[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) w[j] = cv ? W[(int64_t)j * n + col] : T(0);
  __syncthreads();
  if (!cv) return;
  const int rows = min(kTrailRows, m - r0);
  constexpr int kStep = kTrailThreads / kWave;
  for (int i0 = wv; i0 < rows; i0 += kTrailBatch * kStep) {
    T cb[kTrailBatch];
#pragma unroll
    for (int u = 0; u < kTrailBatch; ++u) {
      const int i = i0 + u * kStep;
      cb[u] = i < rows ? A[(int64_t)(r0 + i) * lda + col] : T(0);
    }
#pragma unroll
    for (int u = 0; u < kTrailBatch; ++u) {
      const int i = i0 + u * kStep;
      if (i < rows) {
        T s = T(0);
#pragma unroll
        for (int j = 0; j < NB; ++j) s += sv[i][j] * w[j];
        A[(int64_t)(r0 + i) * lda + col] = cb[u] - s;
      }
    }
  }
}

// float32 forms of (1) and (3) on the matrix cores (v_mfma_f32_16x16x4f32,
// float32 products and accumulation).  Operand layouts (lane l):
//   A operand: A[l % 16][l / 16];  B operand: B[l / 16][l % 16];
//   result:    rows 4 (l / 16) + i (i < 4), column l % 16.
// Each wave of the 256-thread workgroup owns 16 of the block's 64 columns.
typedef float qr_f4 __attribute__((ext_vector_type(4)));

// (1) W(16 x 16 per wave) = V^T C over one 256-row chunk: A = V^T from LDS,
// B = C streamed from HBM, four rows per MFMA.  No cross-wave reduction.
__global__ __launch_bounds__(kTrailThreads) void qr_trail_partial_mfma_kernel(const lxa_qr_desc* __restrict__ descs,
                                                                            int k_all,
                                                                            const int32_t* __restrict__ kcol,
                                                                            const float* __restrict__ a,
                                                                            float* __restrict__ work) {
  constexpr int NB = QrCfg<float>::NB;
  static_assert(NB == 16 && kTrailCols == 64 && kTrailThreads == 256, "one 16-column slab per wave");
  __shared__ float sv[kTrailRows][NB];
  const lxa_qr_desc d = descs[blockIdx.z];
  const int m = d.m, n = d.n, lda = d.lda;
  const int kmax = min(m, n);
  const int k = kcol ? kcol[blockIdx.z] : k_all;
  if (k >= kmax) return;
  const int nb = min(NB, kmax - k);
  const int c0 = k + nb + kTrailCols * (int)blockIdx.x;
  const int r0 = k + kTrailRows * (int)blockIdx.y;
  if (c0 >= n || r0 >= m) return;
  const int tid = threadIdx.x, l = tid & (kWave - 1), wv = tid >> 6;
  const float* A = a + d.a_offset;
  load_v_chunk<float, NB>(sv, A, lda, m, k, nb, r0, tid);
  __syncthreads();
  const int col = c0 + 16 * wv + (l & 15);
  const bool cv = col < n;
  const int rows = min(kTrailRows, m - r0);
  qr_f4 acc = {0.f, 0.f, 0.f, 0.f};
  // rows kk + l / 16 of the chunk; eight MFMA row groups (32 rows) per batch
  constexpr int kGroups = 8;
  for (int kb = 0; kb < rows; kb += 4 * kGroups) {
    float cb[kGroups];
#pragma unroll
    for (int u = 0; u < kGroups; ++u) {
      const int i = kb + 4 * u + (l >> 4);
      cb[u] = (cv && i < rows) ? A[(int64_t)(r0 + i) * lda + col] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < kGroups; ++u) {
      const int i = kb + 4 * u + (l >> 4);
      const float av = i < kTrailRows ? sv[i][l & 15] : 0.f;  // zero past the chunk (and past m)
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, cb[u], acc, 0, 0, 0);
    }
  }
  float* W = work + d.w_offset + (int64_t)blockIdx.y * NB * n;
  if (cv) {
#pragma unroll
    for (int i = 0; i < 4; ++i) W[(int64_t)(4 * (l >> 4) + i) * n + col] = acc[i];
  }
}

// (3) C -= V W: each wave takes its 16 columns down the chunk in 16 x 16
// tiles; the tile is loaded into the accumulator, and V (from LDS) times -W
// (registers) is added by four MFMAs.
__global__ __launch_bounds__(kTrailThreads) void qr_trail_update_mfma_kernel(const lxa_qr_desc* __restrict__ descs,
                                                                           int k_all,
                                                                           const int32_t* __restrict__ kcol,
                                                                           float* __restrict__ a,
                                                                           const float* __restrict__ work) {
  constexpr int NB = QrCfg<float>::NB;
  __shared__ float sv[kTrailRows][NB];
  const lxa_qr_desc d = descs[blockIdx.z];
  const int m = d.m, n = d.n, lda = d.lda;
  const int kmax = min(m, n);
  const int k = kcol ? kcol[blockIdx.z] : k_all;
  if (k >= kmax) return;
  const int nb = min(NB, kmax - k);
  const int c0 = k + nb + kTrailCols * (int)blockIdx.x;
  const int r0 = k + kTrailRows * (int)blockIdx.y;
  if (c0 >= n || r0 >= m) return;
  const int tid = threadIdx.x, l = tid & (kWave - 1), wv = tid >> 6;
  float* A = a + d.a_offset;
  load_v_chunk<float, NB>(sv, A, lda, m, k, nb, r0, tid);
  const int col = c0 + 16 * wv + (l & 15);
  const bool cv = col < n;
  // -W as the B operand: rows j = 4 s + l / 16 of this wave's columns
  const float* W = work + d.w_offset;
  float wb[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) wb[q] = cv ? -W[(int64_t)(4 * q + (l >> 4)) * n + col] : 0.f;
  __syncthreads();
  const int rows = min(kTrailRows, m - r0);
  for (int rt = 0; rt < rows; rt += 16) {
    qr_f4 acc;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = rt + 4 * (l >> 4) + i;
      acc[i] = (cv && row < rows) ? A[(int64_t)(r0 + row) * lda + col] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float av = sv[rt + (l & 15)][4 * q + (l >> 4)];
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, wb[q], acc, 0, 0, 0);
    }
    if (cv) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = rt + 4 * (l >> 4) + i;
        if (row < rows) A[(int64_t)(r0 + row) * lda + col] = acc[i];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// b <- Q^T b (trans) or Q b: one workgroup per (system, right-hand side),
// panel by panel: w = V^T b; w = T^T w (or T w); b -= V w.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBigThreads) void qr_apply_kernel(const lxa_qr_desc* __restrict__ descs,
                                                                 const T* __restrict__ a,
                                                                 const T* __restrict__ tblk, T* __restrict__ b,
                                                                 int trans) {
  constexpr int NB = QrCfg<T>::NB;
  __shared__ T s_stage[kBigWaves * NB];
  __shared__ T s_out[NB];
  __shared__ T s_w[NB];
  const lxa_qr_desc d = descs[blockIdx.x];
  const int m = d.m, n = d.n, lda = d.lda;
  const int kmax = min(m, n);
  if (kmax <= 0) return;
  const int tid = threadIdx.x;
  const T* A = a + d.a_offset;
  T* x = b + d.b_offset + (int64_t)blockIdx.y * d.ldb;
  const int panels = (kmax + NB - 1) / NB;
  for (int q = 0; q < panels; ++q) {
    const int p = trans ? q : panels - 1 - q;
    const int k = p * NB;
    const int nb = min(NB, kmax - k);
    T acc[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[j] = T(0);
    for (int row = k + tid; row < m; row += kBigThreads) {
      const T xv = x[row];
      const T* Ar = A + (int64_t)row * lda + k;
#pragma unroll
      for (int j = 0; j < NB; ++j)
        if (j < nb) acc[j] += vpick(row, k + j, Ar[j]) * xv;
    }
    block_sum<NB, kBigWaves>(acc, s_stage, s_out, tid);
    if (tid < NB) {
      const T* Tp = tblk + d.t_offset + (int64_t)p * NB * NB;
      T s = T(0);
      if (trans) {  // (T^T w)_tid = sum_{i <= tid} T[i][tid] w_i
#pragma unroll
        for (int i = 0; i < NB; ++i)
          if (i <= tid) s += Tp[i * NB + tid] * acc[i];
      } else {  // (T w)_tid = sum_{i >= tid} T[tid][i] w_i
#pragma unroll
        for (int i = 0; i < NB; ++i)
          if (i >= tid) s += Tp[tid * NB + i] * acc[i];
      }
      s_w[tid] = s;
    }
    __syncthreads();
    T w[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) w[j] = s_w[j];
    for (int row = k + tid; row < m; row += kBigThreads) {
      const T* Ar = A + (int64_t)row * lda + k;
      T s = T(0);
#pragma unroll
      for (int j = 0; j < NB; ++j)
        if (j < nb) s += vpick(row, k + j, Ar[j]) * w[j];
      x[row] -= s;
    }
    __syncthreads();  // x rows are re-read by other threads in the next panel
  }
}

// ---------------------------------------------------------------------------
// The same product for systems of up to kApplyMaxRows rows, one vector per
// workgroup: the vector lives in LDS and NB lanes share a row (lane j reads
// column j of the panel, so a wave load covers 64 / NB contiguous row
// segments instead of 64 scattered ones).  RPT > 0: each lane keeps its
// RPT panel entries in registers between the reduction and the update (V is
// read once per panel); RPT == 0 reads them twice.  Per panel:
//   w = V^T x (lanes of a column, then waves, summed in a fixed order);
//   w = T^T w or T w (wave 0, shuffles);  x -= V w (NB-lane row sums).
// ---------------------------------------------------------------------------
constexpr int kApplyMaxRows = 16384;

template <typename T, int RPT>
__global__ __launch_bounds__(kBigThreads) void qr_apply_rows_kernel(const lxa_qr_desc* __restrict__ descs,
                                                                      const T* __restrict__ a,
                                                                      const T* __restrict__ tblk,
                                                                      T* __restrict__ b, int trans) {
  constexpr int NB = QrCfg<T>::NB;
  constexpr int G = kWave / NB;       // rows per wave load
  constexpr int RP = kBigWaves * G;   // rows per pass of the workgroup
  extern __shared__ __align__(16) unsigned char apply_smem[];
  T* xs = reinterpret_cast<T*>(apply_smem);
  __shared__ T s_part[kBigWaves][NB];
  __shared__ T s_w[NB];
  const lxa_qr_desc d = descs[blockIdx.x];
  const int m = d.m, lda = d.lda;
  const int kmax = min(m, d.n);
  if (kmax <= 0) return;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid >> 6;
  const int j = lane % NB, rg = wv * G + lane / NB;
  const T* A = a + d.a_offset;
  T* x = b + d.b_offset + (int64_t)blockIdx.y * d.ldb;
  for (int i = tid; i < m; i += kBigThreads) xs[i] = x[i];
  __syncthreads();
  const int panels = (kmax + NB - 1) / NB;
  T v[RPT > 0 ? RPT : 1];
  for (int q = 0; q < panels; ++q) {
    const int p = trans ? q : panels - 1 - q;
    const int k = p * NB;
    const int nb = min(NB, kmax - k);
    const int jc = min(j, nb - 1);  // clamped column: loads stay inside the panel
    const T* Ak = A + k + jc;
    // wave 0's row (trans) or column of the panel's T block, loaded now so
    // the loads overlap the reduction below instead of one wait per term
    T tv[NB];
    if (wv == 0) {
      const T* Tp = tblk + d.t_offset + (int64_t)p * NB * NB;
      const int lt = min(lane, NB - 1);
#pragma unroll
      for (int i = 0; i < NB; ++i) tv[i] = trans ? Tp[i * NB + lt] : Tp[lt * NB + i];
    }
    T acc = T(0);
    if constexpr (RPT > 0) {
#pragma unroll
      for (int t = 0; t < RPT; ++t) {
        const int row = k + rg + t * RP;
        const T st = Ak[(int64_t)min(row, m - 1) * lda];
        v[t] = (row < m && j < nb) ? vpick(row, k + j, st) : T(0);
      }
#pragma unroll
      for (int t = 0; t < RPT; ++t) {
        const int row = k + rg + t * RP;
        if (row < m) acc += v[t] * xs[row];
      }
    } else {
      // eight rows per round, loads issued together (clamped rows, masked)
      for (int row0 = k + rg; row0 < m; row0 += 8 * RP) {
        T st[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) st[u] = Ak[(int64_t)min(row0 + u * RP, m - 1) * lda];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int row = row0 + u * RP;
          if (row < m && j < nb) acc += vpick(row, k + j, st[u]) * xs[row];
        }
      }
    }
#pragma unroll
    for (int off = NB; off < kWave; off <<= 1) acc += __shfl_xor(acc, off, kWave);
    if (lane < NB) s_part[wv][lane] = acc;
    __syncthreads();
    if (wv == 0) {
      T wsum = T(0);
      if (lane < NB) {
#pragma unroll
        for (int w = 0; w < kBigWaves; ++w) wsum += s_part[w][lane];
      }
      T s = T(0);
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const T wi = __shfl(wsum, i, kWave);
        if (lane < NB && i < nb && lane < nb) {
          if (trans) {  // (T^T w)_lane = sum_{i <= lane} T[i][lane] w_i
            if (i <= lane) s += tv[i] * wi;
          } else {  // (T w)_lane = sum_{i >= lane} T[lane][i] w_i
            if (i >= lane) s += tv[i] * wi;
          }
        }
      }
      if (lane < NB) s_w[lane] = s;
    }
    __syncthreads();
    const T wj = s_w[j];
    if constexpr (RPT > 0) {
#pragma unroll
      for (int t = 0; t < RPT; ++t) {
        const int row = k + rg + t * RP;
        T s = v[t] * wj;
#pragma unroll
        for (int off = 1; off < NB; off <<= 1) s += __shfl_xor(s, off, kWave);
        if (j == 0 && row < m) xs[row] -= s;
      }
    } else {
      for (int row0 = k + rg; row0 < m; row0 += 8 * RP) {
        T st[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) st[u] = Ak[(int64_t)min(row0 + u * RP, m - 1) * lda];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int row = row0 + u * RP;
          T s = (row < m && j < nb) ? vpick(row, k + j, st[u]) * wj : T(0);
#pragma unroll
          for (int off = 1; off < NB; off <<= 1) s += __shfl_xor(s, off, kWave);
          if (j == 0 && row < m) xs[row] -= s;
        }
      }
    }
    __syncthreads();
  }
  for (int i = tid; i < m; i += kBigThreads) x[i] = xs[i];
}

template <typename T, int RPT>
int launch_apply_rows(hipStream_t st, dim3 grid, int max_m, const lxa_qr_desc* desc, const void* a, const void* tblk,
                      void* b, int trans) {
  const size_t smem = sizeof(T) * (size_t)max_m;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&qr_apply_rows_kernel<T, RPT>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) return hip_error(e, "hipFuncSetAttribute(qr_apply_rows)");
  qr_apply_rows_kernel<T, RPT><<<grid, kBigThreads, smem, st>>>(desc, static_cast<const T*>(a),
                                                                 static_cast<const T*>(tblk), static_cast<T*>(b),
                                                                 trans);
  LXA_LAUNCH_CHECK("qr_apply_rows_kernel");
  return LXA_OK;
}

template <typename T>
int apply_rows(hipStream_t st, dim3 grid, int max_m, const lxa_qr_desc* desc, const void* a, const void* tblk,
               void* b, int trans) {
  constexpr int RP = kBigWaves * (kWave / QrCfg<T>::NB);
  constexpr int kMaxRpt = sizeof(T) == 4 ? 32 : 16;  // no spills under the 128-VGPR cap of 1024 threads
  const int rpt = (max_m + RP - 1) / RP;
  if (rpt <= 4) return launch_apply_rows<T, 4>(st, grid, max_m, desc, a, tblk, b, trans);
  if (rpt <= 8) return launch_apply_rows<T, 8>(st, grid, max_m, desc, a, tblk, b, trans);
  if (rpt <= 16) return launch_apply_rows<T, 16>(st, grid, max_m, desc, a, tblk, b, trans);
  if (rpt <= kMaxRpt) return launch_apply_rows<T, kMaxRpt>(st, grid, max_m, desc, a, tblk, b, trans);
  return launch_apply_rows<T, 0>(st, grid, max_m, desc, a, tblk, b, trans);
}

// ---------------------------------------------------------------------------
// Triangular solve with an upper triangular R, R[i][j] = a[i*rs + j*cs]:
// back substitution for R x = y (tr = 0), forward for R^T x = y (tr = 1).
// One workgroup per (system, right-hand side); x lives in LDS; 64-row
// blocks: a parallel product with the solved part, then one wave solves the
// diagonal block from an LDS copy of it.
// ---------------------------------------------------------------------------
constexpr int kTrsvBlock = 64;
constexpr int kTrsvMaxK = 8192;

// v from lane j of the wave (j uniform): v_readlane into a scalar register.
__device__ __forceinline__ float lane_bcast(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ double lane_bcast(double v, int j) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits), j);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), j);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

template <typename T>
__global__ __launch_bounds__(kBigThreads) void trsv_kernel(const lxa_qr_desc* __restrict__ descs,
                                                             const T* __restrict__ a, T* __restrict__ b, int lower,
                                                             int trans, int unit_diag) {
  extern __shared__ __align__(16) unsigned char trsv_smem[];
  T* xs = reinterpret_cast<T*>(trsv_smem);                      // [kTrsvMaxK] (only k used)
  T* tile = xs + ((kTrsvMaxK + 1) & ~1);                        // [64][65]
  T* red = tile + kTrsvBlock * (kTrsvBlock + 1);                // [16][64]
  T* rinv = red + kBigWaves * kWave;                            // [64] reciprocal diagonal
  const lxa_qr_desc d = descs[blockIdx.x];
  const int kk = min(d.m, d.n);
  if (kk <= 0) return;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid >> 6;
  const T* A = a + d.a_offset;
  const int64_t rs = lower ? 1 : d.lda, cs = lower ? d.lda : 1;
  const int tr = trans ^ lower;
  T* y = b + d.b_offset + (int64_t)blockIdx.y * d.ldb;
  for (int i = tid; i < kk; i += kBigThreads) xs[i] = y[i];
  __syncthreads();
  const int blocks = (kk + kTrsvBlock - 1) / kTrsvBlock;
  for (int q = 0; q < blocks; ++q) {
    const int bi = tr ? q : blocks - 1 - q;
    const int r0 = bi * kTrsvBlock;
    const int r1 = min(kk, r0 + kTrsvBlock);
    const int bs = r1 - r0;
    // diagonal block tile[i][j] = R[r0 + i][r0 + j]: loads issued here (clamped
    // indices, masked values), stored to LDS after the product pass
    constexpr int kTileVals = kTrsvBlock * kTrsvBlock / kBigThreads;
    T tv[kTileVals];
#pragma unroll
    for (int u = 0; u < kTileVals; ++u) {
      const int idx = tid + u * kBigThreads;
      const int i = idx >> 6, j = idx & 63;
      const T v = A[(int64_t)(r0 + min(i, bs - 1)) * rs + (int64_t)(r0 + min(j, bs - 1)) * cs];
      tv[u] = (i < bs && j < bs && j >= i) ? v : T(0);
    }
    if (!tr) {
      // z_i = y_i - sum_{c >= r1} R[i][c] x_c: wave wv takes rows r0 + wv + 16 t
      // (eight independent partial sums keep eight loads in flight per lane)
      for (int i = wv; i < bs; i += kBigWaves) {
        T s8[8] = {};
        const T* Ri = A + (int64_t)(r0 + i) * rs;
        int c = r1 + lane;
        for (; c + 7 * kWave < kk; c += 8 * kWave) {
#pragma unroll
          for (int u = 0; u < 8; ++u) s8[u] += Ri[(int64_t)(c + u * kWave) * cs] * xs[c + u * kWave];
        }
        for (; c < kk; c += kWave) s8[0] += Ri[(int64_t)c * cs] * xs[c];
        T s = ((s8[0] + s8[1]) + (s8[2] + s8[3])) + ((s8[4] + s8[5]) + (s8[6] + s8[7]));
        s = wave_sum(s);
        if (lane == 0) red[i] = s;
      }
    } else {
      // z_i = y_i - sum_{c < r0} R[c][i] x_c: thread (column i, group g)
      const int i = lane, g = wv;
      T s8[8] = {};
      if (i < bs) {
        const T* Ai = A + (int64_t)(r0 + i) * cs;
        int c = g;
        for (; c + 7 * kBigWaves < r0; c += 8 * kBigWaves) {
#pragma unroll
          for (int u = 0; u < 8; ++u) s8[u] += Ai[(int64_t)(c + u * kBigWaves) * rs] * xs[c + u * kBigWaves];
        }
        for (; c < r0; c += kBigWaves) s8[0] += Ai[(int64_t)c * rs] * xs[c];
      }
      red[g * kWave + i] = ((s8[0] + s8[1]) + (s8[2] + s8[3])) + ((s8[4] + s8[5]) + (s8[6] + s8[7]));
    }
#pragma unroll
    for (int u = 0; u < kTileVals; ++u) {
      const int idx = tid + u * kBigThreads;
      const int i = idx >> 6, j = idx & 63;
      tile[i * (kTrsvBlock + 1) + j] = tv[u];
      if (i == j) rinv[i] = (unit_diag || i >= bs) ? T(1) : T(1) / tv[u];
    }
    __syncthreads();
    if (wv == 0) {
      T z = T(0);
      if (lane < bs) {
        if (!tr) {
          z = xs[r0 + lane] - red[lane];
        } else {
          T s = T(0);
#pragma unroll
          for (int g = 0; g < kBigWaves; ++g) s += red[g * kWave + lane];
          z = xs[r0 + lane] - s;
        }
      }
      // Substitution through the block, one unknown per step: lane j scales
      // by the reciprocal pivot and readlane broadcasts it (a register read,
      // not an LDS round trip).  Fully unrolled, so the LDS reads of the
      // multipliers do not wait on the chain.  Steps j >= bs are no-ops
      // (zero padding, x_j = 0).
      const T rl = rinv[lane];
      if (!tr) {
#pragma unroll
        for (int j = kTrsvBlock - 1; j >= 0; --j) {
          const T mult = tile[lane * (kTrsvBlock + 1) + j];
          if (lane == j) z *= rl;
          const T xj = j < bs ? lane_bcast(z, j) : T(0);
          if (lane < j) z -= mult * xj;
        }
      } else {
#pragma unroll
        for (int j = 0; j < kTrsvBlock; ++j) {
          const T mult = tile[j * (kTrsvBlock + 1) + lane];
          if (lane == j) z *= rl;
          const T xj = j < bs ? lane_bcast(z, j) : T(0);
          if (lane > j) z -= mult * xj;
        }
      }
      if (lane < bs) xs[r0 + lane] = z;
    }
    __syncthreads();
  }
  for (int i = tid; i < kk; i += kBigThreads) y[i] = xs[i];
}

size_t trsv_smem_bytes(int dtype) {
  const size_t e = dtype == LXA_F64 ? 8 : 4;
  return e * (((kTrsvMaxK + 1) & ~1) + kTrsvBlock * (kTrsvBlock + 1) + kBigWaves * kWave + kTrsvBlock);
}

// One panel step for every system: column k_all for all (kcol == nullptr)
// or kcol[s] for system s.  max_rows bounds m - k and max_trail_cols bounds
// n - k - 1 over the systems that still have columns.
template <typename T>
int qr_step_impl(hipStream_t st, int64_t count, const lxa_qr_desc* desc, int k_all, const int32_t* kcol,
                 int max_rows, int max_trail_cols, T* a, T* tau, T* tblk, T* work) {
  const int rpt = (max_rows + kPanelThreads - 1) / kPanelThreads;
  const dim3 pg((unsigned)count);
  if (rpt <= 1)
    qr_panel_kernel<T, 1><<<pg, kPanelThreads, 0, st>>>(desc, k_all, kcol, a, tau, tblk);
  else if (rpt <= 2)
    qr_panel_kernel<T, 2><<<pg, kPanelThreads, 0, st>>>(desc, k_all, kcol, a, tau, tblk);
  else if (rpt <= 4)
    qr_panel_kernel<T, 4><<<pg, kPanelThreads, 0, st>>>(desc, k_all, kcol, a, tau, tblk);
  else if (rpt <= kPanelMaxRpt)
    qr_panel_kernel<T, kPanelMaxRpt><<<pg, kPanelThreads, 0, st>>>(desc, k_all, kcol, a, tau, tblk);
  else
    qr_panel_kernel<T, 0><<<pg, kPanelThreads, 0, st>>>(desc, k_all, kcol, a, tau, tblk);
  LXA_LAUNCH_CHECK("qr_panel_kernel");
  if (max_trail_cols <= 0) return LXA_OK;
  const unsigned cb = (unsigned)((max_trail_cols + kTrailCols - 1) / kTrailCols);
  const unsigned rc = (unsigned)((max_rows + kTrailRows - 1) / kTrailRows);
  if constexpr (std::is_same_v<T, float>) {
    qr_trail_partial_mfma_kernel<<<dim3(cb, rc, (unsigned)count), kTrailThreads, 0, st>>>(desc, k_all, kcol, a,
                                                                                          work);
  } else {
    qr_trail_partial_kernel<T><<<dim3(cb, rc, (unsigned)count), kTrailThreads, 0, st>>>(desc, k_all, kcol, a, work);
  }
  LXA_LAUNCH_CHECK("qr_trail_partial_kernel");
  qr_trail_reduce_kernel<T><<<dim3(cb, (unsigned)count), kTrailThreads, 0, st>>>(desc, k_all, kcol, tblk, work);
  LXA_LAUNCH_CHECK("qr_trail_reduce_kernel");
  if constexpr (std::is_same_v<T, float>) {
    qr_trail_update_mfma_kernel<<<dim3(cb, rc, (unsigned)count), kTrailThreads, 0, st>>>(desc, k_all, kcol, a,
                                                                                         work);
  } else {
    qr_trail_update_kernel<T><<<dim3(cb, rc, (unsigned)count), kTrailThreads, 0, st>>>(desc, k_all, kcol, a, work);
  }
  LXA_LAUNCH_CHECK("qr_trail_update_kernel");
  return LXA_OK;
}

template <typename T>
int qr_factor_impl(hipStream_t st, int64_t count, const lxa_qr_desc* desc, int max_m, int max_n, int max_k, T* a,
                   T* tau, T* tblk, T* work) {
  constexpr int NB = QrCfg<T>::NB;
  for (int k = 0; k < max_k; k += NB) {
    if (int rc = qr_step_impl<T>(st, count, desc, k, nullptr, max_m - k, max_n - k - 1, a, tau, tblk, work)) return rc;
  }
  return LXA_OK;
}

int check_batch(int dtype, int64_t count, const void* desc) {
  LXA_REQUIRE(dtype == LXA_F32 || dtype == LXA_F64, "dtype must be LXA_F32 or LXA_F64 (got %d)", dtype);
  LXA_REQUIRE(count >= 0 && count <= 0x7fffffff, "count out of range (%lld)", (long long)count);
  LXA_REQUIRE(count == 0 || desc != nullptr, "desc is NULL");
  return LXA_OK;
}

}  // namespace
}  // namespace lxa

using namespace lxa;

extern "C" int lxa_qr_block(int dtype) { return dtype == LXA_F64 ? QrCfg<double>::NB : QrCfg<float>::NB; }

extern "C" int lxa_qr_work_elems(int dtype, int64_t m, int64_t n, int64_t* elems) {
  LXA_REQUIRE(dtype == LXA_F32 || dtype == LXA_F64, "dtype must be LXA_F32 or LXA_F64 (got %d)", dtype);
  LXA_REQUIRE(m >= 0 && n >= 0 && elems != nullptr, "invalid arguments");
  const int64_t nb = lxa_qr_block(dtype);
  *elems = ((m + kTrailRows - 1) / kTrailRows) * nb * n;
  return LXA_OK;
}

extern "C" int lxa_qr_factor(void* stream, int dtype, int64_t count, const lxa_qr_desc* desc, int max_m, int max_n,
                             int max_k, void* a, void* tau, void* tblk, void* work) {
  if (int rc = check_batch(dtype, count, desc)) return rc;
  LXA_REQUIRE(max_m >= 0 && max_n >= 0 && max_k >= 0 && max_k <= max_m && max_k <= max_n,
              "inconsistent bounds max_m=%d max_n=%d max_k=%d", max_m, max_n, max_k);
  if (count == 0 || max_k == 0) return LXA_OK;
  LXA_REQUIRE(a && tau && tblk && work, "a, tau, tblk and work must be non-NULL");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == LXA_F32)
    return qr_factor_impl<float>(st, count, desc, max_m, max_n, max_k, static_cast<float*>(a),
                                 static_cast<float*>(tau), static_cast<float*>(tblk), static_cast<float*>(work));
  return qr_factor_impl<double>(st, count, desc, max_m, max_n, max_k, static_cast<double*>(a),
                                static_cast<double*>(tau), static_cast<double*>(tblk), static_cast<double*>(work));
}

extern "C" int lxa_qr_factor_step(void* stream, int dtype, int64_t count, const lxa_qr_desc* desc,
                                  const int32_t* kcol, int max_rows, int max_trail_cols, void* a, void* tau,
                                  void* tblk, void* work) {
  if (int rc = check_batch(dtype, count, desc)) return rc;
  LXA_REQUIRE(kcol != nullptr, "kcol must be non-NULL");
  LXA_REQUIRE(max_rows >= 0, "max_rows must be >= 0");
  if (count == 0 || max_rows == 0) return LXA_OK;
  LXA_REQUIRE(a && tau && tblk && work, "a, tau, tblk and work must be non-NULL");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == LXA_F32)
    return qr_step_impl<float>(st, count, desc, 0, kcol, max_rows, max_trail_cols, static_cast<float*>(a),
                               static_cast<float*>(tau), static_cast<float*>(tblk), static_cast<float*>(work));
  return qr_step_impl<double>(st, count, desc, 0, kcol, max_rows, max_trail_cols, static_cast<double*>(a),
                              static_cast<double*>(tau), static_cast<double*>(tblk), static_cast<double*>(work));
}

extern "C" int lxa_qr_apply(void* stream, int dtype, int64_t count, const lxa_qr_desc* desc, int max_m, int max_k,
                            const void* a, const void* tblk, int nrhs, void* b, int trans) {
  if (int rc = check_batch(dtype, count, desc)) return rc;
  LXA_REQUIRE(nrhs >= 0 && nrhs <= 65535, "nrhs out of range (%d)", nrhs);
  LXA_REQUIRE(max_k >= 0 && max_m >= max_k, "inconsistent bounds max_m=%d max_k=%d", max_m, max_k);
  if (count == 0 || nrhs == 0 || max_k == 0) return LXA_OK;
  LXA_REQUIRE(a && tblk && b, "a, tblk and b must be non-NULL");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid((unsigned)count, (unsigned)nrhs);
  if (max_m <= kApplyMaxRows) {
    if (dtype == LXA_F32) return apply_rows<float>(st, grid, max_m, desc, a, tblk, b, trans ? 1 : 0);
    return apply_rows<double>(st, grid, max_m, desc, a, tblk, b, trans ? 1 : 0);
  }
  if (dtype == LXA_F32)
    qr_apply_kernel<float><<<grid, kBigThreads, 0, st>>>(desc, static_cast<const float*>(a),
                                                          static_cast<const float*>(tblk), static_cast<float*>(b),
                                                          trans ? 1 : 0);
  else
    qr_apply_kernel<double><<<grid, kBigThreads, 0, st>>>(desc, static_cast<const double*>(a),
                                                           static_cast<const double*>(tblk), static_cast<double*>(b),
                                                           trans ? 1 : 0);
  LXA_LAUNCH_CHECK("qr_apply_kernel");
  return LXA_OK;
}

extern "C" int lxa_trsv(void* stream, int dtype, int64_t count, const lxa_qr_desc* desc, int max_k, const void* a,
                        int nrhs, void* b, int lower, int trans, int unit_diag) {
  if (int rc = check_batch(dtype, count, desc)) return rc;
  LXA_REQUIRE(nrhs >= 0 && nrhs <= 65535, "nrhs out of range (%d)", nrhs);
  if (max_k > kTrsvMaxK)
    return set_error(LXA_EUNSUPPORTED, "triangular solves support order <= %d (got %d)", kTrsvMaxK, max_k);
  if (count == 0 || nrhs == 0 || max_k <= 0) return LXA_OK;
  LXA_REQUIRE(a && b, "a and b must be non-NULL");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid((unsigned)count, (unsigned)nrhs);
  const size_t smem = trsv_smem_bytes(dtype);
  if (dtype == LXA_F32) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&trsv_kernel<float>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return hip_error(e, "hipFuncSetAttribute(trsv)");
    trsv_kernel<float><<<grid, kBigThreads, smem, st>>>(desc, static_cast<const float*>(a), static_cast<float*>(b),
                                                          lower ? 1 : 0, trans ? 1 : 0, unit_diag ? 1 : 0);
  } else {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&trsv_kernel<double>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return hip_error(e, "hipFuncSetAttribute(trsv)");
    trsv_kernel<double><<<grid, kBigThreads, smem, st>>>(desc, static_cast<const double*>(a),
                                                           static_cast<double*>(b), lower ? 1 : 0, trans ? 1 : 0,
                                                           unit_diag ? 1 : 0);
  }
  LXA_LAUNCH_CHECK("trsv_kernel");
  return LXA_OK;
}
