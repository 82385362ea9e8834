// Blocked Cholesky (upper, A = U^T U) for uneven batches of order <= 1024,
// one 256-thread workgroup per system, float32; and the matching solve.
//
// Per 32-column panel at k0 (kb = min(32, n - k0) columns):
//   (1) the kb x kb diagonal block is factored in LDS by one wave (lane j
//       owns column j; readlane-free: the pivot row is read from LDS);
//   (2) the row block U12 = U11^-T A12 is staged in LDS (32 x (n - k0 - kb)
//       floats), solved column-parallel in place, written back to A and kept
//       in LDS for the update;
//   (3) the trailing upper triangle takes the rank-kb update
//       A22 -= U12^T U12 in 16 x 16 tiles on the matrix cores, U12 read
//       from LDS as both MFMA operands.
// The trailing matrix is read and written once per panel (from L2 / HBM).
// A non-positive pivot fills the factor with NaN and sets info = k + 1, as
// lxa_chol_factor does.  Only the upper triangle is read; the strictly lower
// triangle is not touched (the solve reads only the upper triangle).
//
// Reference semantics: lineax/_solver/cholesky.py:43-78 (cho_factor upper,
// cho_solve); this is the uneven-batch form the request loop batches.

#include "lxa_common.h"

namespace lxa {
namespace {

constexpr int kP = 32;           // panel width
constexpr int kThreads = 256;
constexpr int kMaxN = 1024;
constexpr int kPanelLd = kMaxN + 16;  // LDS row stride of the U12 panel (rows start 16 banks apart)
constexpr int kDLd = kP + 4;     // LDS row stride of the diagonal block (16-byte rows)

typedef float cb_f4 __attribute__((ext_vector_type(4)));

template <int B, int E, typename F>
__device__ __forceinline__ void cb_static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    cb_static_for<B + 1, E>(f);
  }
}

struct Geometry {
  int n;
  int64_t a_off, b_off;
  int lda;
};

__device__ __forceinline__ Geometry geometry(const lxa_qr_desc* descs, int n_all, int64_t a_stride, int64_t b_stride) {
  Geometry g;
  if (descs) {
    const lxa_qr_desc d = descs[blockIdx.x];
    g.n = d.n;
    g.a_off = d.a_offset;
    g.b_off = d.b_offset;
    g.lda = d.lda;
  } else {
    g.n = n_all;
    g.a_off = (int64_t)blockIdx.x * a_stride;
    g.b_off = (int64_t)blockIdx.x * b_stride;
    g.lda = n_all;
  }
  return g;
}

__global__ __launch_bounds__(kThreads) void chol_blocked_factor_kernel(const lxa_qr_desc* __restrict__ descs, int n_all,
                                                                       int64_t a_stride, float* __restrict__ a,
                                                                       int32_t* __restrict__ info) {
  extern __shared__ __align__(16) float cbs[];
  float* D = cbs;                      // [32][36]
  float* P = cbs + kP * kDLd;          // [32][kPanelLd]
  __shared__ int s_fail;
  const Geometry g = geometry(descs, n_all, a_stride, 0);
  const int n = g.n, lda = g.lda;
  float* A = a + g.a_off;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid == 0) s_fail = 0;
  __syncthreads();
  if (n <= 0 || n > kMaxN) {
    if (tid == 0) info[blockIdx.x] = n > kMaxN ? -1 : 0;
    return;
  }
  for (int k0 = 0; k0 < n; k0 += kP) {
    const int kb = min(kP, n - k0);
    const int rest = n - k0 - kb;
    {
      // every load issued before any store (indices clamped onto the upper
      // triangle of the block; the values outside it are replaced below)
      constexpr int kPer = kP * kP / kThreads;
      float v[kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int idx = tid + u * kThreads, i = idx / kP, j = idx % kP;
        const int jc = min(j, kb - 1), ic = min(min(i, kb - 1), jc);
        v[u] = A[(int64_t)(k0 + ic) * lda + k0 + jc];
      }
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int idx = tid + u * kThreads, i = idx / kP, j = idx % kP;
        D[i * kDLd + j] = (i < kb && j < kb) ? (j >= i ? v[u] : 0.f) : (i == j ? 1.f : 0.f);
      }
    }
    __syncthreads();
    if (wv == 0) {
      // right-looking on the upper triangle in registers: lane j holds
      // column j (c[i] = D[i][j]); the pivot and row t come by readlane.
      // Entries below the diagonal are updated too and never read.  The
      // padding (rows and columns past kb) is the identity, so its pivots
      // are 1 and its rows stay 0 off the diagonal.
      float c[kP];
      const int jl = lane < kP ? lane : kP - 1;
#pragma unroll
      for (int i = 0; i < kP; ++i) c[i] = D[i * kDLd + jl];
      int fail = 0;
      cb_static_for<0, kP>([&](auto T_) {
        constexpr int t = decltype(T_)::value;
        if (fail == 0 && t < kb) {
          const float dtt = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c[t]), t));
          if (!(dtt > 0.f)) {
            fail = k0 + t + 1;
          } else {
            const float utt = sqrtf(dtt);
            c[t] = lane > t ? c[t] / utt : (lane == t ? utt : c[t]);
            cb_static_for<t + 1, kP>([&](auto I_) {
              constexpr int i = decltype(I_)::value;
              const float uti = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c[t]), i));
              c[i] -= uti * c[t];
            });
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      });
      if (lane < kP) {
#pragma unroll
        for (int i = 0; i < kP; ++i)
          if (i <= lane) D[i * kDLd + lane] = c[i];
      }
      if (lane == 0 && fail) s_fail = fail;
    }
    __syncthreads();
    if (s_fail) break;
    for (int idx = tid; idx < kb * kb; idx += kThreads) {
      const int i = idx / kb, j = idx % kb;
      if (j >= i) A[(int64_t)(k0 + i) * lda + k0 + j] = D[i * kDLd + j];
    }
    if (rest <= 0) break;
    // (2) U12 = U11^-T A12: the row block is staged in the LDS panel
    // (coalesced), swept in place one column per thread, and written back.
    // (thread tid takes column c = tid + kThreads m: its kP loads go out
    // together, coalesced along the rows; the forward substitution with
    // U11^T runs in registers, D read from LDS by broadcast; rows past kb
    // are zero and the identity padding of D keeps them zero)
    for (int c0 = 0; c0 < rest; c0 += kThreads) {
      const int c = c0 + tid, cc = min(c, rest - 1);
      // D's offset, opaque per column pass: otherwise D's 528 entries are
      // loaded once before the loop and held in registers
      int doff = 0;
      asm volatile("" : "+v"(doff));
      const float* Dc = D + doff;
      float v[kP];
#pragma unroll
      for (int i = 0; i < kP; ++i) v[i] = A[(int64_t)(k0 + min(i, kb - 1)) * lda + k0 + kb + cc];
#pragma unroll
      for (int i = 0; i < kP; ++i) v[i] = i < kb ? v[i] : 0.f;
      cb_static_for<0, kP>([&](auto T_) {
        constexpr int t = decltype(T_)::value;
        const float xt = v[t] / Dc[t * kDLd + t];
        v[t] = xt;
        cb_static_for<t + 1, kP>([&](auto I_) {
          constexpr int i = decltype(I_)::value;
          v[i] -= Dc[t * kDLd + i] * xt;
        });
        // one row of D in flight at a time (hoisted, every row's broadcast
        // reads would be live together)
        __builtin_amdgcn_sched_barrier(0);
      });
      if (c < rest) {
#pragma unroll
        for (int i = 0; i < kP; ++i) {
          P[i * kPanelLd + c] = v[i];
          if (i < kb) A[(int64_t)(k0 + i) * lda + k0 + kb + c] = v[i];
        }
      }
    }
    __syncthreads();  // the update reads every thread's columns of P
    // (3) A22 -= U12^T U12 on the upper triangle in 16 x 16 tiles on the
    // matrix cores (v_mfma_f32_16x16x4f32, float32 products and sums; lane l:
    // A operand [l % 16][l / 16], B operand [l / 16][l % 16], result rows
    // 4 (l / 16) + i, column l % 16).  The upper tiles (bi <= bj), row by
    // row, go to the waves in turn; the K dimension is the panel's 32 rows
    // of U12 (rows past kb are zero).  Only entries on or above the diagonal
    // are read or written.  A tile's four loads are issued before the
    // previous tile's products (indices clamped onto the upper triangle;
    // entries outside it are masked when the tile is used), so a wave does
    // not wait for memory once per tile.
    const int T16 = (rest + 15) / 16;
    const int g = lane >> 4, q = lane & 15;
    const int base = k0 + kb;
    constexpr int kWaves = kThreads / 64;
    auto advance = [&](int& ti, int& tj, int step) {  // the step-th upper tile after (ti, tj)
      tj += step;
      while (ti < T16 && tj >= T16) {
        ++ti;
        tj = ti + (tj - T16);
      }
    };
    auto load_raw = [&](int ti, int tj, float (&v)[4]) {
      const int j = min(16 * tj + q, rest - 1);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = min(min(16 * ti + 4 * g + i, rest - 1), j);
        v[i] = A[(int64_t)(base + r) * lda + base + j];
      }
    };
    int bi = 0, bj = 0;
    advance(bi, bj, wv);
    float cur[4] = {0.f, 0.f, 0.f, 0.f};
    if (bi < T16) load_raw(bi, bj, cur);
    while (bi < T16) {
      int ni = bi, nj = bj;
      advance(ni, nj, kWaves);
      float nxt[4] = {0.f, 0.f, 0.f, 0.f};
      if (ni < T16) load_raw(ni, nj, nxt);
      const int j = 16 * bj + q;
      cb_f4 acc;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = 16 * bi + 4 * g + i;
        acc[i] = (j < rest && j >= r) ? cur[i] : 0.f;
      }
#pragma unroll
      for (int t0 = 0; t0 < kP; t0 += 4) {
        const float* Pt = P + (t0 + g) * kPanelLd;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(-Pt[16 * bi + q], Pt[j], acc, 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = 16 * bi + 4 * g + i;
        if (j < rest && j >= r) A[(int64_t)(base + r) * lda + base + j] = acc[i];
      }
      bi = ni;
      bj = nj;
#pragma unroll
      for (int i = 0; i < 4; ++i) cur[i] = nxt[i];
    }
    __syncthreads();
  }
  __syncthreads();
  const int fail = s_fail;
  if (fail) {
    const float nan = quiet_nan<float>();
    for (int64_t t = tid; t < (int64_t)n * n; t += kThreads) A[(t / n) * lda + (t % n)] = nan;
  }
  if (tid == 0) info[blockIdx.x] = fail;
}

// Solves U^T U x = b for one right-hand side per system (in place in b).
__global__ __launch_bounds__(kThreads) void chol_blocked_solve_kernel(const lxa_qr_desc* __restrict__ descs, int n_all,
                                                                      int64_t a_stride, int64_t b_stride,
                                                                      const float* __restrict__ a,
                                                                      float* __restrict__ b) {
  __shared__ float y[kMaxN];
  __shared__ float red[kThreads / kWave][kP];
  __shared__ float part[8][kP];
  const Geometry g = geometry(descs, n_all, a_stride, b_stride);
  const int n = g.n, lda = g.lda;
  if (n <= 0 || n > kMaxN) return;
  const float* U = a + g.a_off;
  float* B = b + g.b_off;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int i = tid; i < n; i += kThreads) y[i] = B[i];
  __syncthreads();
  const int nblk = (n + kP - 1) / kP;
  // forward: U^T y = b
  for (int q = 0; q < nblk; ++q) {
    const int k0 = q * kP, kb = min(kP, n - k0);
    {  // thread (column i = tid % 32, group gp = tid / 32) over rows t < k0
      const int i = tid & (kP - 1), gp = tid >> 5;
      float s = 0.f;
      if (i < kb) {
#pragma unroll 8
        for (int t = gp; t < k0; t += 8) s += U[(int64_t)t * lda + k0 + i] * y[t];
      }
      part[gp][i] = s;
    }
    __syncthreads();
    if (wv == 0) {
      // lane l holds column l of the diagonal block (rows t <= l; the
      // indices are clamped onto the upper triangle), its loads issued
      // together instead of one per step
      const int lc = min(lane, kb - 1);
      float col[kP];
#pragma unroll
      for (int t = 0; t < kP; ++t) col[t] = U[(int64_t)(k0 + min(min(t, kb - 1), lc)) * lda + k0 + lc];
      float z = 0.f;
      if (lane < kb) {
        float s = 0.f;
#pragma unroll
        for (int gp = 0; gp < 8; ++gp) s += part[gp][lane];
        z = y[k0 + lane] - s;
      }
#pragma unroll
      for (int t = 0; t < kP; ++t) {
        if (t < kb) {
          if (lane == t) z = z / col[t];
          const float yt = __shfl(z, t, kWave);
          if (lane > t && lane < kb) z -= col[t] * yt;
        }
      }
      if (lane < kb) y[k0 + lane] = z;
    }
    __syncthreads();
  }
  // backward: U x = y
  for (int q = nblk - 1; q >= 0; --q) {
    const int k0 = q * kP, kb = min(kP, n - k0), k1 = k0 + kb;
    for (int i = wv; i < kb; i += kThreads / kWave) {
      float s = 0.f;
      const float* Ur = U + (int64_t)(k0 + i) * lda;
#pragma unroll 8
      for (int t = k1 + lane; t < n; t += kWave) s += Ur[t] * y[t];
      s = wave_sum(s);
      if (lane == 0) red[0][i] = s;
    }
    __syncthreads();
    if (wv == 0) {
      // lane l holds row l of the diagonal block (columns t >= l), loaded
      // together as above
      const int lr = min(lane, kb - 1);
      float row[kP];
#pragma unroll
      for (int t = 0; t < kP; ++t) row[t] = U[(int64_t)(k0 + lr) * lda + k0 + max(min(t, kb - 1), lr)];
      float z = lane < kb ? y[k0 + lane] - red[0][lane] : 0.f;
#pragma unroll
      for (int t = kP - 1; t >= 0; --t) {
        if (t < kb) {
          if (lane == t) z = z / row[t];
          const float xt = __shfl(z, t, kWave);
          if (lane < t) z -= row[t] * xt;
        }
      }
      if (lane < kb) y[k0 + lane] = z;
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += kThreads) B[i] = y[i];
}

size_t factor_smem() { return sizeof(float) * (kP * kDLd + kP * kPanelLd); }

}  // namespace

int chol_blocked_factor(hipStream_t st, int64_t count, const lxa_qr_desc* descs, int n_all, int64_t a_stride, float* a,
                        int32_t* info) {
  if (count <= 0) return LXA_OK;
  const size_t smem = factor_smem();
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&chol_blocked_factor_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) return hip_error(e, "hipFuncSetAttribute(chol_blocked_factor_kernel)");
  chol_blocked_factor_kernel<<<(unsigned)count, kThreads, smem, st>>>(descs, n_all, a_stride, a, info);
  LXA_LAUNCH_CHECK("chol_blocked_factor_kernel");
  return LXA_OK;
}

int chol_blocked_solve(hipStream_t st, int64_t count, const lxa_qr_desc* descs, int n_all, int64_t a_stride,
                       int64_t b_stride, const float* a, float* b) {
  if (count <= 0) return LXA_OK;
  chol_blocked_solve_kernel<<<(unsigned)count, kThreads, 0, st>>>(descs, n_all, a_stride, b_stride, a, b);
  LXA_LAUNCH_CHECK("chol_blocked_solve_kernel");
  return LXA_OK;
}

}  // namespace lxa

extern "C" int lxa_chol_blocked_factor(void* stream, int64_t count, const lxa_qr_desc* desc, float* a,
                                       int32_t* info) {
  LXA_REQUIRE(count >= 0 && count <= 0x7fffffff, "count out of range");
  if (count == 0) return LXA_OK;
  LXA_REQUIRE(desc && a && info, "desc, a and info must be non-NULL");
  return lxa::chol_blocked_factor(static_cast<hipStream_t>(stream), count, desc, 0, 0, a, info);
}

extern "C" int lxa_chol_blocked_solve(void* stream, int64_t count, const lxa_qr_desc* desc, const float* u,
                                      float* b) {
  LXA_REQUIRE(count >= 0 && count <= 0x7fffffff, "count out of range");
  if (count == 0) return LXA_OK;
  LXA_REQUIRE(desc && u && b, "desc, u and b must be non-NULL");
  return lxa::chol_blocked_solve(static_cast<hipStream_t>(stream), count, desc, 0, 0, 0, u, b);
}
