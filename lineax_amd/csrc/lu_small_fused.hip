// Batched LU solve of small systems (n <= 64) in one launch: getrf with
// partial pivoting and getrs fused, one wavefront per system.
//
// Lane r holds row r of [A | b] in registers (rows n..N-1 are identity rows,
// which never win a pivot search over a real column).  Step k finds the
// pivot with a wave maximum of |a(i,k)| and a ballot (ties to the smallest
// LAPACK position, as idamax), broadcasts the pivot row with readlane (the
// pivot lane is a scalar), and eliminates below it; the right-hand side is eliminated with the same multipliers, so when
// the factorisation ends every lane holds its entry of y = L^-1 P b.  Back
// substitution then runs over the positions in reverse, again by readlane.
// Row interchanges are implicit: each lane tracks the LAPACK position of its
// row, exactly as lxa_lu_factor's small kernel does, so the factors, the
// interchange sequence and the solution equal lxa_lu_factor + lxa_lu_solve's
// bit for bit (the same operations in the same order).  A is only read; the
// factors and pivots are written only when the caller asks for them (the
// solver state).  When n is the tile order, A and the factors move between
// HBM and registers through LDS in contiguous 16-byte chunks.
#include <climits>

#include "lxa_common.h"

namespace lxa {
namespace {

__device__ __forceinline__ uint32_t mag_key(float v) { return __float_as_uint(v) & 0x7fffffffu; }
__device__ __forceinline__ uint32_t mag_key(double v) {
  // |v| ordered as unsigned: the top 32 bits of the magnitude, then the rest
  // compared only on ties (second pass below)
  return static_cast<uint32_t>((static_cast<uint64_t>(__double_as_longlong(v)) & 0x7fffffffffffffffull) >> 32);
}
__device__ __forceinline__ uint32_t mag_lo(float) { return 0u; }
__device__ __forceinline__ uint32_t mag_lo(double v) {
  return static_cast<uint32_t>(static_cast<uint64_t>(__double_as_longlong(v)));
}

template <typename T>
__device__ __forceinline__ T lane_value(T v, int lane) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
  } else {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
  }
}

// Wave maximum of unsigned 32-bit values over the first N lanes: DPP
// quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror and row_mirror
// leave each 16-lane row's maximum in all of its lanes (register crossbar
// moves, no LDS round trip); the rows are combined from scalar reads.  The
// result is wave-uniform.
template <int N>
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0xB1, 0xF, 0xF, false)));
  v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x4E, 0xF, 0xF, false)));
  if constexpr (N > 4) v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x141, 0xF, 0xF, false)));
  if constexpr (N > 8) v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x140, 0xF, 0xF, false)));
  uint32_t best = 0;
#pragma unroll
  for (int q = 0; q < (N + 15) / 16; ++q) {
    const uint32_t h = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 16 * q));
    best = h > best ? h : best;
  }
  return best;
}

__device__ __forceinline__ uint64_t wave_ballot(bool pred) { return __builtin_amdgcn_ballot_w64(pred); }

// The pivot lane of column k: the largest |a(i,k)| among rows not yet used,
// ties to the smallest LAPACK position (idamax's rule).  The magnitude's bit
// pattern orders as an unsigned integer; the common case (one lane holds the
// maximum) costs one 32-bit maximum and a ballot.  Doubles compare their
// high words first and the low words only among lanes tied on them.
template <typename T, int N>
__device__ __forceinline__ int pivot_lane(T v, bool done, int pos, int r) {
  const uint32_t hi = done ? 0u : mag_key(v);
  const uint32_t top = wave_max_u32<N>(hi);
  uint64_t cand = wave_ballot(!done && hi == top);
  if constexpr (sizeof(T) == 8) {
    if (__builtin_popcountll(cand) > 1) {
      const bool c = (cand >> r) & 1ull;
      const uint32_t lo = c ? mag_lo(v) : 0u;
      const uint32_t best = wave_max_u32<N>(lo);
      cand = wave_ballot(c && lo == best);
    }
  }
  if (__builtin_popcountll(cand) > 1) {
    const bool c = (cand >> r) & 1ull;
    const uint32_t rank = c ? static_cast<uint32_t>(64 - pos) : 0u;
    const uint32_t best = wave_max_u32<N>(rank);
    cand = wave_ballot(c && rank == best);
  }
  return __builtin_ctzll(cand);
}

// Swizzled LDS tile of one N x N system, in 16-byte chunks: chunk c of row i
// sits at slot i * kCpr + (c ^ (i mod kCpr)), so a row read by one lane per
// row and a contiguous run read by consecutive lanes both spread over the
// banks.
template <typename T, int N>
struct Tile {
  static constexpr int kVec = 16 / static_cast<int>(sizeof(T));
  static constexpr int kCpr = N / kVec;  // chunks per row (a power of two)
  static constexpr int kChunks = N * kCpr;
  __device__ static int slot(int i, int c) { return i * kCpr + (c ^ (i & (kCpr - 1))); }
};

// x = A^-1 b for `batch` systems of order n <= N.  lu/piv may be null.
// vec_in / vec_out: n == N and the operator (factor) storage is 16-byte
// aligned per system, so it moves through the swizzled LDS tile in
// contiguous 16-byte chunks instead of one row per lane.
template <typename T, int N>
__global__ __launch_bounds__(64) void lu_small_fused_kernel(int64_t batch, int n, const T* __restrict__ a,
                                                            int64_t a_stride, const T* __restrict__ b,
                                                            int64_t b_stride, T* __restrict__ x, int64_t x_stride,
                                                            T* __restrict__ lu, int32_t* __restrict__ piv,
                                                            int32_t* __restrict__ info, int vec_in, int vec_out) {
  using Tl = Tile<T, N>;
  using V = __attribute__((ext_vector_type(4))) uint32_t;
  __shared__ V tile[Tl::kChunks];
  __shared__ int lane_at[N];
  const int r = threadIdx.x;
  const int64_t sys = blockIdx.x;
  if (sys >= batch) return;
  const bool real_row = r < n;
  const bool owns_row = r < N;
  T row[N];
  T bv = T(0);
  if (vec_in) {
    const V* src = reinterpret_cast<const V*>(a + sys * a_stride);
#pragma unroll
    for (int g0 = 0; g0 < Tl::kChunks; g0 += 64) {
      const int g = g0 + r;
      if (Tl::kChunks % 64 == 0 || g < Tl::kChunks) tile[Tl::slot(g / Tl::kCpr, g % Tl::kCpr)] = src[g];
    }
    if (real_row) bv = b[sys * b_stride + r];
    __syncthreads();
    if (owns_row) {
#pragma unroll
      for (int c = 0; c < Tl::kCpr; ++c) {
        const V v = tile[Tl::slot(r, c)];
        const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
        for (int u = 0; u < Tl::kVec; ++u) row[c * Tl::kVec + u] = e[u];
      }
    }
  } else if (real_row) {
    const T* src = a + sys * a_stride + static_cast<int64_t>(r) * n;
#pragma unroll
    for (int j = 0; j < N; ++j) row[j] = (j < n) ? src[j] : T(0);
    bv = b[sys * b_stride + r];
  }
  if (!real_row) {
#pragma unroll
    for (int j = 0; j < N; ++j) row[j] = (j == r) ? T(1) : T(0);
  }
  int pos = r;
  bool done = !owns_row;
  int mypiv = r;
  int zero_at = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int p = pivot_lane<T, N>(row[k], done, pos, r);
    const int cp = __builtin_amdgcn_readlane(pos, p);  // the pivot row's position
    const bool is_piv = r == p;
    const T pval = lane_value(row[k], p);
    const T pb = lane_value(bv, p);
    if (r == k) mypiv = cp;
    if (k < n && pval == T(0) && zero_at == 0) zero_at = k + 1;
    if (r == 0) lane_at[k] = p;
    if (is_piv) {
      pos = k;
      done = true;
    } else if (!done && pos == k) {
      pos = cp;  // the row at position k moves to the pivot's old position
    }
    // every lane computes the update (the pivot row's entries are scalars);
    // rows already pivoted keep their values by selection, not by a branch
    const T l = (pval != T(0)) ? row[k] / pval : row[k];
#pragma unroll
    for (int j = k + 1; j < N; ++j) {
      const T upd = row[j] - l * lane_value(row[j], p);
      row[j] = done ? row[j] : upd;
    }
    bv = done ? bv : bv - l * pb;
    row[k] = done ? row[k] : l;
  }
  __syncthreads();  // lane_at (one wave)
  // U x = y over positions N-1..0; lane_at[k] holds row k of U and y_k
#pragma unroll
  for (int k = N - 1; k >= 0; --k) {
    const int p = __builtin_amdgcn_readfirstlane(lane_at[k]);
    const T t = bv / row[k];
    const T xk = lane_value(t, p);
    if (pos < k) {
      bv -= row[k] * xk;
    } else if (pos == k) {
      bv = xk;
    }
  }
  // real rows end at real positions (padding rows never pivot for k < n)
  if (real_row) {
    x[sys * x_stride + pos] = bv;
    if (piv) piv[sys * n + r] = mypiv;
  }
  if (r == 0 && info) info[sys] = zero_at;
  if (!lu) return;
  if (vec_out) {
    if (owns_row) {
#pragma unroll
      for (int c = 0; c < Tl::kCpr; ++c) {
        V v;
        T* e = reinterpret_cast<T*>(&v);
#pragma unroll
        for (int u = 0; u < Tl::kVec; ++u) e[u] = row[c * Tl::kVec + u];
        tile[Tl::slot(pos, c)] = v;
      }
    }
    __syncthreads();
    V* dst = reinterpret_cast<V*>(lu + sys * static_cast<int64_t>(N) * N);
#pragma unroll
    for (int g0 = 0; g0 < Tl::kChunks; g0 += 64) {
      const int g = g0 + r;
      if (Tl::kChunks % 64 == 0 || g < Tl::kChunks) dst[g] = tile[Tl::slot(g / Tl::kCpr, g % Tl::kCpr)];
    }
  } else if (real_row) {
    T* dst = lu + sys * static_cast<int64_t>(n) * n + static_cast<int64_t>(pos) * n;
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (j < n) dst[j] = row[j];
  }
}

template <typename T, int N>
int launch_fused(hipStream_t s, int64_t batch, int n, const T* a, int64_t as, const T* b, int64_t bs, T* x,
                 int64_t xs, T* lu, int32_t* piv, int32_t* info) {
  // the tile path needs whole 16-byte chunks per system: n == N and aligned
  // storage (N * N * sizeof(T) is a multiple of 16 for every N used here)
  const int vec_in = n == N && (reinterpret_cast<uintptr_t>(a) & 15) == 0 && (as * sizeof(T)) % 16 == 0;
  const int vec_out = lu != nullptr && n == N && (reinterpret_cast<uintptr_t>(lu) & 15) == 0;
  hipLaunchKernelGGL((lu_small_fused_kernel<T, N>), dim3(static_cast<unsigned>(batch)), dim3(64), 0, s, batch, n, a,
                     as, b, bs, x, xs, lu, piv, info, vec_in, vec_out);
  LXA_LAUNCH_CHECK("lu_small_fused_kernel launch");
  return LXA_OK;
}

template <typename T>
int fused_t(hipStream_t s, int64_t batch, int n, const T* a, int64_t as, const T* b, int64_t bs, T* x, int64_t xs,
            T* lu, int32_t* piv, int32_t* info) {
  if (n <= 8) return launch_fused<T, 8>(s, batch, n, a, as, b, bs, x, xs, lu, piv, info);
  if (n <= 16) return launch_fused<T, 16>(s, batch, n, a, as, b, bs, x, xs, lu, piv, info);
  if (n <= 32) return launch_fused<T, 32>(s, batch, n, a, as, b, bs, x, xs, lu, piv, info);
  return launch_fused<T, 64>(s, batch, n, a, as, b, bs, x, xs, lu, piv, info);
}

}  // namespace
}  // namespace lxa

extern "C" int lxa_lu_small_solve(void* stream, int dtype, int64_t batch, int64_t n, const void* a,
                                  int64_t a_batch_stride, const void* b, int64_t b_batch_stride, void* x,
                                  int64_t x_batch_stride, void* lu, int32_t* piv, int32_t* info) {
  LXA_REQUIRE(batch >= 0 && batch <= 2147483647LL, "lxa_lu_small_solve: bad batch %lld", (long long)batch);
  LXA_REQUIRE(n >= 1 && n <= 64, "lxa_lu_small_solve: n must be in [1, 64] (got %lld)", (long long)n);
  LXA_REQUIRE(a && b && x, "lxa_lu_small_solve: null pointer");
  LXA_REQUIRE((lu == nullptr) == (piv == nullptr), "lxa_lu_small_solve: lu and piv go together");
  LXA_REQUIRE(a_batch_stride >= 0 && b_batch_stride >= 0 && x_batch_stride >= n,
              "lxa_lu_small_solve: bad strides");
  if (batch == 0) return LXA_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nn = static_cast<int>(n);
  if (dtype == LXA_F32)
    return lxa::fused_t<float>(s, batch, nn, static_cast<const float*>(a), a_batch_stride,
                               static_cast<const float*>(b), b_batch_stride, static_cast<float*>(x), x_batch_stride,
                               static_cast<float*>(lu), piv, info);
  if (dtype == LXA_F64)
    return lxa::fused_t<double>(s, batch, nn, static_cast<const double*>(a), a_batch_stride,
                                static_cast<const double*>(b), b_batch_stride, static_cast<double*>(x),
                                x_batch_stride, static_cast<double*>(lu), piv, info);
  return lxa::set_error(LXA_EINVAL, "lxa_lu_small_solve: bad dtype %d", dtype);
}
