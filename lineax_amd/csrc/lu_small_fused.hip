// Batched LU solve of small systems (n <= 64) in one launch: getrf with
// partial pivoting and getrs fused, one wavefront per system.
//
// Lane r holds row r of [A | b] in registers (rows n..N-1 are identity rows,
// which never win a pivot search over a real column).  Step k finds the
// pivot with one wave-wide reduction of (|a(i,k)|, position) packed into 64
// bits (ties to the smallest LAPACK position, as idamax), broadcasts the
// pivot row with readlane (the pivot lane is a scalar), and eliminates below
// it; the right-hand side is eliminated with the same multipliers, so when
// the factorisation ends every lane holds its entry of y = L^-1 P b.  Back
// substitution then runs over the positions in reverse, again by readlane.
// Row interchanges are implicit: each lane tracks the LAPACK position of its
// row, exactly as lxa_lu_factor's small kernel does, so the factors, the
// interchange sequence and the solution equal lxa_lu_factor + lxa_lu_solve's
// bit for bit (the same operations in the same order).  A is only read; the
// factors and pivots are written only when the caller asks for them (the
// solver state).
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

// One step of a wave maximum of 64-bit keys with a DPP partner (register
// crossbar moves: no LDS round trip, unlike a shuffle).
template <int kCtrl>
__device__ __forceinline__ void key_max_dpp(uint32_t& hi, uint32_t& lo) {
  const uint32_t oh = (uint32_t)__builtin_amdgcn_mov_dpp((int)hi, kCtrl, 0xF, 0xF, false);
  const uint32_t ol = (uint32_t)__builtin_amdgcn_mov_dpp((int)lo, kCtrl, 0xF, 0xF, false);
  const bool gt = oh > hi || (oh == hi && ol > lo);
  hi = gt ? oh : hi;
  lo = gt ? ol : lo;
}

// Wave maximum: quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror
// and row_mirror leave each 16-lane row's maximum in all of its lanes; the
// rows (the first N / 16 hold every real row) are combined from scalar reads.
// The result is wave-uniform.
template <int N>
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
  uint32_t hi = static_cast<uint32_t>(v >> 32), lo = static_cast<uint32_t>(v);
  key_max_dpp<0xB1>(hi, lo);
  key_max_dpp<0x4E>(hi, lo);
  if constexpr (N > 4) key_max_dpp<0x141>(hi, lo);
  if constexpr (N > 8) key_max_dpp<0x140>(hi, lo);
  uint64_t best = 0;
#pragma unroll
  for (int q = 0; q < (N + 15) / 16; ++q) {
    const uint32_t h = (uint32_t)__builtin_amdgcn_readlane((int)hi, 16 * q);
    const uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)lo, 16 * q);
    const uint64_t k = (static_cast<uint64_t>(h) << 32) | l;
    best = k > best ? k : best;
  }
  return best;
}

// x = A^-1 b for `batch` systems of order n <= N.  lu/piv may be null.
template <typename T, int N>
__global__ __launch_bounds__(64) void lu_small_fused_kernel(int64_t batch, int n, const T* __restrict__ a,
                                                            int64_t a_stride, const T* __restrict__ b,
                                                            int64_t b_stride, T* __restrict__ x, int64_t x_stride,
                                                            T* __restrict__ lu, int32_t* __restrict__ piv,
                                                            int32_t* __restrict__ info) {
  __shared__ int lane_at[N];
  const int r = threadIdx.x;
  const int64_t sys = blockIdx.x;
  if (sys >= batch) return;
  const bool real_row = r < n;
  const bool owns_row = r < N;
  T row[N];
  T bv = T(0);
  if (real_row) {
    const T* src = a + sys * a_stride + static_cast<int64_t>(r) * n;
#pragma unroll
    for (int j = 0; j < N; ++j) row[j] = (j < n) ? src[j] : T(0);
    bv = b[sys * b_stride + r];
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) row[j] = (j == r) ? T(1) : T(0);
  }
  int pos = r;
  bool done = !owns_row;
  int mypiv = r;
  int zero_at = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    // pivot: largest |a(i,k)| among rows not yet used, smallest position on ties
    const uint32_t hi = done ? 0u : mag_key(row[k]);
    // low word: position (smallest wins ties), then the lane (a pure label)
    const uint32_t lo_pos = done ? 0u : ((static_cast<uint32_t>(63 - pos) << 8) | static_cast<uint32_t>(r));
    uint64_t key = (static_cast<uint64_t>(hi) << 32) | lo_pos;
    key = wave_max_u64<N>(key);
    if constexpr (sizeof(T) == 8) {
      // doubles: break ties of the top 32 magnitude bits on the low 32 bits
      const uint32_t top = static_cast<uint32_t>(key >> 32);
      const bool cand = !done && mag_key(row[k]) == top;
      uint64_t k2 = cand ? ((static_cast<uint64_t>(mag_lo(row[k])) << 32) |
                            ((static_cast<uint32_t>(63 - pos) << 8) | static_cast<uint32_t>(r)))
                         : 0;
      k2 = wave_max_u64<N>(k2);
      key = (static_cast<uint64_t>(top) << 32) | static_cast<uint32_t>(k2);
    }
    const int cp = 63 - static_cast<int>(static_cast<uint32_t>(key) >> 8);
    const int p = static_cast<int>(static_cast<uint32_t>(key) & 0xffu);  // the pivot's lane
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
  if (real_row) {
    // real rows end at real positions (padding rows never pivot for k < n)
    x[sys * x_stride + pos] = bv;
    if (lu) {
      T* dst = lu + sys * static_cast<int64_t>(n) * n + static_cast<int64_t>(pos) * n;
#pragma unroll
      for (int j = 0; j < N; ++j)
        if (j < n) dst[j] = row[j];
    }
    if (piv) piv[sys * n + r] = mypiv;
  }
  if (r == 0 && info) info[sys] = zero_at;
}

template <typename T, int N>
int launch_fused(hipStream_t s, int64_t batch, int n, const T* a, int64_t as, const T* b, int64_t bs, T* x,
                 int64_t xs, T* lu, int32_t* piv, int32_t* info) {
  hipLaunchKernelGGL((lu_small_fused_kernel<T, N>), dim3(static_cast<unsigned>(batch)), dim3(64), 0, s, batch, n, a,
                     as, b, bs, x, xs, lu, piv, info);
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
