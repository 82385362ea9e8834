// Shared helpers for the lineax_amd HIP kernels (gfx950 / CDNA4, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/lineax_amd.h"

namespace lxa {

// Records a printf-style message in the thread-local last-error slot that
// lxa_last_error() returns. Always returns `code`.
int set_error(int code, const char* fmt, ...);

// Converts a HIP error into LXA_EHIP with the HIP error string attached.
int hip_error(hipError_t e, const char* what);

// Number of compute units of the current device (cached per device).
int num_cus();

constexpr int kWave = 64;

// C <- beta C + sign A B, float32 storage, products in the precision `mode`
// (LXA_GEMM_*); see gemm.hip.
int gemm_update(hipStream_t st, int mode, int M, int N, int K, const float* A, int64_t lda, const float* B,
                int64_t ldb, float* C, int64_t ldc, int beta, float sign);

template <typename T>
__device__ __forceinline__ T quiet_nan();
template <>
__device__ __forceinline__ float quiet_nan<float>() {
  return __builtin_nanf("");
}
template <>
__device__ __forceinline__ double quiet_nan<double>() {
  return __builtin_nan("");
}

// Orders LDS traffic between lanes of one wavefront.  LDS instructions of a
// wave execute in issue order; this only stops the compiler from moving
// loads/stores across the point.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroup barrier that orders LDS traffic only.  __syncthreads() also
// fences global memory, which makes every wave wait for all of its global
// loads before the barrier; loads prefetched for later rows stay in flight
// across this one.  Use only where no thread reads global data another
// thread of the block wrote before the barrier.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

template <int K>
struct Log2 {
  static constexpr int value = 1 + Log2<K / 2>::value;
};
template <>
struct Log2<1> {
  static constexpr int value = 0;
};

// The value of lane (this lane xor OFF), for a power-of-two OFF < 64, by
// register crossbar moves instead of an LDS exchange: v_permlane32_swap /
// v_permlane16_swap for 32 and 16 (swapping a register with itself leaves
// the values of lanes l & ~OFF and l | OFF), DPP row_ror:8 for 8, row_shr:4 /
// row_shl:4 for 4, quad_perm for 2 and 1.  Every lane of the wave must be
// active (the moves read source lanes whatever EXEC says).
template <int OFF>
__device__ __forceinline__ unsigned lane_xor_u32(unsigned v) {
  static_assert(OFF == 32 || OFF == 16 || OFF == 8 || OFF == 4 || OFF == 2 || OFF == 1, "power of two < 64");
  const int lane = __lane_id();
  if constexpr (OFF == 32) {
    const auto s = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane & 32) ? s[0] : s[1];
  } else if constexpr (OFF == 16) {
    const auto s = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane & 16) ? s[0] : s[1];
  } else if constexpr (OFF == 8) {
    return __builtin_amdgcn_update_dpp(0u, v, 0x128, 0xF, 0xF, false);  // row_ror:8
  } else if constexpr (OFF == 4) {
    const unsigned from_below = __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xF, 0xF, false);  // row_shr:4
    const unsigned from_above = __builtin_amdgcn_update_dpp(0u, v, 0x104, 0xF, 0xF, false);  // row_shl:4
    return (lane & 4) ? from_below : from_above;
  } else if constexpr (OFF == 2) {
    return __builtin_amdgcn_update_dpp(0u, v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
  } else {
    return __builtin_amdgcn_update_dpp(0u, v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
  }
}

template <int OFF, typename T>
__device__ __forceinline__ T lane_xor(T v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, lane_xor_u32<OFF>(__builtin_bit_cast(unsigned, v)));
  } else {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = lane_xor_u32<OFF>(static_cast<unsigned>(u));
    const unsigned hi = lane_xor_u32<OFF>(static_cast<unsigned>(u >> 32));
    return __builtin_bit_cast(T, (static_cast<unsigned long long>(hi) << 32) | lo);
  }
}

// One halving level of wave_reduce_scatter, and the levels after it: lanes
// with bit `off` set keep the upper half of v and add their partner's upper
// half; the others keep the lower half.  The level count is a template
// parameter so that every register index is a compile-time constant (a
// runtime loop bound here turns each v[i + half] into a select chain over
// all of v).
template <int L, int K, typename R>
__device__ __forceinline__ void reduce_scatter_levels(R (&v)[K], int lane) {
  constexpr int half = K >> (L + 1);
  if constexpr (half >= 1) {
    constexpr int off = 32 >> L;
    const bool upper = (lane & off) != 0;
#pragma unroll
    for (int i = 0; i < half; ++i) {
      const R keep = upper ? v[i + half] : v[i];
      const R send = upper ? v[i] : v[i + half];
      v[i] = keep + lane_xor<off>(send);
    }
    reduce_scatter_levels<L + 1, K>(v, lane);
  }
}

// Sums K values (K a power of two <= 64) over the 64 lanes of a wave.  On
// return, every lane holds the complete wave sum of value number
// lane >> (6 - log2 K).  v is destroyed.  Every lane must be active.
template <int K, typename R>
__device__ __forceinline__ R wave_reduce_scatter(R (&v)[K], int lane) {
  static_assert(K >= 1 && K <= 64 && (K & (K - 1)) == 0, "K must be a power of two <= 64");
  reduce_scatter_levels<0, K>(v, lane);
  R x = v[0];
  // the remaining butterfly levels, offsets 32 / K down to 1
  if constexpr (32 / K >= 32) x += lane_xor<32>(x);
  if constexpr (32 / K >= 16) x += lane_xor<16>(x);
  if constexpr (32 / K >= 8) x += lane_xor<8>(x);
  if constexpr (32 / K >= 4) x += lane_xor<4>(x);
  if constexpr (32 / K >= 2) x += lane_xor<2>(x);
  if constexpr (32 / K >= 1) x += lane_xor<1>(x);
  return x;
}

// Index of element (i, j), j >= i, in a row-packed upper triangle of order n.
__device__ __forceinline__ int packed_upper(int n, int i, int j) {
  return i * n - ((i * (i - 1)) >> 1) + (j - i);
}

// Register-tiled Cholesky factor of a uniform float32 batch, orders 1..256
// (chol_uneven.hip); same layout, NaN-on-failure and info conventions as
// lxa_chol_factor.
int chol_factor_regs_f32(hipStream_t s, int64_t batch, int n, float* a, int64_t stride, int negate, int32_t* info);

}  // namespace lxa

#define LXA_REQUIRE(cond, ...)                                   \
  do {                                                           \
    if (!(cond)) return ::lxa::set_error(LXA_EINVAL, __VA_ARGS__); \
  } while (0)

#define LXA_LAUNCH_CHECK(what)                                   \
  do {                                                           \
    hipError_t lxa_e_ = hipGetLastError();                       \
    if (lxa_e_ != hipSuccess) return ::lxa::hip_error(lxa_e_, what); \
  } while (0)
