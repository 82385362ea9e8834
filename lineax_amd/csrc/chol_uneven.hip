// Register-tiled float32 Cholesky factor for uniform batches of order
// 65..256 (lxa_chol_factor), and the entry points of the uneven-batch solve
// (whose kernels are in chol_mfma.hip): the workspace and no-allocation
// variants, the autotuner's hook and the autotune table's query.
//
// The factor kernel: one 256-thread workgroup (a 16 x 16 thread grid, four
// waves) per matrix.  The trailing matrix lives in registers, distributed
// cyclically: thread (ty, tx) holds A(ty + 16a, tx + 16b) for a <= b, i.e.
// the upper triangle in up to 136 registers.  The factorisation runs in
// 16-row panels:
//   (a) the owners copy panel rows K into an LDS triangle T (row-packed,
//       order N16 = n rounded up to 16; padding rows are the identity);
//   (b) wave 0 factors the 16x16 diagonal block in registers (lane q holds
//       column q; broadcasts by readlane) and stores U_KK in T; then every
//       wave solves U_KK^T U_K,J = A_K,J for its own 64-column chunk of the
//       panel (U_KK read by LDS broadcast) and writes the result to T and to
//       a transposed, bank-conflict-free copy PT[j][p];
//   (c) all threads apply the rank-16 update to their register slots below
//       the panel, reading U values from PT two at a time.
// Three barriers per panel.  T ends up holding U, which is written out.
#include <mutex>

#include "lxa_common.h"
#include "uneven_tune.h"

namespace lxa {
namespace {

constexpr int kT = 256;
#ifndef LXA_SMALL_WAVES
#define LXA_SMALL_WAVES 3
#endif
constexpr int kMaxN = 256;
constexpr int kPTStride = 18;  // 16 panel values + 2 pad: b64 reads hit distinct banks
constexpr size_t kRinvBytes = 256 * sizeof(float);
constexpr size_t kItemBytes = 16;

__host__ __device__ inline int round16(int n) { return (n + 15) & ~15; }

// LDS: the transposed panel PT, the reciprocal pivots, a flag word and the
// packed triangle T, all sized by the order.
__host__ __device__ inline size_t panel_bytes(int n) {
  return static_cast<size_t>(round16(n)) * kPTStride * sizeof(float);
}

__host__ __device__ inline size_t uneven_lds_bytes(int n) {
  const size_t n16 = static_cast<size_t>(round16(n));
  return panel_bytes(n) + kRinvBytes + kItemBytes + n16 * (n16 + 1) / 2 * sizeof(float);
}

__device__ __forceinline__ float bcast(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ int pku(int n, int i, int j) { return i * n - ((i * (i - 1)) >> 1) + (j - i); }

struct Lds {
  float* PT;    // [N16][18] transposed panel
  float* rinv;  // [256] 1 / U_kk
  int* item;    // flag words
  float* T;     // packed U, order N16
};

__device__ Lds carve(unsigned char* base, int n) {
  const size_t pb = panel_bytes(n);
  Lds s;
  s.PT = reinterpret_cast<float*>(base);
  s.rinv = reinterpret_cast<float*>(base + pb);
  s.item = reinterpret_cast<int*>(base + pb + kRinvBytes);
  s.T = reinterpret_cast<float*>(base + pb + kRinvBytes + kItemBytes);
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
  // every load unconditional (outside the upper triangle from element 0, then
  // discarded), so that all of them are in flight before the first use; in
  // branches the compiler waited for each one
#pragma unroll
  for (int aa = 0; aa < KR; ++aa) {
#pragma unroll
    for (int bb = aa; bb < KR; ++bb) {
      const int i = ty + 16 * aa, j = tx + 16 * bb;
      const bool in = aa < nk && bb < nk && i < n && j < n && j >= i;
      R[aa][bb] = A[in ? static_cast<int64_t>(i) * n + j : 0];
    }
  }
#pragma unroll
  for (int aa = 0; aa < KR; ++aa) {
#pragma unroll
    for (int bb = aa; bb < KR; ++bb) {
      const int i = ty + 16 * aa, j = tx + 16 * bb;
      const bool tile = aa < nk && bb < nk;
      const bool in = tile && i < n && j < n && j >= i;
      const float pad = (tile && (i >= n || j >= n) && i == j) ? 1.0f : 0.0f;  // identity padding
      R[aa][bb] = in ? (negate ? -R[aa][bb] : R[aa][bb]) : pad;
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

template <typename K>
int allow_lds(K kernel, size_t bytes) {
  if (bytes <= 64 * 1024) return LXA_OK;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
  return e == hipSuccess ? LXA_OK : hip_error(e, "hipFuncSetAttribute(MaxDynamicSharedMemorySize)");
}

constexpr int64_t kMaxGrid = 2147483647;

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

int chol_mfma_launch(hipStream_t s, int64_t count, int max_n, const int32_t* n, const int64_t* a_offset,
                     const float* a, const int64_t* b_offset, const float* b, void* x, int x_dtype,
                     const int32_t* order, bool persistent, bool sorted, int max_refine, double tol, int32_t* info,
                     int32_t* iters, double* eta, unsigned int* queue, int max4, int max8);

namespace {

// Work-queue counters of the persistent launches: one per tile class, and the
// flag of the sortedness check.
constexpr int kQueueCounters = 4;
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

// The launches of one uneven solve (chol_mfma.hip); `queue` holds
// kQueueCounters counters (persistent launches clear them with a fill kernel).
int uneven_launch(hipStream_t s, int64_t count, int max_n, const int32_t* n, const int64_t* a_offset,
                  const float* a, const int64_t* b_offset, const float* b, void* x, int x_dtype,
                  const int32_t* order, int flags, int max_refine, double tol, int32_t* info, int32_t* iters,
                  double* eta, unsigned int* queue, int max4 = kTunedMax4, int max8 = kTunedMax8) {
  return chol_mfma_launch(s, count, max_n, n, a_offset, a, b_offset, b, x, x_dtype, order,
                          (flags & LXA_UNEVEN_PERSISTENT) != 0, (flags & LXA_UNEVEN_SORTED) != 0, max_refine, tol,
                          info, iters, eta, queue, max4, max8);
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

extern "C" int lxa_chol_uneven_tune_solve(void* stream, int64_t count, int max_n, const int32_t* n,
                                          const int64_t* a_offset, const float* a, const int64_t* b_offset,
                                          const float* b, void* x, int x_dtype, const int32_t* order, int flags,
                                          int max_refine, double tol, int32_t* info, int32_t* iters, double* eta,
                                          void* workspace, size_t workspace_bytes, int max4, int max8) {
  if (int rc = lxa::uneven_check(count, max_n, n, a_offset, a, b_offset, b, x, x_dtype, order, max_refine, info))
    return rc;
  const bool persistent = (flags & LXA_UNEVEN_PERSISTENT) != 0;
  LXA_REQUIRE(!persistent || (workspace && workspace_bytes >= lxa::kUnevenWorkspaceBytes),
              "lxa_chol_uneven_tune_solve: the persistent launch needs a workspace of %zu bytes (got %zu)",
              lxa::kUnevenWorkspaceBytes, workspace ? workspace_bytes : static_cast<size_t>(0));
  if (count == 0) return LXA_OK;
  return lxa::uneven_launch(static_cast<hipStream_t>(stream), count, max_n, n, a_offset, a, b_offset, b, x, x_dtype,
                            order, flags, max_refine, tol, info, iters, eta, static_cast<unsigned int*>(workspace),
                            max4, max8);
}

extern "C" int lxa_chol_uneven_tuned_classes(int* max4, int* max8) {
  LXA_REQUIRE(max4 && max8, "lxa_chol_uneven_tuned_classes: null pointer");
  *max4 = lxa::kTunedMax4;
  *max8 = lxa::kTunedMax8;
  return LXA_OK;
}
