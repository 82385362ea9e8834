// Large dense LU for one system of order n (thousands), gfx950.
//
// Pivoting.  getrf's partial pivoting puts one device-wide decision (the
// arg-max of a column of up to n entries) on the critical path of every one
// of the n columns; on a 256-CU device each costs a grid-wide exchange of a
// few microseconds, more than the whole Schur-complement work at n = 16384.
// This path instead randomises the system with a two-sided depth-2 random
// butterfly transform (Parker, "Random butterfly transformations with
// applications in computational linear algebra", 1995; Baboulin et al.,
// "Accelerating linear system solutions using randomization techniques",
// 2013), A' = U^T A V, factors P A' = L U with pivoting confined to each
// 128 x 128 diagonal block (no device-wide exchange), and recovers
// accuracy by refinement against the original A (the host drives GMRES on
// A preconditioned by these factors).  The solution is the same as with
// partial pivoting to the refinement tolerance; the factors differ.
//
// Factorisation (blocked right-looking, two levels):
//   outer panels of 512 columns: the panel is factored by 128-column inner
//   steps, each  (1) LU with in-block partial pivoting of the 128 x 128
//   diagonal block in LDS, with the inverses of its unit-lower and upper
//   triangles (kept for the block triangular solves), and the block's row
//   swaps applied across the matrix, (2) L21 = A21 U11^-1 and U12 = L11^-1 A12
//   as float32 matrix-core products, (3) the panel's Schur update in float32.
//   Then the outer U12 row block (128 rows at a time) and the outer Schur
//   update A22 -= L21 U12, whose products run in bfloat16, float16, split
//   bfloat16 (bf16x3) or float32, with float32 accumulation.
//
// Solves: block triangular solves with 128-row blocks in one persistent
// launch each.  Workgroups take blocks in ticket order (an atomic counter),
// so a workgroup only waits for blocks already owned by running workgroups;
// each block subtracts the products with the solved blocks as their flags
// appear, multiplies by the stored inverse of its diagonal block, and
// publishes its flag.

#include <mutex>

#include "lxa_common.h"

namespace lxa {
namespace {

constexpr int kB = 128;          // diagonal / triangular-solve block
constexpr int kOuter = 512;      // outer panel width
constexpr int kTrsvThreads = 256;
constexpr int kMaxRhs = 8;
constexpr int kSpinLimit = 1 << 26;

// ---------------------------------------------------------------------------
// Random butterflies.  One level with half-width s mixes index pairs
// (b 2s + i, b 2s + s + i), i < s, with diagonals d0 = d[b 2s + i] and
// d1 = d[b 2s + s + i]:
//   B   form: (x, y) -> ((d0 x + d1 y), (d0 x - d1 y)) / sqrt 2
//   B^T form: (x, y) -> (d0 (x + y), d1 (x - y)) / sqrt 2
// A' = U^T A V applies, per level, B^T with the row diagonals to row pairs
// and B with the column diagonals to column pairs (the product of the two is
// formed from the four entries of each 2 x 2 quad in one pass).
// ---------------------------------------------------------------------------
__global__ void rbt_matrix_kernel(int n, int s, float* __restrict__ a, int64_t lda, const float* __restrict__ du,
                                  const float* __restrict__ dv) {
  const int64_t half = n / 2;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= half * half) return;
  const int qi = (int)(q / half), qj = (int)(q % half);
  const int i = (qi / s) * 2 * s + (qi % s);
  const int j = (qj / s) * 2 * s + (qj % s);
  const float r0 = du[i], r1 = du[i + s], c0 = dv[j], c1 = dv[j + s];
  float* p = a + (int64_t)i * lda + j;
  const float x = p[0], y = p[s], z = p[(int64_t)s * lda], w = p[(int64_t)s * lda + s];
  // rows: (x y ; z w) -> (r0 (x+z), r0 (y+w) ; r1 (x-z), r1 (y-w)) / sqrt 2
  const float u0 = x + z, u1 = y + w, v0 = x - z, v1 = y - w;
  // columns: (p, q) -> (c0 (p + q) ... ) per B form on the right: (p c0 + q c0, p c1 - q c1)
  p[0] = 0.5f * r0 * c0 * (u0 + u1);
  p[s] = 0.5f * r0 * c1 * (u0 - u1);
  p[(int64_t)s * lda] = 0.5f * r1 * c0 * (v0 + v1);
  p[(int64_t)s * lda + s] = 0.5f * r1 * c1 * (v0 - v1);
}

// Vectors x[r][i] (r < nrhs, row stride ldx), one level, form 0 = B, 1 = B^T.
template <typename T>
__global__ void rbt_vec_kernel(int n, int s, int nrhs, T* __restrict__ x, int64_t ldx, const float* __restrict__ d,
                               int form) {
  const int half = n / 2;
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (int64_t)half * nrhs) return;
  const int r = (int)(q / half), qi = (int)(q % half);
  const int i = (qi / s) * 2 * s + (qi % s);
  T* p = x + (int64_t)r * ldx;
  const T u = p[i], v = p[i + s];
  const T d0 = (T)d[i], d1 = (T)d[i + s];
  const T h = (T)0.70710678118654752440;
  if (form == 0) {
    p[i] = h * (d0 * u + d1 * v);
    p[i + s] = h * (d0 * u - d1 * v);
  } else {
    p[i] = h * d0 * (u + v);
    p[i + s] = h * d1 * (u - v);
  }
}

// ---------------------------------------------------------------------------
// LU with partial pivoting of the pivot window at (k, k): the nb x nb
// diagonal block (nb <= 128) and the rows below it, win rows in all
// (win <= 256), in place, plus the inverses of the block's unit-lower and
// upper triangles as dense 128 x 128 blocks (zeros outside the triangle;
// identity beyond nb when nb < 128).  Pivoting over the window keeps every
// multiplier of the block column (the diagonal block's L and the next
// block's L21) bounded by 1; over the diagonal block alone, the multipliers
// below it are unbounded and the growth reaches ~1e3 on Gaussian operators
// (~1e-5 backward error instead of ~1e-7).  piv[k + j] receives the 0-based
// global row swapped with row k + j (always inside the window); the caller
// applies the swaps to the columns outside the block.  info[0] is set to
// 1 + the global index of the first exactly-zero pivot.
//
// One workgroup of 256 threads; the window lives in LDS, thread r owning
// window row r (a "physical" row that never moves: each thread tracks the
// position its row has reached, and phys_at[] maps positions back, as the
// row interchanges of getrf would move them).  Blocked right-looking over
// 16-column leaves:
//   * leaf columns are factored one column per barrier: every wave finds
//     its arg-max |a(i, j)| (ties: lowest position) with DPP moves, the four
//     wave winners meet in LDS, each thread reads the pivot row's leaf
//     entries from LDS and updates its own row's leaf entries in registers,
//     then writes them back for the next column's pivot;
//   * the leaf's U rows over the trailing columns: U = L_leaf^-1 A(pivot
//     rows), one thread per column;
//   * the trailing columns of the rows still below the leaf: A -= L U, each
//     thread its own row, the U rows read from LDS as broadcasts.
// Finally the factors are written in position order, and the triangles of
// the diagonal block are inverted by recursive doubling in LDS:
//   inv [A 0; C B] = [A^-1 0; -B^-1 C A^-1  B^-1]
// from 16 x 16 diagonal blocks (forward / backward substitution, one thread
// per column) up to 128 x 128.
// ---------------------------------------------------------------------------
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kPivWin = 2 * kB;     // rows of the pivot window
constexpr int kPanelThreads = 256;  // one window row per thread
constexpr int kLeaf = 16;
constexpr int kPLd = kB + 4;  // LDS row stride (16-byte rows)
// dynamic LDS: the window [256][kPLd] and the leaf's U rows by column [128][16]
constexpr size_t kPanelLds = sizeof(float) * ((size_t)kPivWin * kPLd + (size_t)kB * kLeaf);

// One step of a wave maximum of 64-bit keys (hi, lo) with a DPP partner.
template <int kCtrl>
__device__ __forceinline__ void key_max_dpp(uint32_t& hi, uint32_t& lo) {
  const uint32_t oh = (uint32_t)__builtin_amdgcn_mov_dpp((int)hi, kCtrl, 0xF, 0xF, false);
  const uint32_t ol = (uint32_t)__builtin_amdgcn_mov_dpp((int)lo, kCtrl, 0xF, 0xF, false);
  const bool gt = oh > hi || (oh == hi && ol > lo);
  hi = gt ? oh : hi;
  lo = gt ? ol : lo;
}

// Wave maximum of (hi, lo): quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror and row_mirror leave each 16-lane row's maximum in all of
// its lanes; the four rows are combined from scalar reads (wave-uniform).
__device__ __forceinline__ uint64_t wave_key_max(uint32_t hi, uint32_t lo) {
  key_max_dpp<0xB1>(hi, lo);
  key_max_dpp<0x4E>(hi, lo);
  key_max_dpp<0x141>(hi, lo);
  key_max_dpp<0x140>(hi, lo);
  uint64_t best = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t h = (uint32_t)__builtin_amdgcn_readlane((int)hi, 16 * q);
    const uint32_t l = (uint32_t)__builtin_amdgcn_readlane((int)lo, 16 * q);
    const uint64_t v = ((uint64_t)h << 32) | l;
    best = v > best ? v : best;
  }
  return best;
}

__global__ __launch_bounds__(kPanelThreads) void panel_lu_kernel(int k, int nb, int win, float* __restrict__ a,
                                                                 int64_t lda, float* __restrict__ linv,
                                                                 float* __restrict__ uinv, int32_t* __restrict__ piv,
                                                                 int32_t* __restrict__ info, int s_lo, int s_hi) {
  extern __shared__ __align__(16) float psm[];
  float* P = psm;                              // [256][kPLd]: the window, by physical row
  float* UB = psm + (size_t)kPivWin * kPLd;    // [128][16]: the leaf's U rows, by column
  __shared__ uint32_t keys[2][4][2];           // per column parity: the four wave winners
  __shared__ int phys_at[kPivWin];             // physical row at each position
  __shared__ int piv_s[kB];                    // the interchanges, written out at the end
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = tid;

  // the window: real rows r < win (columns >= nb zero); when nb < 128 (then
  // win = nb) rows nb..127 are identity padding, fixed at their positions
  constexpr int kBatch = 16;  // loads in flight per thread
  // float4 rows when the matrix allows (lda and the base 16-byte aligned;
  // k is a multiple of 128)
  const bool vec = (lda & 3) == 0 && (reinterpret_cast<uintptr_t>(a) & 15) == 0;
  if (vec) {
    constexpr int kQuads = kB / 4;
    for (int base = 0; base < kPivWin * kQuads; base += 8 * kPanelThreads) {
      float4 v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int idx = base + tid + u * kPanelThreads;
        const int i = idx / kQuads, c = 4 * (idx % kQuads);
        v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i < win) {
          const float* src = a + (int64_t)(k + i) * lda + k + c;
          if (c + 4 <= nb) {
            v[u] = *reinterpret_cast<const float4*>(src);
          } else if (c < nb) {
            v[u].x = src[0];
            if (c + 1 < nb) v[u].y = src[1];
            if (c + 2 < nb) v[u].z = src[2];
          }
        } else if (i < kB) {
          v[u].x = (i == c) ? 1.f : 0.f;
          v[u].y = (i == c + 1) ? 1.f : 0.f;
          v[u].z = (i == c + 2) ? 1.f : 0.f;
          v[u].w = (i == c + 3) ? 1.f : 0.f;
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int idx = base + tid + u * kPanelThreads;
        *reinterpret_cast<float4*>(P + (idx / kQuads) * kPLd + 4 * (idx % kQuads)) = v[u];
      }
    }
  }
  for (int base = 0; !vec && base < kPivWin * kB; base += kBatch * kPanelThreads) {
    float v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int idx = base + tid + u * kPanelThreads;
      const int i = idx / kB, c = idx % kB;
      v[u] = 0.f;
      if (i < win) {
        if (c < nb) v[u] = a[(int64_t)(k + i) * lda + k + c];
      } else if (i < kB) {
        v[u] = (i == c) ? 1.f : 0.f;
      }
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int idx = base + tid + u * kPanelThreads;
      P[(idx / kB) * kPLd + idx % kB] = v[u];
    }
  }
  phys_at[tid] = tid;
  int zero_at = 0;  // 1 + the first column with a zero pivot (thread 0)
  int pos = r;
  bool done = r >= win;  // padding and absent rows never pivot
  __syncthreads();

  const int nleaf = (nb + kLeaf - 1) / kLeaf;
  for (int J = 0; J < nleaf; ++J) {
    const int c0 = J * kLeaf;
    float* prow_mine = P + r * kPLd + c0;
    float x[kLeaf];
#pragma unroll
    for (int q = 0; q < kLeaf / 4; ++q) {
      const float4 v = *reinterpret_cast<const float4*>(prow_mine + 4 * q);
      x[4 * q] = v.x;
      x[4 * q + 1] = v.y;
      x[4 * q + 2] = v.z;
      x[4 * q + 3] = v.w;
    }
#pragma unroll
    for (int j = 0; j < kLeaf; ++j) {
      const int c = c0 + j;
      if (c >= nb) continue;  // (uniform; a guard rather than a break keeps the loop unrolled)
      const int buf = j & 1;
      // key: |a(r, c)| (NaN above every number), then the lowest position,
      // then the physical row; rows already pivoted have key 0
      uint32_t hi = 0, lo = 0;
      if (!done) {
        hi = __float_as_uint(x[j]) & 0x7fffffffu;
        lo = 0x80000000u | ((uint32_t)(kPivWin - 1 - pos) << 8) | (uint32_t)r;
      }
      const uint64_t wbest = wave_key_max(hi, lo);
      if (lane == 0) {
        keys[buf][wave][0] = (uint32_t)(wbest >> 32);
        keys[buf][wave][1] = (uint32_t)wbest;
      }
      __syncthreads();
      uint64_t best = 0;
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const uint64_t v = ((uint64_t)keys[buf][w][0] << 32) | keys[buf][w][1];
        best = v > best ? v : best;
      }
      const int p = (int)(best & 0xffu);                              // pivot's physical row
      const int cp = kPivWin - 1 - (int)((best >> 8) & 0xffu);        // and its position
      const float* pr = P + p * kPLd + c0;
      float prow[kLeaf];
#pragma unroll
      for (int q = j / 4; q < kLeaf / 4; ++q) {
        const float4 v = *reinterpret_cast<const float4*>(pr + 4 * q);
        prow[4 * q] = v.x;
        prow[4 * q + 1] = v.y;
        prow[4 * q + 2] = v.z;
        prow[4 * q + 3] = v.w;
      }
      const float pv = prow[j];
      if (tid == 0) {
        piv_s[c] = k + cp;
        if (pv == 0.f && zero_at == 0) zero_at = k + c + 1;
      }
      if (r == p) {
        pos = c;
        done = true;
        phys_at[c] = r;
      } else if (!done) {
        if (pos == c) {  // the row at position c takes the pivot's old position
          pos = cp;
          phys_at[cp] = r;
        }
        const float l = (pv != 0.f) ? x[j] / pv : x[j];
        x[j] = l;
#pragma unroll
        for (int jj = j + 1; jj < kLeaf; ++jj) x[jj] -= l * prow[jj];
        // the updated leaf entries, for the next column's pivot row
#pragma unroll
        for (int q = j / 4; q < kLeaf / 4; ++q)
          *reinterpret_cast<float4*>(prow_mine + 4 * q) =
              make_float4(x[4 * q], x[4 * q + 1], x[4 * q + 2], x[4 * q + 3]);
      }
    }
    __syncthreads();  // the leaf's factors and phys_at are final
    const int t0 = c0 + kLeaf;
    if (t0 >= nb) break;
    // U rows of the leaf over the trailing columns (columns beyond nb are
    // zero in every real row and come out zero)
    if (tid >= t0 && tid < kB) {
      const int col = tid;
      float s[kLeaf];
      int q_at[kLeaf];
#pragma unroll
      for (int i = 0; i < kLeaf; ++i) {
        q_at[i] = phys_at[c0 + i];
        s[i] = P[q_at[i] * kPLd + col];
      }
#pragma unroll
      for (int m = 0; m < kLeaf; ++m) {
        // s[m] is final: u_m; remove it from the rows below
        const float um = s[m];
        P[q_at[m] * kPLd + col] = um;
        UB[col * kLeaf + m] = um;
#pragma unroll
        for (int i = m + 1; i < kLeaf; ++i) s[i] -= P[q_at[i] * kPLd + c0 + m] * um;
      }
    }
    __syncthreads();
    // the trailing columns of the rows still below the leaf
    if (!done) {
      const int cend = (nb + 3) & ~3;
      float* row = P + r * kPLd;
#pragma unroll 2
      for (int col = t0; col < cend; col += 4) {
        float4 v = *reinterpret_cast<const float4*>(row + col);
        const float* u = UB + col * kLeaf;
#pragma unroll
        for (int m4 = 0; m4 < kLeaf / 4; ++m4) {
          const float4 u0 = *reinterpret_cast<const float4*>(u + 4 * m4);
          const float4 u1 = *reinterpret_cast<const float4*>(u + kLeaf + 4 * m4);
          const float4 u2 = *reinterpret_cast<const float4*>(u + 2 * kLeaf + 4 * m4);
          const float4 u3 = *reinterpret_cast<const float4*>(u + 3 * kLeaf + 4 * m4);
          const float l0 = x[4 * m4], l1 = x[4 * m4 + 1], l2 = x[4 * m4 + 2], l3 = x[4 * m4 + 3];
          v.x -= l0 * u0.x + l1 * u0.y + l2 * u0.z + l3 * u0.w;
          v.y -= l0 * u1.x + l1 * u1.y + l2 * u1.z + l3 * u1.w;
          v.z -= l0 * u2.x + l1 * u2.y + l2 * u2.z + l3 * u2.w;
          v.w -= l0 * u3.x + l1 * u3.y + l2 * u3.z + l3 * u3.w;
        }
        *reinterpret_cast<float4*>(row + col) = v;
      }
    }
    __syncthreads();
  }

  // the factors to global memory in position order; the window rows below
  // the block leave with their final multipliers (their L21 block)
  if (vec) {
    constexpr int kQuads = kB / 4;
    const int nq = (nb + 3) / 4;
#pragma unroll 4
    for (int idx = tid; idx < win * nq; idx += kPanelThreads) {
      const int i = idx / nq, c = 4 * (idx % nq);
      const float4 v = *reinterpret_cast<const float4*>(P + phys_at[i] * kPLd + c);
      float* dst = a + (int64_t)(k + i) * lda + k + c;
      if (c + 4 <= nb) {
        *reinterpret_cast<float4*>(dst) = v;
      } else {
        dst[0] = v.x;
        if (c + 1 < nb) dst[1] = v.y;
        if (c + 2 < nb) dst[2] = v.z;
      }
    }
    (void)kQuads;
  } else {
#pragma unroll 8
    for (int idx = tid; idx < win * nb; idx += kPanelThreads) {
      const int i = idx / nb, c = idx % nb;
      a[(int64_t)(k + i) * lda + k + c] = P[phys_at[i] * kPLd + c];
    }
  }
  if (tid < nb) piv[k + tid] = piv_s[tid];
  if (tid == 0 && zero_at != 0 && info[0] == 0) info[0] = zero_at;
  // the diagonal block in position order: S = P rows 0..127 (through
  // registers, as the two overlap), X (the inverse being formed) after it
  constexpr int kSv = kB * kB / kPanelThreads;
  float stage[kSv];
#pragma unroll
  for (int u = 0; u < kSv; ++u) {
    const int idx = tid + u * kPanelThreads;
    const int i = idx / kB, c = idx % kB;
    stage[u] = P[phys_at[i] * kPLd + c];
  }
  __syncthreads();
  float* S = psm;
  float* X = psm + kB * kPLd;
#pragma unroll
  for (int u = 0; u < kSv; ++u) {
    const int idx = tid + u * kPanelThreads;
    S[(idx / kB) * kPLd + idx % kB] = stage[u];
  }
  __syncthreads();

  // inverses: lower (unit) first, then upper, each formed in X and written out
  for (int tri = 0; tri < 2; ++tri) {
    const bool lower = tri == 0;
    for (int idx = tid; idx < kB * kB; idx += kPanelThreads) X[(idx / kB) * kPLd + idx % kB] = 0.f;
    __syncthreads();
    // 16 x 16 diagonal blocks: thread t < 128 forms column t % 16 of block t / 16
    if (tid < kB) {
      const int d0 = (tid / kLeaf) * kLeaf, cc = tid % kLeaf;
      float y[kLeaf];
      if (lower) {
#pragma unroll
        for (int i = 0; i < kLeaf; ++i) {
          float s = (i == cc) ? 1.f : 0.f;
#pragma unroll
          for (int m = 0; m < i; ++m) s -= S[(d0 + i) * kPLd + d0 + m] * y[m];
          y[i] = s;
        }
      } else {
#pragma unroll
        for (int i = kLeaf - 1; i >= 0; --i) {
          float s = (i == cc) ? 1.f : 0.f;
#pragma unroll
          for (int m = i + 1; m < kLeaf; ++m) s -= S[(d0 + i) * kPLd + d0 + m] * y[m];
          y[i] = s / S[(d0 + i) * kPLd + d0 + i];
        }
      }
#pragma unroll
      for (int i = 0; i < kLeaf; ++i) X[(d0 + i) * kPLd + d0 + cc] = y[i];
    }
    __syncthreads();
    // doubling: pairs of s x s blocks (A first, B second along the diagonal)
    //   lower: X_BA = -X_BB (L_BA X_AA)      upper: X_AB = -X_AA (U_AB X_BB)
    // on the matrix cores (v_mfma_f32_16x16x4f32): the s x s off-diagonal
    // blocks of all pairs split into 16 x 16 tiles over the four waves; first
    // T = (the middle product) into the target block, then X = -(.) T.
    // Operand k index of lane group g at step j of a 16-wide chunk: 4 g + j,
    // so a lane's four A entries are one 16-byte LDS read.
    {
      const int g = lane >> 4, q = lane & 15;
#pragma unroll
      for (int lev = 0; lev < 3; ++lev) {
        const int sz = kLeaf << lev;              // 16, 32, 64
        const int tdim = sz / 16;                 // tiles per block side
        constexpr int kTilesPerWave = 4;          // at most: 16 tiles / 4 waves
        const int per_wave = (kB / (2 * sz)) * tdim * tdim / 4;  // 1, 2, 4
        f4 acc[kTilesPerWave];
        int tr0[kTilesPerWave], tc0[kTilesPerWave], mb[kTilesPerWave], ob[kTilesPerWave];
#pragma unroll
        for (int t = 0; t < kTilesPerWave; ++t) {
          if (t >= per_wave) break;
          const int tile = wave + 4 * t;
          const int pair = tile / (tdim * tdim), ti = (tile / tdim) % tdim, tj = tile % tdim;
          const int A0 = 2 * sz * pair, B0 = A0 + sz;
          tr0[t] = (lower ? B0 : A0) + 16 * ti;
          tc0[t] = (lower ? A0 : B0) + 16 * tj;
          mb[t] = lower ? A0 : B0;  // step 1: k runs over these rows of Xd / columns of M
          ob[t] = lower ? B0 : A0;  // step 2: k runs over these rows of T / columns of Xo
        }
        // step 1: T = M Xd, M from the factors S
#pragma unroll
        for (int t = 0; t < kTilesPerWave; ++t) {
          if (t >= per_wave) break;
          acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int m = 0; m < sz; m += 16) {
            const float4 av = *reinterpret_cast<const float4*>(S + (tr0[t] + q) * kPLd + mb[t] + m + 4 * g);
            const float* bp = X + (mb[t] + m + 4 * g) * kPLd + tc0[t] + q;
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bp[0], acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bp[kPLd], acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bp[2 * kPLd], acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bp[3 * kPLd], acc[t], 0, 0, 0);
          }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kTilesPerWave; ++t) {
          if (t >= per_wave) break;
#pragma unroll
          for (int i = 0; i < 4; ++i) X[(tr0[t] + 4 * g + i) * kPLd + tc0[t] + q] = acc[t][i];
        }
        __syncthreads();
        // step 2: X_target = -Xo T, T held in the target block
#pragma unroll
        for (int t = 0; t < kTilesPerWave; ++t) {
          if (t >= per_wave) break;
          acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int m = 0; m < sz; m += 16) {
            const float4 av = *reinterpret_cast<const float4*>(X + (tr0[t] + q) * kPLd + ob[t] + m + 4 * g);
            const float* bp = X + (ob[t] + m + 4 * g) * kPLd + tc0[t] + q;
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bp[0], acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bp[kPLd], acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bp[2 * kPLd], acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bp[3 * kPLd], acc[t], 0, 0, 0);
          }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < kTilesPerWave; ++t) {
          if (t >= per_wave) break;
#pragma unroll
          for (int i = 0; i < 4; ++i) X[(tr0[t] + 4 * g + i) * kPLd + tc0[t] + q] = -acc[t][i];
        }
        __syncthreads();
      }
    }
    float* out = lower ? linv : uinv;
#pragma unroll 8
    for (int idx = tid; idx < kB * kB / 4; idx += kPanelThreads)
      reinterpret_cast<float4*>(out)[idx] = *reinterpret_cast<const float4*>(X + (idx / (kB / 4)) * kPLd + 4 * (idx % (kB / 4)));
    __syncthreads();
  }

  // the window's interchanges over the other columns [s_lo, s_hi) of the
  // caller's panel (128-column chunks aligned with this block, whose own
  // chunk is skipped), staged through LDS: row position i takes window row
  // phys_at[i]; rows that stay are not rewritten
  float* W = psm;  // [256][kPLd]
  for (int cb = s_lo; cb < s_hi; cb += kB) {
    if (cb == k) continue;
    const int cw = min(kB, s_hi - cb);
    const bool full = vec && cw == kB;
    if (full) {
      constexpr int kQuads = kB / 4;
      for (int base = 0; base < win * kQuads; base += 8 * kPanelThreads) {
        float4 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int idx = base + tid + u * kPanelThreads;
          v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
          if (idx < win * kQuads)
            v[u] = *reinterpret_cast<const float4*>(a + (int64_t)(k + idx / kQuads) * lda + cb + 4 * (idx % kQuads));
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int idx = base + tid + u * kPanelThreads;
          if (idx < win * kQuads) *reinterpret_cast<float4*>(W + (idx / kQuads) * kPLd + 4 * (idx % kQuads)) = v[u];
        }
      }
      __syncthreads();
#pragma unroll 4
      for (int idx = tid; idx < win * kQuads; idx += kPanelThreads) {
        const int i = idx / kQuads, q = idx % kQuads;
        const int src = phys_at[i];
        if (src != i)
          *reinterpret_cast<float4*>(a + (int64_t)(k + i) * lda + cb + 4 * q) =
              *reinterpret_cast<const float4*>(W + src * kPLd + 4 * q);
      }
    } else {
#pragma unroll 8
      for (int idx = tid; idx < win * cw; idx += kPanelThreads)
        W[(idx / cw) * kPLd + idx % cw] = a[(int64_t)(k + idx / cw) * lda + cb + idx % cw];
      __syncthreads();
#pragma unroll 8
      for (int idx = tid; idx < win * cw; idx += kPanelThreads) {
        const int i = idx / cw, c = idx % cw;
        const int src = phys_at[i];
        if (src != i) a[(int64_t)(k + i) * lda + cb + c] = W[src * kPLd + c];
      }
    }
    __syncthreads();
  }
}

// The interchanges of a whole outer panel [K0, K1) (its 128-column blocks in
// order, each within its 256-row window) applied to the columns outside the
// panel, [0, K0) and [K1, n): a workgroup stages the rows the panel's windows
// span (K0 .. K1 + 127, at most 640) of 32 columns in LDS, then, block by
// block, each window row finds where the block's interchanges take it (one
// thread per row following its own position through the swap sequence) and
// moves there, and the rows are written back.
constexpr int kOsCols = 32;
constexpr int kOsRows = kOuter + kB;  // rows a panel's windows span
constexpr int kOsLd = kOsCols + 1;
constexpr size_t kOsLds = sizeof(float) * (size_t)kOsRows * kOsLd;

__global__ __launch_bounds__(256) void outer_swaps_kernel(int n, int K0, int K1, float* __restrict__ a, int64_t lda,
                                                          const int32_t* __restrict__ piv) {
  extern __shared__ __align__(16) float T[];  // [kOsRows][kOsLd]
  __shared__ int pv[kOuter];
  const int tid = threadIdx.x;
  const int rend = min(n, K1 + kB), rows = rend - K0;
  const int c0 = blockIdx.x * kOsCols;  // in the column sequence [0, K0) ++ [K1, n)
  for (int j = tid; j < K1 - K0; j += 256) pv[j] = piv[K0 + j] - K0;
  // stage: 32 consecutive global columns (remapped past the panel; the
  // group never straddles it, K0 and K1 - K0 being multiples of 32) as
  // 8 column quads per row, four loads in flight per thread
  int cg = c0;
  if (cg >= K0) cg += K1 - K0;
  const bool vec = (lda & 3) == 0 && (reinterpret_cast<uintptr_t>(a) & 15) == 0 && cg + kOsCols <= n;
  for (int base = 0; base < rows * 8; base += 4 * 256) {
    float4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int idx = base + tid + 256 * u, r = idx >> 3, c = cg + 4 * (idx & 7);
      v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (r < rows) {
        const float* src = a + (int64_t)(K0 + r) * lda + c;
        if (vec) {
          v[u] = *reinterpret_cast<const float4*>(src);
        } else {
          if (c < n) v[u].x = src[0];
          if (c + 1 < n) v[u].y = src[1];
          if (c + 2 < n) v[u].z = src[2];
          if (c + 3 < n) v[u].w = src[3];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int idx = base + tid + 256 * u, r = idx >> 3, q = 4 * (idx & 7);
      if (r < rows) {
        float* t = T + r * kOsLd + q;
        t[0] = v[u].x;
        t[1] = v[u].y;
        t[2] = v[u].z;
        t[3] = v[u].w;
      }
    }
  }
  __syncthreads();
  for (int k = K0; k < K1; k += kB) {
    const int nb = min(K1, k + kB) - k;
    const int win = nb < kB ? nb : min(kPivWin, n - k);
    const int w0 = k - K0;  // window start, relative to K0
    int dest = tid;
    float v[kOsCols];
    if (tid < win) {
      // the position window row tid reaches (steps k .. k + nb - 1)
      for (int j = 0; j < nb; ++j) {
        const int p = pv[w0 + j] - w0;
        dest = (dest == j) ? p : (dest == p ? j : dest);
      }
#pragma unroll
      for (int q = 0; q < kOsCols; ++q) v[q] = T[(w0 + tid) * kOsLd + q];
    }
    __syncthreads();
    if (tid < win && dest != tid) {
#pragma unroll
      for (int q = 0; q < kOsCols; ++q) T[(w0 + dest) * kOsLd + q] = v[q];
    }
    __syncthreads();
  }
#pragma unroll 4
  for (int idx = tid; idx < rows * 8; idx += 256) {
    const int r = idx >> 3, q = 4 * (idx & 7), c = cg + q;
    const float* t = T + r * kOsLd + q;
    float* dst = a + (int64_t)(K0 + r) * lda + c;
    if (vec) {
      *reinterpret_cast<float4*>(dst) = make_float4(t[0], t[1], t[2], t[3]);
    } else {
      if (c < n) dst[0] = t[0];
      if (c + 1 < n) dst[1] = t[1];
      if (c + 2 < n) dst[2] = t[2];
      if (c + 3 < n) dst[3] = t[3];
    }
  }
}

// The row permutation of the whole factorisation: perm[i] = the original
// row now at position i (the swap sequence piv applied in order to the
// identity).  The sequence is windowed (step j exchanges positions j and
// piv[j] in [j, j + 256), as the factorisation writes it), so each row
// follows its own position: step j moves a row at position j to piv[j] and
// one at piv[j] to j, steps before pos - 255 cannot reach position pos, and
// a row whose position is below the current step is final.  One thread per
// row, the swap steps its workgroup's rows meet staged in LDS.
constexpr int kPermBefore = kPivWin;       // steps staged before the first row
constexpr int kPermSpan = kPermBefore + 4 * 256;

__global__ __launch_bounds__(256) void perm_from_swaps_kernel(int n, const int32_t* __restrict__ piv,
                                                              int32_t* __restrict__ perm) {
  __shared__ int32_t ps[kPermSpan];
  const int r0 = blockIdx.x * 256, base = max(0, r0 - kPermBefore);
  for (int i = threadIdx.x; i < kPermSpan; i += 256) ps[i] = base + i < n ? piv[base + i] : 0;
  __syncthreads();
  const int r = r0 + threadIdx.x;
  if (r >= n) return;
  int pos = r;
  for (int j = max(0, r - (kPivWin - 1)); j <= pos && j < n; ++j) {
    const int q = j - base < kPermSpan ? ps[j - base] : piv[j];
    pos = (pos == j) ? q : (pos == q ? j : pos);
  }
  perm[pos] = r;
}

// y = P x (y[i] = x[perm[i]], inverse = 0) or y = P^T x (y[perm[i]] = x[i]),
// nrhs vectors with row strides ldx / ldy.
__global__ void permute_rows_kernel(int n, int nrhs, const float* __restrict__ x, int64_t ldx, float* __restrict__ y,
                                    int64_t ldy, const int32_t* __restrict__ perm, int inverse) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= static_cast<int64_t>(n) * nrhs) return;
  const int64_t r = q / n;
  const int i = static_cast<int>(q % n);
  if (!inverse)
    y[r * ldy + i] = x[r * ldx + perm[i]];
  else
    y[r * ldy + perm[i]] = x[r * ldx + i];
}

// ---------------------------------------------------------------------------
// Block triangular solve, one persistent launch.
//   kind 0: L y = x  (unit lower, forward)      kind 1: U y = x (upper, backward)
//   kind 2: U^T y = x (lower, forward)          kind 3: L^T y = x (unit upper, backward)
// inv holds the inverses of the diagonal blocks of L (kinds 0, 3) or U
// (kinds 1, 2), 128 x 128 each; transposed kinds read blocks transposed.
// ---------------------------------------------------------------------------
constexpr int kTileLd = kB + 4;  // LDS row stride of a staged 128 x 128 block (16-byte rows)
// Workspace of a block triangular solve (int32 words): [0] the ticket
// counter, [1] the failure word, [2..3] padding, then the channel: one
// 64-bit (tag, value) word per solution entry, nblk x 128 x kMaxRhs.
constexpr int kTrsvHeader = 4;
constexpr uint32_t kChanTag = 1;
constexpr int kChanPer = kMaxRhs * kB / kTrsvThreads;  // channel entries per thread and block

// Clears the ticket, the failure word and the channel's tags (a fill kernel,
// not a memset node, so that a solve can sit in a captured graph).
__global__ void trsv_arm_kernel(int32_t* __restrict__ ws, int64_t chan_words) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < kTrsvHeader) ws[i] = 0;
  uint64_t* chan = reinterpret_cast<uint64_t*>(ws + kTrsvHeader);
  for (int64_t j = i; j < chan_words; j += static_cast<int64_t>(gridDim.x) * blockDim.x) chan[j] = 0;
}

// Stages block (bi, bj) of the operator of `kind` into tile[r][c] (r, c <
// 128; zeros outside rows < rows, cols < cols).  Loads are float4 along the
// stored rows: for the transposed kinds the stored rows are the tile's
// columns, so they are written transposed.
// A 128 x 128 operator block staged through registers: tile_load issues all
// of a thread's loads (16 float4, rows of the stored block) and tile_store
// writes them to LDS, transposed for kinds 2/3, so the loads of one block are
// in flight together and can overlap the product of the previous one.
constexpr int kStageVec = kB * (kB / 4) / kTrsvThreads;  // float4 per thread

__device__ void tile_load(float4 (&v)[kStageVec], const float* lu, int64_t lda, int kind, int bi, int bj, int rows,
                          int cols, int tid) {
  const bool tr = kind >= 2;
  const int srow_lim = tr ? cols : rows, scol_lim = tr ? rows : cols;
  const bool full = srow_lim == kB && scol_lim == kB;
#pragma unroll
  for (int u = 0; u < kStageVec; ++u) {
    const int idx = tid + u * kTrsvThreads;
    const int sr = idx >> 5, sc = 4 * (idx & 31);
    const int64_t grow = tr ? (int64_t)bj * kB + sr : (int64_t)bi * kB + sr;
    const int64_t gcol = tr ? (int64_t)bi * kB + sc : (int64_t)bj * kB + sc;
    const float* p = lu + grow * lda + gcol;
    if (full) {
      v[u] = *reinterpret_cast<const float4*>(p);
    } else {
      float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
      if (sr < srow_lim) {
        w.x = sc < scol_lim ? p[0] : 0.f;
        w.y = sc + 1 < scol_lim ? p[1] : 0.f;
        w.z = sc + 2 < scol_lim ? p[2] : 0.f;
        w.w = sc + 3 < scol_lim ? p[3] : 0.f;
      }
      v[u] = w;
    }
  }
}

__device__ void tile_store(float* tile, const float4 (&v)[kStageVec], bool tr, int tid) {
#pragma unroll
  for (int u = 0; u < kStageVec; ++u) {
    const int idx = tid + u * kTrsvThreads;
    const int sr = idx >> 5, sc = 4 * (idx & 31);
    if (!tr) {
      *reinterpret_cast<float4*>(tile + sr * kTileLd + sc) = v[u];
    } else {
      tile[(sc + 0) * kTileLd + sr] = v[u].x;
      tile[(sc + 1) * kTileLd + sr] = v[u].y;
      tile[(sc + 2) * kTileLd + sr] = v[u].z;
      tile[(sc + 3) * kTileLd + sr] = v[u].w;
    }
  }
}

// part[rr] = sum_c tile[row][c] * vec[rr][c] over this thread's 64 columns,
// combined with the partner thread holding the other 64.
template <int R>
__device__ __forceinline__ void tile_product(const float* tile, const float (*vec)[kB], int row, int hcol,
                                             float (&part)[R]) {
#pragma unroll
  for (int rr = 0; rr < R; ++rr) part[rr] = 0.f;
  const float* tr = tile + row * kTileLd + hcol;
#pragma unroll 4
  for (int c = 0; c < 64; c += 4) {
    const float4 e = *reinterpret_cast<const float4*>(tr + c);
#pragma unroll
    for (int rr = 0; rr < R; ++rr)
      part[rr] += e.x * vec[rr][hcol + c] + e.y * vec[rr][hcol + c + 1] + e.z * vec[rr][hcol + c + 2] +
                  e.w * vec[rr][hcol + c + 3];
  }
#pragma unroll
  for (int rr = 0; rr < R; ++rr) part[rr] += __shfl_xor(part[rr], 1, kWave);
}

// R: right-hand sides computed per product (4 or kMaxRhs; nrhs <= R).
//
// Block row bi needs every solved block before it; only the last of them
// (bj = the previous block) is on the chain from block to block.  So
// before it waits, the workgroup forms M = D^-1 A(bi, last) (D the diagonal
// block; D^-1 is stored) and z = D^-1 (x_bi - sum over the other solved
// blocks of A(bi, bj) y_bj); once the last block's flag is up, its y is one
// product away: y_bi = z - M y_last.
template <int R>
__global__ __launch_bounds__(kTrsvThreads) void trsv_blocks_kernel(int n, int nrhs, const float* __restrict__ lu,
                                                                   int64_t lda, const float* __restrict__ inv,
                                                                   float* __restrict__ x, int64_t ldx, int kind,
                                                                   int32_t* __restrict__ sync) {
  extern __shared__ __align__(16) float tsm[];
  float* tile = tsm;                                                        // [128][132]: streamed block, D^-1
  float* Mt = tsm + kB * kTileLd;                                           // [128][132]: M
  float(*xs)[kB] = reinterpret_cast<float(*)[kB]>(tsm + 2 * kB * kTileLd);  // [8][128] a solved block
  float(*acc)[kB] = reinterpret_cast<float(*)[kB]>(tsm + 2 * kB * kTileLd + kMaxRhs * kB);  // [8][128]
  __shared__ int s_ticket;
  int32_t* ticket = sync;
  int32_t* failed = sync + 1;
  uint64_t* chan = reinterpret_cast<uint64_t*>(sync + kTrsvHeader);
  const int nblk = (n + kB - 1) / kB;
  const int tid = threadIdx.x;
  if (tid == 0) s_ticket = atomicAdd(ticket, 1);
  __syncthreads();
  const int t = s_ticket;
  if (t >= nblk) return;
  const bool forward = (kind == 0 || kind == 2);
  const int bi = forward ? t : nblk - 1 - t;
  const int r0 = bi * kB, rows = min(kB, n - r0);
  for (int idx = tid; idx < kMaxRhs * kB; idx += kTrsvThreads) {
    const int rr = idx / kB, i = idx % kB;
    acc[rr][i] = (rr < nrhs && i < rows) ? x[(int64_t)rr * ldx + r0 + i] : 0.f;
  }
  // row i of the block handled by threads 2i, 2i + 1 (64 columns each)
  const int row = tid >> 1, hcol = (tid & 1) * 64;
  const bool tr = kind >= 2;
  float4 pre[kStageVec];
  auto block_col = [&](int step) { return forward ? step : nblk - 1 - step; };
  // the inverse of the diagonal block, staged like an operator block
  const float* D = inv + (int64_t)bi * kB * kB;
  auto load_inv = [&]() {
#pragma unroll
    for (int u = 0; u < kStageVec; ++u) {
      const int idx = tid + u * kTrsvThreads;
      pre[u] = *reinterpret_cast<const float4*>(D + (idx >> 5) * kB + 4 * (idx & 31));
    }
  };
  // A solved block arrives through the channel: each value is published
  // with its tag in one 64-bit store, so a reader needs no flag and the
  // writer no fence; every thread polls its own entries until their tags
  // are up (bounded: past the spin limit the failure word is set).
  auto recv_xs = [&](int bj) {
    const int cols = min(kB, n - bj * kB);
    uint64_t v[kChanPer];
    bool need[kChanPer];
#pragma unroll
    for (int u = 0; u < kChanPer; ++u) {
      const int idx = tid + u * kTrsvThreads, rr = idx / kB, i = idx % kB;
      need[u] = rr < nrhs && i < cols;
      v[u] = need[u] ? __hip_atomic_load(&chan[((int64_t)bj * kB + i) * kMaxRhs + rr], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT)
                     : 0;
    }
#pragma unroll
    for (int u = 0; u < kChanPer; ++u) {
      const int idx = tid + u * kTrsvThreads, rr = idx / kB, i = idx % kB;
      int spin = 0;
      while (need[u] && (v[u] >> 32) != kChanTag) {
        if (++spin > kSpinLimit) {
          __hip_atomic_store(failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        v[u] = __hip_atomic_load(&chan[((int64_t)bj * kB + i) * kMaxRhs + rr], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
      }
      xs[rr][i] = need[u] ? __uint_as_float(static_cast<uint32_t>(v[u])) : 0.f;
    }
  };
  if (t > 0) {
    // M = D^-1 A(bi, last): D^-1 into tile, A(bi, last) into Mt, then each
    // thread its 64 entries of M in registers, written over Mt
    const int bl = block_col(t - 1);
    load_inv();
    tile_store(tile, pre, tr, tid);
    tile_load(pre, lu, lda, kind, bi, bl, rows, min(kB, n - bl * kB), tid);
    tile_store(Mt, pre, tr, tid);
    __syncthreads();
    float m[64];
#pragma unroll
    for (int c = 0; c < 64; ++c) m[c] = 0.f;
    const float* drow = tile + row * kTileLd;
    for (int kk = 0; kk < kB; kk += 4) {
      const float4 d = *reinterpret_cast<const float4*>(drow + kk);
      const float dv[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float* lr = Mt + (kk + j) * kTileLd + hcol;
#pragma unroll
        for (int c = 0; c < 64; c += 4) {
          const float4 v = *reinterpret_cast<const float4*>(lr + c);
          m[c] += dv[j] * v.x;
          m[c + 1] += dv[j] * v.y;
          m[c + 2] += dv[j] * v.z;
          m[c + 3] += dv[j] * v.w;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 64; c += 4)
      *reinterpret_cast<float4*>(Mt + row * kTileLd + hcol + c) = make_float4(m[c], m[c + 1], m[c + 2], m[c + 3]);
  }
  // the solved blocks before the last: operator blocks do not depend on the
  // solution, so block `step + 1` is loaded while block `step` waits for its
  // x and is applied
  if (t > 1) tile_load(pre, lu, lda, kind, bi, block_col(0), rows, min(kB, n - block_col(0) * kB), tid);
  for (int step = 0; step + 1 < t; ++step) {
    const int bj = block_col(step);
    tile_store(tile, pre, tr, tid);
    if (step + 2 < t) {
      const int bn = block_col(step + 1);
      tile_load(pre, lu, lda, kind, bi, bn, rows, min(kB, n - bn * kB), tid);
    }
    recv_xs(bj);
    __syncthreads();
    float part[R];
    tile_product<R>(tile, xs, row, hcol, part);
    if ((tid & 1) == 0 && row < rows) {
#pragma unroll
      for (int rr = 0; rr < R; ++rr)
        if (rr < nrhs) acc[rr][row] -= part[rr];
    }
    __syncthreads();
  }
  // z = D^-1 acc (D^-1 staged like an operator block; transposed kinds use
  // its transpose), kept in registers by the even thread of each row pair
  load_inv();
  tile_store(tile, pre, tr, tid);
  __syncthreads();
  float z[R];
  tile_product<R>(tile, acc, row, hcol, z);
  if (t > 0) {
    recv_xs(block_col(t - 1));
    __syncthreads();
    float part[R];
    tile_product<R>(Mt, xs, row, hcol, part);
#pragma unroll
    for (int rr = 0; rr < R; ++rr) z[rr] -= part[rr];
  }
  if ((tid & 1) == 0 && row < rows) {
#pragma unroll
    for (int rr = 0; rr < R; ++rr)
      if (rr < nrhs) {
        x[(int64_t)rr * ldx + r0 + row] = z[rr];
        const uint64_t word = (static_cast<uint64_t>(kChanTag) << 32) | __float_as_uint(z[rr]);
        __hip_atomic_store(&chan[((int64_t)bi * kB + row) * kMaxRhs + rr], word, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      }
  }
}

// ---------------------------------------------------------------------------
// y = A x (b == null) or y = b - A x, A float32 [m, k] row-major; x, b, y of
// either dtype, NR vectors each (row stride ldx / ldb / ldy); FP64
// accumulation.  A wave owns 4 consecutive rows, so each x load serves four
// rows; lanes stride the columns (coalesced rows of A).
// ---------------------------------------------------------------------------
constexpr int kGemvRows = 4;

template <typename TX, typename TY, int NR>
__global__ __launch_bounds__(256) void gemv_f32a_kernel(int m, int k, const float* __restrict__ a, int64_t lda,
                                                        const TX* __restrict__ x, int64_t ldx,
                                                        const TY* __restrict__ b, int64_t ldb, TY* __restrict__ y,
                                                        int64_t ldy) {
  const int lane = threadIdx.x & 63;
  const int row0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * kGemvRows;
  if (row0 >= m) return;
  double s[kGemvRows][NR];
#pragma unroll
  for (int i = 0; i < kGemvRows; ++i)
#pragma unroll
    for (int r = 0; r < NR; ++r) s[i][r] = 0.0;
  for (int c = lane; c < k; c += kWave) {
    double xv[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) xv[r] = (double)x[(int64_t)r * ldx + c];
#pragma unroll
    for (int i = 0; i < kGemvRows; ++i) {
      if (row0 + i < m) {
        const double av = (double)a[(int64_t)(row0 + i) * lda + c];
#pragma unroll
        for (int r = 0; r < NR; ++r) s[i][r] += av * xv[r];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kGemvRows; ++i)
#pragma unroll
    for (int r = 0; r < NR; ++r) s[i][r] = wave_sum(s[i][r]);
  if (lane < kGemvRows * NR) {
    const int i = lane / NR, r = lane % NR;
    double v = 0.0;
#pragma unroll
    for (int ii = 0; ii < kGemvRows; ++ii)
#pragma unroll
      for (int rr = 0; rr < NR; ++rr)
        if (ii == i && rr == r) v = s[ii][rr];
    const int row = row0 + i;
    if (row < m) {
      const double out = b ? (double)b[(int64_t)r * ldb + row] - v : v;
      y[(int64_t)r * ldy + row] = (TY)out;
    }
  }
}

// The same product through 16-byte loads, for 16-byte aligned rows of A and
// x with k a multiple of four: each lane takes four consecutive columns of
// the wave's four rows per step (4 KB of A per wave per step, as against
// 1 KB with a float per lane).
__device__ __forceinline__ void gemv_x4(const float* p, double (&o)[4]) {
  const float4 q = *reinterpret_cast<const float4*>(p);
  o[0] = q.x;
  o[1] = q.y;
  o[2] = q.z;
  o[3] = q.w;
}
__device__ __forceinline__ void gemv_x4(const double* p, double (&o)[4]) {
  const double2 q0 = *reinterpret_cast<const double2*>(p);
  const double2 q1 = *reinterpret_cast<const double2*>(p + 2);
  o[0] = q0.x;
  o[1] = q0.y;
  o[2] = q1.x;
  o[3] = q1.y;
}

template <typename TX, typename TY, int NR>
__global__ __launch_bounds__(256) void gemv_f32a4_kernel(int m, int k, const float* __restrict__ a, int64_t lda,
                                                         const TX* __restrict__ x, int64_t ldx,
                                                         const TY* __restrict__ b, int64_t ldb, TY* __restrict__ y,
                                                         int64_t ldy) {
  const int lane = threadIdx.x & 63;
  const int row0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * kGemvRows;
  if (row0 >= m) return;
  double s[kGemvRows][NR];
#pragma unroll
  for (int i = 0; i < kGemvRows; ++i)
#pragma unroll
    for (int r = 0; r < NR; ++r) s[i][r] = 0.0;
  // rows past m read the last row; their sums are never written
  const float* arow[kGemvRows];
#pragma unroll
  for (int i = 0; i < kGemvRows; ++i) arow[i] = a + (int64_t)(row0 + i < m ? row0 + i : m - 1) * lda;
  for (int c = 4 * lane; c < k; c += 4 * kWave) {
    float4 av[kGemvRows];
#pragma unroll
    for (int i = 0; i < kGemvRows; ++i) av[i] = *reinterpret_cast<const float4*>(arow[i] + c);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      double xv[4];
      gemv_x4(x + (int64_t)r * ldx + c, xv);
#pragma unroll
      for (int i = 0; i < kGemvRows; ++i) {
        s[i][r] = fma((double)av[i].x, xv[0], s[i][r]);
        s[i][r] = fma((double)av[i].y, xv[1], s[i][r]);
        s[i][r] = fma((double)av[i].z, xv[2], s[i][r]);
        s[i][r] = fma((double)av[i].w, xv[3], s[i][r]);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kGemvRows; ++i)
#pragma unroll
    for (int r = 0; r < NR; ++r) s[i][r] = wave_sum(s[i][r]);
  if (lane < kGemvRows * NR) {
    const int i = lane / NR, r = lane % NR;
    double v = 0.0;
#pragma unroll
    for (int ii = 0; ii < kGemvRows; ++ii)
#pragma unroll
      for (int rr = 0; rr < NR; ++rr)
        if (ii == i && rr == r) v = s[ii][rr];
    const int row = row0 + i;
    if (row < m) {
      const double out = b ? (double)b[(int64_t)r * ldb + row] - v : v;
      y[(int64_t)r * ldy + row] = (TY)out;
    }
  }
}

template <typename TX, typename TY>
int gemv_f32a_launch(hipStream_t st, int m, int k, int nrhs, const float* a, int64_t lda, const TX* x, int64_t ldx,
                     const TY* b, int64_t ldb, TY* y, int64_t ldy) {
  const unsigned blocks = (unsigned)((m + 4 * kGemvRows - 1) / (4 * kGemvRows));
  // 16-byte loads when every row of A and of x starts 16-byte aligned
  const bool vec4 = k % 4 == 0 && lda % 4 == 0 && ldx % 4 == 0 && (reinterpret_cast<uintptr_t>(a) & 15) == 0 &&
                    (reinterpret_cast<uintptr_t>(x) & 15) == 0;
  // split into launches of 8, 4, 2, 1 right-hand sides
  int done = 0;
  while (done < nrhs) {
    const int left = nrhs - done;
    const TX* xp = x + (int64_t)done * ldx;
    const TY* bp = b ? b + (int64_t)done * ldb : nullptr;
    TY* yp = y + (int64_t)done * ldy;
    int take;
    if (vec4) {
      if (left >= 8) {
        take = 8;
        gemv_f32a4_kernel<TX, TY, 8><<<blocks, 256, 0, st>>>(m, k, a, lda, xp, ldx, bp, ldb, yp, ldy);
      } else if (left >= 4) {
        take = 4;
        gemv_f32a4_kernel<TX, TY, 4><<<blocks, 256, 0, st>>>(m, k, a, lda, xp, ldx, bp, ldb, yp, ldy);
      } else if (left >= 2) {
        take = 2;
        gemv_f32a4_kernel<TX, TY, 2><<<blocks, 256, 0, st>>>(m, k, a, lda, xp, ldx, bp, ldb, yp, ldy);
      } else {
        take = 1;
        gemv_f32a4_kernel<TX, TY, 1><<<blocks, 256, 0, st>>>(m, k, a, lda, xp, ldx, bp, ldb, yp, ldy);
      }
    } else if (left >= 8) {
      take = 8;
      gemv_f32a_kernel<TX, TY, 8><<<blocks, 256, 0, st>>>(m, k, a, lda, xp, ldx, bp, ldb, yp, ldy);
    } else if (left >= 4) {
      take = 4;
      gemv_f32a_kernel<TX, TY, 4><<<blocks, 256, 0, st>>>(m, k, a, lda, xp, ldx, bp, ldb, yp, ldy);
    } else if (left >= 2) {
      take = 2;
      gemv_f32a_kernel<TX, TY, 2><<<blocks, 256, 0, st>>>(m, k, a, lda, xp, ldx, bp, ldb, yp, ldy);
    } else {
      take = 1;
      gemv_f32a_kernel<TX, TY, 1><<<blocks, 256, 0, st>>>(m, k, a, lda, xp, ldx, bp, ldb, yp, ldy);
    }
    LXA_LAUNCH_CHECK("gemv_f32a_kernel");
    done += take;
  }
  return LXA_OK;
}

// Per device: the side stream that runs the far part of each outer Schur
// update, and the two events that order it against the main stream.  Made
// once, kept for the process; the mutex serialises factorisations that are
// enqueued from several host threads (they would share these).
struct LookAhead {
  hipStream_t side = nullptr;
  hipEvent_t u12_done = nullptr, far_done = nullptr;
};
constexpr int kMaxDevices = 64;
constexpr int kFarChunk = 1024;  // columns per launch of the far update
std::mutex g_look_mu[kMaxDevices];
LookAhead g_look[kMaxDevices];

int look_ahead_for(int dev, LookAhead** out) {
  LookAhead& la = g_look[dev];
  if (!la.side) {
    // non-blocking: the caller's stream may be the null stream
    hipError_t e = hipStreamCreateWithFlags(&la.side, hipStreamNonBlocking);
    if (e != hipSuccess) {
      la.side = nullptr;
      return hip_error(e, "hipStreamCreateWithFlags");
    }
    e = hipEventCreateWithFlags(&la.u12_done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&la.far_done, hipEventDisableTiming);
    if (e != hipSuccess) return hip_error(e, "hipEventCreateWithFlags");
  }
  *out = &la;
  return LXA_OK;
}

// Blocked right-looking LU with look-ahead.  Outer panels of 512 columns
// [K0, K1), each factored by 128-column inner steps that touch only the
// panel's columns (row interchanges included).  The outer Schur update of
// the trailing matrix is split: its first 512 columns (the next panel,
// "near") on the caller's stream, the rest ("far") on a side stream, where
// it runs while the next panel is factored.  A panel's interchanges reach
// the columns outside it only after the previous far update is done (that
// update reads and writes those rows); as the far update acts on whole rows
// of the trailing matrix and of its multipliers alike, applying the
// interchanges after it gives the same matrix as applying them before.
int lu_large_factor_impl(hipStream_t st, int mode, int n, float* a, int64_t lda, float* linv, float* uinv,
                         int32_t* piv, int32_t* info) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_error(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return set_error(LXA_EINVAL, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_look_mu[dev]);
  LookAhead* la = nullptr;
  if (int rc = look_ahead_for(dev, &la)) return rc;
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(&panel_lu_kernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPanelLds);
  if (e != hipSuccess) return hip_error(e, "hipFuncSetAttribute(panel_lu_kernel)");
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(&outer_swaps_kernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)kOsLds);
  if (e != hipSuccess) return hip_error(e, "hipFuncSetAttribute(outer_swaps_kernel)");
  bool far_pending = false;
  for (int K0 = 0; K0 < n; K0 += kOuter) {
    const int K1 = min(n, K0 + kOuter);
    for (int k = K0; k < K1; k += kB) {
      const int k2 = min(K1, k + kB), nb = k2 - k;
      // pivot window: the block and the 128 rows below it (all rows below
      // are current in these columns: the inner updates are right-looking)
      const int win = nb < kB ? nb : min(kPivWin, n - k);
      float* Li = linv + (int64_t)(k / kB) * kB * kB;
      float* Ui = uinv + (int64_t)(k / kB) * kB * kB;
      // (the kernel also applies the block's interchanges to the rest of
      // the panel's columns)
      panel_lu_kernel<<<1, kPanelThreads, kPanelLds, st>>>(k, nb, win, a, lda, Li, Ui, piv, info, K0, K1);
      LXA_LAUNCH_CHECK("panel_lu_kernel");
      float* A11 = a + (int64_t)k * lda + k;
      // L21 = A21 U11^-1 (in place, N = nb <= 128) for the rows below the
      // window (the window's rows hold theirs from panel_lu_kernel)
      if (n - k - win > 0) {
        if (int rc = gemm_update(st, LXA_GEMM_F32, n - k - win, nb, nb, A11 + (int64_t)win * lda, lda, Ui, kB,
                                 A11 + (int64_t)win * lda, lda, 0, 1.f))
          return rc;
      }
      // panel U12 = L11^-1 A12 (in place, M = nb <= 128), then the panel's Schur update
      if (K1 - k2 > 0) {
        if (int rc = gemm_update(st, LXA_GEMM_F32, nb, K1 - k2, nb, Li, kB, A11 + nb, lda, A11 + nb, lda, 0, 1.f))
          return rc;
        if (n - k2 > 0) {
          if (int rc = gemm_update(st, LXA_GEMM_F32, n - k2, K1 - k2, nb, A11 + (int64_t)nb * lda, lda, A11 + nb, lda,
                                   A11 + (int64_t)nb * lda + nb, lda, 1, -1.f))
            return rc;
        }
      }
    }
    // the panel's interchanges outside its columns, once the previous far
    // update is done with those rows
    if (far_pending) {
      e = hipStreamWaitEvent(st, la->far_done, 0);
      if (e != hipSuccess) return hip_error(e, "hipStreamWaitEvent");
      far_pending = false;
    }
    if (const int cols = n - (K1 - K0); cols > 0) {
      outer_swaps_kernel<<<(cols + kOsCols - 1) / kOsCols, 256, kOsLds, st>>>(n, K0, K1, a, lda, piv);
      LXA_LAUNCH_CHECK("outer_swaps_kernel");
    }
    if (n - K1 <= 0) continue;
    // outer U12 = L11^-1 A12 over the panel's 128-row blocks
    for (int k = K0; k < K1; k += kB) {
      const int nb = min(K1, k + kB) - k;
      float* rowblk = a + (int64_t)k * lda + K1;
      if (k > K0) {
        if (int rc = gemm_update(st, LXA_GEMM_F32, nb, n - K1, k - K0, a + (int64_t)k * lda + K0, lda,
                                 a + (int64_t)K0 * lda + K1, lda, rowblk, lda, 1, -1.f))
          return rc;
      }
      if (int rc = gemm_update(st, LXA_GEMM_F32, nb, n - K1, nb, linv + (int64_t)(k / kB) * kB * kB, kB, rowblk, lda,
                               rowblk, lda, 0, 1.f))
        return rc;
    }
    // outer Schur update: near columns here, far columns on the side stream
    const float* L21 = a + (int64_t)K1 * lda + K0;
    const float* U12 = a + (int64_t)K0 * lda + K1;
    float* A22 = a + (int64_t)K1 * lda + K1;
    const int near = min(kOuter, n - K1);
    if (int rc = gemm_update(st, mode, n - K1, near, K1 - K0, L21, lda, U12, lda, A22, lda, 1, -1.f)) return rc;
    if (n - K1 - near > 0) {
      e = hipEventRecord(la->u12_done, st);
      if (e == hipSuccess) e = hipStreamWaitEvent(la->side, la->u12_done, 0);
      if (e != hipSuccess) return hip_error(e, "look-ahead event");
      // in column chunks, so that the next panel's kernels find compute
      // units between them
      for (int c = near; c < n - K1; c += kFarChunk) {
        const int w = min(kFarChunk, n - K1 - c);
        if (int rc = gemm_update(la->side, mode, n - K1, w, K1 - K0, L21, lda, U12 + c, lda, A22 + c, lda, 1, -1.f))
          return rc;
      }
      e = hipEventRecord(la->far_done, la->side);
      if (e != hipSuccess) return hip_error(e, "hipEventRecord");
      far_pending = true;
    }
  }
  if (far_pending) {
    e = hipStreamWaitEvent(st, la->far_done, 0);
    if (e != hipSuccess) return hip_error(e, "hipStreamWaitEvent");
  }
  return LXA_OK;
}

}  // namespace
}  // namespace lxa

using namespace lxa;

extern "C" int lxa_lu_large_block(void) { return kB; }

extern "C" int lxa_lu_large_factor(void* stream, int mode, int64_t n, float* a, int64_t lda, float* linv, float* uinv,
                                   int32_t* piv, int32_t* info) {
  LXA_REQUIRE(n >= 0 && n <= (1 << 20), "n out of range (%lld)", (long long)n);
  LXA_REQUIRE(lda >= n, "lda (%lld) < n (%lld)", (long long)lda, (long long)n);
  LXA_REQUIRE(mode == LXA_GEMM_F32 || mode == LXA_GEMM_BF16 || mode == LXA_GEMM_F16 || mode == LXA_GEMM_BF16X3,
              "unknown mode %d", mode);
  if (n == 0) return LXA_OK;
  LXA_REQUIRE(a && linv && uinv && piv && info, "a, linv, uinv, piv and info must be non-NULL");
  return lu_large_factor_impl(static_cast<hipStream_t>(stream), mode, (int)n, a, lda, linv, uinv, piv, info);
}

extern "C" int lxa_lu_large_perm(void* stream, int64_t n, const int32_t* piv, int32_t* perm) {
  LXA_REQUIRE(n >= 0 && n <= (1 << 20), "invalid n");
  if (n == 0) return LXA_OK;
  LXA_REQUIRE(piv && perm, "piv and perm must be non-NULL");
  perm_from_swaps_kernel<<<(unsigned)((n + 255) / 256), 256, 0, static_cast<hipStream_t>(stream)>>>((int)n, piv,
                                                                                                      perm);
  LXA_LAUNCH_CHECK("perm_from_swaps_kernel");
  return LXA_OK;
}

extern "C" int lxa_lu_large_permute(void* stream, int64_t n, int nrhs, const float* x, int64_t ldx, float* y,
                                    int64_t ldy, const int32_t* perm, int inverse) {
  LXA_REQUIRE(n >= 0 && n <= (1 << 20) && nrhs >= 0 && ldx >= n && ldy >= n, "invalid n / nrhs / ldx / ldy");
  if (n == 0 || nrhs == 0) return LXA_OK;
  LXA_REQUIRE(x && y && perm && x != y, "x, y and perm must be non-NULL, y must not alias x");
  const int64_t threads = n * nrhs;
  permute_rows_kernel<<<(unsigned)((threads + 255) / 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
      (int)n, nrhs, x, ldx, y, ldy, perm, inverse ? 1 : 0);
  LXA_LAUNCH_CHECK("permute_rows_kernel");
  return LXA_OK;
}

extern "C" int lxa_lu_large_trsv_workspace(int64_t n, int64_t* words) {
  LXA_REQUIRE(n >= 0 && n <= (1 << 20) && words, "lxa_lu_large_trsv_workspace: bad argument");
  const int64_t nblk = (n + kB - 1) / kB;
  *words = kTrsvHeader + 2 * nblk * kB * kMaxRhs;
  return LXA_OK;
}

extern "C" int lxa_lu_large_trsv(void* stream, int kind, int64_t n, int nrhs, const float* lu, int64_t lda,
                                 const float* inv, float* x, int64_t ldx, int32_t* sync) {
  LXA_REQUIRE(kind >= 0 && kind <= 3, "kind must be 0..3 (got %d)", kind);
  LXA_REQUIRE(n >= 0 && n <= (1 << 20) && lda >= n, "invalid n / lda");
  LXA_REQUIRE(nrhs >= 0 && nrhs <= kMaxRhs, "nrhs must be in [0, %d] (got %d)", kMaxRhs, nrhs);
  if (n == 0 || nrhs == 0) return LXA_OK;
  LXA_REQUIRE(lu && inv && x && sync && ldx >= n, "invalid pointers or ldx");
  LXA_REQUIRE((reinterpret_cast<uintptr_t>(sync) & 15) == 0, "the workspace must be 16-byte aligned");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nblk = (int)((n + kB - 1) / kB);
  const int64_t chan_words = (int64_t)nblk * kB * kMaxRhs;
  const int64_t arm_blocks = (chan_words + 255) / 256 < 1024 ? (chan_words + 255) / 256 : 1024;
  trsv_arm_kernel<<<(unsigned)arm_blocks, 256, 0, st>>>(sync, chan_words);
  LXA_LAUNCH_CHECK("trsv_arm_kernel");
  const size_t smem = sizeof(float) * (2 * kB * kTileLd + 2 * kMaxRhs * kB);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&trsv_blocks_kernel<4>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e == hipSuccess)
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&trsv_blocks_kernel<kMaxRhs>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) return hip_error(e, "hipFuncSetAttribute(trsv_blocks_kernel)");
  if (nrhs <= 4)
    trsv_blocks_kernel<4><<<nblk, kTrsvThreads, smem, st>>>((int)n, nrhs, lu, lda, inv, x, ldx, kind, sync);
  else
    trsv_blocks_kernel<kMaxRhs><<<nblk, kTrsvThreads, smem, st>>>((int)n, nrhs, lu, lda, inv, x, ldx, kind, sync);
  LXA_LAUNCH_CHECK("trsv_blocks_kernel");
  return LXA_OK;
}

extern "C" int lxa_rbt_matrix(void* stream, int64_t n, float* a, int64_t lda, const float* du, const float* dv) {
  LXA_REQUIRE(n >= 0 && n % 4 == 0 && lda >= n, "butterflies need n divisible by 4 (got %lld)", (long long)n);
  if (n == 0) return LXA_OK;
  LXA_REQUIRE(a && du && dv, "a, du and dv must be non-NULL");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t quads = (n / 2) * (n / 2);
  const unsigned blocks = (unsigned)((quads + 255) / 256);
  // inner level (two butterflies of order n/2), then the outer level
  rbt_matrix_kernel<<<blocks, 256, 0, st>>>((int)n, (int)(n / 4), a, lda, du + n, dv + n);
  LXA_LAUNCH_CHECK("rbt_matrix_kernel");
  rbt_matrix_kernel<<<blocks, 256, 0, st>>>((int)n, (int)(n / 2), a, lda, du, dv);
  LXA_LAUNCH_CHECK("rbt_matrix_kernel");
  return LXA_OK;
}

extern "C" int lxa_rbt_vectors(void* stream, int dtype, int64_t n, int nrhs, void* x, int64_t ldx, const float* d,
                               int form) {
  LXA_REQUIRE(dtype == LXA_F32 || dtype == LXA_F64, "dtype must be LXA_F32 or LXA_F64");
  LXA_REQUIRE(n >= 0 && n % 4 == 0 && ldx >= n && nrhs >= 0, "invalid n / ldx / nrhs");
  LXA_REQUIRE(form == 0 || form == 1, "form must be 0 (B) or 1 (B^T)");
  if (n == 0 || nrhs == 0) return LXA_OK;
  LXA_REQUIRE(x && d, "x and d must be non-NULL");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const unsigned blocks = (unsigned)(((n / 2) * nrhs + 255) / 256);
  // B^T (applied for U^T b): inner level then outer; B (for V y): outer then inner
  const int order[2][2] = {{(int)(n / 2), (int)(n / 4)}, {(int)(n / 4), (int)(n / 2)}};
  for (int t = 0; t < 2; ++t) {
    const int s = order[form][t];
    const float* dd = (s == n / 2) ? d : d + n;
    if (dtype == LXA_F32)
      rbt_vec_kernel<float><<<blocks, 256, 0, st>>>((int)n, s, nrhs, static_cast<float*>(x), ldx, dd, form);
    else
      rbt_vec_kernel<double><<<blocks, 256, 0, st>>>((int)n, s, nrhs, static_cast<double*>(x), ldx, dd, form);
    LXA_LAUNCH_CHECK("rbt_vec_kernel");
  }
  return LXA_OK;
}

extern "C" int lxa_gemv_f32a(void* stream, int64_t m, int64_t k, int nrhs, const float* a, int64_t lda,
                             int x_dtype, const void* x, int64_t ldx, const void* b, int64_t ldb, int y_dtype, void* y,
                             int64_t ldy) {
  LXA_REQUIRE(m >= 0 && k >= 0 && lda >= k && m < (1 << 30) && k < (1 << 30), "invalid m / k / lda");
  LXA_REQUIRE(nrhs >= 0 && nrhs <= 65535, "nrhs out of range (%d)", nrhs);
  LXA_REQUIRE((x_dtype == LXA_F32 || x_dtype == LXA_F64) && (y_dtype == LXA_F32 || y_dtype == LXA_F64),
              "dtypes must be LXA_F32 or LXA_F64");
  if (m == 0 || nrhs == 0) return LXA_OK;
  LXA_REQUIRE(a && x && y && ldx >= k && ldy >= m && (!b || ldb >= m), "invalid pointers or strides");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (x_dtype == LXA_F32 && y_dtype == LXA_F32)
    return gemv_f32a_launch<float, float>(st, (int)m, (int)k, nrhs, a, lda, static_cast<const float*>(x), ldx,
                                          static_cast<const float*>(b), ldb, static_cast<float*>(y), ldy);
  if (x_dtype == LXA_F32)
    return gemv_f32a_launch<float, double>(st, (int)m, (int)k, nrhs, a, lda, static_cast<const float*>(x), ldx,
                                           static_cast<const double*>(b), ldb, static_cast<double*>(y), ldy);
  if (y_dtype == LXA_F32)
    return gemv_f32a_launch<double, float>(st, (int)m, (int)k, nrhs, a, lda, static_cast<const double*>(x), ldx,
                                           static_cast<const float*>(b), ldb, static_cast<float*>(y), ldy);
  return gemv_f32a_launch<double, double>(st, (int)m, (int)k, nrhs, a, lda, static_cast<const double*>(x), ldx,
                                          static_cast<const double*>(b), ldb, static_cast<double*>(y), ldy);
}
