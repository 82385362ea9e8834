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

#include "lxa_common.h"

namespace lxa {
namespace {

constexpr int kB = 128;          // diagonal / triangular-solve block
constexpr int kOuter = 512;      // outer panel width
constexpr int kDiagThreads = 1024;
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
// LU with partial pivoting restricted to the nb x nb diagonal block at (k, k)
// (nb <= 128), in place, and the inverses of its unit-lower and upper
// triangles written as dense 128 x 128 blocks (zeros outside the triangle).
// Pivoting inside the block keeps its multipliers <= 1 in magnitude, so the
// block's L is well conditioned and its explicit inverse accurate; the
// butterflies keep the growth across blocks in check.  piv[k + j] receives
// the 0-based global row swapped with row k + j (always inside the block);
// the caller applies the swaps to the rest of the block's rows.  info[0] is
// set to 1 + the global index of the first exactly-zero pivot.
// ---------------------------------------------------------------------------
// 1024 threads as a 32 x 32 grid; thread (ty, tx) keeps the 4 x 4 elements
// (ty + 32 a, tx + 32 b) in registers.  Per column: the owners of column j
// publish it, every wave finds the pivot from it redundantly, the owners of
// rows j and p publish those rows, and each thread swaps and updates its own
// elements -- two barriers per column, LDS buffers alternating by column.
// The inverses run the same way on register tiles with the factors staged
// in LDS, one barrier per column.
// One step of a wave arg-max (value, lowest index on ties) with a DPP
// partner: a register-crossbar move, no LDS round trip.
template <int kCtrl>
__device__ __forceinline__ void amax_dpp(float& v, int& idx) {
  const float ov = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), kCtrl, 0xF, 0xF, false));
  const int oi = __builtin_amdgcn_mov_dpp(idx, kCtrl, 0xF, 0xF, false);
  if (ov > v || (ov == v && oi < idx)) {
    v = ov;
    idx = oi;
  }
}

// Arg-max over the wave: quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror and row_mirror leave each 16-lane row's arg-max in all of
// its lanes; the four rows are combined from scalar reads.  The result is
// wave-uniform.
__device__ __forceinline__ int wave_argmax(float v, int idx) {
  amax_dpp<0xB1>(v, idx);
  amax_dpp<0x4E>(v, idx);
  amax_dpp<0x141>(v, idx);
  amax_dpp<0x140>(v, idx);
  float bv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
  int bi = __builtin_amdgcn_readlane(idx, 0);
#pragma unroll
  for (int r = 1; r < 4; ++r) {
    const float rv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16 * r));
    const int ri = __builtin_amdgcn_readlane(idx, 16 * r);
    if (rv > bv || (rv == bv && ri < bi)) {
      bv = rv;
      bi = ri;
    }
  }
  return bi;
}

// Rows of the pivot window: the diagonal block and the block below it.
// Pivoting over 256 rows keeps every multiplier of the block column (the
// diagonal block's L and the next block's L21) bounded by 1; over the
// diagonal block alone, the multipliers below it are unbounded and the
// growth reaches ~1e3 on Gaussian operators (~1e-5 backward error instead
// of ~1e-7).
constexpr int kPivWin = 2 * kB;

__global__ __launch_bounds__(kDiagThreads) void diag_lu_kernel(int k, int nb, int win, float* __restrict__ a,
                                                               int64_t lda, float* __restrict__ linv,
                                                               float* __restrict__ uinv, int32_t* __restrict__ piv,
                                                               int32_t* __restrict__ info) {
  extern __shared__ __align__(16) float dsm[];
  float* S = dsm;  // [128][129]: the factors, for the inverse sweeps
  constexpr int LD = kB + 1;
  constexpr int kRowT = kPivWin / 32;  // register rows per thread
  __shared__ float colv[2][kPivWin];
  __shared__ float rowj[2][kB];
  __shared__ float rowp[2][kB];
  const int tid = threadIdx.x, tx = tid & 31, ty = tid >> 5, lane = tid & 63;
  // R[ai][bi]: window row ty + 32 ai, column tx + 32 bi.  Rows of the
  // diagonal block beyond nb are identity padding; window rows beyond win
  // are zero and never chosen.
  float R[kRowT][4];
#pragma unroll
  for (int ai = 0; ai < kRowT; ++ai)
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
      const int r = ty + 32 * ai, c = tx + 32 * bi;
      if (r < kB)
        R[ai][bi] = (r < nb && c < nb) ? a[(int64_t)(k + r) * lda + k + c] : (r == c ? 1.f : 0.f);
      else
        R[ai][bi] = (r < win && c < nb) ? a[(int64_t)(k + r) * lda + k + c] : 0.f;
    }
  for (int j = 0; j < kB; ++j) {
    const int buf = j & 1, jb = j >> 5, jt = j & 31;
    if (tx == jt) {
#pragma unroll
      for (int ai = 0; ai < kRowT; ++ai)
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
          if (bi == jb) colv[buf][ty + 32 * ai] = R[ai][bi];
    }
    __syncthreads();
    // arg-max of |column j| over window rows >= j (NaN counts as largest;
    // ties: lowest row), computed by every wave
    float v = -1.f;
    int idx = lane;
#pragma unroll
    for (int q = 0; q < kPivWin / kWave; ++q) {
      const int row = lane + kWave * q;
      float w = -1.f;
      if (row >= j && (row < kB || row < win)) {
        w = fabsf(colv[buf][row]);
        if (w != w) w = INFINITY;
      }
      if (w > v) {
        v = w;
        idx = row;
      }
    }
    const int p = wave_argmax(v, idx);
    const float pv = colv[buf][p];
    // rows j and p to LDS (before the swap)
    if (ty == jt) {
#pragma unroll
      for (int ai = 0; ai < 4; ++ai)
        if (ai == jb) {
#pragma unroll
          for (int bi = 0; bi < 4; ++bi) rowj[buf][tx + 32 * bi] = R[ai][bi];
        }
    }
    if (ty == (p & 31)) {
#pragma unroll
      for (int ai = 0; ai < kRowT; ++ai)
        if (ai == (p >> 5)) {
#pragma unroll
          for (int bi = 0; bi < 4; ++bi) rowp[buf][tx + 32 * bi] = R[ai][bi];
        }
    }
    if (tid == 0 && j < nb) {
      piv[k + j] = k + p;
      if (pv == 0.f && info[0] == 0) info[0] = k + j + 1;
    }
    __syncthreads();
    // swap rows j and p, multipliers into column j, rank-1 update
#pragma unroll
    for (int ai = 0; ai < kRowT; ++ai) {
      const int r = ty + 32 * ai;
      if (r == j) {
#pragma unroll
        for (int bi = 0; bi < 4; ++bi) R[ai][bi] = rowp[buf][tx + 32 * bi];
      } else if (r > j) {
        const bool isp = r == p;
        const float l = (isp ? colv[buf][j] : colv[buf][r]) / pv;
#pragma unroll
        for (int bi = 0; bi < 4; ++bi) {
          const int c = tx + 32 * bi;
          const float old = isp ? rowj[buf][c] : R[ai][bi];
          R[ai][bi] = c == j ? l : (c > j ? old - l * rowp[buf][c] : old);
        }
      }
    }
  }
  // the factors to global memory and to LDS; the window rows below the
  // block leave with their final multipliers (their L21 block)
#pragma unroll
  for (int ai = 0; ai < kRowT; ++ai)
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
      const int r = ty + 32 * ai, c = tx + 32 * bi;
      if (r < kB) {
        S[r * LD + c] = R[ai][bi];
        if (r < nb && c < nb) a[(int64_t)(k + r) * lda + k + c] = R[ai][bi];
      } else if (r < win && c < nb) {
        a[(int64_t)(k + r) * lda + k + c] = R[ai][bi];
      }
    }
  __syncthreads();
  // X = L^-1 (unit lower): for j: X[i][:] -= L[i][j] X[j][:], i > j.
  // Y = (U^T)^-1 (lower): for j: Y[j][:] /= U[j][j]; Y[i][:] -= U[j][i] Y[j][:].
  // uinv = Y^T.  Row j of each is final when column j is processed.
  float X[4][4], Y[4][4];
#pragma unroll
  for (int ai = 0; ai < 4; ++ai)
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) X[ai][bi] = Y[ai][bi] = (ty + 32 * ai == tx + 32 * bi) ? 1.f : 0.f;
  for (int j = 0; j < kB; ++j) {
    const int buf = j & 1, jb = j >> 5, jt = j & 31;
    if (ty == jt) {
      const float d = S[j * LD + j];
#pragma unroll
      for (int ai = 0; ai < 4; ++ai)
        if (ai == jb) {
#pragma unroll
          for (int bi = 0; bi < 4; ++bi) {
            Y[ai][bi] /= d;
            rowj[buf][tx + 32 * bi] = X[ai][bi];
            rowp[buf][tx + 32 * bi] = Y[ai][bi];
          }
        }
    }
    __syncthreads();
#pragma unroll
    for (int ai = 0; ai < 4; ++ai) {
      const int i = ty + 32 * ai;
      if (i > j) {
        const float lij = S[i * LD + j], uji = S[j * LD + i];
#pragma unroll
        for (int bi = 0; bi < 4; ++bi) {
          const int c = tx + 32 * bi;
          if (c <= j) {
            X[ai][bi] -= lij * rowj[buf][c];
            Y[ai][bi] -= uji * rowp[buf][c];
          }
        }
      }
    }
  }
#pragma unroll
  for (int ai = 0; ai < 4; ++ai)
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
      const int r = ty + 32 * ai, c = tx + 32 * bi;
      linv[r * kB + c] = X[ai][bi];
      uinv[c * kB + r] = Y[ai][bi];
    }
}

__global__ void zero_i32_kernel(int32_t* __restrict__ p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}

// The row swaps of block column k (pivot steps k .. k + nb - 1, in order,
// each within the window rows k .. k + win - 1) applied to every column
// outside the block column: one thread per column, the window's rows staged
// in LDS.
constexpr int kSwapCols = 128;

__global__ __launch_bounds__(kSwapCols) void block_row_swaps_kernel(int n, int k, int nb, int win,
                                                                    float* __restrict__ a, int64_t lda,
                                                                    const int32_t* __restrict__ piv) {
  extern __shared__ __align__(16) float col[];  // [win][kSwapCols]
  const int t = threadIdx.x;
  const int c = blockIdx.x * kSwapCols + t;
  const bool live = c < n && (c < k || c >= k + nb);
  if (live)
    for (int r = 0; r < win; ++r) col[r * kSwapCols + t] = a[(int64_t)(k + r) * lda + c];
  if (!live) return;
  for (int j = 0; j < nb; ++j) {
    const int p = piv[k + j] - k;
    if (p != j) {
      const float v = col[j * kSwapCols + t];
      col[j * kSwapCols + t] = col[p * kSwapCols + t];
      col[p * kSwapCols + t] = v;
    }
  }
  for (int r = 0; r < win; ++r) a[(int64_t)(k + r) * lda + c] = col[r * kSwapCols + t];
}

// The row permutation of the whole factorisation: perm[i] = the original
// row now at position i (the swap sequence piv applied in order to the
// identity).  One workgroup; the sequence is inherently sequential, so one
// thread walks it over an LDS copy (n <= kPermLds) or over global memory.
constexpr int kPermLds = 32768;

__global__ __launch_bounds__(256) void perm_from_swaps_kernel(int n, const int32_t* __restrict__ piv,
                                                              int32_t* __restrict__ perm) {
  extern __shared__ int32_t pl[];
  const bool in_lds = n <= kPermLds;
  int32_t* p = in_lds ? pl : perm;
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 0; j < n; ++j) {
      const int q = piv[j];
      if (q != j) {
        const int32_t t = p[j];
        p[j] = p[q];
        p[q] = t;
      }
    }
  }
  __syncthreads();
  if (in_lds)
    for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = pl[i];
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
template <int R>
__global__ __launch_bounds__(kTrsvThreads) void trsv_blocks_kernel(int n, int nrhs, const float* __restrict__ lu,
                                                                   int64_t lda, const float* __restrict__ inv,
                                                                   float* __restrict__ x, int64_t ldx, int kind,
                                                                   int32_t* __restrict__ sync) {
  extern __shared__ __align__(16) float tsm[];
  float* tile = tsm;                                                     // [128][132]
  float(*xs)[kB] = reinterpret_cast<float(*)[kB]>(tsm + kB * kTileLd);   // [8][128] solved block
  float(*acc)[kB] = reinterpret_cast<float(*)[kB]>(tsm + kB * kTileLd + kMaxRhs * kB);  // [8][128]
  __shared__ int s_ticket;
  int32_t* ticket = sync;
  int32_t* flags = sync + 1;
  int32_t* failed = sync + 1 + ((n + kB - 1) / kB);
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
  // operator blocks do not depend on the solution: block `step + 1` is
  // loaded while block `step` waits for its x and is applied
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
  if (t > 0) {
    const int bj = block_col(0);
    tile_load(pre, lu, lda, kind, bi, bj, rows, min(kB, n - bj * kB), tid);
  } else {
    load_inv();
  }
  for (int step = 0; step < t; ++step) {
    const int bj = block_col(step);
    const int c0 = bj * kB, cols = min(kB, n - c0);
    tile_store(tile, pre, tr, tid);
    if (step + 1 < t) {
      const int bn = block_col(step + 1);
      tile_load(pre, lu, lda, kind, bi, bn, rows, min(kB, n - bn * kB), tid);
    } else {
      load_inv();  // in flight while the last dependency is awaited
    }
    if (tid == 0) {
      int spin = 0;
      while (__hip_atomic_load(&flags[bj], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        if (++spin > kSpinLimit) {
          __hip_atomic_store(failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __syncthreads();
    (void)__hip_atomic_load(&flags[bj], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    for (int idx = tid; idx < kMaxRhs * kB; idx += kTrsvThreads) {
      const int rr = idx / kB, i = idx % kB;
      xs[rr][i] = (rr < nrhs && i < cols) ? x[(int64_t)rr * ldx + c0 + i] : 0.f;
    }
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
  // y = inv(D) acc, with D staged like an operator block (transposed kinds
  // use inv(D)^T); its loads were issued before the last wait
  tile_store(tile, pre, tr, tid);
  __syncthreads();
  float part[R];
  tile_product<R>(tile, acc, row, hcol, part);
  if ((tid & 1) == 0 && row < rows) {
#pragma unroll
    for (int rr = 0; rr < R; ++rr)
      if (rr < nrhs) x[(int64_t)rr * ldx + r0 + row] = part[rr];
  }
  // Every wave makes its own stores visible device-wide before the flag is
  // raised (a fence orders only the issuing wave's memory operations).
  __threadfence();
  __syncthreads();
  if (tid == 0) __hip_atomic_store(&flags[bi], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
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

template <typename TX, typename TY>
int gemv_f32a_launch(hipStream_t st, int m, int k, int nrhs, const float* a, int64_t lda, const TX* x, int64_t ldx,
                     const TY* b, int64_t ldb, TY* y, int64_t ldy) {
  const unsigned blocks = (unsigned)((m + 4 * kGemvRows - 1) / (4 * kGemvRows));
  // split into launches of 8, 4, 2, 1 right-hand sides
  int done = 0;
  while (done < nrhs) {
    const int left = nrhs - done;
    const TX* xp = x + (int64_t)done * ldx;
    const TY* bp = b ? b + (int64_t)done * ldb : nullptr;
    TY* yp = y + (int64_t)done * ldy;
    int take;
    if (left >= 8) {
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

int lu_large_factor_impl(hipStream_t st, int mode, int n, float* a, int64_t lda, float* linv, float* uinv,
                         int32_t* piv, int32_t* info) {
  const size_t dsmem = sizeof(float) * kB * (kB + 1);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&diag_lu_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)dsmem);
  if (e != hipSuccess) return hip_error(e, "hipFuncSetAttribute(diag_lu_kernel)");
  e = hipFuncSetAttribute(reinterpret_cast<const void*>(&block_row_swaps_kernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)(sizeof(float) * kPivWin * kSwapCols));
  if (e != hipSuccess) return hip_error(e, "hipFuncSetAttribute(block_row_swaps_kernel)");
  for (int K0 = 0; K0 < n; K0 += kOuter) {
    const int K1 = min(n, K0 + kOuter);
    for (int k = K0; k < K1; k += kB) {
      const int k2 = min(K1, k + kB), nb = k2 - k;
      // pivot window: the block and the 128 rows below it (all rows below
      // are current in these columns: the inner updates are right-looking)
      const int win = nb < kB ? nb : min(kPivWin, n - k);
      float* Li = linv + (int64_t)(k / kB) * kB * kB;
      float* Ui = uinv + (int64_t)(k / kB) * kB * kB;
      diag_lu_kernel<<<1, kDiagThreads, dsmem, st>>>(k, nb, win, a, lda, Li, Ui, piv, info);
      LXA_LAUNCH_CHECK("diag_lu_kernel");
      block_row_swaps_kernel<<<(n + kSwapCols - 1) / kSwapCols, kSwapCols, sizeof(float) * win * kSwapCols, st>>>(
          n, k, nb, win, a, lda, piv);
      LXA_LAUNCH_CHECK("block_row_swaps_kernel");
      float* A11 = a + (int64_t)k * lda + k;
      // L21 = A21 U11^-1 (in place, N = nb <= 128) for the rows below the
      // window (the window's rows hold theirs from diag_lu_kernel)
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
    // outer Schur update
    if (int rc = gemm_update(st, mode, n - K1, n - K1, K1 - K0, a + (int64_t)K1 * lda + K0, lda,
                             a + (int64_t)K0 * lda + K1, lda, a + (int64_t)K1 * lda + K1, lda, 1, -1.f))
      return rc;
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
  const size_t smem = n <= kPermLds ? sizeof(int32_t) * (size_t)n : 0;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&perm_from_swaps_kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)(sizeof(int32_t) * kPermLds));
  if (e != hipSuccess) return hip_error(e, "hipFuncSetAttribute(perm_from_swaps_kernel)");
  perm_from_swaps_kernel<<<1, 256, smem, static_cast<hipStream_t>(stream)>>>((int)n, piv, perm);
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

extern "C" int lxa_lu_large_trsv(void* stream, int kind, int64_t n, int nrhs, const float* lu, int64_t lda,
                                 const float* inv, float* x, int64_t ldx, int32_t* sync) {
  LXA_REQUIRE(kind >= 0 && kind <= 3, "kind must be 0..3 (got %d)", kind);
  LXA_REQUIRE(n >= 0 && n <= (1 << 20) && lda >= n, "invalid n / lda");
  LXA_REQUIRE(nrhs >= 0 && nrhs <= kMaxRhs, "nrhs must be in [0, %d] (got %d)", kMaxRhs, nrhs);
  if (n == 0 || nrhs == 0) return LXA_OK;
  LXA_REQUIRE(lu && inv && x && sync && ldx >= n, "invalid pointers or ldx");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nblk = (int)((n + kB - 1) / kB);
  // a fill kernel rather than a memset node, so the call can sit in a captured graph
  zero_i32_kernel<<<(nblk + 2 + 255) / 256, 256, 0, st>>>(sync, nblk + 2);
  LXA_LAUNCH_CHECK("zero_i32_kernel");
  const size_t smem = sizeof(float) * (kB * kTileLd + 2 * kMaxRhs * kB);
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
