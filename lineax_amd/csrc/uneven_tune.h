// Autotune table of the uneven-batch Cholesky launch (chol_mfma.hip): the
// largest order each tile class runs.  Orders up to kTunedMax4 run on 4 x 4
// blocks of one wave, up to kTunedMax8 on 8 x 8 blocks of four waves, larger
// ones on 16 x 16 blocks of four waves.  Written by lineax_amd/tune_uneven.py
// on an MI355X (4096 systems per bucket); measured batch times per class and order bucket:
//   n =   8: 16x16 0.217 ms, 8x8 0.100 ms, 4x4 0.109 ms
//   n =  16: 16x16 0.218 ms, 8x8 0.100 ms, 4x4 0.111 ms
//   n =  24: 16x16 0.320 ms, 8x8 0.133 ms, 4x4 0.132 ms
//   n =  32: 16x16 0.320 ms, 8x8 0.136 ms, 4x4 0.129 ms
//   n =  40: 16x16 0.429 ms, 8x8 0.171 ms, 4x4 0.160 ms
//   n =  48: 16x16 0.432 ms, 8x8 0.174 ms, 4x4 0.162 ms
//   n =  56: 16x16 0.548 ms, 8x8 0.217 ms, 4x4 0.198 ms
//   n =  64: 16x16 0.548 ms, 8x8 0.216 ms, 4x4 0.194 ms
//   n =  72: 16x16 0.673 ms, 8x8 0.266 ms
//   n =  80: 16x16 0.678 ms, 8x8 0.269 ms
//   n =  88: 16x16 0.800 ms, 8x8 0.323 ms
//   n =  96: 16x16 0.801 ms, 8x8 0.322 ms
//   n = 104: 16x16 0.951 ms, 8x8 0.393 ms
//   n = 112: 16x16 0.955 ms, 8x8 0.390 ms
//   n = 120: 16x16 1.112 ms, 8x8 0.462 ms
//   n = 128: 16x16 1.102 ms, 8x8 0.461 ms
#pragma once

namespace lxa {
constexpr int kTunedMax4 = 64;
constexpr int kTunedMax8 = 128;
}  // namespace lxa
