// Autotune table of the uneven-batch Cholesky launch (chol_mfma.hip): the
// largest order each tile class runs.  Orders up to kTunedMax4 run on 4 x 4
// blocks of one wave, up to kTunedMax8 on 8 x 8 blocks of four waves, larger
// ones on 16 x 16 blocks of four waves.  lineax_amd/tune_uneven.py measures
// every eligible class per order bucket on the GPU and rewrites this file;
// the values below are the capacity limits of the two small classes until a
// tuning run replaces them.
#pragma once

namespace lxa {
constexpr int kTunedMax4 = 64;
constexpr int kTunedMax8 = 128;
}  // namespace lxa
