// densepc.hpp -- resident FP32 inverse of the blob-blob mobility (CFM_PC_DENSE).
#pragma once
#include "ctx.hpp"

namespace cfm {
// Assemble M_bb (FP64), invert it by Cholesky (rocSOLVER) and keep it as FP32.
Status dense_build(cfm_ctx* c, DenseInverse& d);
// y = P x over blob-system vectors [3][n_blob].
Status dense_apply(cfm_ctx* c, const DenseInverse& d, const double* x, double* y);
}  // namespace cfm
