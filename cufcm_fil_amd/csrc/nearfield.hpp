// nearfield.hpp -- near-field block preconditioner of the blob system.
#pragma once
#include "ctx.hpp"

namespace cfm {
Status nearfield_build(cfm_ctx* c, NearField& nf);
// y = (W + C) x over the blob system vectors [3][m]; C may be null.
Status nearfield_apply(cfm_ctx* c, const NearField& nf, const double* C, const double* x, double* y);
}  // namespace cfm
