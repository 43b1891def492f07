// coarse.hpp -- coarse correction of the near-field preconditioner (coarse.hip).
#pragma once
#include <vector>

#include "ctx.hpp"

namespace cfm {
constexpr int kCoarseLinear = 12;  // 3 directions x {1, y_x, y_y, y_z}
// Builds the coarse space of the current mode (analytic fields, learned vectors).
Status coarse_build(cfm_ctx* c, CoarseSpace& cs);
// y += Z c with c from r and y = W r (see CoarseSpace); builds the space if needed.
Status coarse_apply(cfm_ctx* c, CoarseSpace& cs, const double* r, double* y);
// The 12 linear fields of the blob positions into Z ([12][3][m]).
Status coarse_linear_fields(cfm_ctx* c, double* Z);
// (G + delta I)^-1 of a k x k row-major matrix on the host; false if singular.
bool invert_gram(std::vector<double> A, int k, std::vector<double>& inv, double delta);
// Learned coarse vectors of the network on the current blob positions: *zl (device,
// [kl][3][m]) stays valid until the network changes.  Defined in gnn.hip.
Status gnn_coarse_vectors(cfm_ctx* c, const double** zl, int* kl);
}  // namespace cfm
