// gnn.hip -- learned initial guess for the blob-force GMRES (placeholder until the
// network kernels land: reports that no guess was produced, so the previous
// solution is used, which is the reference's behaviour).
#include "ctx.hpp"

namespace cfm {
Status gnn_guess(cfm_ctx* c, const double* rhs, double* x0, bool* used) {
  (void)c;
  (void)rhs;
  (void)x0;
  *used = false;
  return {};
}
}  // namespace cfm

void cfm_gnn_release(cfm_ctx* c) { c->gnn = nullptr; }
