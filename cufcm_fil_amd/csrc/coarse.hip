// coarse.hip -- two-level (coarse) correction of the near-field blob preconditioner.
//
// CFM_PC_NEAR's local inverse W of each blob's neighbourhood leaves the collective
// modes of the blob system untouched: the rigid and linear motions of a whole body
// (the largest eigenvalues of M W) and the "squeeze" of a closed surface (a near-null
// mode of M that W inflates).  A coarse space Z of k vectors of the blob system
// corrects them after W:
//     y = W r + Z c,  c = argmin_c ||s - M Z c||^2 + delta ||c||^2,  s = r - M W r,
// i.e. (G + delta I) c = (M Z)^T r - (M^2 Z)^T W r with G = (M Z)^T (M Z), using that
// M is symmetric.  Z, M Z and M^2 Z are kept (3 k vectors) and (G + delta I)^-1 is
// formed on the host once per blob geometry (and network state), so one application
// costs 2k dot products, a k x k product and k axpys -- no mobility product.
// Vectors are in the blob-system layout [3][m].  No counterpart in the reference.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ctx.hpp"
#include "coarse.hpp"
#include "krylov.hpp"

namespace cfm {

// Z[3 p + d][d][i] = phi_p(y_i), phi = (1, y_x, y_y, y_z); y = (x - centroid) / radius.
__global__ void coarse_linear_kernel(int m, int off, const double* __restrict__ x,
                                     const double* __restrict__ y, const double* __restrict__ z,
                                     double cx, double cy, double cz, double inv_r,
                                     double* __restrict__ Z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double phi[4] = {1.0, (x[off + i] - cx) * inv_r, (y[off + i] - cy) * inv_r,
                         (z[off + i] - cz) * inv_r};
  const size_t L = 3 * (size_t)m;
  for (int p = 0; p < 4; ++p)
    for (int d = 0; d < 3; ++d) {
      double* v = Z + (size_t)(3 * p + d) * L;
      for (int e = 0; e < 3; ++e) v[(size_t)e * m + i] = e == d ? phi[p] : 0.0;
    }
}

// coef[2k + a] = -sum_b Ginv[a][b] (coef[b] - coef[k + b])
__global__ void coarse_coef_kernel(int k, const double* __restrict__ Ginv, double* __restrict__ coef) {
  const int a = threadIdx.x;
  if (a >= k) return;
  double s = 0.0;
  for (int b = 0; b < k; ++b) s += Ginv[(size_t)a * k + b] * (coef[b] - coef[k + b]);
  coef[2 * k + a] = -s;
}

Status coarse_linear_fields(cfm_ctx* c, double* Z) {
  const int m = c->p.n_blob;
  const std::vector<double>& bp = c->blob_pos;
  if ((int)bp.size() != 3 * m || m == 0) return make_status(CFM_ESTATE, "blob positions not set");
  double cx = 0, cy = 0, cz = 0;
  for (int i = 0; i < m; ++i) {
    cx += bp[3 * i];
    cy += bp[3 * i + 1];
    cz += bp[3 * i + 2];
  }
  cx /= m;
  cy /= m;
  cz /= m;
  double r2 = 0.0;
  for (int i = 0; i < m; ++i) {
    const double dx = bp[3 * i] - cx, dy = bp[3 * i + 1] - cy, dz = bp[3 * i + 2] - cz;
    r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
  }
  const double inv_r = r2 > 0.0 ? 1.0 / std::sqrt(r2) : 1.0;
  hipLaunchKernelGGL(coarse_linear_kernel, dim3((m + 255) / 256), dim3(256), 0, c->stream, m, c->p.n_seg,
                     c->p.x.p, c->p.y.p, c->p.z.p, cx, cy, cz, inv_r, Z);
  CFM_CHECK_LAUNCH();
  return {};
}

// (G + delta I)^-1 by Gauss-Jordan with partial pivoting (small k, on the host).
bool invert_gram(std::vector<double> A, int k, std::vector<double>& inv, double delta) {
  inv.assign((size_t)k * k, 0.0);
  for (int i = 0; i < k; ++i) {
    A[(size_t)i * k + i] += delta;
    inv[(size_t)i * k + i] = 1.0;
  }
  for (int col = 0; col < k; ++col) {
    int p = col;
    for (int r = col + 1; r < k; ++r)
      if (std::fabs(A[(size_t)r * k + col]) > std::fabs(A[(size_t)p * k + col])) p = r;
    const double piv = A[(size_t)p * k + col];
    if (!(std::fabs(piv) > 0.0) || !std::isfinite(piv)) return false;
    if (p != col)
      for (int j = 0; j < k; ++j) {
        std::swap(A[(size_t)p * k + j], A[(size_t)col * k + j]);
        std::swap(inv[(size_t)p * k + j], inv[(size_t)col * k + j]);
      }
    const double ip = 1.0 / piv;
    for (int j = 0; j < k; ++j) {
      A[(size_t)col * k + j] *= ip;
      inv[(size_t)col * k + j] *= ip;
    }
    for (int r = 0; r < k; ++r) {
      if (r == col) continue;
      const double f = A[(size_t)r * k + col];
      if (f == 0.0) continue;
      for (int j = 0; j < k; ++j) {
        A[(size_t)r * k + j] -= f * A[(size_t)col * k + j];
        inv[(size_t)r * k + j] -= f * inv[(size_t)col * k + j];
      }
    }
  }
  return true;
}

Status coarse_build(cfm_ctx* c, CoarseSpace& cs) {
  const SystemDesc s = system_desc(c, CFM_SYS_BLOB);
  const int L = s.L;
  const double* zl = nullptr;
  int kl = 0;
  if (cs.mode == CFM_COARSE_LEARNED) CFM_TRY(gnn_coarse_vectors(c, &zl, &kl));
  const int k = kCoarseLinear + kl;
  CFM_TRY(cs.Z.alloc((size_t)k * L));
  CFM_TRY(cs.MZ.alloc((size_t)k * L));
  CFM_TRY(cs.M2Z.alloc((size_t)k * L));
  CFM_TRY(cs.Ginv.alloc((size_t)k * k));
  CFM_TRY(cs.coef.alloc((size_t)3 * k));
  CFM_TRY(coarse_linear_fields(c, cs.Z.p));
  if (kl)
    CFM_HIP(hipMemcpyAsync(cs.Z.p + (size_t)kCoarseLinear * L, zl, sizeof(double) * kl * L,
                           hipMemcpyDeviceToDevice, c->stream));
  for (int j = 0; j < k; ++j) {
    CFM_TRY(system_apply(c, s, CFM_FP64, cs.Z.p + (size_t)j * L, cs.MZ.p + (size_t)j * L));
    CFM_TRY(system_apply(c, s, CFM_FP64, cs.MZ.p + (size_t)j * L, cs.M2Z.p + (size_t)j * L));
  }
  // Gram matrix of M Z, one row per dot-product launch
  DevBuf<double> gdev, part;
  CFM_TRY(gdev.alloc((size_t)k * k));
  CFM_TRY(part.alloc(dev_dots_partial_size(L, k)));
  for (int j = 0; j < k; ++j)
    CFM_TRY(dev_dots(c, part.p, cs.MZ.p, (size_t)L, k, cs.MZ.p + (size_t)j * L, L, gdev.p + (size_t)j * k));
  cs.G.assign((size_t)k * k, 0.0);
  CFM_HIP(copy_to_host(c->stream, cs.G.data(), gdev.p, sizeof(double) * k * k));
  gdev.release();
  part.release();
  double tr = 0.0;
  for (int i = 0; i < k; ++i) tr += cs.G[(size_t)i * k + i];
  std::vector<double> inv;
  if (!invert_gram(cs.G, k, inv, 1e-12 * tr / k))
    return make_status(CFM_ENONFINITE, "singular Gram matrix of the coarse space");
  CFM_HIP(copy_to_device(c->stream, cs.Ginv.p, inv.data(), sizeof(double) * k * k));
  cs.k = k;
  cs.L = L;
  cs.valid = true;
  return {};
}

Status coarse_apply(cfm_ctx* c, CoarseSpace& cs, const double* r, double* y) {
  if (!cs.valid) CFM_TRY(coarse_build(c, cs));
  const int k = cs.k, L = cs.L;
  // reductions use a partial buffer of their own (the GMRES workspace is in use)
  DevBuf<double>& part = cs.part;
  CFM_TRY(part.alloc(dev_dots_partial_size(L, k)));
  CFM_TRY(dev_dots(c, part.p, cs.MZ.p, (size_t)L, k, r, L, cs.coef.p));
  CFM_TRY(dev_dots(c, part.p, cs.M2Z.p, (size_t)L, k, y, L, cs.coef.p + k));
  hipLaunchKernelGGL(coarse_coef_kernel, dim3(1), dim3(64), 0, c->stream, k, cs.Ginv.p, cs.coef.p);
  CFM_CHECK_LAUNCH();
  return dev_gemv_sub(c, cs.Z.p, (size_t)L, k, cs.coef.p + 2 * k, y, L);
}

}  // namespace cfm

using namespace cfm;

extern "C" int cfm_nearfield_coarse(cfm_ctx* c, int32_t mode) {
  if (!c) return CFM_EINVAL;
  if (mode < CFM_COARSE_NONE || mode > CFM_COARSE_LEARNED) {
    c->err = "unknown coarse mode";
    return CFM_EINVAL;
  }
  if (mode == CFM_COARSE_LEARNED && !c->gnn) {
    c->err = "the learned coarse space needs cfm_gnn_init";
    return CFM_ESTATE;
  }
  if (mode != CFM_COARSE_NONE && c->p.n_blob == 0) {
    c->err = "the coarse correction needs blobs";
    return CFM_EINVAL;
  }
  if (mode != c->coarse.mode) c->coarse.valid = false;
  c->coarse.mode = mode;
  return CFM_OK;
}
