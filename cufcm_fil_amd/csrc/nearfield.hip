// nearfield.hip -- near-field block preconditioner of the blob system.
//
// For every blob i the k nearest blobs N(i) (k = 16) are found by a brute-force
// GPU search; the 3(k+1) x 3(k+1) wall-RPY mobility of {i} + N(i) is assembled in
// LDS, inverted in place (Gauss-Jordan, partial pivoting), and the three rows of
// blob i are kept:  (P^-1 x)_i = sum_{j in {i} + N(i)} W_ij x_j.  This "local
// inverse" resolves the near-field (lubrication-free but overlapping) coupling of
// neighbouring blobs that a per-particle diagonal leaves to the Krylov iteration.
// An optional second set of blocks C_ij (the learned correction) is added in the
// same sparse product.  The operator is rebuilt whenever positions change.
//
// Blocks are stored [m][k+1][3][3] (slot 0 = the blob itself).
#include <algorithm>

#include "ctx.hpp"
#include "krylov.hpp"
#include "nearfield.hpp"
#include "rpy_pair.hpp"

namespace cfm {

constexpr int kNbr = NF_K;          // neighbours per blob
constexpr int kSlots = NF_K + 1;    // with the blob itself
constexpr int kLN = 3 * kSlots;     // local system size (51)

// k nearest (excluding itself) by brute force; ties broken by index.
__global__ __launch_bounds__(256) void knn_kernel(const double* __restrict__ x,
                                                  const double* __restrict__ y,
                                                  const double* __restrict__ z, int off, int m,
                                                  int* __restrict__ nbr) {
  __shared__ double sx[256], sy[256], sz[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < m;
  const int ic = valid ? i : 0;
  const double xi = x[off + ic], yi = y[off + ic], zi = z[off + ic];
  double bd[kNbr];
  int bi[kNbr];
  for (int q = 0; q < kNbr; ++q) {
    bd[q] = 1e300;
    bi[q] = -1;
  }
  for (int t0 = 0; t0 < m; t0 += 256) {
    __syncthreads();
    const int j = t0 + threadIdx.x;
    if (j < m) {
      sx[threadIdx.x] = x[off + j];
      sy[threadIdx.x] = y[off + j];
      sz[threadIdx.x] = z[off + j];
    }
    __syncthreads();
    const int cnt = min(256, m - t0);
    for (int jj = 0; jj < cnt; ++jj) {
      const int j2 = t0 + jj;
      if (j2 == ic) continue;
      const double dx = xi - sx[jj], dy = yi - sy[jj], dz = zi - sz[jj];
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < bd[kNbr - 1] || (d2 == bd[kNbr - 1] && j2 < bi[kNbr - 1])) {
        int q = kNbr - 1;
        while (q > 0 && (d2 < bd[q - 1] || (d2 == bd[q - 1] && j2 < bi[q - 1]))) {
          bd[q] = bd[q - 1];
          bi[q] = bi[q - 1];
          --q;
        }
        bd[q] = d2;
        bi[q] = j2;
      }
    }
  }
  if (valid)
    for (int q = 0; q < kNbr; ++q) nbr[(size_t)i * kNbr + q] = bi[q];
}

// In-place Gauss-Jordan inversion of an N x N row-major matrix in LDS by one
// workgroup (partial pivoting, column unscrambling).  piv/colk: N-entry scratch.
template <int N>
__device__ void gj_invert(double* J, double* colk, int* piv, int* status) {
  __shared__ int prow;
  __shared__ double pval;
  const int tid = threadIdx.x, nth = blockDim.x;
  for (int k = 0; k < N; ++k) {
    if (tid < 64) {
      double best = -1.0;
      int bi = k;
      for (int r = k + tid; r < N; r += 64) {
        const double a = fabs(J[r * N + k]);
        if (a > best) {
          best = a;
          bi = r;
        }
      }
      for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (tid == 0) {
        prow = bi;
        piv[k] = bi;
      }
    }
    __syncthreads();
    const int p = prow;
    if (p != k)
      for (int c = tid; c < N; c += nth) {
        const double t = J[k * N + c];
        J[k * N + c] = J[p * N + c];
        J[p * N + c] = t;
      }
    __syncthreads();
    if (tid == 0) pval = J[k * N + k];
    for (int r = tid; r < N; r += nth) colk[r] = J[r * N + k];
    __syncthreads();
    const double d = pval;
    if (tid == 0 && !(fabs(d) > 0.0)) status[0] = 1;
    for (int c = tid; c < N; c += nth) J[k * N + c] = (c == k) ? 1.0 / d : J[k * N + c] / d;
    __syncthreads();
    const double inv = J[k * N + k];
    for (int e = tid; e < N * N; e += nth) {
      const int r = e / N, c = e - r * N;
      if (r == k) continue;
      J[e] = (c == k) ? -colk[r] * inv : J[e] - colk[r] * J[k * N + c];
    }
    __syncthreads();
  }
  for (int k = N - 1; k >= 0; --k) {
    const int p = piv[k];
    if (p != k)
      for (int r = tid; r < N; r += nth) {
        const double t = J[r * N + k];
        J[r * N + k] = J[r * N + p];
        J[r * N + p] = t;
      }
    __syncthreads();
  }
}

// Local inverse of blob i's neighbourhood; keeps its three rows as 3x3 blocks.
__global__ __launch_bounds__(128) void local_inverse_kernel(const double* __restrict__ x,
                                                            const double* __restrict__ y,
                                                            const double* __restrict__ z,
                                                            int off, int m, int wall, double mu,
                                                            double a, const int* __restrict__ nbr,
                                                            double* __restrict__ W,
                                                            int* __restrict__ status) {
  __shared__ double J[kLN * kLN];
  __shared__ double colk[kLN];
  __shared__ int piv[kLN];
  __shared__ int ids[kSlots];
  const int i = blockIdx.x;
  if (threadIdx.x == 0) ids[0] = i;
  if (threadIdx.x >= 1 && threadIdx.x < kSlots) ids[threadIdx.x] = nbr[(size_t)i * kNbr + threadIdx.x - 1];
  __syncthreads();
  // column (slot b, component c) = mobility response of all slots to a unit load
  for (int col = threadIdx.x; col < kLN; col += blockDim.x) {
    const int b = col / 3, c = col - 3 * b;
    const int jb = off + ids[b];
    double F[3] = {0.0, 0.0, 0.0};
    F[c] = 1.0;
    for (int r = 0; r < kSlots; ++r) {
      const int ja = off + ids[r];
      double V[3] = {0.0, 0.0, 0.0};
      if (wall) {
        if (r == b) {
          wall_self<3, 3>(z[ja], F, V);
        } else {
          wall_pair<double, 3, 3>(x[ja] - x[jb], y[ja] - y[jb], z[ja] - z[jb], z[ja] + z[jb],
                                  z[jb], F, V);
        }
        V[0] /= mu;
        V[1] /= mu;
        V[2] /= mu;
      } else {
        free_pair<double, 3, 3>(x[ja] - x[jb], y[ja] - y[jb], z[ja] - z[jb], a, a, mu, F, V);
      }
      for (int q = 0; q < 3; ++q) J[(3 * r + q) * kLN + col] = V[q];
    }
  }
  __syncthreads();
  gj_invert<kLN>(J, colk, piv, status);
  // rows 0..2 (blob i): W[i][slot][p][q] = Jinv[p][3 slot + q]
  double* out = W + (size_t)i * kSlots * 9;
  for (int e = threadIdx.x; e < kSlots * 9; e += blockDim.x) {
    const int s = e / 9, pq = e - 9 * s, p = pq / 3, q = pq - 3 * p;
    out[e] = J[p * kLN + 3 * s + q];
  }
}

// y = W x (+ C x); vectors [3][m]
__global__ void nearfield_apply_kernel(int m, const int* __restrict__ nbr,
                                       const double* __restrict__ W, const double* __restrict__ C,
                                       const double* __restrict__ xv, double* __restrict__ yv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double o0 = 0.0, o1 = 0.0, o2 = 0.0;
  for (int s = 0; s < kSlots; ++s) {
    const int j = s == 0 ? i : nbr[(size_t)i * kNbr + s - 1];
    const double a = xv[j], b = xv[m + j], c = xv[2 * (size_t)m + j];
    const double* w = W + ((size_t)i * kSlots + s) * 9;
    double w9[9];
    for (int q = 0; q < 9; ++q) w9[q] = w[q] + (C ? C[((size_t)i * kSlots + s) * 9 + q] : 0.0);
    o0 += w9[0] * a + w9[1] * b + w9[2] * c;
    o1 += w9[3] * a + w9[4] * b + w9[5] * c;
    o2 += w9[6] * a + w9[7] * b + w9[8] * c;
  }
  yv[i] = o0;
  yv[m + i] = o1;
  yv[2 * (size_t)m + i] = o2;
}

Status nearfield_build(cfm_ctx* c, NearField& nf) {
  const Particles& p = c->p;
  const int m = p.n_blob, off = p.n_seg;
  if (m < kSlots) return make_status(CFM_EINVAL, "the near-field preconditioner needs >= %d blobs", kSlots);
  CFM_TRY(nf.nbr.alloc((size_t)m * kNbr));
  CFM_TRY(nf.W.alloc((size_t)m * kSlots * 9));
  CFM_TRY(nf.status.alloc(1));
  CFM_HIP(hipMemsetAsync(nf.status.p, 0, sizeof(int), c->stream));
  hipLaunchKernelGGL(knn_kernel, dim3((m + 255) / 256), dim3(256), 0, c->stream, p.x.p, p.y.p,
                     p.z.p, off, m, nf.nbr.p);
  CFM_CHECK_LAUNCH();
  hipLaunchKernelGGL(local_inverse_kernel, dim3(m), dim3(128), 0, c->stream, p.x.p, p.y.p, p.z.p,
                     off, m, c->cfg.wall, c->cfg.mu, c->cfg.a_blob, nf.nbr.p, nf.W.p, nf.status.p);
  CFM_CHECK_LAUNCH();
  int st = 0;
  CFM_HIP(hipMemcpyAsync(&st, nf.status.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  CFM_HIP(hipStreamSynchronize(c->stream));
  if (st) return make_status(CFM_ENONFINITE, "singular near-field block");
  nf.m = m;
  nf.valid = true;
  return {};
}

Status nearfield_apply(cfm_ctx* c, const NearField& nf, const double* C, const double* x, double* y) {
  hipLaunchKernelGGL(nearfield_apply_kernel, dim3((nf.m + 255) / 256), dim3(256), 0, c->stream,
                     nf.m, nf.nbr.p, nf.W.p, C, x, y);
  CFM_CHECK_LAUNCH();
  return {};
}

}  // namespace cfm
