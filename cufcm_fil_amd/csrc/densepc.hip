// densepc.hip -- resident inverse of the blob-blob mobility as the blob-system
// preconditioner (CFM_PC_DENSE).
//
// With the body held fixed the blob positions never change, so M_bb is one matrix
// for the whole run.  It is assembled once in FP64 (the same pair functions as the
// apply), factorised and inverted with rocSOLVER (dpotrf + dpotri: M_bb is
// symmetric positive definite), rounded to FP32 and kept resident in HBM as the upper
// triangle of 128 x 128 tiles (2 n^2 bytes, n = 3 n_blob: 0.65 GB at 6,000 blobs).
// Applying it reads every stored tile once: a tile (I, J) gives P_IJ x_J to row block I
// and, off the diagonal, P_IJ^T x_I to row block J; the partial products are summed
// in a fixed order, in FP64, so the preconditioner is linear to FP64 rounding and the
// FP32 rounding of the entries only perturbs A P^-1 around the identity.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <chrono>
#include <vector>

#include "ctx.hpp"
#include "densepc.hpp"
#include "rpy_pair.hpp"

namespace cfm {

namespace {

constexpr int kT = 128;  // tile side
constexpr int kG = 16;   // row groups of 32 threads in a tile's workgroup

// A (column-major, leading dimension lda) <- M_bb in the [component][blob] order.
// Block row i = blob index over threads, block column j = blockIdx.y.
__global__ __launch_bounds__(256) void dense_assemble_kernel(const double* __restrict__ x,
                                                             const double* __restrict__ y,
                                                             const double* __restrict__ z,
                                                             int off, int m, int wall, double mu,
                                                             double a, double* __restrict__ A,
                                                             size_t lda) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= m) return;
  const int gi = off + i, gj = off + j;
  for (int q = 0; q < 3; ++q) {
    double F[3] = {0.0, 0.0, 0.0};
    double V[3] = {0.0, 0.0, 0.0};
    F[q] = 1.0;
    if (wall) {
      if (i == j)
        wall_self<3, 3>(z[gi], F, V);
      else
        wall_pair<double, 3, 3>(x[gi] - x[gj], y[gi] - y[gj], z[gi] - z[gj], z[gi] + z[gj], z[gj],
                                F, V);
      V[0] /= mu;
      V[1] /= mu;
      V[2] /= mu;
    } else {
      free_pair<double, 3, 3>(x[gi] - x[gj], y[gi] - y[gj], z[gi] - z[gj], a, a, mu, F, V);
    }
    const size_t col = (size_t)q * m + j;
    for (int p = 0; p < 3; ++p) A[(size_t)p * m + i + col * lda] = V[p];
  }
}

// Packed tiles <- the symmetric inverse whose lower triangle dpotri left in A; entries
// beyond n are zero.  Block (J, I) of the grid writes row r = threadIdx.x of tile (I, J)
// when I <= J.
__global__ __launch_bounds__(kT) void dense_pack_kernel(const double* __restrict__ A, size_t lda, int n,
                                                        int nt, float* __restrict__ P) {
  const int J = blockIdx.x, I = blockIdx.y;
  if (I > J) return;
  const size_t t = (size_t)I * nt - (size_t)I * (I - 1) / 2 + (J - I);  // packed index
  float* T = P + t * kT * kT;
  for (int r = 0; r < kT; ++r) {
    const int c = threadIdx.x;
    const int gi = I * kT + r, gj = J * kT + c;
    double v = 0.0;
    if (gi < n && gj < n) v = gi >= gj ? A[gi + (size_t)gj * lda] : A[gj + (size_t)gi * lda];
    T[(size_t)r * kT + c] = (float)v;
  }
}

// One workgroup per stored tile (I, J): part[I][J] = P_IJ x_J and, for I < J,
// part[J][I] = P_IJ^T x_I.  Thread (g, q) reads rows g, g + kG, ... of column quad q
// (coalesced float4 loads), keeps its column sums in registers; row sums go through LDS.
__global__ __launch_bounds__(32 * kG) void dense_symv_kernel(int n, int nt, const int* __restrict__ tiles,
                                                         const float* __restrict__ P,
                                                         const double* __restrict__ xv,
                                                         double* __restrict__ part) {
  __shared__ double rowp[kT][33];
  __shared__ double colp[kG][kT];
  const int t = blockIdx.x;
  const int I = tiles[2 * t], J = tiles[2 * t + 1];
  const float4* T = reinterpret_cast<const float4*>(P + (size_t)t * kT * kT);
  const int q = threadIdx.x & 31, g = threadIdx.x >> 5;
  double xj[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int col = J * kT + 4 * q + k;
    xj[k] = col < n ? xv[col] : 0.0;
  }
  double ca[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int k = 0; k < kT / kG; ++k) {
    const int r = g + kG * k;
    const float4 p = T[r * (kT / 4) + q];
    const int gr = I * kT + r;
    const double xi = gr < n ? xv[gr] : 0.0;
    rowp[r][q] = (double)p.x * xj[0] + (double)p.y * xj[1] + (double)p.z * xj[2] + (double)p.w * xj[3];
    ca[0] += (double)p.x * xi;
    ca[1] += (double)p.y * xi;
    ca[2] += (double)p.z * xi;
    ca[3] += (double)p.w * xi;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) colp[g][4 * q + k] = ca[k];
  __syncthreads();
  const int tid = threadIdx.x;
  if (tid < kT) {
    double s = 0.0;
    for (int k = 0; k < 32; ++k) s += rowp[tid][k];
    part[((size_t)I * nt + J) * kT + tid] = s;
  } else if (I != J && tid < 2 * kT) {
    const int c = tid - kT;
    double s = 0.0;
    for (int k = 0; k < kG; ++k) s += colp[k][c];
    part[((size_t)J * nt + I) * kT + c] = s;
  }
}

// y_r = sum over tile columns J of part[r / kT][J][r % kT]: a workgroup takes 64 rows
// of one tile row, wave w the columns J = w mod 4, the four sums added in order.
__global__ __launch_bounds__(256) void dense_sum_kernel(int n, int nt, const double* __restrict__ part,
                                                        double* __restrict__ yv) {
  __shared__ double red[4][64];
  const int K = blockIdx.x, half = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int rr = half * 64 + lane;
  const double* p = part + (size_t)K * nt * kT + rr;
  double s0 = 0.0, s1 = 0.0;
  int J = w;
  for (; J + 4 < nt; J += 8) {
    s0 += p[(size_t)J * kT];
    s1 += p[(size_t)(J + 4) * kT];
  }
  if (J < nt) s0 += p[(size_t)J * kT];
  red[w][lane] = s0 + s1;
  __syncthreads();
  const int r = K * kT + rr;
  if (w == 0 && r < n) yv[r] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

}  // namespace

Status dense_build(cfm_ctx* c, DenseInverse& d) {
  const Particles& p = c->p;
  const int m = p.n_blob;
  if (m <= 0) return make_status(CFM_EINVAL, "CFM_PC_DENSE needs blobs");
  const int n = 3 * m;
  const size_t lda = (size_t)n;
  const int nt = (n + kT - 1) / kT;
  const size_t ntiles = (size_t)nt * (nt + 1) / 2;
  const size_t need = sizeof(double) * lda * n + sizeof(float) * ntiles * kT * kT;
  size_t free_b = 0, total_b = 0;
  CFM_HIP(hipMemGetInfo(&free_b, &total_b));
  if (need + need / 8 > free_b)
    return make_status(CFM_ENOMEM,
                       "CFM_PC_DENSE at %d blobs needs %.2f GB of device memory (%.2f GB free)",
                       m, need / 1e9, free_b / 1e9);
  const auto t0 = std::chrono::steady_clock::now();
  d.release();
  DevBuf<double> A;
  DevBuf<int> info;
  struct Release {  // the FP64 factor is scratch: freed on every return
    DevBuf<double>& a;
    DevBuf<int>& i;
    ~Release() {
      a.release();
      i.release();
    }
  } release_scratch{A, info};
  CFM_TRY(A.alloc(lda * n));
  CFM_TRY(info.alloc(1));
  hipLaunchKernelGGL(dense_assemble_kernel, dim3((m + 255) / 256, m), dim3(256), 0, c->stream, p.x.p,
                     p.y.p, p.z.p, p.n_seg, m, c->cfg.wall, c->cfg.mu, c->cfg.a_blob, A.p, lda);
  CFM_CHECK_LAUNCH();
  rocblas_handle hb = nullptr;
  CFM_TRY(ctx_rocblas(c, (void**)&hb));
  int hinfo = 0;
  if (rocsolver_dpotrf(hb, rocblas_fill_lower, n, A.p, (rocblas_int)lda, info.p) !=
      rocblas_status_success)
    return make_status(CFM_EHIP, "rocsolver_dpotrf failed");
  CFM_HIP(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  CFM_HIP(hipStreamSynchronize(c->stream));
  if (hinfo != 0)
    return make_status(CFM_ENONFINITE, "blob mobility not positive definite (dpotrf info %d)", hinfo);
  if (rocsolver_dpotri(hb, rocblas_fill_lower, n, A.p, (rocblas_int)lda, info.p) !=
      rocblas_status_success)
    return make_status(CFM_EHIP, "rocsolver_dpotri failed");
  CFM_HIP(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  CFM_HIP(hipStreamSynchronize(c->stream));
  if (hinfo != 0) return make_status(CFM_ENONFINITE, "singular blob mobility (dpotri info %d)", hinfo);
  CFM_TRY(d.P.alloc(ntiles * kT * kT));
  CFM_TRY(d.tiles.alloc(2 * ntiles));
  CFM_TRY(d.part.alloc((size_t)nt * nt * kT));
  std::vector<int> tl(2 * ntiles);
  for (int I = 0, t = 0; I < nt; ++I)
    for (int J = I; J < nt; ++J, ++t) {
      tl[2 * t] = I;
      tl[2 * t + 1] = J;
    }
  CFM_HIP(hipMemcpyAsync(d.tiles.p, tl.data(), sizeof(int) * tl.size(), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(dense_pack_kernel, dim3(nt, nt), dim3(kT), 0, c->stream, A.p, lda, n, nt, d.P.p);
  CFM_CHECK_LAUNCH();
  CFM_HIP(hipStreamSynchronize(c->stream));
  d.n = n;
  d.nt = nt;
  d.valid = true;
  ++d.generation;
  d.build_seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return {};
}

Status dense_apply(cfm_ctx* c, const DenseInverse& d, const double* x, double* y) {
  const unsigned ntiles = (unsigned)((size_t)d.nt * (d.nt + 1) / 2);
  hipLaunchKernelGGL(dense_symv_kernel, dim3(ntiles), dim3(32 * kG), 0, c->stream, d.n, d.nt, d.tiles.p,
                     d.P.p, x, d.part.p);
  hipLaunchKernelGGL(dense_sum_kernel, dim3(d.nt, kT / 64), dim3(256), 0, c->stream, d.n, d.nt, d.part.p, y);
  CFM_CHECK_LAUNCH();
  return {};
}

}  // namespace cfm
