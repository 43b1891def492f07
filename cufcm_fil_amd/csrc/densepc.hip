// densepc.hip -- resident inverse of the blob-blob mobility as the blob-system
// preconditioner (CFM_PC_DENSE).
//
// With the body held fixed the blob positions never change, so M_bb is one matrix
// for the whole run.  It is assembled once in FP64 (the same pair functions as the
// apply), factorised and inverted with rocSOLVER (dpotrf + dpotri: M_bb is
// symmetric positive definite), rounded to FP32 and kept resident in HBM
// (4 n^2 bytes, n = 3 n_blob: 1.3 GB at 6,000 blobs).  Applying it is one GEMV
// with FP64 accumulation, so the preconditioner is linear to FP64 rounding and the
// FP32 rounding of the entries only perturbs A P^-1 around the identity.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <chrono>

#include "ctx.hpp"
#include "densepc.hpp"
#include "rpy_pair.hpp"

namespace cfm {

namespace {

constexpr int kRows = 8;  // matrix rows per GEMV workgroup

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

// P (FP32, rows of length ld, zero padding) <- the full symmetric inverse whose lower
// triangle dpotri left in A.  Thread r of column c writes P[c * ld + r] (P is
// symmetric, so this is also row-major).
__global__ __launch_bounds__(256) void dense_round_kernel(const double* __restrict__ A, size_t lda,
                                                          int n, float* __restrict__ P, size_t ld) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t c = blockIdx.y;
  if (r >= ld) return;
  float v = 0.0f;
  if (r < (size_t)n) v = (float)(r >= c ? A[r + c * lda] : A[c + r * lda]);
  P[c * ld + r] = v;
}

// y = P x: kRows rows per workgroup, float4 loads of P, FP64 products and sums.
__global__ __launch_bounds__(256) void dense_gemv_kernel(int n, size_t ld,
                                                         const float* __restrict__ P,
                                                         const double* __restrict__ xv,
                                                         double* __restrict__ yv) {
  __shared__ double red[4][kRows];
  const int r0 = blockIdx.x * kRows;
  const int nr = min(kRows, n - r0);
  double acc[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) acc[r] = 0.0;
  const int n4 = (int)(ld / 4);
  const float4* P4 = reinterpret_cast<const float4*>(P + (size_t)r0 * ld);
  for (int c4 = threadIdx.x; c4 < n4; c4 += 256) {
    const int c = 4 * c4;
    const double x0 = c < n ? xv[c] : 0.0;
    const double x1 = c + 1 < n ? xv[c + 1] : 0.0;
    const double x2 = c + 2 < n ? xv[c + 2] : 0.0;
    const double x3 = c + 3 < n ? xv[c + 3] : 0.0;
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      if (r < nr) {
        const float4 p = P4[(size_t)r * n4 + c4];
        acc[r] += (double)p.x * x0 + (double)p.y * x1 + (double)p.z * x2 + (double)p.w * x3;
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    double v = acc[r];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wave][r] = v;
  }
  __syncthreads();
  if (threadIdx.x < nr) {
    const int r = threadIdx.x;
    yv[r0 + r] = (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
  }
}

}  // namespace

Status dense_build(cfm_ctx* c, DenseInverse& d) {
  const Particles& p = c->p;
  const int m = p.n_blob;
  if (m <= 0) return make_status(CFM_EINVAL, "CFM_PC_DENSE needs blobs");
  const int n = 3 * m;
  const size_t lda = (size_t)n;
  const size_t ld = ((size_t)n + 63) / 64 * 64;
  const size_t need = sizeof(double) * lda * n + sizeof(float) * ld * n;
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
  CFM_TRY(d.P.alloc(ld * n));
  hipLaunchKernelGGL(dense_round_kernel, dim3((unsigned)((ld + 255) / 256), n), dim3(256), 0,
                     c->stream, A.p, lda, n, d.P.p, ld);
  CFM_CHECK_LAUNCH();
  CFM_HIP(hipStreamSynchronize(c->stream));
  d.n = n;
  d.ld = ld;
  d.valid = true;
  ++d.generation;
  d.build_seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return {};
}

Status dense_apply(cfm_ctx* c, const DenseInverse& d, const double* x, double* y) {
  hipLaunchKernelGGL(dense_gemv_kernel, dim3((d.n + kRows - 1) / kRows), dim3(256), 0, c->stream,
                     d.n, d.ld, d.P.p, x, y);
  CFM_CHECK_LAUNCH();
  return {};
}

}  // namespace cfm
