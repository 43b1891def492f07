// gnn.hip -- learned near-field correction of the blob preconditioner.
//
// A message-passing network on the k-nearest-neighbour graph of the blobs
// (k = 16, the graph of the near-field preconditioner) maps geometry to one 3x3
// block C_ij per edge (slot 0 = the blob itself).  The blob preconditioner then
// applies P^-1 = W + C, where W is the analytic near-field local inverse.
//
//   node features n_i (8):   1/z_i, mean |r_ij|/2, mean r_ij/2 (3), 1, 0, 0
//   edge features e_ij (8):  r_ij/|r_ij| (3), 1/|r_ij|, 1/|r_ij|^2, 2/(z_i+z_j), |r_ij|/4, 1
//                            (self slot: 0, 0, 0, 0, 0, 1/z_i, 0, 1)
//   h^0 = relu(n Wn + bn)                                        (8 -> 64)
//   3 layers:  m_e = relu([h_i, h_j, e_ij] W1 + b1)              (136 -> 64)
//              a_i = mean over the 16 neighbour edges of m_e
//              h_i <- h_i + relu([h_i, a_i] W2 + b2)             (128 -> 64)
//   C_e = [h_i, h_j, e_ij] Wo + bo                               (136 -> 9, zero-initialised)
//
// Every dense layer is a GEMM on the matrix cores: operands are rounded to BF16
// (the weights are kept in FP32 and rounded when loaded), products accumulate in
// FP32 with v_mfma_f32_16x16x16_bf16.  Training minimises the scale-free one-step
// residual of the recorded right-hand sides (GMRES is invariant to a global scale
// of P^-1, so only directions matter):
//   u_t = A (W + C) b_t,  alpha_t = (b_t . u_t)/(u_t . u_t),  r_t = b_t - alpha_t u_t,
//   L = mean_t ||r_t||^2 / ||b_t||^2 ,
// whose gradient with respect to C_e is -2 alpha_t (A r_t)_i b_t,j^T / ||b_t||^2
// (A is symmetric); it is back-propagated through the network and applied with Adam.
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "coarse.hpp"
#include "ctx.hpp"
#include "krylov.hpp"
#include "nearfield.hpp"

namespace cfm {

constexpr int G_K = NF_K;          // neighbours
constexpr int G_S = NF_K + 1;      // edge slots per node (self first)
constexpr int G_H = 64;            // hidden width
constexpr int G_F = 8;             // node / edge feature width
constexpr int G_XE = 2 * G_H + G_F;  // 136
constexpr int G_XN = 2 * G_H;      // 128
constexpr int G_O = 16;            // output width (9 used)
constexpr int G_L = 3;             // message-passing layers
constexpr int G_QF = 16;           // positional monomials of the coarse head (10 used)
constexpr int G_KC = 16;           // learned coarse vectors
constexpr int G_XC = G_H + G_QF;   // 80: coarse-head input [h_i, q_i]
constexpr int G_OC = 3 * G_KC;     // 48: coarse-head output (3 components per vector)
constexpr int G_TMAX = 128;        // samples per coarse training batch

// parameter offsets in one FP32 buffer
struct GnnLayout {
  size_t Wn, bn, W1[G_L], b1[G_L], W2[G_L], b2[G_L], Wo, bo, Wc, bc, total;
  GnnLayout() {
    size_t o = 0;
    Wn = o; o += G_F * G_H;
    bn = o; o += G_H;
    for (int l = 0; l < G_L; ++l) {
      W1[l] = o; o += G_XE * G_H;
      b1[l] = o; o += G_H;
      W2[l] = o; o += G_XN * G_H;
      b2[l] = o; o += G_H;
    }
    Wo = o; o += G_XE * G_O;
    bo = o; o += G_O;
    Wc = o; o += G_XC * G_OC;
    bc = o; o += G_OC;
    total = o;
  }
};

struct Gnn {
  int m = 0;
  size_t E = 0;
  GnnLayout lay;
  DevBuf<float> theta, grad, adam_m, adam_v;
  DevBuf<float> nfeat, efeat, Z0, H[G_L + 1];
  DevBuf<float> Xe[G_L], Ze[G_L], Me[G_L], Xn[G_L], Zn[G_L], Xo, Co;
  DevBuf<float> dH, dXe, dZe, dXn, dZn, dCo, dXo, dA;
  DevBuf<double> corr;                   // C in FP64, [m][G_S][9] (the preconditioner layout)
  DevBuf<double> Gacc;                   // dL/dC accumulated, [m][G_S][9]
  DevBuf<double> y, Ay, r, g, scal, partial;
  DevBuf<double> samples_b, samples_x;   // recorded training pairs [n][3m]
  int nsamples = 0, capacity = 0, next = 0;
  // coarse head: positional monomials, input [h, q], output, learned vectors ([KC][3][m])
  DevBuf<float> qfeat, Xc, Oc, dOc, dXc;
  DevBuf<double> zl;
  // coarse training: s_t = b_t - M W b_t per sample, M Z ([12 + KC][3m]; rows 0..11 the
  // linear fields, fixed), the descent direction D and dL/dZL ([KC][3m]), host copies
  DevBuf<double> S, snorm, MZt, Zp, D, dZ, gram, bmat, cmat, wmat, cpart;
  std::vector<double> s_norm2;
  int s_ready = 0;          // samples whose s_t is computed
  double delta = 0.0;       // Tikhonov term of the coarse objective (geometry only)
  bool enabled = false, trained = false, recording = false;
  int adam_t = 0;
  double last_loss = 0.0;
};

static Gnn* gnn_of(cfm_ctx* c) { return reinterpret_cast<Gnn*>(c->gnn); }
Status upload_aos(cfm_ctx* c, const double* host, int cnt, int dof, double* dst, int ld, int off);

// ---------------------------------------------------------------- BF16 MFMA GEMM
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (short)(u >> 16);
}

// C[M x N] = op(A)[M x K] op(B)[K x N] (+ bias[N]); op(A)(m,k) = A[m*lda+k] or
// A[k*lda+m] (transA); op(B)(k,n) = B[k*ldb+n] or B[n*ldb+k] (transB).  If Cact is
// non-null it also receives relu(C).  One wave computes one 16x16 tile; out-of-range
// rows/columns/k are zero-padded on load and skipped on store.
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, const float* __restrict__ A,
                                                        int lda, int transA, const float* __restrict__ B,
                                                        int ldb, int transB, const float* __restrict__ bias,
                                                        float* __restrict__ C, int ldc,
                                                        float* __restrict__ Cact, int accumulate) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int tm = (blockIdx.x * 4 + wave) * 16;
  const int tn = blockIdx.y * 16;
  if (tm >= M) return;
  const int r16 = lane & 15;
  const int kq = (lane >> 4) * 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const int am = tm + r16;   // row of A this lane loads
  const int bn = tn + r16;   // column of B this lane loads
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
    for (int e = 0; e < 4; ++e) {
      const int k = k0 + kq + e;
      float av = 0.f, bv = 0.f;
      if (am < M && k < K) av = transA ? A[(size_t)k * lda + am] : A[(size_t)am * lda + k];
      if (bn < N && k < K) bv = transB ? B[(size_t)bn * ldb + k] : B[(size_t)k * ldb + bn];
      a[e] = to_bf16(av);
      b[e] = to_bf16(bv);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  // D layout: lane -> column (lane & 15), rows 4*(lane >> 4) + e
  const int col = tn + r16;
  for (int e = 0; e < 4; ++e) {
    const int row = tm + kq + e;
    if (row < M && col < N) {
      float v = acc[e] + (bias ? bias[col] : 0.f);
      float* dst = C + (size_t)row * ldc + col;
      if (accumulate) v += *dst;
      *dst = v;
      if (Cact) Cact[(size_t)row * ldc + col] = v > 0.f ? v : 0.f;
    }
  }
}

static Status gemm(cfm_ctx* c, int M, int N, int K, const float* A, int lda, bool tA, const float* B,
                   int ldb, bool tB, const float* bias, float* C, int ldc, float* Cact = nullptr,
                   bool accumulate = false) {
  dim3 grid((M + 63) / 64, (N + 15) / 16);
  hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, c->stream, M, N, K, A, lda, (int)tA, B, ldb,
                     (int)tB, bias, C, ldc, Cact, (int)accumulate);
  CFM_CHECK_LAUNCH();
  return {};
}

// ---------------------------------------------------------------- features / graph glue
__global__ void gnn_features_kernel(int m, int off, const double* __restrict__ x,
                                    const double* __restrict__ y, const double* __restrict__ z,
                                    const int* __restrict__ nbr, float* __restrict__ nf,
                                    float* __restrict__ ef) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double xi = x[off + i], yi = y[off + i], zi = z[off + i];
  double mx = 0, my = 0, mz = 0, md = 0;
  for (int s = 0; s < G_S; ++s) {
    float* e = ef + ((size_t)i * G_S + s) * G_F;
    if (s == 0) {
      e[0] = e[1] = e[2] = e[3] = e[4] = 0.f;
      e[5] = (float)(1.0 / zi);
      e[6] = 0.f;
      e[7] = 1.f;
      continue;
    }
    const int j = nbr[(size_t)i * G_K + s - 1];
    const double dx = xi - x[off + j], dy = yi - y[off + j], dz = zi - z[off + j];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    mx += dx; my += dy; mz += dz; md += d;
    e[0] = (float)(dx / d);
    e[1] = (float)(dy / d);
    e[2] = (float)(dz / d);
    e[3] = (float)(1.0 / d);
    e[4] = (float)(1.0 / (d * d));
    e[5] = (float)(2.0 / (zi + z[off + j]));
    e[6] = (float)(0.25 * d);
    e[7] = 1.f;
  }
  float* n = nf + (size_t)i * G_F;
  n[0] = (float)(1.0 / zi);
  n[1] = (float)(md / G_K / 2.0);
  n[2] = (float)(mx / G_K / 2.0);
  n[3] = (float)(my / G_K / 2.0);
  n[4] = (float)(mz / G_K / 2.0);
  n[5] = 1.f;
  n[6] = 0.f;
  n[7] = 0.f;
}

// Xe[e] = [h_i, h_j, e_ij] (j = i for the self slot)
__global__ void gnn_edge_input_kernel(int m, const int* __restrict__ nbr, const float* __restrict__ h,
                                      const float* __restrict__ ef, float* __restrict__ Xe) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t E = (size_t)m * G_S;
  if (t >= E * G_XE) return;
  const size_t e = t / G_XE;
  const int c = (int)(t - e * G_XE);
  const int i = (int)(e / G_S), s = (int)(e - (size_t)i * G_S);
  float v;
  if (c < G_H) {
    v = h[(size_t)i * G_H + c];
  } else if (c < 2 * G_H) {
    const int j = s == 0 ? i : nbr[(size_t)i * G_K + s - 1];
    v = h[(size_t)j * G_H + c - G_H];
  } else {
    v = ef[e * G_F + c - 2 * G_H];
  }
  Xe[t] = v;
}

// Xn[i] = [h_i, mean_{s>=1} M[i*S+s]]
__global__ void gnn_node_input_kernel(int m, const float* __restrict__ h, const float* __restrict__ Me,
                                      float* __restrict__ Xn) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)m * G_XN) return;
  const int i = (int)(t / G_XN), c = (int)(t - (size_t)i * G_XN);
  if (c < G_H) {
    Xn[t] = h[(size_t)i * G_H + c];
  } else {
    float s = 0.f;
    for (int q = 1; q < G_S; ++q) s += Me[((size_t)i * G_S + q) * G_H + c - G_H];
    Xn[t] = s * (1.0f / G_K);
  }
}

// h_out = h_in + relu(Zn)
__global__ void gnn_residual_kernel(int n, const float* __restrict__ hin, const float* __restrict__ Zn,
                                    float* __restrict__ hout) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) hout[t] = hin[t] + fmaxf(Zn[t], 0.f);
}

// corr (FP64, [m][S][9]) from Co ([E][16])
__global__ void gnn_output_kernel(size_t E, const float* __restrict__ Co, double* __restrict__ corr) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * 9) return;
  const size_t e = t / 9;
  corr[t] = (double)Co[e * G_O + (t - e * 9)];
}

// ---- coarse head ----------------------------------------------------------------
// q_i: 1, y, y^2 (per component), y_x y_y, y_x y_z, y_y y_z, 0 x 6; y = (x_i - c) / R
__global__ void gnn_positional_kernel(int m, int off, const double* __restrict__ x,
                                      const double* __restrict__ y, const double* __restrict__ z,
                                      double cx, double cy, double cz, double inv_r,
                                      float* __restrict__ q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double a = (x[off + i] - cx) * inv_r, b = (y[off + i] - cy) * inv_r, c = (z[off + i] - cz) * inv_r;
  const double v[G_QF] = {1.0, a, b, c, a * a, b * b, c * c, a * b, a * c, b * c, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < G_QF; ++k) q[(size_t)i * G_QF + k] = (float)v[k];
}

// Xc[i] = [h_i, q_i]
__global__ void gnn_coarse_input_kernel(int m, const float* __restrict__ h, const float* __restrict__ q,
                                        float* __restrict__ Xc) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)m * G_XC) return;
  const int i = (int)(t / G_XC), c = (int)(t - (size_t)i * G_XC);
  Xc[t] = c < G_H ? h[(size_t)i * G_H + c] : q[(size_t)i * G_QF + c - G_H];
}

// zl[k][d][i] = Oc[i][3k + d] (FP64, blob-system layout)
__global__ void gnn_coarse_out_kernel(int m, const float* __restrict__ Oc, double* __restrict__ zl) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)m * G_OC) return;
  const int i = (int)(t / G_OC), c = (int)(t - (size_t)i * G_OC);
  const int k = c / 3, d = c - 3 * k;
  zl[((size_t)k * 3 + d) * m + i] = (double)Oc[t];
}

// dOc[i][3k + d] = scale dZ[k][d][i]
__global__ void gnn_coarse_dout_kernel(int m, const double* __restrict__ dZ, double scale,
                                       float* __restrict__ dOc) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)m * G_OC) return;
  const int i = (int)(t / G_OC), c = (int)(t - (size_t)i * G_OC);
  const int k = c / 3, d = c - 3 * k;
  dOc[t] = (float)(scale * dZ[((size_t)k * 3 + d) * m + i]);
}

// D[kk][e] = sum_t w[t][kk] (S[t][e] - sum_j MZ[j][e] C[t][j]) over the T samples of a
// batch: C[t] the coarse coefficients of sample t (k values), w[t][kk] = C[t][12 + kk] /
// ||s_t||^2 for the learned vectors.
__global__ void gnn_coarse_dir_kernel(int L, int T, int k, const double* __restrict__ S,
                                      const double* __restrict__ MZ, const double* __restrict__ C,
                                      const double* __restrict__ w, double* __restrict__ D) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L) return;
  double acc[G_KC];
  for (int q = 0; q < G_KC; ++q) acc[q] = 0.0;
  for (int t = 0; t < T; ++t) {
    double r = S[(size_t)t * L + e];
    for (int j = 0; j < k; ++j) r -= MZ[(size_t)j * L + e] * C[(size_t)t * k + j];
    for (int q = 0; q < G_KC; ++q) acc[q] += w[(size_t)t * G_KC + q] * r;
  }
  for (int q = 0; q < G_KC; ++q) D[(size_t)q * L + e] = acc[q];
}

// dH[i][c] = dXc[i][c], c < G_H
__global__ void gnn_copy_dh_kernel(int m, const float* __restrict__ dXc, float* __restrict__ dH) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)m * G_H) return;
  const int i = (int)(t / G_H), c = (int)(t - (size_t)i * G_H);
  dH[t] = dXc[(size_t)i * G_XC + c];
}

// ---- training pieces -------------------------------------------------------------
// G[e] += -2 alpha / ||b||^2 * g_i b_j^T   (b, g: [3][m]; e = (i, slot));
// scal = [||b||^2, ., alpha]
__global__ void gnn_grad_corr_kernel(int m, const int* __restrict__ nbr, const double* __restrict__ g,
                                     const double* __restrict__ b, const double* __restrict__ scal,
                                     double* __restrict__ G) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)m * G_S) return;
  const int i = (int)(e / G_S), s = (int)(e - (size_t)i * G_S);
  const int j = s == 0 ? i : nbr[(size_t)i * G_K + s - 1];
  const double w = -2.0 * scal[2] / scal[0];
  const double gi[3] = {g[i], g[m + i], g[2 * (size_t)m + i]};
  const double bj[3] = {b[j], b[m + j], b[2 * (size_t)m + j]};
  double* out = G + e * 9;
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) out[3 * p + q] += w * gi[p] * bj[q];
}

__global__ void gnn_dco_kernel(size_t E, const double* __restrict__ G, double scale, float* __restrict__ dCo) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * G_O) return;
  const size_t e = t / G_O;
  const int c = (int)(t - e * G_O);
  dCo[t] = c < 9 ? (float)(scale * G[e * 9 + c]) : 0.f;
}

// scatter dXe ([E][136]) into dH (node i gets cols 0..63, neighbour j cols 64..127)
__global__ void gnn_scatter_edge_kernel(int m, const int* __restrict__ nbr, const float* __restrict__ dXe,
                                        float* __restrict__ dH) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t E = (size_t)m * G_S;
  if (t >= E * 2 * G_H) return;
  const size_t e = t / (2 * G_H);
  const int c = (int)(t - e * 2 * G_H);
  const int i = (int)(e / G_S), s = (int)(e - (size_t)i * G_S);
  const float v = dXe[e * G_XE + c];
  if (c < G_H) {
    atomicAdd(&dH[(size_t)i * G_H + c], v);
  } else {
    const int j = s == 0 ? i : nbr[(size_t)i * G_K + s - 1];
    atomicAdd(&dH[(size_t)j * G_H + c - G_H], v);
  }
}

// dMe[i*S+s] = dA[i] / K for s >= 1 (0 for the self slot), masked by Ze > 0 -> dZe
__global__ void gnn_mean_back_kernel(int m, const float* __restrict__ dXn, const float* __restrict__ Ze,
                                     float* __restrict__ dZe) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t E = (size_t)m * G_S;
  if (t >= E * G_H) return;
  const size_t e = t / G_H;
  const int c = (int)(t - e * G_H);
  const int i = (int)(e / G_S), s = (int)(e - (size_t)i * G_S);
  const float d = s == 0 ? 0.f : dXn[(size_t)i * G_XN + G_H + c] * (1.0f / G_K);
  dZe[t] = Ze[t] > 0.f ? d : 0.f;
}

__global__ void relu_mask_kernel(size_t n, const float* __restrict__ dY, const float* __restrict__ Z,
                                 float* __restrict__ dZ) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dZ[t] = Z[t] > 0.f ? dY[t] : 0.f;
}

__global__ void add_cols_kernel(int rows, int cols, const float* __restrict__ src, int lds, float* __restrict__ dst,
                                int ldd) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)rows * cols) return;
  const int r = (int)(t / cols), cc = (int)(t - (size_t)r * cols);
  dst[(size_t)r * ldd + cc] += src[(size_t)r * lds + cc];
}

// bias gradient: db[c] = sum_r dZ[r*ld + c]
__global__ void col_sum_kernel(int rows, int cols, const float* __restrict__ dZ, int ld, float* __restrict__ db) {
  const int c = blockIdx.x;
  __shared__ float red[256];
  float s = 0.f;
  for (int r = threadIdx.x; r < rows; r += 256) s += dZ[(size_t)r * ld + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] = red[0];
}

__global__ void adam_kernel(size_t n, float* __restrict__ th, const float* __restrict__ g, float* __restrict__ mm,
                            float* __restrict__ vv, float lr, float b1, float b2, float c1, float c2) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  mm[t] = b1 * mm[t] + (1.f - b1) * gt;
  vv[t] = b2 * vv[t] + (1.f - b2) * gt * gt;
  th[t] -= lr * (mm[t] / c1) / (sqrtf(vv[t] / c2) + 1e-8f);
}

__global__ void fill_f32_kernel(size_t n, float* p, float v) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

static inline dim3 gl(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

// ---------------------------------------------------------------- forward / backward
static Status gnn_alloc(cfm_ctx* c, Gnn& G) {
  const size_t m = G.m, E = G.E;
  CFM_TRY(G.nfeat.alloc(m * G_F));
  CFM_TRY(G.efeat.alloc(E * G_F));
  CFM_TRY(G.Z0.alloc(m * G_H));
  for (int l = 0; l <= G_L; ++l) CFM_TRY(G.H[l].alloc(m * G_H));
  for (int l = 0; l < G_L; ++l) {
    CFM_TRY(G.Xe[l].alloc(E * G_XE));
    CFM_TRY(G.Ze[l].alloc(E * G_H));
    CFM_TRY(G.Me[l].alloc(E * G_H));
    CFM_TRY(G.Xn[l].alloc(m * G_XN));
    CFM_TRY(G.Zn[l].alloc(m * G_H));
  }
  CFM_TRY(G.Xo.alloc(E * G_XE));
  CFM_TRY(G.Co.alloc(E * G_O));
  CFM_TRY(G.dH.alloc(m * G_H));
  CFM_TRY(G.dXe.alloc(E * G_XE));
  CFM_TRY(G.dZe.alloc(E * G_H));
  CFM_TRY(G.dXn.alloc(m * G_XN));
  CFM_TRY(G.dZn.alloc(m * G_H));
  CFM_TRY(G.dCo.alloc(E * G_O));
  CFM_TRY(G.dXo.alloc(E * G_XE));
  CFM_TRY(G.dA.alloc(m * G_H));
  CFM_TRY(G.corr.alloc(E * 9));
  CFM_TRY(G.Gacc.alloc(E * 9));
  CFM_TRY(G.qfeat.alloc(m * G_QF));
  CFM_TRY(G.Xc.alloc(m * G_XC));
  CFM_TRY(G.Oc.alloc(m * G_OC));
  CFM_TRY(G.dOc.alloc(m * G_OC));
  CFM_TRY(G.dXc.alloc(m * G_XC));
  CFM_TRY(G.zl.alloc((size_t)G_KC * 3 * m));
  const int L = 3 * G.m;
  CFM_TRY(G.y.alloc(L));
  CFM_TRY(G.Ay.alloc(L));
  CFM_TRY(G.r.alloc(L));
  CFM_TRY(G.g.alloc(L));
  CFM_TRY(G.scal.alloc(8));
  CFM_TRY(G.partial.alloc(dev_dots_partial_size(L, 2)));
  (void)c;
  return {};
}

static Status gnn_forward(cfm_ctx* c, Gnn& G) {
  const int m = G.m;
  const size_t E = G.E;
  const float* th = G.theta.p;
  const NearField& nf = c->nf;
  const Particles& p = c->p;
  hipLaunchKernelGGL(gnn_features_kernel, gl(m), dim3(256), 0, c->stream, m, p.n_seg, p.x.p, p.y.p, p.z.p,
                     nf.nbr.p, G.nfeat.p, G.efeat.p);
  CFM_CHECK_LAUNCH();
  CFM_TRY(gemm(c, m, G_H, G_F, G.nfeat.p, G_F, false, th + G.lay.Wn, G_H, false, th + G.lay.bn, G.Z0.p,
               G_H, G.H[0].p));
  for (int l = 0; l < G_L; ++l) {
    hipLaunchKernelGGL(gnn_edge_input_kernel, gl(E * G_XE), dim3(256), 0, c->stream, m, nf.nbr.p, G.H[l].p,
                       G.efeat.p, G.Xe[l].p);
    CFM_CHECK_LAUNCH();
    CFM_TRY(gemm(c, (int)E, G_H, G_XE, G.Xe[l].p, G_XE, false, th + G.lay.W1[l], G_H, false,
                 th + G.lay.b1[l], G.Ze[l].p, G_H, G.Me[l].p));
    hipLaunchKernelGGL(gnn_node_input_kernel, gl((size_t)m * G_XN), dim3(256), 0, c->stream, m, G.H[l].p,
                       G.Me[l].p, G.Xn[l].p);
    CFM_CHECK_LAUNCH();
    CFM_TRY(gemm(c, m, G_H, G_XN, G.Xn[l].p, G_XN, false, th + G.lay.W2[l], G_H, false, th + G.lay.b2[l],
                 G.Zn[l].p, G_H));
    hipLaunchKernelGGL(gnn_residual_kernel, gl((size_t)m * G_H), dim3(256), 0, c->stream, m * G_H, G.H[l].p,
                       G.Zn[l].p, G.H[l + 1].p);
    CFM_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(gnn_edge_input_kernel, gl(E * G_XE), dim3(256), 0, c->stream, m, nf.nbr.p, G.H[G_L].p,
                     G.efeat.p, G.Xo.p);
  CFM_CHECK_LAUNCH();
  CFM_TRY(gemm(c, (int)E, G_O, G_XE, G.Xo.p, G_XE, false, th + G.lay.Wo, G_O, false, th + G.lay.bo, G.Co.p,
               G_O));
  hipLaunchKernelGGL(gnn_output_kernel, gl(E * 9), dim3(256), 0, c->stream, E, G.Co.p, G.corr.p);
  CFM_CHECK_LAUNCH();
  // coarse head: positional monomials about the blob centroid, [h, q] Wc + bc
  const std::vector<double>& bp = c->blob_pos;
  if ((int)bp.size() != 3 * m) return make_status(CFM_ESTATE, "blob positions not set");
  double cx = 0, cy = 0, cz = 0, r2 = 0;
  for (int i = 0; i < m; ++i) {
    cx += bp[3 * i];
    cy += bp[3 * i + 1];
    cz += bp[3 * i + 2];
  }
  cx /= m;
  cy /= m;
  cz /= m;
  for (int i = 0; i < m; ++i) {
    const double dx = bp[3 * i] - cx, dy = bp[3 * i + 1] - cy, dz = bp[3 * i + 2] - cz;
    r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
  }
  hipLaunchKernelGGL(gnn_positional_kernel, gl(m), dim3(256), 0, c->stream, m, p.n_seg, p.x.p, p.y.p, p.z.p,
                     cx, cy, cz, r2 > 0 ? 1.0 / std::sqrt(r2) : 1.0, G.qfeat.p);
  hipLaunchKernelGGL(gnn_coarse_input_kernel, gl((size_t)m * G_XC), dim3(256), 0, c->stream, m, G.H[G_L].p,
                     G.qfeat.p, G.Xc.p);
  CFM_CHECK_LAUNCH();
  CFM_TRY(gemm(c, m, G_OC, G_XC, G.Xc.p, G_XC, false, th + G.lay.Wc, G_OC, false, th + G.lay.bc, G.Oc.p,
               G_OC));
  hipLaunchKernelGGL(gnn_coarse_out_kernel, gl((size_t)m * G_OC), dim3(256), 0, c->stream, m, G.Oc.p, G.zl.p);
  CFM_CHECK_LAUNCH();
  return {};
}

static Status gnn_trunk_backward(cfm_ctx* c, Gnn& G);

// Back-propagates dL/dC (G.Gacc, FP64 [E][9], scaled by `scale`) into G.grad.
static Status gnn_backward(cfm_ctx* c, Gnn& G, double scale) {
  const int m = G.m;
  const size_t E = G.E;
  const float* th = G.theta.p;
  float* gr = G.grad.p;
  const NearField& nf = c->nf;
  hipLaunchKernelGGL(gnn_dco_kernel, gl(E * G_O), dim3(256), 0, c->stream, E, G.Gacc.p, scale, G.dCo.p);
  CFM_CHECK_LAUNCH();
  // output layer
  CFM_TRY(gemm(c, G_XE, G_O, (int)E, G.Xo.p, G_XE, true, G.dCo.p, G_O, false, nullptr, gr + G.lay.Wo, G_O));
  hipLaunchKernelGGL(col_sum_kernel, dim3(G_O), dim3(256), 0, c->stream, (int)E, G_O, G.dCo.p, G_O, gr + G.lay.bo);
  CFM_TRY(gemm(c, (int)E, G_XE, G_O, G.dCo.p, G_O, false, th + G.lay.Wo, G_O, true, nullptr, G.dXo.p, G_XE));
  hipLaunchKernelGGL(fill_f32_kernel, gl((size_t)m * G_H), dim3(256), 0, c->stream, (size_t)m * G_H, G.dH.p, 0.f);
  hipLaunchKernelGGL(gnn_scatter_edge_kernel, gl(E * 2 * G_H), dim3(256), 0, c->stream, m, nf.nbr.p, G.dXo.p, G.dH.p);
  CFM_CHECK_LAUNCH();
  return gnn_trunk_backward(c, G);
}

// Back-propagates dL/dZL (FP64 [KC][3][m], scaled by `scale`) through the coarse head and
// the trunk into G.grad.
static Status gnn_backward_coarse(cfm_ctx* c, Gnn& G, const double* dZ, double scale) {
  const int m = G.m;
  const float* th = G.theta.p;
  float* gr = G.grad.p;
  hipLaunchKernelGGL(gnn_coarse_dout_kernel, gl((size_t)m * G_OC), dim3(256), 0, c->stream, m, dZ, scale, G.dOc.p);
  CFM_CHECK_LAUNCH();
  CFM_TRY(gemm(c, G_XC, G_OC, m, G.Xc.p, G_XC, true, G.dOc.p, G_OC, false, nullptr, gr + G.lay.Wc, G_OC));
  hipLaunchKernelGGL(col_sum_kernel, dim3(G_OC), dim3(256), 0, c->stream, m, G_OC, G.dOc.p, G_OC, gr + G.lay.bc);
  CFM_TRY(gemm(c, m, G_XC, G_OC, G.dOc.p, G_OC, false, th + G.lay.Wc, G_OC, true, nullptr, G.dXc.p, G_XC));
  hipLaunchKernelGGL(gnn_copy_dh_kernel, gl((size_t)m * G_H), dim3(256), 0, c->stream, m, G.dXc.p, G.dH.p);
  CFM_CHECK_LAUNCH();
  return gnn_trunk_backward(c, G);
}

// Back-propagates dL/dh of the last layer (G.dH) through the message-passing trunk and
// the embedding into G.grad.
static Status gnn_trunk_backward(cfm_ctx* c, Gnn& G) {
  const int m = G.m;
  const size_t E = G.E;
  const float* th = G.theta.p;
  float* gr = G.grad.p;
  const NearField& nf = c->nf;
  for (int l = G_L - 1; l >= 0; --l) {
    // h_{l+1} = h_l + relu(Zn): dZn = dH (masked); dH stays (identity path)
    hipLaunchKernelGGL(relu_mask_kernel, gl((size_t)m * G_H), dim3(256), 0, c->stream, (size_t)m * G_H, G.dH.p,
                       G.Zn[l].p, G.dZn.p);
    CFM_CHECK_LAUNCH();
    CFM_TRY(gemm(c, G_XN, G_H, m, G.Xn[l].p, G_XN, true, G.dZn.p, G_H, false, nullptr, gr + G.lay.W2[l], G_H));
    hipLaunchKernelGGL(col_sum_kernel, dim3(G_H), dim3(256), 0, c->stream, m, G_H, G.dZn.p, G_H, gr + G.lay.b2[l]);
    CFM_TRY(gemm(c, m, G_XN, G_H, G.dZn.p, G_H, false, th + G.lay.W2[l], G_H, true, nullptr, G.dXn.p, G_XN));
    // dH += dXn[:, :64]
    hipLaunchKernelGGL(add_cols_kernel, gl((size_t)m * G_H), dim3(256), 0, c->stream, m, G_H, G.dXn.p, G_XN,
                       G.dH.p, G_H);
    // mean aggregation and relu of the edge MLP
    hipLaunchKernelGGL(gnn_mean_back_kernel, gl(E * G_H), dim3(256), 0, c->stream, m, G.dXn.p, G.Ze[l].p, G.dZe.p);
    CFM_CHECK_LAUNCH();
    CFM_TRY(gemm(c, G_XE, G_H, (int)E, G.Xe[l].p, G_XE, true, G.dZe.p, G_H, false, nullptr, gr + G.lay.W1[l], G_H));
    hipLaunchKernelGGL(col_sum_kernel, dim3(G_H), dim3(256), 0, c->stream, (int)E, G_H, G.dZe.p, G_H, gr + G.lay.b1[l]);
    CFM_TRY(gemm(c, (int)E, G_XE, G_H, G.dZe.p, G_H, false, th + G.lay.W1[l], G_H, true, nullptr, G.dXe.p, G_XE));
    hipLaunchKernelGGL(gnn_scatter_edge_kernel, gl(E * 2 * G_H), dim3(256), 0, c->stream, m, nf.nbr.p, G.dXe.p, G.dH.p);
    CFM_CHECK_LAUNCH();
  }
  // embedding
  hipLaunchKernelGGL(relu_mask_kernel, gl((size_t)m * G_H), dim3(256), 0, c->stream, (size_t)m * G_H, G.dH.p,
                     G.Z0.p, G.dZn.p);
  CFM_CHECK_LAUNCH();
  CFM_TRY(gemm(c, G_F, G_H, m, G.nfeat.p, G_F, true, G.dZn.p, G_H, false, nullptr, gr + G.lay.Wn, G_H));
  hipLaunchKernelGGL(col_sum_kernel, dim3(G_H), dim3(256), 0, c->stream, m, G_H, G.dZn.p, G_H, gr + G.lay.bn);
  CFM_CHECK_LAUNCH();
  return {};
}

// Loss of the current correction over samples [s0, s1) and (if acc) dL/dC into Gacc.
// Returns the mean relative squared residual.
static Status gnn_loss_grad(cfm_ctx* c, Gnn& G, int s0, int s1, bool acc, double* loss) {
  const SystemDesc s = system_desc(c, CFM_SYS_BLOB);
  const int L = s.L;
  double tot = 0.0;
  if (acc) CFM_TRY(dev_fill(c, G.Gacc.p, 0.0, (int)(G.E * 9)));
  double* hb = nullptr;
  CFM_HIP(hipHostMalloc((void**)&hb, sizeof(double) * 8, hipHostMallocDefault));
  for (int t = s0; t < s1; ++t) {
    const double* b = G.samples_b.p + (size_t)t * L;
    CFM_TRY(nearfield_apply(c, c->nf, G.corr.p, b, G.y.p));
    CFM_TRY(system_apply(c, s, CFM_FP32C, G.y.p, G.Ay.p));
    // scal: [b.b, u.u, alpha, b.u, r.r]
    CFM_TRY(dev_dots(c, G.partial.p, b, 0, 1, b, L, G.scal.p));
    CFM_TRY(dev_dots(c, G.partial.p, G.Ay.p, 0, 1, G.Ay.p, L, G.scal.p + 1));
    CFM_TRY(dev_dots(c, G.partial.p, b, 0, 1, G.Ay.p, L, G.scal.p + 3));
    CFM_HIP(hipMemcpyAsync(hb, G.scal.p, sizeof(double) * 4, hipMemcpyDeviceToHost, c->stream));
    CFM_HIP(hipStreamSynchronize(c->stream));
    const double alpha = hb[1] > 0.0 ? hb[3] / hb[1] : 0.0;
    CFM_TRY(dev_axpby(c, 1.0, b, -alpha, G.Ay.p, G.r.p, L));
    CFM_TRY(dev_dots(c, G.partial.p, G.r.p, 0, 1, G.r.p, L, G.scal.p + 4));
    if (acc) {
      hb[2] = alpha;
      CFM_HIP(hipMemcpyAsync(G.scal.p + 2, hb + 2, sizeof(double), hipMemcpyHostToDevice, c->stream));
      CFM_TRY(system_apply(c, s, CFM_FP32C, G.r.p, G.g.p));
      hipLaunchKernelGGL(gnn_grad_corr_kernel, gl(G.E), dim3(256), 0, c->stream, G.m, c->nf.nbr.p, G.g.p, b,
                         G.scal.p, G.Gacc.p);
      CFM_CHECK_LAUNCH();
    }
    CFM_HIP(hipMemcpyAsync(hb + 4, G.scal.p + 4, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CFM_HIP(hipStreamSynchronize(c->stream));
    tot += hb[4] / hb[0];
  }
  (void)hipHostFree(hb);
  *loss = tot / std::max(1, s1 - s0);
  return {};
}

// ---- coarse objective --------------------------------------------------------------
// Training data of the coarse head: s_t = b_t - M W b_t for every recorded right-hand
// side (FP32C products), ||s_t||^2, the linear fields Zp with M Zp (rows 0..11 of MZt),
// and the Tikhonov term delta = 1e-10 tr((M Zp)^T M Zp) / 12, which depends on the
// geometry only.
static Status coarse_prepare(cfm_ctx* c, Gnn& G) {
  const SystemDesc s = system_desc(c, CFM_SYS_BLOB);
  const int L = s.L;
  const int k = kCoarseLinear + G_KC;
  if (!G.MZt.p) {
    CFM_TRY(G.Zp.alloc((size_t)kCoarseLinear * L));
    CFM_TRY(G.MZt.alloc((size_t)k * L));
    CFM_TRY(G.D.alloc((size_t)G_KC * L));
    CFM_TRY(G.dZ.alloc((size_t)G_KC * L));
    CFM_TRY(G.gram.alloc((size_t)k * k));
    CFM_TRY(G.bmat.alloc((size_t)G_TMAX * k));
    CFM_TRY(G.cmat.alloc((size_t)G_TMAX * k));
    CFM_TRY(G.wmat.alloc((size_t)G_TMAX * G_KC));
    CFM_TRY(G.cpart.alloc(dev_dots_partial_size(L, k)));
    CFM_TRY(coarse_linear_fields(c, G.Zp.p));
    for (int j = 0; j < kCoarseLinear; ++j)
      CFM_TRY(system_apply(c, s, CFM_FP32C, G.Zp.p + (size_t)j * L, G.MZt.p + (size_t)j * L));
    for (int j = 0; j < kCoarseLinear; ++j)
      CFM_TRY(dev_dots(c, G.cpart.p, G.MZt.p + (size_t)j * L, 0, 1, G.MZt.p + (size_t)j * L, L, G.gram.p + j));
    std::vector<double> d(kCoarseLinear);
    CFM_HIP(copy_to_host(c->stream, d.data(), G.gram.p, sizeof(double) * kCoarseLinear));
    double tr = 0.0;
    for (double v : d) tr += v;
    G.delta = 1e-10 * tr / kCoarseLinear;
  }
  CFM_TRY(G.S.alloc((size_t)G.capacity * L));
  CFM_TRY(G.snorm.alloc((size_t)G.capacity));
  for (int t = 0; t < G.nsamples; ++t) {
    const double* b = G.samples_b.p + (size_t)t * L;
    double* st = G.S.p + (size_t)t * L;
    CFM_TRY(nearfield_apply(c, c->nf, nullptr, b, G.y.p));
    CFM_TRY(system_apply(c, s, CFM_FP32C, G.y.p, G.Ay.p));
    CFM_TRY(dev_axpby(c, 1.0, b, -1.0, G.Ay.p, st, L));
    CFM_TRY(dev_dots(c, G.cpart.p, st, 0, 1, st, L, G.snorm.p + t));
  }
  G.s_norm2.assign(G.nsamples, 0.0);
  CFM_HIP(copy_to_host(c->stream, G.s_norm2.data(), G.snorm.p, sizeof(double) * G.nsamples));
  G.s_ready = G.nsamples;
  return {};
}

// Coarse objective over samples [s0, s1) (s1 - s0 <= G_TMAX):
//   L = mean_t (||s_t - M Z c_t||^2 + delta ||c_t||^2) / ||s_t||^2,  Z = [Zp | ZL],
//   (G + delta I) c_t = (M Z)^T s_t,  G = (M Z)^T M Z,
// which at the optimum equals 1 - c_t . (M Z)^T s_t / ||s_t||^2.  With acc, dL/dZL =
// -2/T M sum_t (s_t - M Z c_t) c_t,L^T / ||s_t||^2 (envelope theorem) is back-propagated
// into G.grad (zero-filled by the caller).
static Status coarse_loss_grad(cfm_ctx* c, Gnn& G, int s0, int s1, bool acc, double* loss) {
  const SystemDesc s = system_desc(c, CFM_SYS_BLOB);
  const int L = s.L;
  const int k = kCoarseLinear + G_KC;
  const int T = s1 - s0;
  CFM_TRY(gnn_forward(c, G));
  for (int q = 0; q < G_KC; ++q)
    CFM_TRY(system_apply(c, s, CFM_FP32C, G.zl.p + (size_t)q * L, G.MZt.p + (size_t)(kCoarseLinear + q) * L));
  for (int j = 0; j < k; ++j)
    CFM_TRY(dev_dots(c, G.cpart.p, G.MZt.p, (size_t)L, k, G.MZt.p + (size_t)j * L, L, G.gram.p + (size_t)j * k));
  for (int t = 0; t < T; ++t)
    CFM_TRY(dev_dots(c, G.cpart.p, G.MZt.p, (size_t)L, k, G.S.p + (size_t)(s0 + t) * L, L,
                     G.bmat.p + (size_t)t * k));
  std::vector<double> gram((size_t)k * k), bm((size_t)T * k), inv;
  CFM_HIP(hipMemcpyAsync(gram.data(), G.gram.p, sizeof(double) * k * k, hipMemcpyDeviceToHost, c->stream));
  CFM_HIP(copy_to_host(c->stream, bm.data(), G.bmat.p, sizeof(double) * T * k));
  if (!invert_gram(gram, k, inv, G.delta)) return make_status(CFM_ENONFINITE, "singular coarse Gram matrix");
  std::vector<double> cm((size_t)T * k), wm((size_t)T * G_KC);
  double tot = 0.0;
  for (int t = 0; t < T; ++t) {
    const double n2 = G.s_norm2[s0 + t];
    double cb = 0.0;
    for (int a = 0; a < k; ++a) {
      double v = 0.0;
      for (int b = 0; b < k; ++b) v += inv[(size_t)a * k + b] * bm[(size_t)t * k + b];
      cm[(size_t)t * k + a] = v;
      cb += v * bm[(size_t)t * k + a];
    }
    tot += n2 > 0.0 ? 1.0 - cb / n2 : 0.0;
    for (int q = 0; q < G_KC; ++q) wm[(size_t)t * G_KC + q] = n2 > 0.0 ? cm[(size_t)t * k + kCoarseLinear + q] / n2 : 0.0;
  }
  *loss = tot / std::max(1, T);
  if (!acc) return {};
  CFM_HIP(hipMemcpyAsync(G.cmat.p, cm.data(), sizeof(double) * cm.size(), hipMemcpyHostToDevice, c->stream));
  CFM_HIP(copy_to_device(c->stream, G.wmat.p, wm.data(), sizeof(double) * wm.size()));
  hipLaunchKernelGGL(gnn_coarse_dir_kernel, gl(L), dim3(256), 0, c->stream, L, T, k, G.S.p + (size_t)s0 * L, G.MZt.p,
                     G.cmat.p, G.wmat.p, G.D.p);
  CFM_CHECK_LAUNCH();
  for (int q = 0; q < G_KC; ++q)
    CFM_TRY(system_apply(c, s, CFM_FP32C, G.D.p + (size_t)q * L, G.dZ.p + (size_t)q * L));
  return gnn_backward_coarse(c, G, G.dZ.p, -2.0 / T);
}

Status gnn_coarse_vectors(cfm_ctx* c, const double** zl, int* kl) {
  Gnn* G = reinterpret_cast<Gnn*>(c->gnn);
  if (!G) return make_status(CFM_ESTATE, "cfm_gnn_init has not been called");
  CFM_TRY(gnn_forward(c, *G));
  *zl = G->zl.p;
  *kl = G_KC;
  return {};
}

static void init_params(Gnn& G, unsigned seed, std::vector<float>& h) {
  h.assign(G.lay.total, 0.f);
  std::mt19937 rng(seed);
  auto fill = [&](size_t off, int fan_in, int n) {
    std::normal_distribution<float> d(0.f, std::sqrt(2.0f / fan_in));
    for (int i = 0; i < n; ++i) h[off + i] = d(rng);
  };
  fill(G.lay.Wn, G_F, G_F * G_H);
  for (int l = 0; l < G_L; ++l) {
    fill(G.lay.W1[l], G_XE, G_XE * G_H);
    fill(G.lay.W2[l], G_XN, G_XN * G_H);
    // keep the residual branches small at the start
    for (int i = 0; i < G_XN * G_H; ++i) h[G.lay.W2[l] + i] *= 0.1f;
  }
  // edge output layer zero: the network starts as "no correction"; the coarse head
  // starts random (its vectors only extend the linear fields' coarse space)
  fill(G.lay.Wc, G_XC, G_XC * G_OC);
}

// Stores one (rhs, solution) pair of a blob solve in the training ring (device copies).
Status gnn_record(cfm_ctx* c, const double* b, const double* x) {
  Gnn* G = gnn_of(c);
  if (!G || !G->recording) return {};
  const size_t L = 3 * (size_t)G->m;
  const int slot = G->next;
  CFM_HIP(hipMemcpyAsync(G->samples_b.p + slot * L, b, sizeof(double) * L, hipMemcpyDeviceToDevice, c->stream));
  CFM_HIP(hipMemcpyAsync(G->samples_x.p + slot * L, x, sizeof(double) * L, hipMemcpyDeviceToDevice, c->stream));
  G->next = (G->next + 1) % G->capacity;
  G->nsamples = std::min(G->nsamples + 1, G->capacity);
  return {};
}

// The network's initial guess of a blob solve: the two-level preconditioner with the
// learned coarse space applied to the right-hand side, x0 = (W + Z c) b, when that
// space is in use (CFM_COARSE_LEARNED); the solver still checks its FP64 residual and
// rejects it when it is worse than a zero guess.
Status gnn_guess(cfm_ctx* c, const double* rhs, double* x0, bool* used) {
  *used = false;
  if (!gnn_of(c) || c->coarse.mode != CFM_COARSE_LEARNED) return {};
  CFM_TRY(system_precond(c, system_desc(c, CFM_SYS_BLOB), CFM_PC_NEAR, rhs, x0));
  *used = true;
  return {};
}

}  // namespace cfm

using namespace cfm;

static int gfail(cfm_ctx* c, const Status& s) {
  c->err = s.msg;
  return s.code;
}

void cfm_gnn_release(cfm_ctx* c) {
  Gnn* G = gnn_of(c);
  if (!G) return;
  DevBuf<float>* fb[] = {&G->theta, &G->grad, &G->adam_m, &G->adam_v, &G->nfeat, &G->efeat, &G->Z0,
                         &G->Xo, &G->Co, &G->dH, &G->dXe, &G->dZe, &G->dXn, &G->dZn, &G->dCo, &G->dXo, &G->dA,
                         &G->qfeat, &G->Xc, &G->Oc, &G->dOc, &G->dXc};
  for (auto* b : fb) b->release();
  for (int l = 0; l <= G_L; ++l) G->H[l].release();
  for (int l = 0; l < G_L; ++l) {
    G->Xe[l].release(); G->Ze[l].release(); G->Me[l].release(); G->Xn[l].release(); G->Zn[l].release();
  }
  DevBuf<double>* db[] = {&G->corr, &G->Gacc, &G->y, &G->Ay, &G->r, &G->g, &G->scal, &G->partial,
                          &G->samples_b, &G->samples_x, &G->zl, &G->S, &G->snorm, &G->MZt, &G->Zp, &G->D,
                          &G->dZ, &G->gram, &G->bmat, &G->cmat, &G->wmat, &G->cpart};
  for (auto* b : db) b->release();
  delete G;
  c->gnn = nullptr;
  c->nf_corr = nullptr;
  if (c->coarse.mode == CFM_COARSE_LEARNED) {
    c->coarse.mode = CFM_COARSE_NONE;
    c->coarse.valid = false;
  }
}

extern "C" {

int cfm_gnn_init(cfm_ctx* c, uint32_t seed, int32_t capacity) {
  if (!c) return CFM_EINVAL;
  if (!c->have_positions || c->p.n_blob == 0)
    return gfail(c, make_status(CFM_ESTATE, "the learned correction needs blob positions"));
  if (capacity < 1) return gfail(c, make_status(CFM_EINVAL, "capacity must be >= 1"));
  (void)hipSetDevice(c->device);
  cfm_gnn_release(c);
  Gnn* G = new Gnn();
  c->gnn = G;
  auto run = [&]() -> Status {
    if (!c->nf.valid) CFM_TRY(nearfield_build(c, c->nf));
    G->m = c->p.n_blob;
    G->E = (size_t)G->m * G_S;
    const size_t P = G->lay.total;
    CFM_TRY(G->theta.alloc(P));
    CFM_TRY(G->grad.alloc(P));
    CFM_TRY(G->adam_m.alloc(P));
    CFM_TRY(G->adam_v.alloc(P));
    CFM_TRY(gnn_alloc(c, *G));
    std::vector<float> h;
    init_params(*G, seed, h);
    CFM_HIP(copy_to_device(c->stream, G->theta.p, h.data(), sizeof(float) * P));
    hipLaunchKernelGGL(fill_f32_kernel, gl(P), dim3(256), 0, c->stream, P, G->adam_m.p, 0.f);
    hipLaunchKernelGGL(fill_f32_kernel, gl(P), dim3(256), 0, c->stream, P, G->adam_v.p, 0.f);
    CFM_CHECK_LAUNCH();
    G->capacity = capacity;
    CFM_TRY(G->samples_b.alloc((size_t)capacity * 3 * G->m));
    CFM_TRY(G->samples_x.alloc((size_t)capacity * 3 * G->m));
    CFM_TRY(gnn_forward(c, *G));
    CFM_HIP(hipStreamSynchronize(c->stream));
    return {};
  };
  Status s = run();
  if (!s.ok()) {
    cfm_gnn_release(c);
    return gfail(c, s);
  }
  return CFM_OK;
}

int cfm_gnn_num_params(cfm_ctx* c) {
  (void)c;
  return (int)GnnLayout().total;
}

int cfm_gnn_add_samples(cfm_ctx* c, int32_t n, const double* rhs, const double* sol) {
  if (!c || !rhs) return CFM_EINVAL;
  Gnn* G = gnn_of(c);
  if (!G) return gfail(c, make_status(CFM_ESTATE, "cfm_gnn_init has not been called"));
  const int m = G->m;
  (void)hipSetDevice(c->device);
  for (int t = 0; t < n; ++t) {
    const int slot = G->next;
    Status s = upload_aos(c, rhs + (size_t)t * 3 * m, m, 3, G->samples_b.p + (size_t)slot * 3 * m, m, 0);
    if (s.ok() && sol) s = upload_aos(c, sol + (size_t)t * 3 * m, m, 3, G->samples_x.p + (size_t)slot * 3 * m, m, 0);
    if (!s.ok()) return gfail(c, s);
    G->next = (G->next + 1) % G->capacity;
    G->nsamples = std::min(G->nsamples + 1, G->capacity);
  }
  return CFM_OK;
}

int cfm_gnn_train(cfm_ctx* c, int32_t epochs, int32_t batch, double lr, double* loss_before,
                  double* loss_after) {
  if (!c) return CFM_EINVAL;
  Gnn* G = gnn_of(c);
  if (!G) return gfail(c, make_status(CFM_ESTATE, "cfm_gnn_init has not been called"));
  if (G->nsamples < 1) return gfail(c, make_status(CFM_ESTATE, "no training samples"));
  if (epochs < 0 || batch < 1 || !(lr > 0)) return gfail(c, make_status(CFM_EINVAL, "bad training options"));
  (void)hipSetDevice(c->device);
  auto run = [&]() -> Status {
    double l0 = 0.0;
    CFM_TRY(gnn_forward(c, *G));
    CFM_TRY(gnn_loss_grad(c, *G, 0, G->nsamples, false, &l0));
    if (loss_before) *loss_before = l0;
    const size_t P = G->lay.total;
    for (int ep = 0; ep < epochs; ++ep) {
      for (int s0 = 0; s0 < G->nsamples; s0 += batch) {
        const int s1 = std::min(G->nsamples, s0 + batch);
        double l = 0.0;
        CFM_TRY(gnn_forward(c, *G));
        CFM_TRY(gnn_loss_grad(c, *G, s0, s1, true, &l));
        hipLaunchKernelGGL(fill_f32_kernel, gl(P), dim3(256), 0, c->stream, P, G->grad.p, 0.f);
        CFM_CHECK_LAUNCH();
        CFM_TRY(gnn_backward(c, *G, 1.0 / (s1 - s0)));
        ++G->adam_t;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)G->adam_t), c2 = 1.f - std::pow(b2, (float)G->adam_t);
        hipLaunchKernelGGL(adam_kernel, gl(P), dim3(256), 0, c->stream, P, G->theta.p, G->grad.p, G->adam_m.p,
                           G->adam_v.p, (float)lr, b1, b2, c1, c2);
        CFM_CHECK_LAUNCH();
      }
    }
    CFM_TRY(gnn_forward(c, *G));
    double l1 = 0.0;
    CFM_TRY(gnn_loss_grad(c, *G, 0, G->nsamples, false, &l1));
    if (loss_after) *loss_after = l1;
    G->last_loss = l1;
    G->trained = true;
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : gfail(c, s);
}

int cfm_gnn_record(cfm_ctx* c, int32_t on) {
  if (!c) return CFM_EINVAL;
  Gnn* G = gnn_of(c);
  if (!G) return gfail(c, make_status(CFM_ESTATE, "cfm_gnn_init has not been called"));
  G->recording = on != 0;
  return CFM_OK;
}

int cfm_gnn_samples(cfm_ctx* c) {
  Gnn* G = c ? gnn_of(c) : nullptr;
  return G ? G->nsamples : 0;
}

int cfm_gnn_enable(cfm_ctx* c, int32_t on) {
  if (!c) return CFM_EINVAL;
  Gnn* G = gnn_of(c);
  if (on && !G) return gfail(c, make_status(CFM_ESTATE, "cfm_gnn_init has not been called"));
  c->nf_corr = (on && G) ? G->corr.p : nullptr;
  if (G) G->enabled = on != 0;
  return CFM_OK;
}

// Parameters (FP32) in and out, for checks against a CPU restatement.
int cfm_gnn_get_params(cfm_ctx* c, float* out) {
  Gnn* G = c ? gnn_of(c) : nullptr;
  if (!G || !out) return c ? gfail(c, make_status(CFM_ESTATE, "no network")) : CFM_EINVAL;
  (void)hipSetDevice(c->device);
  hipError_t e = copy_to_host(c->stream, out, G->theta.p, sizeof(float) * G->lay.total);
  return e == hipSuccess ? CFM_OK : gfail(c, make_status(CFM_EHIP, "%s", hipGetErrorString(e)));
}

int cfm_gnn_set_params(cfm_ctx* c, const float* in) {
  Gnn* G = c ? gnn_of(c) : nullptr;
  if (!G || !in) return c ? gfail(c, make_status(CFM_ESTATE, "no network")) : CFM_EINVAL;
  (void)hipSetDevice(c->device);
  hipError_t e = copy_to_device(c->stream, G->theta.p, in, sizeof(float) * G->lay.total);
  if (e != hipSuccess) return gfail(c, make_status(CFM_EHIP, "%s", hipGetErrorString(e)));
  if (c->coarse.mode == CFM_COARSE_LEARNED) c->coarse.valid = false;
  Status s = gnn_forward(c, *G);
  if (s.ok() && hipStreamSynchronize(c->stream) != hipSuccess) s = make_status(CFM_EHIP, "sync failed");
  return s.ok() ? CFM_OK : gfail(c, s);
}

// Correction blocks C ([m][17][9], FP64) of the current parameters; the graph
// (neighbour ids [m][16]) is returned alongside.
int cfm_gnn_correction(cfm_ctx* c, double* corr, int32_t* nbr) {
  Gnn* G = c ? gnn_of(c) : nullptr;
  if (!G) return c ? gfail(c, make_status(CFM_ESTATE, "no network")) : CFM_EINVAL;
  (void)hipSetDevice(c->device);
  Status s = gnn_forward(c, *G);
  if (!s.ok()) return gfail(c, s);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return gfail(c, make_status(CFM_EHIP, "sync failed"));
  if (corr && copy_to_host(c->stream, corr, G->corr.p, sizeof(double) * G->E * 9) != hipSuccess)
    return gfail(c, make_status(CFM_EHIP, "copy failed"));
  if (nbr && copy_to_host(c->stream, nbr, c->nf.nbr.p, sizeof(int) * (size_t)G->m * G_K) != hipSuccess)
    return gfail(c, make_status(CFM_EHIP, "copy failed"));
  return CFM_OK;
}

// Gradient of the training loss over samples [s0, s1) with respect to the
// parameters (FP32 out) and the loss itself.
int cfm_gnn_gradient(cfm_ctx* c, int32_t s0, int32_t s1, float* grad, double* loss) {
  Gnn* G = c ? gnn_of(c) : nullptr;
  if (!G) return c ? gfail(c, make_status(CFM_ESTATE, "no network")) : CFM_EINVAL;
  if (s0 < 0 || s1 > G->nsamples || s1 <= s0) return gfail(c, make_status(CFM_EINVAL, "bad sample range"));
  (void)hipSetDevice(c->device);
  auto run = [&]() -> Status {
    double l = 0.0;
    CFM_TRY(gnn_forward(c, *G));
    CFM_TRY(gnn_loss_grad(c, *G, s0, s1, true, &l));
    hipLaunchKernelGGL(fill_f32_kernel, gl(G->lay.total), dim3(256), 0, c->stream, G->lay.total, G->grad.p, 0.f);
    CFM_CHECK_LAUNCH();
    CFM_TRY(gnn_backward(c, *G, 1.0 / (s1 - s0)));
    CFM_HIP(hipStreamSynchronize(c->stream));
    if (grad) CFM_HIP(copy_to_host(c->stream, grad, G->grad.p, sizeof(float) * G->lay.total));
    if (loss) *loss = l;
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : gfail(c, s);
}

// Trains the coarse head on the recorded right-hand sides (coarse objective above);
// the learned coarse space of CFM_COARSE_LEARNED is rebuilt on its next use.
int cfm_gnn_train_coarse(cfm_ctx* c, int32_t epochs, int32_t batch, double lr, double* loss_before,
                         double* loss_after) {
  if (!c) return CFM_EINVAL;
  Gnn* G = gnn_of(c);
  if (!G) return gfail(c, make_status(CFM_ESTATE, "cfm_gnn_init has not been called"));
  if (G->nsamples < 1) return gfail(c, make_status(CFM_ESTATE, "no training samples"));
  if (epochs < 0 || batch < 1 || batch > G_TMAX || !(lr > 0))
    return gfail(c, make_status(CFM_EINVAL, "bad training options (batch 1..%d)", G_TMAX));
  (void)hipSetDevice(c->device);
  auto full_loss = [&](double* out) -> Status {
    double tot = 0.0;
    for (int s0 = 0; s0 < G->nsamples; s0 += G_TMAX) {
      const int s1 = std::min(G->nsamples, s0 + G_TMAX);
      double l = 0.0;
      CFM_TRY(coarse_loss_grad(c, *G, s0, s1, false, &l));
      tot += l * (s1 - s0);
    }
    *out = tot / G->nsamples;
    return {};
  };
  auto run = [&]() -> Status {
    CFM_TRY(coarse_prepare(c, *G));
    double l0 = 0.0;
    CFM_TRY(full_loss(&l0));
    if (loss_before) *loss_before = l0;
    const size_t P = G->lay.total;
    for (int ep = 0; ep < epochs; ++ep) {
      for (int s0 = 0; s0 < G->nsamples; s0 += batch) {
        const int s1 = std::min(G->nsamples, s0 + batch);
        double l = 0.0;
        hipLaunchKernelGGL(fill_f32_kernel, gl(P), dim3(256), 0, c->stream, P, G->grad.p, 0.f);
        CFM_CHECK_LAUNCH();
        CFM_TRY(coarse_loss_grad(c, *G, s0, s1, true, &l));
        ++G->adam_t;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)G->adam_t), c2 = 1.f - std::pow(b2, (float)G->adam_t);
        hipLaunchKernelGGL(adam_kernel, gl(P), dim3(256), 0, c->stream, P, G->theta.p, G->grad.p, G->adam_m.p,
                           G->adam_v.p, (float)lr, b1, b2, c1, c2);
        CFM_CHECK_LAUNCH();
      }
    }
    double l1 = 0.0;
    CFM_TRY(full_loss(&l1));
    if (loss_after) *loss_after = l1;
    if (c->coarse.mode == CFM_COARSE_LEARNED) c->coarse.valid = false;
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : gfail(c, s);
}

// Gradient of the coarse objective over samples [s0, s1) with respect to the parameters
// (FP32 out), the objective and its Tikhonov term delta.
int cfm_gnn_coarse_gradient(cfm_ctx* c, int32_t s0, int32_t s1, float* grad, double* loss, double* delta) {
  Gnn* G = c ? gnn_of(c) : nullptr;
  if (!G) return c ? gfail(c, make_status(CFM_ESTATE, "no network")) : CFM_EINVAL;
  if (s0 < 0 || s1 > G->nsamples || s1 <= s0 || s1 - s0 > G_TMAX)
    return gfail(c, make_status(CFM_EINVAL, "bad sample range"));
  (void)hipSetDevice(c->device);
  auto run = [&]() -> Status {
    CFM_TRY(coarse_prepare(c, *G));
    hipLaunchKernelGGL(fill_f32_kernel, gl(G->lay.total), dim3(256), 0, c->stream, G->lay.total, G->grad.p, 0.f);
    CFM_CHECK_LAUNCH();
    double l = 0.0;
    CFM_TRY(coarse_loss_grad(c, *G, s0, s1, true, &l));
    CFM_HIP(hipStreamSynchronize(c->stream));
    if (grad) CFM_HIP(copy_to_host(c->stream, grad, G->grad.p, sizeof(float) * G->lay.total));
    if (loss) *loss = l;
    if (delta) *delta = G->delta;
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : gfail(c, s);
}

// The learned coarse vectors of the current parameters: G_KC vectors of the blob system,
// each [3][n_blob] (component-major), one after another.
int cfm_gnn_coarse_vectors(cfm_ctx* c, double* zl) {
  Gnn* G = c ? gnn_of(c) : nullptr;
  if (!G || !zl) return c ? gfail(c, make_status(CFM_ESTATE, "no network")) : CFM_EINVAL;
  (void)hipSetDevice(c->device);
  const double* d = nullptr;
  int kl = 0;
  Status s = gnn_coarse_vectors(c, &d, &kl);
  if (s.ok()) {
    hipError_t e = copy_to_host(c->stream, zl, d, sizeof(double) * kl * 3 * (size_t)G->m);
    if (e != hipSuccess) s = make_status(CFM_EHIP, "%s", hipGetErrorString(e));
  }
  return s.ok() ? CFM_OK : gfail(c, s);
}

// Test hook: C = op(A) op(B) on the matrix cores (BF16 operands, FP32 accumulate).
int cfm_gemm_bf16(cfm_ctx* c, int32_t M, int32_t N, int32_t K, const float* A, int32_t transA,
                  const float* B, int32_t transB, float* C) {
  if (!c || !A || !B || !C || M <= 0 || N <= 0 || K <= 0) return CFM_EINVAL;
  (void)hipSetDevice(c->device);
  DevBuf<float> a, b, cc;
  auto run = [&]() -> Status {
    CFM_TRY(a.alloc((size_t)M * K));
    CFM_TRY(b.alloc((size_t)K * N));
    CFM_TRY(cc.alloc((size_t)M * N));
    CFM_HIP(copy_to_device(c->stream, a.p, A, sizeof(float) * M * K));
    CFM_HIP(copy_to_device(c->stream, b.p, B, sizeof(float) * K * N));
    CFM_TRY(gemm(c, M, N, K, a.p, transA ? M : K, transA != 0, b.p, transB ? K : N, transB != 0, nullptr, cc.p, N));
    CFM_HIP(hipStreamSynchronize(c->stream));
    CFM_HIP(copy_to_host(c->stream, C, cc.p, sizeof(float) * M * N));
    return {};
  };
  Status s = run();
  a.release();
  b.release();
  cc.release();
  return s.ok() ? CFM_OK : gfail(c, s);
}

}  // extern "C"
