// gmres.hip -- right-preconditioned restarted GMRES on the device, FP64 Krylov
// arithmetic, with FP64 iterative refinement around an FP32C inner solve.
//
// Restates the algorithm of mobility_solver::solve_linear_system
// (src/cilia/mobility/mobility_solver.cpp:2187-2669): right preconditioning
// (config.hpp:377), initial guess rejected when its residual exceeds ||rhs||
// (:2406-2417), Givens rotations in the reference's form (:2467-2506), back
// substitution and x += P^-1 Q y (:2520-2559).  Differences: Krylov cycles are
// restarted after `restart` vectors (the reference keeps min(n, 350) vectors),
// orthogonalisation is classical Gram-Schmidt applied twice (two block products
// per step instead of j sequential dot/axpy pairs), and every reduction runs on
// the GPU in a fixed order.  The residual estimate |beta_{j+1}| is the only value
// read back per step (to decide termination).
//
// Vectors are structure-of-arrays [dof][m] over the particles of the system.
#include <algorithm>
#include <cmath>

#include "ctx.hpp"
#include "krylov.hpp"
#include "nearfield.hpp"
#include "densepc.hpp"
#include "coarse.hpp"
#include "precond.hpp"

namespace cfm {

Status launch_apply(cfm_ctx* c, const ApplyPlan& plan);
int choose_split(cfm_ctx* c, const ApplyPlan& pl);
ApplyPlan plan_for(const cfm_ctx* c, int which);


constexpr int kRB = 256;  // threads per reduction block
constexpr int kChunk = kRB * 8;

// partial[b*k + i] = sum over chunk b of Q[i][r] * w[r]
__global__ __launch_bounds__(kRB) void dots_partial_kernel(const double* __restrict__ Q,
                                                           size_t ldq, int k,
                                                           const double* __restrict__ w, int L,
                                                           double* __restrict__ partial) {
  __shared__ double red[kRB];
  const int i = blockIdx.x;
  const int b = blockIdx.y;
  const double* q = Q + (size_t)i * ldq;
  const int r0 = b * kChunk;
  const int r1 = min(L, r0 + kChunk);
  double s = 0.0;
  for (int r = r0 + threadIdx.x; r < r1; r += kRB) s += q[r] * w[r];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = kRB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(size_t)b * k + i] = red[0];
}

// Q[0..k)^T w in one launch.  Block b covers rows [b kChunk, (b+1) kChunk) of every
// column, each thread the eight rows r0 + 256 j with its entries of w in registers.
// Columns go four at a time: a thread's four running sums take all 32 loads at once,
// each sum is reduced by a butterfly within the wave, and the four wave sums of a column
// are added in wave order through LDS by wave 0, which stores the block's partials.  The
// block then publishes them with an agent-scope release fence (by wave 0, the writer)
// and takes a ticket from `counter`; the block holding the last ticket (after an acquire
// fence) reduces every column's partials -- a wave per column: lane l adds chunks l,
// l + 64, ... in order, then the 64 lane sums are added by a butterfly in a fixed
// pattern -- and resets the counter for the next reduction queued on the stream.  No
// block waits for another, one ticket is taken per block whatever k is, and the order of
// every addition is fixed (deterministic).
// mode 0: out[i] = s; 1: out[i] += s; 2: out[i] = sqrt(max(s, 0)).  out2, when given,
// receives the same value at out2[i * ld2].
// Members (blockIdx.z = e, independent reductions of equal length L): Q + e mq, w + e mw,
// out + e ldo, counter e, their own partials; members with active[e] == 0 are skipped.
constexpr int kPerThread = kChunk / kRB;
constexpr int kColBatch = 4;

// NC columns of Q (stride ldq) against this thread's rows r0 + 256 j: the products of
// all columns are issued together, each column's sums reduced by a butterfly within the
// wave; lane 0 writes the wave's sums to red_w[0 .. NC).  FULL: every row is below L.
template <int NC, bool FULL>
__device__ __forceinline__ void dots_columns(const double* __restrict__ Q, size_t ldq,
                                             const double (&wv)[kPerThread], int r0, int L, int lane,
                                             double* red_w) {
  double v[NC][kPerThread];
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    const int r = r0 + j * kRB;
#pragma unroll
    for (int i = 0; i < NC; ++i) v[i][j] = (FULL || r < L) ? Q[(size_t)i * ldq + r] : 0.0;
  }
  double s[NC];
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    s[i] = 0.0;
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) s[i] += v[i][j] * wv[j];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int i = 0; i < NC; ++i) s[i] += __shfl_xor(s[i], off, 64);
  }
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NC; ++i) red_w[i] = s[i];
  }
}
__global__ __launch_bounds__(kRB) void dots_fused_kernel(const double* __restrict__ Q, size_t ldq,
                                                         int k, const double* __restrict__ w, int L,
                                                         double* partial, unsigned* counter,
                                                         double* out, int mode, double* out2,
                                                         int ld2, size_t mq, size_t mw, int ldo,
                                                         const int* __restrict__ active) {
  __shared__ double red[kRB / 64][64];
  __shared__ int last;
  const int e = blockIdx.z;
  if (active && !active[e]) return;
  Q += e * mq;
  w += e * mw;
  out += (size_t)e * ldo;
  counter += e;
  partial += (size_t)e * gridDim.x * k;
  const int b = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r0 = b * kChunk + threadIdx.x;
  const bool full = (b + 1) * kChunk <= L;
  double wv[kPerThread];
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    const int r = r0 + j * kRB;
    wv[j] = r < L ? w[r] : 0.0;
  }
  for (int g0 = 0; g0 < k; g0 += 64) {
    const int gn = min(64, k - g0);
    int c0 = 0;
    if (full) {
      for (; c0 + kColBatch <= gn; c0 += kColBatch)
        dots_columns<kColBatch, true>(Q + (size_t)(g0 + c0) * ldq, ldq, wv, r0, L, lane, red[wave] + c0);
      for (; c0 < gn; ++c0)
        dots_columns<1, true>(Q + (size_t)(g0 + c0) * ldq, ldq, wv, r0, L, lane, red[wave] + c0);
    } else {
      for (; c0 + kColBatch <= gn; c0 += kColBatch)
        dots_columns<kColBatch, false>(Q + (size_t)(g0 + c0) * ldq, ldq, wv, r0, L, lane, red[wave] + c0);
      for (; c0 < gn; ++c0)
        dots_columns<1, false>(Q + (size_t)(g0 + c0) * ldq, ldq, wv, r0, L, lane, red[wave] + c0);
    }
    __syncthreads();
    if ((int)threadIdx.x < gn) {
      const int i = threadIdx.x;
      partial[(size_t)b * k + g0 + i] = (red[0][i] + red[1][i]) + (red[2][i] + red[3][i]);
    }
    __syncthreads();  // red is rewritten by the next group of columns
  }
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    last = atomicAdd(counter, 1u) == gridDim.x - 1u;
  }
  __syncthreads();
  if (!last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  const int nb = gridDim.x;
  for (int col = wave; col < k; col += kRB / 64) {
    double t = 0.0;
    for (int bb = lane; bb < nb; bb += 64) t += partial[(size_t)bb * k + col];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (lane == 0) {
      if (mode == 2) t = sqrt(fmax(t, 0.0));
      const double v = mode == 1 ? out[col] + t : t;
      out[col] = v;
      if (out2) out2[(size_t)col * ld2] = v;
    }
  }
  if (threadIdx.x == 0) *counter = 0u;
}

// w[r] -= sum_{i<k} h[i] Q[i][r]
__global__ void gemv_sub_kernel(const double* __restrict__ Q, size_t ldq, int k,
                                const double* __restrict__ h, double* __restrict__ w, int L) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L) return;
  double s = 0.0;
  for (int i = 0; i < k; ++i) s += h[i] * Q[(size_t)i * ldq + r];
  w[r] -= s;
}

// out[r] = sum_{i<k} y[i] Q[i][r]
__global__ void gemv_kernel(const double* __restrict__ Q, size_t ldq, int k,
                            const double* __restrict__ y, double* __restrict__ out, int L) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L) return;
  double s = 0.0;
  for (int i = 0; i < k; ++i) s += y[i] * Q[(size_t)i * ldq + r];
  out[r] = s;
}

__global__ void axpby_kernel(double a, const double* __restrict__ x, double b,
                             const double* __restrict__ y, double* __restrict__ out, int L) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < L) out[r] = a * x[r] + b * y[r];
}

// out = x / s[0] (s on device)
__global__ void scale_by_inv_kernel(const double* __restrict__ x, const double* __restrict__ s,
                                    double* __restrict__ out, int L) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const double d = s[0];
  if (r < L) out[r] = d != 0.0 ? x[r] / d : 0.0;
}

__global__ void fill_kernel(double* __restrict__ x, double v, int L) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < L) x[r] = v;
}

// Copies `dof` rows of length m between two row-major blocks.
__global__ void copy_rows_kernel(const double* __restrict__ src, size_t src_ld, int src_off,
                                 double* __restrict__ dst, size_t dst_ld, int dst_off, int dof,
                                 int m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  for (int k = 0; k < dof; ++k) dst[(size_t)k * dst_ld + dst_off + i] = src[(size_t)k * src_ld + src_off + i];
}

// h[i] (+)= sum_b partial[b*k+i] (every block reduces the partials in the same
// order), then w[r] -= sum_{i<k} h_i Q[i][r]; block 0 stores h.  The column passes
// through LDS in chunks of kRB entries, so any restart length fits (the reference
// keeps up to 350 vectors); the sum over i keeps ascending order.  Launched with
// kRB threads per block.
__global__ __launch_bounds__(kRB) void cgs_finish_sub_kernel(const double* __restrict__ Q,
                                                             size_t ldq, int k,
                                                             const double* __restrict__ partial,
                                                             int nb, double* __restrict__ h_out,
                                                             const double* __restrict__ h_prev,
                                                             double* __restrict__ w, int L) {
  __shared__ double hs[kRB];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  for (int c0 = 0; c0 < k; c0 += kRB) {
    const int cn = min(kRB, k - c0);
    __syncthreads();  // the previous chunk has been read by every thread
    if ((int)threadIdx.x < cn) {
      const int i = c0 + (int)threadIdx.x;
      double a = 0.0;
      for (int b = 0; b < nb; ++b) a += partial[(size_t)b * k + i];
      hs[threadIdx.x] = a;
      if (blockIdx.x == 0) h_out[i] = h_prev ? h_prev[i] + a : a;
    }
    __syncthreads();
    if (r < L)
      for (int i = 0; i < cn; ++i) s += hs[i] * Q[(size_t)(c0 + i) * ldq + r];
  }
  if (r < L) w[r] -= s;
}

// Finishes ||w|| from its partials, then one Givens step on column j (below).
__global__ void norm_givens_kernel(const double* __restrict__ partial, int nb,
                                   double* __restrict__ H, int ldh, double* __restrict__ cs,
                                   double* __restrict__ sn, double* __restrict__ beta,
                                   const double* __restrict__ hcol, double* __restrict__ hnorm,
                                   int j, double* __restrict__ res);

// One Givens step on column j of H (ld = m+1 rows, column-major), the reference's
// formulas (mobility_solver.cpp:2467-2506); writes |beta_{j+1}| to res.
__global__ void givens_kernel(double* __restrict__ H, int ldh, double* __restrict__ cs,
                              double* __restrict__ sn, double* __restrict__ beta,
                              const double* __restrict__ hcol, const double* __restrict__ hnorm,
                              int j, double* __restrict__ res) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double* col = H + (size_t)j * ldh;
  for (int i = 0; i <= j; ++i) col[i] = hcol[i];
  col[j + 1] = hnorm[0];
  for (int i = 1; i <= j; ++i) {
    const double t = cs[i - 1] * col[i - 1] + sn[i - 1] * col[i];
    col[i] = cs[i - 1] * col[i] - sn[i - 1] * col[i - 1];
    col[i - 1] = t;
  }
  const double a = col[j], b = col[j + 1];
  double c, s;
  if (b == 0.0) {
    c = 1.0;
    s = 0.0;
  } else if (fabs(b) > fabs(a)) {
    const double t = a / b;
    s = ((b > 0.0) ? 1.0 : -1.0) / sqrt(1.0 + t * t);
    c = t * s;
  } else {
    const double t = b / a;
    c = ((a > 0.0) ? 1.0 : -1.0) / sqrt(1.0 + t * t);
    s = t * c;
  }
  cs[j] = c;
  sn[j] = s;
  col[j] = c * a + s * b;
  col[j + 1] = 0.0;
  beta[j + 1] = -s * beta[j];
  beta[j] = c * beta[j];
  res[0] = fabs(beta[j + 1]);
}

__global__ void norm_givens_kernel(const double* __restrict__ partial, int nb,
                                   double* __restrict__ H, int ldh, double* __restrict__ cs,
                                   double* __restrict__ sn, double* __restrict__ beta,
                                   const double* __restrict__ hcol, double* __restrict__ hnorm,
                                   int j, double* __restrict__ res) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += partial[b];
  hnorm[0] = sqrt(fmax(s, 0.0));
  double* col = H + (size_t)j * ldh;
  for (int i = 0; i <= j; ++i) col[i] = hcol[i];
  col[j + 1] = hnorm[0];
  for (int i = 1; i <= j; ++i) {
    const double t = cs[i - 1] * col[i - 1] + sn[i - 1] * col[i];
    col[i] = cs[i - 1] * col[i] - sn[i - 1] * col[i - 1];
    col[i - 1] = t;
  }
  const double a = col[j], b = col[j + 1];
  double c, sg;
  if (b == 0.0) {
    c = 1.0;
    sg = 0.0;
  } else if (fabs(b) > fabs(a)) {
    const double t = a / b;
    sg = ((b > 0.0) ? 1.0 : -1.0) / sqrt(1.0 + t * t);
    c = t * sg;
  } else {
    const double t = b / a;
    c = ((a > 0.0) ? 1.0 : -1.0) / sqrt(1.0 + t * t);
    sg = t * c;
  }
  cs[j] = c;
  sn[j] = sg;
  col[j] = c * a + sg * b;
  col[j + 1] = 0.0;
  beta[j + 1] = -sg * beta[j];
  beta[j] = c * beta[j];
  res[0] = fabs(beta[j + 1]);
}

__global__ void backsub_kernel(const double* __restrict__ H, int ldh,
                               const double* __restrict__ beta, int k, double* __restrict__ y) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int i = k - 1; i >= 0; --i) {
    double s = beta[i];
    for (int jj = i + 1; jj < k; ++jj) s -= H[(size_t)jj * ldh + i] * y[jj];
    const double d = H[(size_t)i * ldh + i];
    y[i] = d != 0.0 ? s / d : 0.0;
  }
}

__global__ void beta_init_kernel(double* __restrict__ beta, int n, const double* __restrict__ b0) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    beta[i] = (i == 0) ? b0[0] : 0.0;
}

// ---- preconditioners --------------------------------------------------------
// y = P^-1 x over the particles [off, off+m) of the combined index space.
__global__ void precond_kernel(int kind, int wall, int dof, int m, int off,
                               const double* __restrict__ z, double a, double mu,
                               const double* __restrict__ x, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k < dof; ++k) v[k] = x[(size_t)k * m + i];
  double o[6] = {0, 0, 0, 0, 0, 0};
  particle_precond(kind, wall, dof, z[off + i], a, mu, v, o);
  for (int k = 0; k < dof; ++k) y[(size_t)k * m + i] = o[k];
}

static inline dim3 g1(int L) { return dim3((L + 255) / 256); }

// ---------------------------------------------------------------------------
Status Krylov::reserve(int L_, int m_) {
  if (L_ <= L && m_ <= m && Q.p) return {};
  L = std::max(L, L_);
  m = std::max(m, m_);
  CFM_TRY(Q.alloc((size_t)(m + 1) * L));
  CFM_TRY(Z.alloc((size_t)m * L));
  CFM_TRY(H.alloc((size_t)(m + 1) * (m + 1)));
  CFM_TRY(cs.alloc(m + 1));
  CFM_TRY(sn.alloc(m + 1));
  CFM_TRY(beta.alloc(m + 2));
  CFM_TRY(h.alloc(m + 2));
  CFM_TRY(h2.alloc(m + 2));
  CFM_TRY(y.alloc(m + 2));
  CFM_TRY(scal.alloc(8));
  CFM_TRY(r.alloc(L));
  CFM_TRY(w.alloc(L));
  CFM_TRY(z.alloc(L));
  CFM_TRY(t.alloc(L));
  CFM_TRY(d.alloc(L));
  const int nb = (L + kChunk - 1) / kChunk;
  CFM_TRY(partial.alloc((size_t)nb * (m + 2)));
  CFM_TRY(res_ring.alloc(4));
  if (!host_res) CFM_HIP(hipHostMalloc((void**)&host_res, sizeof(double) * 8, hipHostMallocDefault));
  if (!ev[0])
    for (int i = 0; i < 4; ++i) CFM_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
  return {};
}

void Krylov::release() {
  Q.release(); Z.release(); H.release(); cs.release(); sn.release(); beta.release(); h.release();
  h2.release(); y.release(); scal.release(); r.release(); w.release(); z.release();
  t.release(); d.release(); dx.release(); rhs.release(); partial.release(); res_ring.release();
  for (int i = 0; i < 4; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  if (host_res) (void)hipHostFree(host_res);
  host_res = nullptr;
  L = m = 0;
}

// The context's reduction counters (at least n), created zeroed on the context's stream.
static Status red_counter(cfm_ctx* c, unsigned** out, int n = 1) {
  if (!c->red_count.p || c->red_count.n < (size_t)n) {
    CFM_TRY(c->red_count.alloc((size_t)std::max(n, 1)));
    CFM_HIP(hipMemsetAsync(c->red_count.p, 0, sizeof(unsigned) * c->red_count.n, c->stream));
  }
  *out = c->red_count.p;
  return {};
}

// One-launch reduction of Q[0..k)^T w into out (mode as dots_fused_kernel); with
// members > 1, one reduction per member (strides mq, mw, ldo; active may be null).
static Status dots_launch(cfm_ctx* c, double* partial, const double* Q, size_t ldq, int k,
                          const double* w, int L, double* out, int mode, double* out2 = nullptr,
                          int ld2 = 0, int members = 1, size_t mq = 0, size_t mw = 0, int ldo = 0,
                          const int* active = nullptr) {
  if (k <= 0 || members <= 0) return {};
  unsigned* cnt = nullptr;
  CFM_TRY(red_counter(c, &cnt, members));
  const int nb = (L + kChunk - 1) / kChunk;
  hipLaunchKernelGGL(dots_fused_kernel, dim3(nb, 1, members), dim3(kRB), 0, c->stream, Q, ldq, k,
                     w, L, partial, cnt, out, mode, out2, ld2, mq, mw, ldo, active);
  CFM_CHECK_LAUNCH();
  return {};
}

// out[0..k) = Q[0..k)^T w
Status dots(cfm_ctx* c, Krylov& K, const double* Q, size_t ldq, int k, const double* w, int L,
            double* out, bool accumulate) {
  return dots_launch(c, K.partial.p, Q, ldq, k, w, L, out, accumulate ? 1 : 0);
}

Status norm2(cfm_ctx* c, Krylov& K, const double* w, int L, double* out) {
  return dots_launch(c, K.partial.p, w, 0, 1, w, L, out, 2);
}

// Copies n <= 4 consecutive device doubles to K.host_res[0..n) with one wait.
Status read_scalars(cfm_ctx* c, Krylov& K, const double* dev, int n) {
  CFM_HIP(hipMemcpyAsync(K.host_res, dev, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  CFM_HIP(hipStreamSynchronize(c->stream));
  return {};
}

// As read_scalars, running hooks->before_first_wait (when given) between queueing the
// copy and waiting for it.
static Status read_scalars_hooked(cfm_ctx* c, Krylov& K, const double* dev, int n,
                                  const GmresHooks* hooks) {
  if (!hooks || !hooks->before_first_wait) return read_scalars(c, K, dev, n);
  CFM_HIP(hipMemcpyAsync(K.host_res, dev, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  CFM_HIP(hipEventRecord(K.ev[0], c->stream));
  CFM_TRY(hooks->before_first_wait(c, hooks->wait_user));
  CFM_HIP(hipEventSynchronize(K.ev[0]));
  return {};
}

Status read_scalar(cfm_ctx* c, Krylov& K, const double* dev, double* host) {
  CFM_HIP(hipMemcpyAsync(K.host_res, dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CFM_HIP(hipStreamSynchronize(c->stream));
  *host = K.host_res[0];
  return {};
}

SystemDesc system_desc(const cfm_ctx* c, int system) {
  SystemDesc s;
  if (system == CFM_SYS_SEG) {
    s.m = c->p.n_seg;
    s.dof = c->p.seg_dof;
    s.off = 0;
    s.which = CFM_SS;
    s.a = c->cfg.a_seg;
  } else {
    s.m = c->p.n_blob;
    s.dof = 3;
    s.off = c->p.n_seg;
    s.which = CFM_BB;
    s.a = c->cfg.a_blob;
  }
  s.L = s.m * s.dof;
  return s;
}

// y = A x for a system vector (layout [dof][m]); precision CFM_FP64 / CFM_FP32C.
// The pair kernel reads x and writes y in place (no staging copies).
Status system_apply(cfm_ctx* c, const SystemDesc& s, int precision, const double* x, double* y) {
  const int saved = c->cfg.precision;
  c->cfg.precision = precision;
  ApplyPlan pl = plan_for(c, s.which);
  pl.f_ext = x;
  pl.f_ld = s.m;
  pl.f_off = s.off;
  pl.v_ext = y;
  pl.v_ld = s.m;
  pl.v_off = s.off;
  c->split = choose_split(c, pl);
  Status st = launch_apply(c, pl);
  c->cfg.precision = saved;
  if (st.ok() && s.cilia) st = cilia_couple_apply(c, x, y);
  if (st.ok() && s.body) st = body_couple_apply(c, x, y);
  return st;
}

Status system_precond(cfm_ctx* c, const SystemDesc& s, int kind, const double* x, double* y) {
  if (s.cilia) return cilia_precond(c, kind, x, y);
  if (s.body) return body_precond(c, kind, x, y);
  if (kind == CFM_PC_NEAR) {
    if (s.which != CFM_BB) return make_status(CFM_EINVAL, "CFM_PC_NEAR applies to the blob system only");
    if (!c->nf.valid) CFM_TRY(nearfield_build(c, c->nf));
    CFM_TRY(nearfield_apply(c, c->nf, c->nf_corr, x, y));
    if (c->coarse.mode != CFM_COARSE_NONE) CFM_TRY(coarse_apply(c, c->coarse, x, y));
    return {};
  }
  if (kind == CFM_PC_DENSE) {
    if (s.which != CFM_BB) return make_status(CFM_EINVAL, "CFM_PC_DENSE applies to the blob system only");
    if (!c->dinv.valid) CFM_TRY(dense_build(c, c->dinv));
    return dense_apply(c, c->dinv, x, y);
  }
  hipLaunchKernelGGL(precond_kernel, g1(s.m), dim3(256), 0, c->stream, kind, c->cfg.wall, s.dof,
                     s.m, s.off, c->p.z.p, s.a, c->cfg.mu, x, y);
  CFM_CHECK_LAUNCH();
  return {};
}

// Restarted GMRES on A (pair precision `prec`) for b (device), x in/out (device).
// Stops when the residual estimate reaches `target_abs`, after max_iter steps in
// total, or after `max_cycles` cycles.  On return K.r holds b - A x (operator
// `prec`) and *res_abs its norm, except after max_cycles: then *res_abs is the last
// residual estimate and K.r is stale.  With r_given the caller has already put
// b - A x into K.r and its norm into *res_abs (no initial product).
Status gmres_core(cfm_ctx* c, Krylov& K, const SystemDesc& s, int prec, int pc, const double* b,
                  double* x, double target_abs, int restart, int max_iter, int max_cycles,
                  int* iters, double* res_abs, const GmresHooks* hooks, bool r_given) {
  const int L = s.L;
  const int m = restart;
  CFM_TRY(K.reserve(L, m));
  const size_t ldq = (size_t)K.L;
  const int ldh = m + 1;
  double nr = 0.0;
  if (r_given) {
    nr = *res_abs;
    CFM_TRY(norm2(c, K, K.r.p, L, K.scal.p));  // device copy of ||r|| for q0
  } else {
    // r = b - A x
    CFM_TRY(system_apply(c, s, prec, x, K.w.p));
    hipLaunchKernelGGL(axpby_kernel, g1(L), dim3(256), 0, c->stream, 1.0, b, -1.0, K.w.p, K.r.p, L);
    CFM_CHECK_LAUNCH();
    CFM_TRY(norm2(c, K, K.r.p, L, K.scal.p));
    CFM_TRY(read_scalar(c, K, K.scal.p, &nr));
  }
  if (!std::isfinite(nr)) return make_status(CFM_ENONFINITE, "non-finite GMRES residual");
  int cycles = 0;
  double last_est = nr;
  const bool sync_check = pc == CFM_PC_DENSE;
  while (nr > target_abs && *iters < max_iter) {
    // q0 = r / ||r||
    hipLaunchKernelGGL(scale_by_inv_kernel, g1(L), dim3(256), 0, c->stream, K.r.p, K.scal.p, K.Q.p, L);
    hipLaunchKernelGGL(beta_init_kernel, dim3(1), dim3(64), 0, c->stream, K.beta.p, m + 1, K.scal.p);
    CFM_CHECK_LAUNCH();
    int k = 0;
    const int nb = (L + kChunk - 1) / kChunk;
    // Residual estimates are copied to a pinned ring and read one step late, so
    // the host never waits for the step it has just queued.
    int pending = -1;
    bool stop = false;
    for (int j = 0; j < m && !stop; ++j) {
      double* qj = K.Q.p + (size_t)j * ldq;
      double* qn = K.Q.p + (size_t)(j + 1) * ldq;
      double* zj = K.Z.p + (size_t)j * ldq;
      CFM_TRY(system_precond(c, s, pc, qj, zj));
      if (hooks && hooks->precond_correction) CFM_TRY(hooks->precond_correction(c, s, qj, zj, hooks->user));
      CFM_TRY(system_apply(c, s, prec, zj, K.w.p));
      // classical Gram-Schmidt, twice
      hipLaunchKernelGGL(dots_partial_kernel, dim3(j + 1, nb), dim3(kRB), 0, c->stream, K.Q.p, ldq, j + 1, K.w.p, L, K.partial.p);
      hipLaunchKernelGGL(cgs_finish_sub_kernel, g1(L), dim3(256), 0, c->stream, K.Q.p, ldq, j + 1, K.partial.p, nb, K.h.p, (const double*)nullptr, K.w.p, L);
      hipLaunchKernelGGL(dots_partial_kernel, dim3(j + 1, nb), dim3(kRB), 0, c->stream, K.Q.p, ldq, j + 1, K.w.p, L, K.partial.p);
      hipLaunchKernelGGL(cgs_finish_sub_kernel, g1(L), dim3(256), 0, c->stream, K.Q.p, ldq, j + 1, K.partial.p, nb, K.h2.p, (const double*)K.h.p, K.w.p, L);
      hipLaunchKernelGGL(dots_partial_kernel, dim3(1, nb), dim3(kRB), 0, c->stream, K.w.p, (size_t)0, 1, K.w.p, L, K.partial.p);
      hipLaunchKernelGGL(norm_givens_kernel, dim3(1), dim3(1), 0, c->stream, K.partial.p, nb, K.H.p, ldh, K.cs.p, K.sn.p,
                         K.beta.p, K.h2.p, K.scal.p + 1, j, K.res_ring.p + (j % 4));
      hipLaunchKernelGGL(scale_by_inv_kernel, g1(L), dim3(256), 0, c->stream, K.w.p, K.scal.p + 1, qn, L);
      CFM_CHECK_LAUNCH();
      CFM_HIP(hipMemcpyAsync(K.host_res + 4 + (j % 4), K.res_ring.p + (j % 4), sizeof(double), hipMemcpyDeviceToHost, c->stream));
      CFM_HIP(hipEventRecord(K.ev[j % 4], c->stream));
      ++*iters;
      k = j + 1;
      if (sync_check) {
        // few, expensive steps (resident inverse): read this step's estimate now
        CFM_HIP(hipEventSynchronize(K.ev[j % 4]));
        const double est = K.host_res[4 + j % 4];
        if (!std::isfinite(est)) return make_status(CFM_ENONFINITE, "non-finite GMRES iterate");
        last_est = est;
        if (est <= target_abs) stop = true;
      } else if (pending >= 0) {
        CFM_HIP(hipEventSynchronize(K.ev[pending % 4]));
        const double est = K.host_res[4 + pending % 4];
        if (!std::isfinite(est)) return make_status(CFM_ENONFINITE, "non-finite GMRES iterate");
        last_est = est;
        if (est <= target_abs) stop = true;
      }
      pending = j;
      if (*iters >= max_iter) stop = true;
    }
    // x += P^-1 Q y = Z y
    hipLaunchKernelGGL(backsub_kernel, dim3(1), dim3(1), 0, c->stream, K.H.p, ldh, K.beta.p, k, K.y.p);
    hipLaunchKernelGGL(gemv_kernel, g1(L), dim3(256), 0, c->stream, K.Z.p, ldq, k, K.y.p, K.z.p, L);
    hipLaunchKernelGGL(axpby_kernel, g1(L), dim3(256), 0, c->stream, 1.0, x, 1.0, K.z.p, x, L);
    CFM_CHECK_LAUNCH();
    if (++cycles >= max_cycles) {  // caller recomputes the residual itself
      nr = last_est;
      break;
    }
    // With the FP64 operator and the estimate of the final step in hand (read
    // synchronously), an estimate within the target (with a 10% margin) ends the
    // solve without another product: with CGS2 the recurrence residual and b - A x
    // agree to rounding, far below the target (the reference stops on the estimate
    // alone, :2510).
    if (sync_check && prec == CFM_FP64 && last_est <= 0.9 * target_abs) {
      nr = last_est;
      break;
    }
    // true residual with this operator
    CFM_TRY(system_apply(c, s, prec, x, K.w.p));
    hipLaunchKernelGGL(axpby_kernel, g1(L), dim3(256), 0, c->stream, 1.0, b, -1.0, K.w.p, K.r.p, L);
    CFM_CHECK_LAUNCH();
    CFM_TRY(norm2(c, K, K.r.p, L, K.scal.p));
    CFM_TRY(read_scalar(c, K, K.scal.p, &nr));
    if (!std::isfinite(nr)) return make_status(CFM_ENONFINITE, "non-finite GMRES residual");
  }
  *res_abs = nr;
  return {};
}

// Full solve with the reference's initial-guess rule and optional mixed precision.
Status solve(cfm_ctx* c, Krylov& K, const SystemDesc& s, const cfm_solve_opts& o, const double* b,
             double* x, cfm_solve_stats* st, const GmresHooks* hooks) {
  const int L = s.L;
  const int restart = o.restart > 0 ? o.restart : 50;
  const int max_iter = o.max_iter > 0 ? o.max_iter : 1000;
  CFM_TRY(K.reserve(L, restart));
  cfm_solve_stats S{};
  // ||b|| and the FP64 residual of the guess come back with one wait
  double nb = 0.0, nr = 0.0;
  CFM_TRY(norm2(c, K, b, L, K.scal.p + 3));
  if (o.use_guess) {
    CFM_TRY(system_apply(c, s, CFM_FP64, x, K.w.p));
    hipLaunchKernelGGL(axpby_kernel, g1(L), dim3(256), 0, c->stream, 1.0, b, -1.0, K.w.p, K.d.p, L);
    CFM_CHECK_LAUNCH();
    CFM_TRY(norm2(c, K, K.d.p, L, K.scal.p));
    CFM_TRY(read_scalars_hooked(c, K, K.scal.p, 4, hooks));
    nr = K.host_res[0];
    nb = K.host_res[3];
  } else {
    CFM_TRY(read_scalars_hooked(c, K, K.scal.p + 3, 1, hooks));
    nb = K.host_res[0];
    nr = nb;
  }
  if (!std::isfinite(nb)) return make_status(CFM_ENONFINITE, "non-finite right-hand side");
  if (nb == 0.0 || !o.use_guess) {
    hipLaunchKernelGGL(fill_kernel, g1(L), dim3(256), 0, c->stream, x, 0.0, L);
    CFM_CHECK_LAUNCH();
  }
  if (nb == 0.0) {  // zero rhs: zero solution (mobility_solver.cpp:2221-2294)
    S.converged = 1;
    if (st) *st = S;
    return {};
  }
  const double target = o.rtol * nb;
  if (!o.use_guess) {  // x = 0: the residual is b itself
    CFM_HIP(hipMemcpyAsync(K.d.p, b, sizeof(double) * L, hipMemcpyDeviceToDevice, c->stream));
  }
  // reject a guess worse than zero
  if (!std::isfinite(nr)) return make_status(CFM_ENONFINITE, "non-finite residual of the guess");
  if (nr > nb) {
    hipLaunchKernelGGL(fill_kernel, g1(L), dim3(256), 0, c->stream, x, 0.0, L);
    hipLaunchKernelGGL(axpby_kernel, g1(L), dim3(256), 0, c->stream, 1.0, b, 0.0, b, K.d.p, L);
    CFM_CHECK_LAUNCH();
    nr = nb;
  }
  S.initial_rel_residual = nr / nb;
  int iters = 0;
  if (o.mode == CFM_SOLVE_FP64) {
    double res = nr;
    if (nr > target) {
      // K.d holds b - A x of the (possibly rejected) guess
      CFM_HIP(hipMemcpyAsync(K.r.p, K.d.p, sizeof(double) * L, hipMemcpyDeviceToDevice, c->stream));
      CFM_TRY(gmres_core(c, K, s, CFM_FP64, o.precond, b, x, target, restart, max_iter, 1 << 30,
                         &iters, &res, hooks, true));
    }
    nr = res;
  } else {
    // iterative refinement: d solves A32 d = r to a relative 5e-7 (or what is left
    // to reach the target), x += d, r = b - A64 x.
    int refinements = 0;
    while (nr > target && iters < max_iter && refinements < 40) {
      const double rel_in = std::max(5e-7, 0.5 * target / nr);
      hipLaunchKernelGGL(fill_kernel, g1(L), dim3(256), 0, c->stream, K.t.p, 0.0, L);
      CFM_CHECK_LAUNCH();
      // K.d holds r; the inner solve overwrites K.r/K.w/K.z, so keep r in K.d
      double inner_res = 0.0;
      DevBuf<double>& dvec = K.dx;
      CFM_TRY(dvec.alloc(L));
      hipLaunchKernelGGL(fill_kernel, g1(L), dim3(256), 0, c->stream, dvec.p, 0.0, L);
      CFM_CHECK_LAUNCH();
      CFM_TRY(K.rhs.alloc(L));
      CFM_HIP(hipMemcpyAsync(K.rhs.p, K.d.p, sizeof(double) * L, hipMemcpyDeviceToDevice, c->stream));
      // one restart cycle per refinement: the FP32C operator is linear only to its
      // rounding, so an inner solve can stall short of rel_in; the FP64 residual
      // below decides what is left.
      // d starts at zero, so the inner residual is r itself
      CFM_HIP(hipMemcpyAsync(K.r.p, K.d.p, sizeof(double) * L, hipMemcpyDeviceToDevice, c->stream));
      inner_res = nr;
      CFM_TRY(gmres_core(c, K, s, CFM_FP32C, o.precond, K.rhs.p, dvec.p, rel_in * nr, restart,
                         max_iter, 1, &iters, &inner_res, hooks, true));
      hipLaunchKernelGGL(axpby_kernel, g1(L), dim3(256), 0, c->stream, 1.0, x, 1.0, dvec.p, x, L);
      CFM_CHECK_LAUNCH();
      CFM_TRY(system_apply(c, s, CFM_FP64, x, K.w.p));
      hipLaunchKernelGGL(axpby_kernel, g1(L), dim3(256), 0, c->stream, 1.0, b, -1.0, K.w.p, K.d.p, L);
      CFM_CHECK_LAUNCH();
      CFM_TRY(norm2(c, K, K.d.p, L, K.scal.p));
      CFM_TRY(read_scalar(c, K, K.scal.p, &nr));
      if (!std::isfinite(nr)) return make_status(CFM_ENONFINITE, "non-finite refined residual");
      ++refinements;
    }
    S.refinements = refinements;
  }
  S.iterations = iters;
  S.rel_residual = nr / nb;
  S.converged = S.rel_residual <= o.rtol;
  if (st) *st = S;
  if (!S.converged)
    return make_status(CFM_ENOCONV, "GMRES stopped at relative residual %.3e > %.3e after %d steps",
                       S.rel_residual, o.rtol, iters);
  return {};
}

}  // namespace cfm

namespace cfm {

// ---- BLAS-1/2 helpers shared with the Broyden driver -------------------------
Status dev_dots(cfm_ctx* c, double* partial, const double* Q, size_t ldq, int k, const double* w,
                int L, double* out) {
  return dots_launch(c, partial, Q, ldq, k, w, L, out, 0);
}
Status dev_dots2(cfm_ctx* c, double* partial, const double* Q, size_t ldq, int k, const double* w,
                 int L, double* out, double* out2, int ld2) {
  return dots_launch(c, partial, Q, ldq, k, w, L, out, 0, out2, ld2);
}
size_t dev_dots_partial_size(int L, int k) { return (size_t)((L + kChunk - 1) / kChunk) * k; }

Status dev_gemv_sub(cfm_ctx* c, const double* Q, size_t ldq, int k, const double* h, double* w, int L) {
  hipLaunchKernelGGL(gemv_sub_kernel, g1(L), dim3(256), 0, c->stream, Q, ldq, k, h, w, L);
  CFM_CHECK_LAUNCH();
  return {};
}
Status dev_axpby(cfm_ctx* c, double a, const double* x, double b, const double* y, double* out, int L) {
  hipLaunchKernelGGL(axpby_kernel, g1(L), dim3(256), 0, c->stream, a, x, b, y, out, L);
  CFM_CHECK_LAUNCH();
  return {};
}
Status dev_gemv(cfm_ctx* c, const double* Q, size_t ldq, int k, const double* y, double* out, int L) {
  hipLaunchKernelGGL(gemv_kernel, g1(L), dim3(256), 0, c->stream, Q, ldq, k, y, out, L);
  CFM_CHECK_LAUNCH();
  return {};
}
Status dev_fill(cfm_ctx* c, double* x, double v, int L) {
  hipLaunchKernelGGL(fill_kernel, g1(L), dim3(256), 0, c->stream, x, v, L);
  CFM_CHECK_LAUNCH();
  return {};
}
Status dev_scale_by_inv(cfm_ctx* c, const double* x, const double* s, double* out, int L) {
  hipLaunchKernelGGL(scale_by_inv_kernel, g1(L), dim3(256), 0, c->stream, x, s, out, L);
  CFM_CHECK_LAUNCH();
  return {};
}
Status dev_norm2(cfm_ctx* c, double* partial, const double* w, int L, double* out) {
  return dots_launch(c, partial, w, 0, 1, w, L, out, 2);
}

// ---- members: E independent vectors of length Lm laid end to end ------------------
// w[r] -= sum_{i<k} h[m ldh + i] Q[i][r], m = r / Lm (the sum in the order of
// gemv_sub_kernel)
__global__ void member_gemv_sub_kernel(const double* __restrict__ Q, size_t ldq, int k,
                                       const double* __restrict__ h, int ldh,
                                       double* __restrict__ w, int L, int Lm) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L) return;
  const double* hm = h + (size_t)(r / Lm) * ldh;
  double s = 0.0;
  for (int i = 0; i < k; ++i) s += hm[i] * Q[(size_t)i * ldq + r];
  w[r] -= s;
}
// out[r] = x[r] / s[m] (0 when s[m] is 0)
__global__ void member_scale_by_inv_kernel(const double* __restrict__ x, const double* __restrict__ s,
                                           double* __restrict__ out, int L, int Lm) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L) return;
  const double d = s[r / Lm];
  out[r] = d != 0.0 ? x[r] / d : 0.0;
}
// x[r] = 0 where member m is inactive
__global__ void member_mask_kernel(double* __restrict__ x, int L, int Lm, const int* __restrict__ active) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < L && !active[r / Lm]) x[r] = 0.0;
}

Status dev_member_dots(cfm_ctx* c, double* partial, const double* Q, size_t ldq, int k,
                       const double* w, int Lm, int E, double* out, int ldo, bool norm,
                       const int* active) {
  return dots_launch(c, partial, Q, ldq, k, w, Lm, out, norm ? 2 : 0, nullptr, 0, E, (size_t)Lm,
                     (size_t)Lm, ldo, active);
}
size_t dev_member_dots_partial_size(int Lm, int k, int E) {
  return (size_t)E * dev_dots_partial_size(Lm, k);
}
Status dev_member_gemv_sub(cfm_ctx* c, const double* Q, size_t ldq, int k, const double* h, int ldh,
                           double* w, int L, int Lm) {
  if (k <= 0) return {};
  hipLaunchKernelGGL(member_gemv_sub_kernel, g1(L), dim3(256), 0, c->stream, Q, ldq, k, h, ldh, w, L, Lm);
  CFM_CHECK_LAUNCH();
  return {};
}
Status dev_member_scale_by_inv(cfm_ctx* c, const double* x, const double* s, double* out, int L, int Lm) {
  hipLaunchKernelGGL(member_scale_by_inv_kernel, g1(L), dim3(256), 0, c->stream, x, s, out, L, Lm);
  CFM_CHECK_LAUNCH();
  return {};
}
Status dev_member_mask(cfm_ctx* c, double* x, int L, int Lm, const int* active) {
  hipLaunchKernelGGL(member_mask_kernel, g1(L), dim3(256), 0, c->stream, x, L, Lm, active);
  CFM_CHECK_LAUNCH();
  return {};
}
Status dev_copy_rows(cfm_ctx* c, const double* src, size_t src_ld, int src_off, double* dst,
                     size_t dst_ld, int dst_off, int dof, int m) {
  hipLaunchKernelGGL(copy_rows_kernel, g1(m), dim3(256), 0, c->stream, src, src_ld, src_off, dst,
                     dst_ld, dst_off, dof, m);
  CFM_CHECK_LAUNCH();
  return {};
}

}  // namespace cfm
