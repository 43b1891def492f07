// batch.hip -- many independent wall systems of uneven size solved together.
//
// An ensemble (parameter sweeps, bisection, periodic-orbit searches) is a set of
// independent filament systems.  The reference runs each one as its own process
// (pyfile/driver/func.py:544 launches one executable per parameter set) with its own
// mobility_solver.  Here the systems share one context: their particles are
// concatenated, every mobility product is one launch of a persistent kernel that
// pulls (system, 64-target tile, source slice) work items from an atomic queue
// (largest first), and GMRES runs all systems in lockstep with per-system Krylov
// bases, Hessenbergs, Givens rotations and convergence flags.  One Arnoldi step of
// the whole batch is a fixed kernel sequence whose step index lives in device
// memory, so it is captured once into a hipGraph and replayed.
//
// Vectors of the batch are AoS over the concatenated particles (dof values per
// particle), so system s owns the contiguous slice [dof off_s, dof (off_s + n_s)).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "apply_tile.hpp"
#include "ctx.hpp"
#include "precond.hpp"

namespace cfm {

struct BItem {
  int sys, t_lo, t_hi, s_lo, s_hi, slice, pad0, pad1;
};
struct BChunk {
  int lo, hi, sys, pad;
};

constexpr int kBChunk = 2048;  // vector elements per reduction chunk (8 per thread)

// system status
constexpr int ST_DONE = 0;     // converged (or not part of this solve)
constexpr int ST_RUN = 1;      // taking Arnoldi steps
constexpr int ST_CHECK = 2;    // solution updated; true residual to be computed

// per-system scalars (double): slots of sinfo[s * 8 + k]
constexpr int SI_TARGET = 0, SI_EST = 1, SI_INV = 2, SI_NR = 3, SI_NB = 4;
// per-system ints: slots of iinfo[s * 8 + k]
constexpr int II_ST = 0, II_K = 1, II_ITERS = 2, II_UPD = 3, II_DONE_AT = 4;

struct BArgs {
  const double* __restrict__ x;
  const double* __restrict__ y;
  const double* __restrict__ z;
  const double* __restrict__ f;   // AoS, dof per particle
  double* __restrict__ partial;   // [slice][dof * N]
  const BItem* __restrict__ items;
  int* counters;                  // [0] next item, [1] finished workgroups
  const int* __restrict__ iinfo;  // per-system status (null: every system)
  int want;                       // status a system must have to be evaluated
  int n_items, N;
  double mu, a;
};

__device__ __forceinline__ bool sys_on(const int* iinfo, int s, int want) {
  return !iinfo || iinfo[s * 8 + II_ST] == want;
}

// Persistent mobility product: each workgroup takes items until the queue is empty;
// the last workgroup to leave resets the queue, so the kernel can be replayed in a
// graph without a separate reset.
template <typename T, int D, bool WALL>
__global__ __launch_bounds__(kWG) void batch_apply_kernel(BArgs a) {
  __shared__ SrcRec<T> tile[kTS];
  __shared__ double red[kWaves * 6][kTT];
  __shared__ int s_item;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(a.counters, 1);
    __syncthreads();
    const int it = s_item;
    if (it >= a.n_items) break;
    const BItem I = a.items[it];
    if (!sys_on(a.iinfo, I.sys, a.want)) continue;
    const int i = I.t_lo + lane;
    const bool valid = i < I.t_hi;
    const int ic = valid ? i : I.t_lo;
    const double xi = a.x[ic], yi = a.y[ic], zi = a.z[ic];
    const T zif = T(zi);
    double accd[D];
    float s32[D], c32[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      accd[k] = 0.0;
      s32[k] = 0.0f;
      c32[k] = 0.0f;
    }
    for (int t0 = I.s_lo; t0 < I.s_hi; t0 += kTS) {
      const int j = t0 + threadIdx.x;
      __syncthreads();
      if (j < I.s_hi) {
        SrcRec<T> r;
        r.x = a.x[j];
        r.y = a.y[j];
        r.z = a.z[j];
        if constexpr (sizeof(T) == 4) r.zf = float(r.z);
#pragma unroll
        for (int k = 0; k < D; ++k) r.f[k] = T((WALL ? kStageScale : 1.0) * a.f[(size_t)D * j + k]);
#pragma unroll
        for (int k = D; k < 6; ++k) r.f[k] = T(0);
        tile[threadIdx.x] = r;
      }
      __syncthreads();
      const int w0 = t0 + wave * 64;
      const int cnt = min(64, I.s_hi - w0);
      if (cnt <= 0) continue;
      const SrcRec<T>* rec = tile + wave * 64;
      const bool check = WALL && (w0 < I.t_hi) && (w0 + cnt > I.t_lo);
      const int iid = valid ? i : -1;
      T part[D];
#pragma unroll
      for (int k = 0; k < D; ++k) part[k] = T(0);
      if (check)
        tile_sweep<T, D, D, WALL, true>(rec, cnt, w0, iid, xi, yi, zi, zif, T(a.a), T(a.a), T(a.mu), part);
      else
        tile_sweep<T, D, D, WALL, false>(rec, cnt, w0, iid, xi, yi, zi, zif, T(a.a), T(a.a), T(a.mu), part);
      if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int k = 0; k < D; ++k) two_sum_add(s32[k], c32[k], part[k]);
      } else {
#pragma unroll
        for (int k = 0; k < D; ++k) accd[k] += double(part[k]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < D; ++k)
      red[wave * 6 + k][lane] = (sizeof(T) == 4) ? double(s32[k]) + double(c32[k]) : accd[k];
    __syncthreads();
    if (wave == 0 && valid) {
      double* out = a.partial + (size_t)I.slice * D * a.N + (size_t)D * i;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        double s = red[k][lane];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) s += red[w * 6 + k][lane];
        out[k] = s;
      }
    }
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(a.counters + 1, 1) == (int)gridDim.x - 1) {
      atomicExch(a.counters, 0);
      atomicExch(a.counters + 1, 0);
    }
  }
}

// out = scale (sum of the slices + self term); with rhs: out = rhs - that.
template <int D>
__global__ void batch_finalize_kernel(BArgs a, const int* __restrict__ sys_of,
                                      const int* __restrict__ nslice, int wall,
                                      const double* __restrict__ rhs, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const int s = sys_of[i];
  if (!sys_on(a.iinfo, s, a.want)) return;
  double o[6] = {0, 0, 0, 0, 0, 0};
  const int ns = nslice[s];
  for (int sl = 0; sl < ns; ++sl) {
    const double* p = a.partial + (size_t)sl * D * a.N + (size_t)D * i;
#pragma unroll
    for (int k = 0; k < D; ++k) o[k] += p[k];
  }
  if (wall) {
    double F[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < D; ++k) F[k] = a.f[(size_t)D * i + k];
    wall_self<D, D>(a.z[i], F, o);
  }
  const double scale = wall ? 1.0 / a.mu : 1.0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const double v = scale * o[k];
    out[(size_t)D * i + k] = rhs ? rhs[(size_t)D * i + k] - v : v;
  }
}

// ---- batched Krylov kernels ---------------------------------------------------------
struct BK {
  int L, m, n_sys, nchunk, dof, N;
  const BChunk* chunks;
  const int* chunk_begin;  // n_sys + 1: chunks of system s are [begin[s], begin[s+1])
  const int* sys_of;       // particle -> system
  double* Q;               // (m + 1) x L
  double* H;               // n_sys x (m + 1) x (m + 1), column j at H[s][j][0..m]
  double *cs, *sn, *beta, *h, *h2, *y;  // n_sys x (m + 2)
  double* partial;         // (m + 1) x nchunk
  double* sinfo;           // n_sys x 8
  int* iinfo;              // n_sys x 8
  int* jdev;               // [0] step j, [1] steps taken by the batch, [2] systems running
};

__device__ __forceinline__ int bk_sys_of_elem(const BK& k, int e) { return k.sys_of[e / k.dof]; }

// y = P^-1 x over particles of running systems; x = Q_j when from_q.
__global__ void bprecond_kernel(BK k, int kind, int wall, double a, double mu,
                                const double* __restrict__ z, const double* __restrict__ xin,
                                int from_q, int want, double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= k.N) return;
  const int s = k.sys_of[p];
  if (k.iinfo[s * 8 + II_ST] != want) return;
  const double* x = from_q ? k.Q + (size_t)k.jdev[0] * k.L : xin;
  double v[6] = {0, 0, 0, 0, 0, 0}, o[6] = {0, 0, 0, 0, 0, 0};
  for (int c = 0; c < k.dof; ++c) v[c] = x[(size_t)k.dof * p + c];
  particle_precond(kind, wall, k.dof, z[p], a, mu, v, o);
  for (int c = 0; c < k.dof; ++c) out[(size_t)k.dof * p + c] = o[c];
}

// partial[i][chunk] = <Q_i, w> over the chunk for i <= j (or <w, w> with norm != 0).
__global__ __launch_bounds__(256) void bdots_kernel(BK k, const double* __restrict__ w, int norm,
                                                    int want) {
  __shared__ double red[4];
  const int c = blockIdx.x;
  const BChunk ch = k.chunks[c];
  if (k.iinfo[ch.sys * 8 + II_ST] != want) return;
  double wv[kBChunk / 256];
#pragma unroll
  for (int e = 0; e < kBChunk / 256; ++e) {
    const int idx = ch.lo + e * 256 + (int)threadIdx.x;
    wv[e] = idx < ch.hi ? w[idx] : 0.0;
  }
  const int cnt = norm ? 1 : k.jdev[0] + 1;
  for (int i = 0; i < cnt; ++i) {
    const double* q = norm ? w : k.Q + (size_t)i * k.L;
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < kBChunk / 256; ++e) {
      const int idx = ch.lo + e * 256 + (int)threadIdx.x;
      if (idx < ch.hi) s += q[idx] * wv[e];
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) k.partial[(size_t)i * k.nchunk + c] = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
  }
}

// h[s][i] = sum over the system's chunks (fixed order); second pass: H[:, j] = h + h2.
__global__ void bfinish_kernel(BK k, int pass) {
  const int s = blockIdx.x;
  if (k.iinfo[s * 8 + II_ST] != ST_RUN) return;
  const int j = k.jdev[0];
  const int c0 = k.chunk_begin[s], c1 = k.chunk_begin[s + 1];
  const int ld = k.m + 2;
  for (int i = threadIdx.x; i <= j; i += blockDim.x) {
    double acc = 0.0;
    for (int c = c0; c < c1; ++c) acc += k.partial[(size_t)i * k.nchunk + c];
    if (pass == 0) {
      k.h[(size_t)s * ld + i] = acc;
    } else {
      k.h2[(size_t)s * ld + i] = acc;
      k.H[((size_t)s * (k.m + 1) + j) * (k.m + 1) + i] = k.h[(size_t)s * ld + i] + acc;
    }
  }
}

// w -= sum_{i <= j} h[s][i] Q_i over running systems.
__global__ void bsub_kernel(BK k, double* __restrict__ w, int pass) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= k.L) return;
  const int s = bk_sys_of_elem(k, e);
  if (k.iinfo[s * 8 + II_ST] != ST_RUN) return;
  const int j = k.jdev[0];
  const double* hh = (pass == 0 ? k.h : k.h2) + (size_t)s * (k.m + 2);
  double acc = 0.0;
  for (int i = 0; i <= j; ++i) acc += hh[i] * k.Q[(size_t)i * k.L + e];
  w[e] -= acc;
}

// Upper-triangular solve R y = beta over the first kk columns of system s.
__device__ void bbacksub(const BK& k, int s, int kk) {
  const int ldh = k.m + 1;
  const double* H = k.H + (size_t)s * ldh * ldh;
  const double* beta = k.beta + (size_t)s * (k.m + 2);
  double* y = k.y + (size_t)s * (k.m + 2);
  for (int i = kk - 1; i >= 0; --i) {
    double acc = beta[i];
    for (int c = i + 1; c < kk; ++c) acc -= H[(size_t)c * ldh + i] * y[c];
    y[i] = acc / H[(size_t)i * ldh + i];
  }
}

// Norm of w, Givens rotations of column j, residual estimate, and the per-system
// decisions of this step (mobility_solver.cpp:2467-2510 rotation formulas).  One
// workgroup; thread t handles systems t, t + blockDim, ...; thread 0 then advances
// the step counters.
__global__ __launch_bounds__(256) void bgivens_kernel(BK k) {
  __shared__ int run_count;
  if (threadIdx.x == 0) run_count = 0;
  __syncthreads();
  const int j = k.jdev[0];
  const int ldh = k.m + 1, ld = k.m + 2;
  for (int s = threadIdx.x; s < k.n_sys; s += blockDim.x) {
    int* ii = k.iinfo + s * 8;
    ii[II_UPD] = 0;
    if (ii[II_ST] != ST_RUN) continue;
    double nrm2 = 0.0;
    for (int c = k.chunk_begin[s]; c < k.chunk_begin[s + 1]; ++c) nrm2 += k.partial[c];
    const double nrm = sqrt(nrm2);
    double* col = k.H + ((size_t)s * ldh + j) * ldh;
    col[j + 1] = nrm;
    double* cs = k.cs + (size_t)s * ld;
    double* sn = k.sn + (size_t)s * ld;
    for (int i = 0; i < j; ++i) {
      const double t = cs[i] * col[i] + sn[i] * col[i + 1];
      col[i + 1] = -sn[i] * col[i] + cs[i] * col[i + 1];
      col[i] = t;
    }
    const double a = col[j], b = col[j + 1];
    const double den = sqrt(a * a + b * b);
    const double c_ = den > 0.0 ? a / den : 1.0, s_ = den > 0.0 ? b / den : 0.0;
    cs[j] = c_;
    sn[j] = s_;
    col[j] = c_ * a + s_ * b;
    col[j + 1] = 0.0;
    double* beta = k.beta + (size_t)s * ld;
    beta[j + 1] = -s_ * beta[j];
    beta[j] = c_ * beta[j];
    const double est = fabs(beta[j + 1]);
    double* si = k.sinfo + s * 8;
    si[SI_EST] = est;
    si[SI_INV] = nrm > 0.0 ? 1.0 / nrm : 0.0;
    ii[II_K] = j + 1;
    ii[II_ITERS] += 1;
    const bool conv = est <= si[SI_TARGET];
    if (conv) ii[II_DONE_AT] = k.jdev[1];
    if (conv || j + 1 >= k.m || !(nrm > 0.0) || !isfinite(est)) {
      bbacksub(k, s, j + 1);
      ii[II_UPD] = 1;
      ii[II_ST] = ST_CHECK;
    } else {
      atomicAdd(&run_count, 1);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    k.jdev[0] = j + 1;
    k.jdev[1] += 1;
    k.jdev[2] = run_count;
  }
}

// Q_{j+1} = w / ||w|| for systems still running (j already advanced).
__global__ void bscale_kernel(BK k, const double* __restrict__ w) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= k.L) return;
  const int s = bk_sys_of_elem(k, e);
  if (k.iinfo[s * 8 + II_ST] != ST_RUN) return;
  k.Q[(size_t)k.jdev[0] * k.L + e] = w[e] * k.sinfo[s * 8 + SI_INV];
}

// x += P^-1 Q y for systems whose cycle ended at this step.
__global__ void bupdate_kernel(BK k, int kind, int wall, double a, double mu,
                               const double* __restrict__ z, double* __restrict__ x) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= k.N) return;
  const int s = k.sys_of[p];
  const int* ii = k.iinfo + s * 8;
  if (!ii[II_UPD]) return;
  const int kk = ii[II_K];
  const double* y = k.y + (size_t)s * (k.m + 2);
  double t[6] = {0, 0, 0, 0, 0, 0}, o[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < kk; ++i)
    for (int c = 0; c < k.dof; ++c) t[c] += y[i] * k.Q[(size_t)i * k.L + (size_t)k.dof * p + c];
  particle_precond(kind, wall, k.dof, z[p], a, mu, t, o);
  for (int c = 0; c < k.dof; ++c) x[(size_t)k.dof * p + c] += o[c];
}

// After a residual pass over ST_CHECK systems (norms in partial[0][chunk]):
// converged -> ST_DONE, else restart: q0 = r / ||r||, beta = (||r||, 0, ...).
// mode 0: FP64 solve (target absolute in SI_TARGET); the step counter j resets.
__global__ __launch_bounds__(256) void brestart_kernel(BK k, int max_iter) {
  __shared__ int run_count;
  if (threadIdx.x == 0) run_count = 0;
  __syncthreads();
  for (int s = threadIdx.x; s < k.n_sys; s += blockDim.x) {
    int* ii = k.iinfo + s * 8;
    if (ii[II_ST] != ST_CHECK) continue;
    double nr2 = 0.0;
    for (int c = k.chunk_begin[s]; c < k.chunk_begin[s + 1]; ++c) nr2 += k.partial[c];
    const double nr = sqrt(nr2);
    double* si = k.sinfo + s * 8;
    si[SI_NR] = nr;
    if (nr <= si[SI_TARGET] || ii[II_ITERS] >= max_iter || !isfinite(nr)) {
      ii[II_ST] = ST_DONE;
      continue;
    }
    si[SI_INV] = 1.0 / nr;
    double* beta = k.beta + (size_t)s * (k.m + 2);
    beta[0] = nr;
    for (int i = 1; i < k.m + 2; ++i) beta[i] = 0.0;
    ii[II_K] = 0;
    ii[II_ST] = ST_RUN;
    atomicAdd(&run_count, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    k.jdev[0] = 0;
    k.jdev[2] = run_count;
  }
}

// Q_0 = r / ||r|| for systems that restarted.
__global__ void bq0_kernel(BK k, const double* __restrict__ r) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= k.L) return;
  const int s = bk_sys_of_elem(k, e);
  if (k.iinfo[s * 8 + II_ST] != ST_RUN) return;
  k.Q[e] = r[e] * k.sinfo[s * 8 + SI_INV];
}

// Norms of b per system -> SI_NB; all systems enter ST_CHECK with target rtol ||b||.
__global__ __launch_bounds__(256) void bsetup_kernel(BK k, double rtol) {
  for (int s = threadIdx.x; s < k.n_sys; s += blockDim.x) {
    double nb2 = 0.0;
    for (int c = k.chunk_begin[s]; c < k.chunk_begin[s + 1]; ++c) nb2 += k.partial[c];
    double* si = k.sinfo + s * 8;
    si[SI_NB] = sqrt(nb2);
    si[SI_TARGET] = rtol * si[SI_NB];
    int* ii = k.iinfo + s * 8;
    ii[II_ST] = ST_CHECK;
    ii[II_ITERS] = 0;
    ii[II_UPD] = 0;
    ii[II_K] = 0;
    ii[II_DONE_AT] = -1;
  }
}

// Mixed-precision bookkeeping.  Outer state lives in the slots of the outer solve:
//   set the inner targets from the outer residual norms (after a residual pass over
//   ST_CHECK systems): converged systems -> ST_DONE (outer), the rest get an inner
//   solve with target max(5e-7, 0.5 target / nr) * nr.
__global__ __launch_bounds__(256) void bouter_kernel(BK k, double* __restrict__ outer_target,
                                                     int max_iter) {
  __shared__ int run_count;
  if (threadIdx.x == 0) run_count = 0;
  __syncthreads();
  for (int s = threadIdx.x; s < k.n_sys; s += blockDim.x) {
    int* ii = k.iinfo + s * 8;
    if (ii[II_ST] != ST_CHECK) continue;
    double nr2 = 0.0;
    for (int c = k.chunk_begin[s]; c < k.chunk_begin[s + 1]; ++c) nr2 += k.partial[c];
    const double nr = sqrt(nr2);
    double* si = k.sinfo + s * 8;
    si[SI_NR] = nr;
    const double tgt = outer_target[s];
    if (nr <= tgt || ii[II_ITERS] >= max_iter || !isfinite(nr)) {
      ii[II_ST] = ST_DONE;
      continue;
    }
    si[SI_TARGET] = fmax(5e-7, 0.5 * tgt / nr) * nr;
    si[SI_INV] = 1.0 / nr;
    double* beta = k.beta + (size_t)s * (k.m + 2);
    beta[0] = nr;
    for (int i = 1; i < k.m + 2; ++i) beta[i] = 0.0;
    ii[II_K] = 0;
    ii[II_ST] = ST_RUN;
    atomicAdd(&run_count, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    k.jdev[0] = 0;
    k.jdev[2] = run_count;
  }
}

__global__ void bfill_kernel(double* __restrict__ x, double v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// x += d over systems with a status in {ST_CHECK}; d zeroed there afterwards.
__global__ void badd_kernel(BK k, double* __restrict__ x, double* __restrict__ d) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= k.L) return;
  const int s = bk_sys_of_elem(k, e);
  if (k.iinfo[s * 8 + II_ST] != ST_CHECK) return;
  x[e] += d[e];
  d[e] = 0.0;
}

static inline dim3 gb(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace cfm

using namespace cfm;

struct cfm_batch {
  cfm_batch_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  int n_sys = 0, N = 0, dof = 6, L = 0;
  std::vector<int> n_h, off_h;
  DevBuf<int> sys_of, nslice;
  DevBuf<double> x, y, z;
  DevBuf<double> f, v, partial;
  DevBuf<BItem> items;
  DevBuf<int> counters;
  int n_items = 0;
  cfm_batch_tuning tuning{};
  bool have_positions = false;
  double* host = nullptr;  // pinned staging
  size_t host_n = 0;
  // Krylov workspace
  int m = 0;
  DevBuf<BChunk> chunks;
  DevBuf<int> chunk_begin;
  int nchunk = 0;
  DevBuf<double> Q, H, cs, sn, beta, h, h2, yv, kpart, sinfo, r, w, zt, xs, dv, rhs, otarget;
  DevBuf<int> iinfo, jdev;
  // pinned: [0, 4) the step counters after a residual pass, [4, 8) a ring of the
  // running-system count after each graph replay (slot = step % 4)
  int* host_flags = nullptr;
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // one Arnoldi step, [FP64, FP32C]
  int graph_pc = -1;
  std::vector<hipEvent_t> ev;
  std::string err;
};

namespace {

int bfail(cfm_batch* b, const Status& s) {
  if (b) b->err = s.msg;
  return s.code;
}

Status stage_host(cfm_batch* b, size_t n) {
  if (b->host_n >= n) return {};
  if (b->host) (void)hipHostFree(b->host);
  b->host = nullptr;
  b->host_n = 0;
  CFM_HIP(hipHostMalloc((void**)&b->host, sizeof(double) * n, hipHostMallocDefault));
  b->host_n = n;
  return {};
}

Status upload(cfm_batch* b, const double* h, double* d, size_t n) {
  CFM_TRY(stage_host(b, n));
  std::memcpy(b->host, h, sizeof(double) * n);
  CFM_HIP(hipMemcpyAsync(d, b->host, sizeof(double) * n, hipMemcpyHostToDevice, b->stream));
  CFM_HIP(hipStreamSynchronize(b->stream));
  return {};
}

Status download(cfm_batch* b, const double* d, double* h, size_t n) {
  CFM_TRY(stage_host(b, n));
  CFM_HIP(hipMemcpyAsync(b->host, d, sizeof(double) * n, hipMemcpyDeviceToHost, b->stream));
  CFM_HIP(hipStreamSynchronize(b->stream));
  std::memcpy(h, b->host, sizeof(double) * n);
  return {};
}

void drop_graphs(cfm_batch* b) {
  for (auto& g : b->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  b->graph_pc = -1;
}

// Work items for the tuning: 64-target tiles x source slices of at most max_src
// sources (0: the whole system), largest first.
Status build_items(cfm_batch* b) {
  const int ms = b->tuning.max_sources;
  std::vector<BItem> it;
  std::vector<int> nsl(b->n_sys);
  int max_slices = 1;
  for (int s = 0; s < b->n_sys; ++s) {
    const int n = b->n_h[s], off = b->off_h[s];
    const int per = ms > 0 ? ms : std::max(n, 1);
    const int slices = std::max(1, (n + per - 1) / per);
    nsl[s] = slices;
    max_slices = std::max(max_slices, slices);
    for (int t = 0; t < n; t += kTT)
      for (int sl = 0; sl < slices; ++sl) {
        BItem I{};
        I.sys = s;
        I.t_lo = off + t;
        I.t_hi = off + std::min(n, t + kTT);
        I.s_lo = off + sl * per;
        I.s_hi = off + std::min(n, (sl + 1) * per);
        I.slice = sl;
        it.push_back(I);
      }
  }
  std::stable_sort(it.begin(), it.end(), [](const BItem& p, const BItem& q) {
    return (long long)(p.t_hi - p.t_lo) * (p.s_hi - p.s_lo) >
           (long long)(q.t_hi - q.t_lo) * (q.s_hi - q.s_lo);
  });
  b->n_items = (int)it.size();
  CFM_TRY(b->items.alloc(std::max<size_t>(1, it.size())));
  CFM_TRY(b->nslice.alloc(b->n_sys));
  CFM_TRY(b->partial.alloc((size_t)max_slices * b->dof * b->N));
  CFM_HIP(hipMemcpyAsync(b->items.p, it.data(), sizeof(BItem) * it.size(), hipMemcpyHostToDevice, b->stream));
  CFM_HIP(hipMemcpyAsync(b->nslice.p, nsl.data(), sizeof(int) * nsl.size(), hipMemcpyHostToDevice, b->stream));
  CFM_HIP(hipStreamSynchronize(b->stream));
  drop_graphs(b);
  return {};
}

BArgs bargs(cfm_batch* b, const double* f, const int* iinfo, int want) {
  BArgs a{};
  a.x = b->x.p;
  a.y = b->y.p;
  a.z = b->z.p;
  a.f = f;
  a.partial = b->partial.p;
  a.items = b->items.p;
  a.counters = b->counters.p;
  a.iinfo = iinfo;
  a.want = want;
  a.n_items = b->n_items;
  a.N = b->N;
  a.mu = b->cfg.mu;
  a.a = b->cfg.a_seg;
  return a;
}

// out = A f (or rhs - A f) over the systems with status `want` (all when iinfo null).
Status batch_apply(cfm_batch* b, int prec, const double* f, const int* iinfo, int want,
                   const double* rhs, double* out) {
  const BArgs a = bargs(b, f, iinfo, want);
  const dim3 grid(b->tuning.grid), blk(kWG);
  const bool wall = b->cfg.wall != 0;
  const bool s6 = b->dof == 6;
#define CFM_BLAUNCH(T, D, W) hipLaunchKernelGGL((batch_apply_kernel<T, D, W>), grid, blk, 0, b->stream, a)
  if (prec == CFM_FP32C) {
    if (s6) { if (wall) CFM_BLAUNCH(float, 6, true); else CFM_BLAUNCH(float, 6, false); }
    else { if (wall) CFM_BLAUNCH(float, 3, true); else CFM_BLAUNCH(float, 3, false); }
  } else {
    if (s6) { if (wall) CFM_BLAUNCH(double, 6, true); else CFM_BLAUNCH(double, 6, false); }
    else { if (wall) CFM_BLAUNCH(double, 3, true); else CFM_BLAUNCH(double, 3, false); }
  }
#undef CFM_BLAUNCH
  if (s6)
    hipLaunchKernelGGL(batch_finalize_kernel<6>, gb(b->N), dim3(256), 0, b->stream, a, b->sys_of.p,
                       b->nslice.p, (int)wall, rhs, out);
  else
    hipLaunchKernelGGL(batch_finalize_kernel<3>, gb(b->N), dim3(256), 0, b->stream, a, b->sys_of.p,
                       b->nslice.p, (int)wall, rhs, out);
  CFM_CHECK_LAUNCH();
  return {};
}

BK bk_of(cfm_batch* b) {
  BK k{};
  k.L = b->L;
  k.m = b->m;
  k.n_sys = b->n_sys;
  k.nchunk = b->nchunk;
  k.dof = b->dof;
  k.N = b->N;
  k.chunks = b->chunks.p;
  k.chunk_begin = b->chunk_begin.p;
  k.sys_of = b->sys_of.p;
  k.Q = b->Q.p;
  k.H = b->H.p;
  k.cs = b->cs.p;
  k.sn = b->sn.p;
  k.beta = b->beta.p;
  k.h = b->h.p;
  k.h2 = b->h2.p;
  k.y = b->yv.p;
  k.partial = b->kpart.p;
  k.sinfo = b->sinfo.p;
  k.iinfo = b->iinfo.p;
  k.jdev = b->jdev.p;
  return k;
}

// Workspace for restart length m (the cycle length is baked into the step graph, so
// a different m rebuilds it; buffers only grow).
Status reserve_krylov(cfm_batch* b, int m) {
  if (b->m == m && b->Q.p) return {};
  b->m = m;
  const int S = b->n_sys;
  const size_t L = b->L;
  // reduction chunks never straddle two systems
  std::vector<BChunk> ch;
  std::vector<int> cb(S + 1, 0);
  for (int s = 0; s < S; ++s) {
    cb[s] = (int)ch.size();
    const int lo = b->dof * b->off_h[s], hi = lo + b->dof * b->n_h[s];
    for (int c = lo; c < hi; c += kBChunk) ch.push_back(BChunk{c, std::min(hi, c + kBChunk), s, 0});
  }
  cb[S] = (int)ch.size();
  b->nchunk = (int)ch.size();
  CFM_TRY(b->chunks.alloc(std::max<size_t>(1, ch.size())));
  CFM_TRY(b->chunk_begin.alloc(S + 1));
  CFM_HIP(hipMemcpyAsync(b->chunks.p, ch.data(), sizeof(BChunk) * ch.size(), hipMemcpyHostToDevice, b->stream));
  CFM_HIP(hipMemcpyAsync(b->chunk_begin.p, cb.data(), sizeof(int) * cb.size(), hipMemcpyHostToDevice, b->stream));
  CFM_TRY(b->Q.alloc((size_t)(m + 1) * L));
  CFM_TRY(b->H.alloc((size_t)S * (m + 1) * (m + 1)));
  for (auto* v : {&b->cs, &b->sn, &b->beta, &b->h, &b->h2, &b->yv}) CFM_TRY(v->alloc((size_t)S * (m + 2)));
  CFM_TRY(b->kpart.alloc((size_t)(m + 1) * std::max(1, b->nchunk)));
  CFM_TRY(b->sinfo.alloc((size_t)S * 8));
  CFM_TRY(b->otarget.alloc((size_t)S));
  CFM_TRY(b->iinfo.alloc((size_t)S * 8));
  CFM_TRY(b->jdev.alloc(4));
  for (auto* v : {&b->r, &b->w, &b->zt, &b->xs, &b->dv, &b->rhs}) CFM_TRY(v->alloc(L));
  if (!b->host_flags) CFM_HIP(hipHostMalloc((void**)&b->host_flags, sizeof(int) * 8, hipHostMallocDefault));
  CFM_HIP(hipStreamSynchronize(b->stream));
  drop_graphs(b);
  return {};
}

// Enqueue one Arnoldi step of every running system (captured into the graph).
Status enqueue_step(cfm_batch* b, int prec, int pc, double* x) {
  const BK k = bk_of(b);
  const int wall = b->cfg.wall;
  const double a = b->cfg.a_seg, mu = b->cfg.mu;
  hipLaunchKernelGGL(bprecond_kernel, gb(b->N), dim3(256), 0, b->stream, k, pc, wall, a, mu, b->z.p,
                     (const double*)nullptr, 1, ST_RUN, b->zt.p);
  CFM_CHECK_LAUNCH();
  CFM_TRY(batch_apply(b, prec, b->zt.p, b->iinfo.p, ST_RUN, nullptr, b->w.p));
  for (int pass = 0; pass < 2; ++pass) {
    hipLaunchKernelGGL(bdots_kernel, dim3(b->nchunk), dim3(256), 0, b->stream, k, b->w.p, 0, ST_RUN);
    hipLaunchKernelGGL(bfinish_kernel, dim3(b->n_sys), dim3(64), 0, b->stream, k, pass);
    hipLaunchKernelGGL(bsub_kernel, gb(b->L), dim3(256), 0, b->stream, k, b->w.p, pass);
  }
  hipLaunchKernelGGL(bdots_kernel, dim3(b->nchunk), dim3(256), 0, b->stream, k, b->w.p, 1, ST_RUN);
  hipLaunchKernelGGL(bgivens_kernel, dim3(1), dim3(256), 0, b->stream, k);
  hipLaunchKernelGGL(bscale_kernel, gb(b->L), dim3(256), 0, b->stream, k, b->w.p);
  hipLaunchKernelGGL(bupdate_kernel, gb(b->N), dim3(256), 0, b->stream, k, pc, wall, a, mu, b->z.p, x);
  CFM_CHECK_LAUNCH();
  return {};
}

// Graph of one Arnoldi step for (prec, pc); x is the vector the step updates
// (fixed per solve mode: the solution in FP64 mode, the correction in mixed mode).
// The graph holds kernel nodes only: the running count the host polls is copied out
// by run_cycle after each replay, outside the graph.
Status step_graph(cfm_batch* b, int prec, int pc, double* x, hipGraphExec_t* out) {
  const int gi = prec == CFM_FP32C ? 1 : 0;
  if (b->graph_pc != pc) drop_graphs(b);
  if (!b->graph[gi]) {
    hipGraph_t g = nullptr;
    CFM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    Status s = enqueue_step(b, prec, pc, x);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (!s.ok()) {
      if (g) (void)hipGraphDestroy(g);
      return s;
    }
    if (e != hipSuccess) return make_status(CFM_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph[gi], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return make_status(CFM_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    b->graph_pc = pc;
  }
  *out = b->graph[gi];
  return {};
}

Status event_at(cfm_batch* b, size_t i, hipEvent_t* e) {
  while (b->ev.size() <= i) {
    hipEvent_t t;
    CFM_HIP(hipEventCreate(&t));
    b->ev.push_back(t);
  }
  *e = b->ev[i];
  return {};
}

// Run Arnoldi steps (graph replays) until no system is running or the cycle is
// full.  After each replay the running count is copied to its own pinned ring slot
// (step j -> host_flags[4 + j % 4]) and read one step late, after the event behind
// that copy: a slot is rewritten only four replays later, by which time the host has
// read it.
Status run_cycle(cfm_batch* b, hipGraphExec_t g, size_t* nev, int max_steps) {
  int pending = -1;
  int pending_slot = 0;
  for (int j = 0; j < b->m && j < max_steps; ++j) {
    CFM_HIP(hipGraphLaunch(g, b->stream));
    const int slot = 4 + (j % 4);
    CFM_HIP(hipMemcpyAsync(b->host_flags + slot, b->jdev.p + 2, sizeof(int), hipMemcpyDeviceToHost,
                           b->stream));
    hipEvent_t e;
    CFM_TRY(event_at(b, *nev, &e));
    CFM_HIP(hipEventRecord(e, b->stream));
    const size_t mine = (*nev)++;
    if (pending >= 0) {
      hipEvent_t pe;
      CFM_TRY(event_at(b, (size_t)pending, &pe));
      CFM_HIP(hipEventSynchronize(pe));
      if (b->host_flags[pending_slot] == 0) break;
    }
    pending = (int)mine;
    pending_slot = slot;
  }
  CFM_HIP(hipStreamSynchronize(b->stream));
  return {};
}

// Norms of `v` per system (status want) into kpart[0][chunk].
Status norms(cfm_batch* b, const double* v, int want) {
  const BK k = bk_of(b);
  hipLaunchKernelGGL(bdots_kernel, dim3(b->nchunk), dim3(256), 0, b->stream, k, v, 1, want);
  CFM_CHECK_LAUNCH();
  return {};
}

}  // namespace

// ---- C ABI ---------------------------------------------------------------------------
extern "C" {

static thread_local std::string g_create_err;  // message of the last failed cfm_batch_create

static int create_fail(int code, const char* msg) {
  g_create_err = msg;
  return code;
}

int cfm_batch_create(const cfm_batch_config* cfg, const int32_t* n_seg, cfm_batch** out) {
  if (!cfg || !n_seg || !out) return create_fail(CFM_EINVAL, "null argument");
  *out = nullptr;
  if (cfg->n_sys <= 0) return create_fail(CFM_EINVAL, "n_sys must be positive");
  if (cfg->seg_dof != 3 && cfg->seg_dof != 6) return create_fail(CFM_EINVAL, "seg_dof must be 3 or 6");
  if (!(cfg->mu > 0.0) || !(cfg->a_seg > 0.0)) return create_fail(CFM_EINVAL, "mu and a_seg must be positive");
  if (cfg->wall && cfg->a_seg != 1.0)
    return create_fail(CFM_EINVAL, "the wall-corrected mobility is defined for a_seg = 1");
  long long total = 0;
  for (int s = 0; s < cfg->n_sys; ++s) {
    if (n_seg[s] <= 0) return create_fail(CFM_EINVAL, "every system needs at least one segment");
    total += n_seg[s];
  }
  if (total * cfg->seg_dof >= (1LL << 31)) return create_fail(CFM_EINVAL, "batch too large");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return create_fail(CFM_EHIP, "no HIP device available");
  if (cfg->device < 0 || cfg->device >= ndev) return create_fail(CFM_EINVAL, "device index out of range");
  cfm_batch* b = new (std::nothrow) cfm_batch();
  if (!b) return CFM_ENOMEM;
  b->cfg = *cfg;
  b->device = cfg->device;
  b->n_sys = cfg->n_sys;
  b->dof = cfg->seg_dof;
  b->N = (int)total;
  b->L = b->dof * b->N;
  b->n_h.assign(n_seg, n_seg + cfg->n_sys);
  b->off_h.resize(b->n_sys);
  std::exclusive_scan(b->n_h.begin(), b->n_h.end(), b->off_h.begin(), 0);
  auto init = [&]() -> Status {
    CFM_HIP(hipSetDevice(b->device));
    CFM_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    CFM_TRY(b->x.alloc(b->N));
    CFM_TRY(b->y.alloc(b->N));
    CFM_TRY(b->z.alloc(b->N));
    CFM_TRY(b->f.alloc(b->L));
    CFM_TRY(b->v.alloc(b->L));
    CFM_TRY(b->sys_of.alloc(b->N));
    CFM_TRY(b->counters.alloc(4));
    CFM_HIP(hipMemsetAsync(b->counters.p, 0, sizeof(int) * 4, b->stream));
    hipLaunchKernelGGL(bfill_kernel, gb(b->L), dim3(256), 0, b->stream, b->f.p, 1.0, b->L);
    CFM_CHECK_LAUNCH();
    std::vector<int> so(b->N);
    for (int s = 0; s < b->n_sys; ++s)
      std::fill(so.begin() + b->off_h[s], so.begin() + b->off_h[s] + b->n_h[s], s);
    CFM_HIP(hipMemcpyAsync(b->sys_of.p, so.data(), sizeof(int) * so.size(), hipMemcpyHostToDevice, b->stream));
    CFM_HIP(hipStreamSynchronize(b->stream));
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess && prop.multiProcessorCount > 0)
      cus = prop.multiProcessorCount;
    b->tuning.max_sources = 2048;
    b->tuning.grid = 4 * cus;
    CFM_TRY(build_items(b));
    return {};
  };
  Status s = init();
  if (!s.ok()) {
    g_create_err = s.msg;
    cfm_batch_destroy(b);
    return s.code;
  }
  *out = b;
  return CFM_OK;
}

int cfm_batch_destroy(cfm_batch* b) {
  if (!b) return CFM_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graphs(b);
  for (auto e : b->ev) (void)hipEventDestroy(e);
  if (b->host) (void)hipHostFree(b->host);
  if (b->host_flags) (void)hipHostFree(b->host_flags);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  for (auto* d : {&b->x, &b->y, &b->z, &b->f, &b->v, &b->partial, &b->Q, &b->H, &b->cs, &b->sn,
                  &b->beta, &b->h, &b->h2, &b->yv, &b->kpart, &b->sinfo, &b->r, &b->w, &b->zt,
                  &b->xs, &b->dv, &b->rhs, &b->otarget})
    d->release();
  for (auto* d : {&b->sys_of, &b->nslice, &b->counters, &b->chunk_begin, &b->iinfo, &b->jdev})
    d->release();
  b->items.release();
  b->chunks.release();
  delete b;
  return CFM_OK;
}

const char* cfm_batch_last_error(const cfm_batch* b) {
  return b ? b->err.c_str() : g_create_err.c_str();
}

int cfm_batch_set_positions(cfm_batch* b, const double* x) {
  if (!b || !x) return CFM_EINVAL;
  const size_t n = b->N;
  for (size_t i = 0; i < 3 * n; ++i)
    if (!std::isfinite(x[i])) return bfail(b, make_status(CFM_EINVAL, "non-finite position"));
  if (b->cfg.wall)
    for (size_t i = 0; i < n; ++i)
      if (!(x[3 * i + 2] > 0.0))
        return bfail(b, make_status(CFM_EINVAL, "particle %zu is not above the wall", i));
  (void)hipSetDevice(b->device);
  std::vector<double> soa(3 * n);
  for (size_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) soa[c * n + i] = x[3 * i + c];
  auto run = [&]() -> Status {
    CFM_TRY(upload(b, soa.data(), b->x.p, n));
    CFM_TRY(upload(b, soa.data() + n, b->y.p, n));
    CFM_TRY(upload(b, soa.data() + 2 * n, b->z.p, n));
    return {};
  };
  Status s = run();
  if (!s.ok()) return bfail(b, s);
  b->have_positions = true;
  return CFM_OK;
}

int cfm_batch_apply(cfm_batch* b, int32_t precision, const double* f, double* v) {
  if (!b || !f || !v) return CFM_EINVAL;
  if (precision != CFM_FP64 && precision != CFM_FP32C)
    return bfail(b, make_status(CFM_EINVAL, "unknown precision %d", precision));
  if (!b->have_positions) return bfail(b, make_status(CFM_ESTATE, "positions not set"));
  (void)hipSetDevice(b->device);
  auto run = [&]() -> Status {
    CFM_TRY(upload(b, f, b->f.p, b->L));
    CFM_TRY(batch_apply(b, precision, b->f.p, nullptr, 0, nullptr, b->v.p));
    return download(b, b->v.p, v, b->L);
  };
  Status s = run();
  return s.ok() ? CFM_OK : bfail(b, s);
}

int cfm_batch_set_tuning(cfm_batch* b, const cfm_batch_tuning* t) {
  if (!b || !t) return CFM_EINVAL;
  if (t->max_sources < 0 || (t->max_sources % kTS) != 0 || t->grid <= 0 || t->grid > 65536)
    return bfail(b, make_status(CFM_EINVAL, "tuning: max_sources must be a multiple of %d, grid in (0, 65536]", kTS));
  (void)hipSetDevice(b->device);
  b->tuning.max_sources = t->max_sources;
  b->tuning.grid = t->grid;
  Status s = build_items(b);
  return s.ok() ? CFM_OK : bfail(b, s);
}

int cfm_batch_get_tuning(cfm_batch* b, cfm_batch_tuning* t) {
  if (!b || !t) return CFM_EINVAL;
  *t = b->tuning;
  return CFM_OK;
}

// Autotuner: times every (max_sources, grid) candidate on the batch's own
// positions with the loads last given to cfm_batch_apply (or unit loads) and keeps
// the fastest.  tried[] (optional) receives each candidate with its time.
int cfm_batch_tune(cfm_batch* b, int32_t precision, cfm_batch_tuning* chosen,
                   cfm_batch_tuning* tried, int32_t max_tried, int32_t* n_tried) {
  if (!b) return CFM_EINVAL;
  if (precision != CFM_FP64 && precision != CFM_FP32C)
    return bfail(b, make_status(CFM_EINVAL, "unknown precision %d", precision));
  if (!b->have_positions) return bfail(b, make_status(CFM_ESTATE, "positions not set"));
  (void)hipSetDevice(b->device);
  int cus = 256;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, b->device) == hipSuccess && prop.multiProcessorCount > 0)
    cus = prop.multiProcessorCount;
  int nmax = 0;
  for (int n : b->n_h) nmax = std::max(nmax, n);
  std::vector<int> srcs = {512, 1024, 2048, 4096, 0};
  std::vector<int> grids = {cus, 2 * cus, 4 * cus, 8 * cus};
  auto run = [&]() -> Status {
    hipEvent_t e0, e1;
    CFM_HIP(hipEventCreate(&e0));
    CFM_HIP(hipEventCreate(&e1));
    cfm_batch_tuning best{};
    best.ms = 1e300;
    int cnt = 0;
    Status st;
    for (int ms : srcs) {
      if (ms > 0 && ms >= nmax && ms != srcs.front()) continue;  // same items as "whole system"
      for (int g : grids) {
        b->tuning.max_sources = ms;
        b->tuning.grid = g;
        st = build_items(b);
        if (!st.ok()) break;
        st = batch_apply(b, precision, b->f.p, nullptr, 0, nullptr, b->v.p);  // warm-up
        if (!st.ok()) break;
        float best_ms = 1e30f;
        for (int rep = 0; rep < 3 && st.ok(); ++rep) {
          if (hipEventRecord(e0, b->stream) != hipSuccess) { st = make_status(CFM_EHIP, "event"); break; }
          st = batch_apply(b, precision, b->f.p, nullptr, 0, nullptr, b->v.p);
          if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            st = make_status(CFM_EHIP, "event");
            break;
          }
          float t = 0.0f;
          (void)hipEventElapsedTime(&t, e0, e1);
          best_ms = std::min(best_ms, t);
        }
        if (!st.ok()) break;
        cfm_batch_tuning c{};
        c.max_sources = ms;
        c.grid = g;
        c.ms = best_ms;
        if (tried && cnt < max_tried) tried[cnt] = c;
        ++cnt;
        if (c.ms < best.ms) best = c;
      }
      if (!st.ok()) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    CFM_TRY(st);
    if (n_tried) *n_tried = cnt;
    b->tuning = best;
    CFM_TRY(build_items(b));
    if (chosen) *chosen = best;
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : bfail(b, s);
}

// All systems solved together: M_s F_s = rhs_s for every s (segment resistance
// problem of each system), right-preconditioned restarted GMRES in lockstep.
int cfm_batch_gmres(cfm_batch* b, const double* rhs, double* xout, const cfm_solve_opts* o,
                    cfm_batch_sys_stats* per_sys, double* total_ms) {
  if (!b || !rhs || !xout || !o) return CFM_EINVAL;
  if (!b->have_positions) return bfail(b, make_status(CFM_ESTATE, "positions not set"));
  if (o->precond < CFM_PC_NONE || o->precond > CFM_PC_SELF)
    return bfail(b, make_status(CFM_EINVAL, "batched solves take CFM_PC_NONE, _DIAG or _SELF"));
  if (o->mode != CFM_SOLVE_FP64 && o->mode != CFM_SOLVE_MIXED)
    return bfail(b, make_status(CFM_EINVAL, "unknown solve mode %d", o->mode));
  if (!(o->rtol > 0.0)) return bfail(b, make_status(CFM_EINVAL, "rtol must be positive"));
  for (size_t i = 0; i < (size_t)b->L; ++i)
    if (!std::isfinite(rhs[i])) return bfail(b, make_status(CFM_EINVAL, "non-finite right-hand side"));
  (void)hipSetDevice(b->device);
  const int restart = o->restart > 0 ? o->restart : 50;
  const int max_iter = o->max_iter > 0 ? o->max_iter : 1000;
  const int pc = o->precond;
  const bool mixed = o->mode == CFM_SOLVE_MIXED;
  double elapsed = 0.0;
  auto solve = [&]() -> Status {
    CFM_TRY(reserve_krylov(b, restart));
    const BK k = bk_of(b);
    CFM_TRY(upload(b, rhs, b->rhs.p, b->L));
    double* x = b->xs.p;
    if (o->use_guess)
      CFM_TRY(upload(b, xout, x, b->L));
    else
      hipLaunchKernelGGL(bfill_kernel, gb(b->L), dim3(256), 0, b->stream, x, 0.0, b->L);
    hipLaunchKernelGGL(bfill_kernel, gb(b->L), dim3(256), 0, b->stream, b->dv.p, 0.0, b->L);
    CFM_CHECK_LAUNCH();
    hipGraphExec_t g = nullptr;
    CFM_TRY(step_graph(b, mixed ? CFM_FP32C : CFM_FP64, pc, mixed ? b->dv.p : x, &g));
    CFM_HIP(hipMemsetAsync(b->jdev.p, 0, sizeof(int) * 4, b->stream));
    // every system: status ST_DONE for the norm of b (bdots filters on status)
    CFM_HIP(hipMemsetAsync(b->iinfo.p, 0, sizeof(int) * 8 * b->n_sys, b->stream));
    size_t nev = 0;
    hipEvent_t e0;
    CFM_TRY(event_at(b, nev++, &e0));
    CFM_HIP(hipEventRecord(e0, b->stream));
    CFM_TRY(norms(b, b->rhs.p, ST_DONE));
    hipLaunchKernelGGL(bsetup_kernel, dim3(1), dim3(256), 0, b->stream, k, o->rtol);
    CFM_CHECK_LAUNCH();
    if (mixed)
      CFM_HIP(hipMemcpy2DAsync(b->otarget.p, sizeof(double), b->sinfo.p + SI_TARGET, 8 * sizeof(double),
                               sizeof(double), b->n_sys, hipMemcpyDeviceToDevice, b->stream));
    for (int pass = 0; pass < 4 * (max_iter / std::max(1, restart) + 2); ++pass) {
      // FP64 residual of the systems to check: r = b - A x
      CFM_TRY(batch_apply(b, CFM_FP64, x, b->iinfo.p, ST_CHECK, b->rhs.p, b->r.p));
      CFM_TRY(norms(b, b->r.p, ST_CHECK));
      if (mixed)
        hipLaunchKernelGGL(bouter_kernel, dim3(1), dim3(256), 0, b->stream, k, b->otarget.p, max_iter);
      else
        hipLaunchKernelGGL(brestart_kernel, dim3(1), dim3(256), 0, b->stream, k, max_iter);
      hipLaunchKernelGGL(bq0_kernel, gb(b->L), dim3(256), 0, b->stream, k, b->r.p);
      CFM_CHECK_LAUNCH();
      int jd[4] = {0, 0, 0, 0};
      CFM_HIP(hipMemcpyAsync(b->host_flags, b->jdev.p, sizeof(int) * 4, hipMemcpyDeviceToHost, b->stream));
      CFM_HIP(hipStreamSynchronize(b->stream));
      std::memcpy(jd, b->host_flags, sizeof(jd));
      if (jd[2] == 0) break;  // nobody left running
      CFM_TRY(run_cycle(b, g, &nev, b->m));
      if (mixed) {
        // x += d for the systems whose inner solve ended (status ST_CHECK)
        hipLaunchKernelGGL(badd_kernel, gb(b->L), dim3(256), 0, b->stream, k, x, b->dv.p);
        CFM_CHECK_LAUNCH();
      }
    }
    hipEvent_t eend;
    CFM_TRY(event_at(b, nev++, &eend));
    CFM_HIP(hipEventRecord(eend, b->stream));
    CFM_HIP(hipEventSynchronize(eend));
    float ms = 0.0f;
    CFM_HIP(hipEventElapsedTime(&ms, e0, eend));
    elapsed = ms;
    CFM_TRY(download(b, x, xout, b->L));
    // per-system results
    std::vector<double> si((size_t)8 * b->n_sys);
    std::vector<int> ii((size_t)8 * b->n_sys);
    CFM_HIP(copy_to_host(b->stream, si.data(), b->sinfo.p, sizeof(double) * si.size()));
    CFM_HIP(copy_to_host(b->stream, ii.data(), b->iinfo.p, sizeof(int) * ii.size()));
    bool all = true;
    for (int s = 0; s < b->n_sys; ++s) {
      const double nb = si[8 * s + SI_NB];
      const double rel = nb > 0.0 ? si[8 * s + SI_NR] / nb : 0.0;
      const bool conv = rel <= o->rtol;
      all = all && conv;
      if (per_sys) {
        cfm_batch_sys_stats& P = per_sys[s];
        P.iterations = ii[8 * s + II_ITERS];
        P.converged = conv ? 1 : 0;
        P.rel_residual = rel;
        // time to the end of the graph step at which the system's residual estimate
        // last met its target (its solution was final from then on)
        const int step = ii[8 * s + II_DONE_AT];
        float t = ms;
        // events [1, nev-1) are the graph steps in order; a system that never took
        // a step reports the whole call
        if (step >= 0 && (size_t)step + 1 < nev - 1) {
          if (hipEventElapsedTime(&t, e0, b->ev[(size_t)step + 1]) != hipSuccess) t = ms;
        }
        P.latency_ms = t;
        P.reserved0 = 0;
        P.reserved1 = 0;
      }
    }
    if (!all) return make_status(CFM_ENOCONV, "batched GMRES: a system stopped above rtol %.3e", o->rtol);
    return {};
  };
  Status s = solve();
  if (total_ms) *total_ms = elapsed;
  return s.ok() ? CFM_OK : bfail(b, s);
}

}  // extern "C"
