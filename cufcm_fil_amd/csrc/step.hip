// step.hip -- implicit (BDF + Broyden) filament step driven on the device.
//
// Host side restates the time-step loop of src/cilia/cilia_sim_main.cu:405-552:
//   initial guess -> loads -> velocities -> residual; if too large: J0^-1, then
//   Broyden iterations (find_update, update, loads, velocities, residual,
//   end_of_iter) until the residual meets TOL.
// Velocities follow mobility_solver::compute_velocities (mobility_solver.cpp:2673-2758)
// for a body held fixed: v_seg = Mss f_seg; v_blob = Mbs f_seg; GMRES for the blob
// forces with Mbb f_blob = -v_blob; v_seg += Msb f_blob.  The Broyden low-rank
// matrices C, D (broyden_solver.cpp:54-174) live in HBM as max_broyden columns of
// length 6 n_seg; every vector operation runs on the GPU and only the scaled
// residual norm (and GMRES residual estimates) come back to the host.
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>

#include "ctx.hpp"
#include "body.hpp"
#include "fil.hpp"
#include "swim.hpp"
#include "krylov.hpp"

namespace cfm {

Status launch_apply(cfm_ctx* c, const ApplyPlan& plan);
int choose_split(cfm_ctx* c, const ApplyPlan& pl);
ApplyPlan plan_for(const cfm_ctx* c, int which);
Status side_fork(cfm_ctx* c);
Status side_join(cfm_ctx* c);
Status upload_aos(cfm_ctx* c, const double* host, int cnt, int dof, double* dst, int ld, int off);
Status download_aos(cfm_ctx* c, double* host, int cnt, int dof, const double* src, int ld, int off);
Status gnn_guess(cfm_ctx* c, const double* rhs, double* x0, bool* used);
Status sym_members_segments(cfm_ctx* c, int members, const int* active);
Status gnn_record(cfm_ctx* c, const double* b, const double* x);

// Ring of past (rhs, solution) pairs of the blob system and, on the device, the Gram
// matrix of the stored right-hand sides (H x H, row-major).  The guess minimises
// ||b - sum_i c_i b_i|| (with a tiny Tikhonov term) and returns sum_i c_i x_i.
// Because A x_i = b_i to the solver tolerance, its residual is the projection residual
// of b.  The coefficients are solved on the device, so the guess costs the host no wait.
struct History {
  static constexpr int H = 16;
  int L = 0, count = 0, head = 0;
  DevBuf<double> B, X, G, g, coef, partial;
  DevBuf<int> used;  // 1 when the last guess replaced x0
};

// Solves (G + d I) c = g, d = 1e-13 tr(G) / k, for the k <= H coefficients of the
// projected guess: Gaussian elimination with partial pivoting (the first row holding
// the largest magnitude), then column-oriented back substitution.  *used = 0,
// coefficients untouched, when a pivot is not positive in magnitude or a coefficient
// is not finite.  One wave: the augmented matrix lives in LDS, lane 0 searches the
// pivot, the lanes share the row updates; every operation has a fixed order.
__global__ __launch_bounds__(64) void history_solve_kernel(const double* __restrict__ G,
                                                           const double* __restrict__ g, int k,
                                                           double* __restrict__ coef,
                                                           int* __restrict__ used) {
  constexpr int H = History::H;
  constexpr int W = H + 1;
  __shared__ double A[H][W];
  __shared__ double fr[H];
  __shared__ double dshift;
  __shared__ int piv, ok;
  const int t = threadIdx.x;
  // every lane issues its loads before any is used (device-memory latency paid once)
  double ld[(H * W + 63) / 64];
#pragma unroll
  for (int q = 0; q < (H * W + 63) / 64; ++q) {
    const int e = t + 64 * q, i = e / W, j = e % W;
    ld[q] = (e < H * W && i < k && (j < k || j == H)) ? (j < k ? G[i * H + j] : g[i]) : 0.0;
  }
#pragma unroll
  for (int q = 0; q < (H * W + 63) / 64; ++q) {
    const int e = t + 64 * q;
    if (e < H * W) A[e / W][e % W] = ld[q];
  }
  __syncthreads();
  if (t == 0) {
    double tr = 0.0;
    for (int i = 0; i < k; ++i) tr += A[i][i];
    dshift = 1e-13 * tr / k;
  }
  __syncthreads();
  if (t < k) A[t][t] += dshift;
  __syncthreads();
  for (int col = 0; col < k; ++col) {
    if (t == 0) {
      int p = col;
      double best = fabs(A[col][col]);
      for (int r = col + 1; r < k; ++r) {
        const double v = fabs(A[r][col]);
        if (v > best) {
          best = v;
          p = r;
        }
      }
      piv = p;
      ok = best > 0.0;
    }
    __syncthreads();
    if (!ok) {
      if (t == 0) *used = 0;
      return;
    }
    const int p = piv;
    if (p != col)
      for (int j = t; j < W; j += 64) {  // columns col..k and the right-hand side (H)
        const double a = A[p][j];
        A[p][j] = A[col][j];
        A[col][j] = a;
      }
    __syncthreads();
    if (t > col && t < k) fr[t] = A[t][col] / A[col][col];
    __syncthreads();
    const int w = k - col;  // columns col..k-1, then the right-hand side
    for (int e = t; e < (k - 1 - col) * (w + 1); e += 64) {
      const int r = col + 1 + e / (w + 1), jj = e % (w + 1);
      const int j = jj < w ? col + jj : H;
      A[r][j] -= fr[r] * A[col][j];
    }
    __syncthreads();
  }
  // back substitution by columns: x_i = b_i / A_ii, then b_r -= A_ri x_i for r < i
  for (int i = k - 1; i >= 0; --i) {
    if (t == 0) fr[i] = A[i][H] / A[i][i];
    __syncthreads();
    if (t < i) A[t][H] -= A[t][i] * fr[i];
    __syncthreads();
  }
  if (t != 0) return;
  for (int i = 0; i < k; ++i)
    if (!isfinite(fr[i])) {
      *used = 0;
      return;
    }
  for (int i = 0; i < k; ++i) coef[i] = fr[i];
  *used = 1;
}

// out[r] = sum_{i<k} y[i] Q[i][r] when *flag is set; out untouched otherwise.
__global__ void gemv_if_kernel(const double* __restrict__ Q, size_t ldq, int k,
                               const double* __restrict__ y, const int* __restrict__ flag,
                               double* __restrict__ out, int L) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L || !flag[0]) return;
  double s = 0.0;
  for (int i = 0; i < k; ++i) s += y[i] * Q[(size_t)i * ldq + r];
  out[r] = s;
}

// dst_a = a, dst_b = b (length L each)
__global__ void copy2_kernel(const double* __restrict__ a, double* __restrict__ dst_a,
                             const double* __restrict__ b, double* __restrict__ dst_b, int L) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L) return;
  dst_a[r] = a[r];
  dst_b[r] = b[r];
}

// out = x (n[0] - s[0]) (the scalars on the device)
__global__ void scale_by_diff_kernel(const double* __restrict__ x, const double* __restrict__ n,
                                     const double* __restrict__ s, double* __restrict__ out, int L) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < L) out[r] = x[r] * (n[0] - s[0]);
}

// out[r] = x[r] (n[m] - s[m]), m = r / Lm
__global__ void member_scale_by_diff_kernel(const double* __restrict__ x, const double* __restrict__ n,
                                            const double* __restrict__ s, double* __restrict__ out,
                                            int L, int Lm) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < L) out[r] = x[r] * (n[r / Lm] - s[r / Lm]);
}

static Status history_reset(History& h, int L) {
  h.L = L;
  h.count = h.head = 0;
  CFM_TRY(h.B.alloc((size_t)History::H * L));
  CFM_TRY(h.X.alloc((size_t)History::H * L));
  CFM_TRY(h.G.alloc((size_t)History::H * History::H));
  CFM_TRY(h.g.alloc(History::H));
  CFM_TRY(h.coef.alloc(History::H));
  CFM_TRY(h.used.alloc(1));
  return h.partial.alloc(dev_dots_partial_size(L, History::H));
}

// x0 = the projected guess once two solves are stored (*queued = true); x0 keeps its
// value (the previous solution) when the device finds the Gram system singular.
static Status history_guess(cfm_ctx* c, History& h, const double* b, double* x0, bool* queued) {
  *queued = false;
  if (h.count < 2) return {};
  const int k = h.count;
  CFM_TRY(dev_dots(c, h.partial.p, h.B.p, (size_t)h.L, k, b, h.L, h.g.p));
  hipLaunchKernelGGL(history_solve_kernel, dim3(1), dim3(64), 0, c->stream, h.G.p, h.g.p, k,
                     h.coef.p, h.used.p);
  CFM_CHECK_LAUNCH();
  hipLaunchKernelGGL(gemv_if_kernel, dim3((h.L + 255) / 256), dim3(256), 0, c->stream, h.X.p,
                     (size_t)h.L, k, h.coef.p, h.used.p, x0, h.L);
  CFM_CHECK_LAUNCH();
  *queued = true;
  return {};
}

// Stores (b, x) in the next slot and its row and column of the Gram matrix.
static Status history_push(cfm_ctx* c, History& h, const double* b, const double* x) {
  const int slot = h.head;
  hipLaunchKernelGGL(copy2_kernel, dim3((h.L + 255) / 256), dim3(256), 0, c->stream, b,
                     h.B.p + (size_t)slot * h.L, x, h.X.p + (size_t)slot * h.L, h.L);
  CFM_CHECK_LAUNCH();
  h.head = (h.head + 1) % History::H;
  h.count = std::min(h.count + 1, History::H);
  return dev_dots2(c, h.partial.p, h.B.p, (size_t)h.L, h.count, h.B.p + (size_t)slot * h.L, h.L,
                   h.G.p + (size_t)slot * History::H, h.G.p + slot, History::H);
}

struct FilState {
  cfm_fil_params prm{};
  FilK K;
  int nbroy = 0;
  DevBuf<double> q, qm1, u, um1, xm1, xm2, lam, lam_m1, lam_m2, clamp, clamp_m1, clamp_m2, base,
      pert, fint, jinv;
  DevBuf<double> err, nerr, upd, C, D, coef, partial, bad_part, scal;
  DevBuf<double> fb_prev, rhs_b;  // blob-force guess and blob rhs
  BarrierGrid grid;               // barrier cell list
  History hist;                   // past blob solves (projected guess)
  DevBuf<int> status;
  bool have_jinv = false;
  double* host = nullptr;  // pinned read-back
  // free swimmers (cfm_fil_attach_bodies): body state among the Broyden unknowns
  bool swim = false;
  SwimK S;
  DevBuf<double> sw_state, sw_fref, sw_fext, sw_mob, sw_BJ, sw_Sf, sw_clamp, sw_rep, sw_tmp;
  // blob forces among the Broyden unknowns (broyden_only): [3][n_blob] and two previous steps
  bool bx = false;
  BxK X;
  int nt = 0;  // the step in progress
  DevBuf<double> bx_fb, bx_fb_m1, bx_fb_m2;
  // equal independent systems stepped together (cfm_fil_set_members): per member the
  // Broyden badness, low-rank coefficients (members x (max_broyden + 1)), norms and the
  // active flag (1 while its iteration continues)
  int members = 1;
  DevBuf<double> m_bad, m_coef, m_norm, m_dot, m_partial;
  DevBuf<int> m_active;
  double* m_host = nullptr;  // pinned read-back of the members' badness
  int* m_host_flags = nullptr;  // pinned staging of the active flags
  std::vector<cfm_step_stats> m_stats;
  // the approximate Jacobian built on the side stream (jacobian_fork), joined by ev_jac
  hipEvent_t ev_jac = nullptr;
  hipEvent_t ev_res = nullptr;  // the Broyden residual's read-back
  bool stepped = false;  // a step of this state has completed
};

static Status ensure_host(FilState* F) {
  if (!F->host) CFM_HIP(hipHostMalloc((void**)&F->host, sizeof(double) * 8, hipHostMallocDefault));
  return {};
}

static Status read_dev(cfm_ctx* c, FilState* F, const double* dev, double* out) {
  CFM_TRY(ensure_host(F));
  CFM_HIP(hipMemcpyAsync(F->host, dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CFM_HIP(hipStreamSynchronize(c->stream));
  *out = F->host[0];
  return {};
}

static Status apply_plan(cfm_ctx* c, const ApplyPlan& pl, int prec) {
  const int saved = c->cfg.precision;
  c->cfg.precision = prec;
  c->split = choose_split(c, pl);
  Status s = launch_apply(c, pl);
  c->cfg.precision = saved;
  return s;
}

// v = M f for the current loads (segments from the filament loads, blob forces by GMRES)
static Status swim_velocities(cfm_ctx* c, FilState* F, cfm_step_stats* st);

// Blob forces among the Broyden unknowns: every product from the current loads
// (mobility_solver.cpp:2697-2730); with free bodies also the body loads of the body rows.
static Status bx_velocities(cfm_ctx* c, FilState* F) {
  Particles& p = c->p;
  const double* rep = nullptr;
  if (F->prm.barrier) {
    CFM_TRY(fil_barrier(c, F->K, F->grid));
    if (F->swim) {
      CFM_TRY(fil_blob_barrier(c, F->K, F->grid, F->S.nb, F->sw_rep.p));
      rep = F->sw_rep.p;
    }
  }
  if (F->swim) CFM_TRY(swim_force(c, F->S, F->K, rep, body_offsets(c)));
  CFM_TRY(dev_copy_rows(c, F->X.fb, (size_t)p.n_blob, 0, p.f.p, (size_t)p.n, p.n_seg, 3, p.n_blob));
  ApplyPlan all;
  all.tgt_begin = 0;
  all.tgt_end = p.n;
  all.src_seg = true;
  all.src_blob = true;
  return apply_plan(c, all, F->prm.pair_precision);
}

static Status velocities(cfm_ctx* c, FilState* F, cfm_step_stats* st) {
  if (F->bx) return bx_velocities(c, F);
  if (F->swim) return swim_velocities(c, F, st);
  Particles& p = c->p;
  const int prec = F->prm.pair_precision;
  if (F->prm.barrier) CFM_TRY(fil_barrier(c, F->K, F->grid));
  if (p.n_blob == 0) {
    ApplyPlan pl;
    pl.tgt_begin = 0;
    pl.tgt_end = p.n_seg;
    pl.src_seg = true;
    return apply_plan(c, pl, prec);
  }
  // With overlap, the blob solve starts from the blob rows of M f_seg alone while the
  // segment rows (the largest product) run on the side stream, joined before
  // v_seg += Msb f_b; the two touch disjoint rows of v, and the blob solve reads and
  // writes only blob rows of f and its own vectors.  Every product's source split is
  // then tuned (first call) on the main stream before the fork, never beside the side
  // launch.
  const SystemDesc s = system_desc(c, CFM_SYS_BLOB);
  ApplyPlan bs;
  bs.tgt_begin = p.n_seg;
  bs.tgt_end = p.n;
  bs.src_seg = true;
  ApplyPlan ss;
  ss.tgt_begin = 0;
  ss.tgt_end = p.n_seg;
  ss.src_seg = true;
  ss.side = true;
  if (F->prm.overlap) {
    const int saved = c->cfg.precision;
    c->cfg.precision = prec;
    (void)choose_split(c, bs);
    (void)choose_split(c, ss);
    c->cfg.precision = CFM_FP64;
    (void)choose_split(c, plan_for(c, CFM_BB));
    if (F->prm.blob_solve.mode == CFM_SOLVE_MIXED) {
      c->cfg.precision = CFM_FP32C;
      (void)choose_split(c, plan_for(c, CFM_BB));
    }
    c->cfg.precision = saved;
  }
  const bool overlap = F->prm.overlap != 0;
  if (overlap) CFM_TRY(side_fork(c));
  // Without overlap the segment rows M_ss f_seg (the largest product) are queued on the
  // main stream behind the blob solve's first read-back, so the device computes them
  // while the host waits for that read; they touch only segment rows of v.
  struct SegRows {
    ApplyPlan pl;
    int prec = 0;
    bool queued = false;
  } seg_rows{ss, prec, false};
  seg_rows.pl.side = false;
  auto queue_seg_rows = [](cfm_ctx* cc, void* u) -> Status {
    SegRows* sr = static_cast<SegRows*>(u);
    if (sr->queued) return {};
    sr->queued = true;
    return apply_plan(cc, sr->pl, sr->prec);
  };
  GmresHooks hooks;
  if (!overlap) {
    hooks.before_first_wait = queue_seg_rows;
    hooks.wait_user = &seg_rows;
  }
  auto blob_part = [&]() -> Status {
    CFM_TRY(apply_plan(c, bs, prec));  // blob rows from the segment loads
    if (overlap) CFM_TRY(apply_plan(c, ss, prec));
    // blob forces: Mbb f_b = -v_b (body held fixed, so K V_body = 0)
    CFM_TRY(F->rhs_b.alloc(s.L));
    CFM_TRY(F->fb_prev.alloc(s.L));
    CFM_TRY(dev_copy_rows(c, p.v.p, (size_t)p.n, p.n_seg, F->rhs_b.p, (size_t)s.m, 0, 3, s.m));
    CFM_TRY(dev_axpby(c, -1.0, F->rhs_b.p, 0.0, F->rhs_b.p, F->rhs_b.p, s.L));
    if (F->prm.guess == CFM_GUESS_PROJECTED || F->prm.guess == CFM_GUESS_LEARNED) {
      History& h = F->hist;
      if (h.L != s.L) CFM_TRY(history_reset(h, s.L));
      bool used = false;
      CFM_TRY(history_guess(c, h, F->rhs_b.p, F->fb_prev.p, &used));
      // cold start (fewer than two stored solves): the network's guess
      if (!used && F->prm.guess == CFM_GUESS_LEARNED) CFM_TRY(gnn_guess(c, F->rhs_b.p, F->fb_prev.p, &used));
    }
    cfm_solve_opts o = F->prm.blob_solve;
    o.system = CFM_SYS_BLOB;
    o.use_guess = 1;
    if (!c->krylov) c->krylov = new Krylov();
    cfm_solve_stats ss_{};
    CFM_TRY(solve(c, *c->krylov, s, o, F->rhs_b.p, F->fb_prev.p, &ss_, overlap ? nullptr : &hooks));
    if (!overlap) CFM_TRY(queue_seg_rows(c, &seg_rows));  // already queued by the solve
    if (F->prm.guess == CFM_GUESS_PROJECTED || F->prm.guess == CFM_GUESS_LEARNED)
      CFM_TRY(history_push(c, F->hist, F->rhs_b.p, F->fb_prev.p));
    CFM_TRY(gnn_record(c, F->rhs_b.p, F->fb_prev.p));
    if (st) {
      st->gmres_iterations += ss_.iterations;
      st->gmres_solves += 1;
      st->max_gmres_rel_residual = std::max(st->max_gmres_rel_residual, ss_.rel_residual);
    }
    return dev_copy_rows(c, F->fb_prev.p, (size_t)s.m, 0, p.f.p, (size_t)p.n, p.n_seg, 3, s.m);
  };
  const Status sb_st = blob_part();
  const Status sj = overlap ? side_join(c) : Status{};  // the side stream has finished M_ss f_seg
  if (!sb_st.ok()) return sb_st;
  CFM_TRY(sj);
  // v_seg += Msb f_b
  ApplyPlan sb;
  sb.tgt_begin = 0;
  sb.tgt_end = p.n_seg;
  sb.src_blob = true;
  sb.acc_seg = true;
  return apply_plan(c, sb, prec);
}

// Free swimmer: the blob system carries the body velocities as unknowns and the
// body load (swimmer.cpp:470-535 plus the blob repulsion of mobility_solver.cpp:215-236)
// as its last rows (mobility_solver.cpp:455-632, 2673-2758).
static Status swim_velocities(cfm_ctx* c, FilState* F, cfm_step_stats* st) {
  Particles& p = c->p;
  const int prec = F->prm.pair_precision;
  const double* rep = nullptr;
  if (F->prm.barrier) {
    CFM_TRY(fil_barrier(c, F->K, F->grid));
    CFM_TRY(fil_blob_barrier(c, F->K, F->grid, F->S.nb, F->sw_rep.p));
    rep = F->sw_rep.p;
  }
  SystemDesc s;
  CFM_TRY(body_system_desc(c, &s));
  CFM_TRY(swim_force(c, F->S, F->K, rep, body_offsets(c)));
  ApplyPlan pl;
  pl.tgt_begin = 0;
  pl.tgt_end = p.n;
  pl.src_seg = true;
  CFM_TRY(apply_plan(c, pl, prec));
  // rhs: minus the blob velocities the segments induce, minus the body load
  CFM_TRY(F->rhs_b.alloc(s.L));
  CFM_TRY(dev_copy_rows(c, p.v.p, (size_t)p.n, p.n_seg, F->rhs_b.p, (size_t)s.m, 0, 3, s.m));
  CFM_HIP(hipMemcpyAsync(F->rhs_b.p + 3 * (size_t)s.m, F->S.force, sizeof(double) * s.extra,
                         hipMemcpyDeviceToDevice, c->stream));
  CFM_TRY(dev_axpby(c, -1.0, F->rhs_b.p, 0.0, F->rhs_b.p, F->rhs_b.p, s.L));
  CFM_TRY(body_prescribed_rhs(c, F->rhs_b.p));  // prescribed rotation: no torque balance
  cfm_solve_opts o = F->prm.blob_solve;
  o.system = CFM_SYS_BODY;
  o.use_guess = 1;  // previous blob forces and body velocities
  if (!c->krylov) c->krylov = new Krylov();
  cfm_solve_stats ss{};
  CFM_TRY(solve(c, *c->krylov, s, o, F->rhs_b.p, F->fb_prev.p, &ss, nullptr));
  if (st) {
    st->gmres_iterations += ss.iterations;
    st->gmres_solves += 1;
    st->max_gmres_rel_residual = std::max(st->max_gmres_rel_residual, ss.rel_residual);
  }
  CFM_HIP(hipMemcpyAsync(F->S.vel, F->fb_prev.p + 3 * (size_t)s.m, sizeof(double) * s.extra,
                         hipMemcpyDeviceToDevice, c->stream));
  // v_seg += Msb f_b, after the segment rows of the side stream
  CFM_TRY(dev_copy_rows(c, F->fb_prev.p, (size_t)s.m, 0, p.f.p, (size_t)p.n, p.n_seg, 3, s.m));
  CFM_TRY(side_join(c));
  ApplyPlan sb;
  sb.tgt_begin = 0;
  sb.tgt_end = p.n_seg;
  sb.src_blob = true;
  sb.acc_seg = true;
  return apply_plan(c, sb, prec);
}

// The residual e of the Broyden rows and its badness read to *bad.  before_wait, when
// given, is queued after the read-back and before the host waits for it (work that the
// outcome may make useless but never wrong).
static Status residual(cfm_ctx* c, FilState* F, int nt, double* e, double* bad,
                       const std::function<Status()>* before_wait = nullptr) {
  int extra = 0;
  if (F->bx) {
    const int nb = (c->p.n_seg + 255) / 256;
    const size_t L0 = 6 * (size_t)c->p.n_seg;
    CFM_TRY(bx_errors(c, F->X, F->S, nt, e + L0, e + L0 + 6 * (size_t)F->X.nbody, F->bad_part.p + nb));
    extra = F->X.nbody + (F->X.m + 255) / 256;
  } else if (F->swim) {
    const int nb = (c->p.n_seg + 255) / 256;
    CFM_TRY(swim_errors(c, F->S, nt, e + 6 * (size_t)c->p.n_seg, F->bad_part.p + nb));
    extra = F->S.nbody;
  }
  CFM_TRY(fil_errors(c, F->K, nt, e, F->bad_part.p, F->scal.p, extra));
  if (!before_wait) return read_dev(c, F, F->scal.p, bad);
  CFM_TRY(ensure_host(F));
  if (!F->ev_res) CFM_HIP(hipEventCreateWithFlags(&F->ev_res, hipEventDisableTiming));
  CFM_HIP(hipMemcpyAsync(F->host, F->scal.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CFM_HIP(hipEventRecord(F->ev_res, c->stream));
  CFM_TRY((*before_wait)());
  CFM_HIP(hipEventSynchronize(F->ev_res));
  *bad = F->host[0];
  return {};
}

// out = scale J0^-1 in (block diagonal over filaments, plus the body Schur complement)
static Status jinv_mult(cfm_ctx* c, FilState* F, const double* in, double* out, double scale) {
  if (F->bx) return bx_jinv_mult(c, F->X, F->S, F->K, F->nt, in, out, scale, F->sw_tmp.p);
  if (F->swim) return swim_jinv_mult(c, F->S, F->K, in, out, scale, F->sw_tmp.p);
  return fil_jinv_mult(c, F->K, in, out, scale);
}

static Status jacobian(cfm_ctx* c, FilState* F, int nt) {
  CFM_HIP(hipMemsetAsync(F->status.p, 0, sizeof(int), c->stream));
  CFM_TRY(fil_jacobian(c, F->K, nt, F->status.p));
  if (F->bx) return bx_jacobian(c, F->X, F->S, F->K, nt, F->status.p);
  if (F->swim) CFM_TRY(swim_jacobian(c, F->S, F->K, nt, F->status.p));
  return {};
}

// The approximate Jacobian depends only on the state after the initial guess and its
// internal loads, so for plain filaments (no free bodies, no blob forces among the
// unknowns) it is built and inverted on the side stream while the main stream evaluates
// the first velocities and residual: the inversion is bound by its per-pivot latency and
// leaves most issue slots to the pair kernels beside it.  The main stream waits for
// ev_jac before J0^-1 is first used (and before the step ends when no iteration is
// needed).  Not on the first step of a state, whose products tune their source splits
// by timing launches, which a concurrent launch would disturb.
static Status jacobian_fork(cfm_ctx* c, FilState* F, int nt, bool* forked) {
  *forked = false;
  if (F->swim || F->bx || !F->stepped) return {};
  if (!F->ev_jac) CFM_HIP(hipEventCreateWithFlags(&F->ev_jac, hipEventDisableTiming));
  CFM_TRY(side_fork(c));
  CFM_HIP(hipMemsetAsync(F->status.p, 0, sizeof(int), c->side));
  CFM_TRY(fil_jacobian(c, F->K, nt, F->status.p, c->side));
  CFM_HIP(hipEventRecord(F->ev_jac, c->side));
  *forked = true;
  return {};
}

static Status update(cfm_ctx* c, FilState* F, const double* upd) {
  if (F->swim) CFM_TRY(swim_update(c, F->S, upd + 6 * (size_t)c->p.n_seg));
  if (F->bx) CFM_TRY(bx_update(c, F->X, upd + 6 * ((size_t)c->p.n_seg + F->X.nbody)));
  CFM_TRY(fil_update(c, F->K, upd));
  if (F->swim) CFM_TRY(body_place_device(c, F->S.x, F->S.q));
  return {};
}

static Status initial_guess(cfm_ctx* c, FilState* F, int nt) {
  if (F->swim) CFM_TRY(swim_guess(c, F->S, nt));
  if (F->bx) CFM_TRY(bx_guess(c, F->X, nt));
  CFM_TRY(fil_initial_guess(c, F->K, nt));
  if (F->swim) CFM_TRY(body_place_device(c, F->S.x, F->S.q));
  return {};
}

// out = -gamma J0^-1 e - sum_i (e . D_i) C_i over the first `it` columns
static Status low_rank_apply(cfm_ctx* c, FilState* F, const double* e, double* out, int it) {
  const int L = F->nbroy;
  CFM_TRY(jinv_mult(c, F, e, out, -F->prm.gamma));
  if (it > 0) {
    CFM_TRY(dev_dots(c, F->partial.p, F->D.p, (size_t)L, it, e, L, F->coef.p));
    CFM_TRY(dev_gemv_sub(c, F->C.p, (size_t)L, it, F->coef.p, out, L));
  }
  return {};
}

static Status step(cfm_ctx* c, FilState* F, int nt, cfm_step_stats* st) {
  const auto t0 = std::chrono::steady_clock::now();
  if (F->bx && F->swim && body_rotation_prescribed(c))
    return make_status(CFM_EINVAL, "broyden_only takes free bodies, not a prescribed rotation");
  F->nt = nt;
  cfm_step_stats S{};
  const int L = F->nbroy;
  const int maxit = F->prm.max_broyden;
  CFM_TRY(initial_guess(c, F, nt));
  CFM_TRY(fil_forces(c, F->K, nt));
  bool jac_side = false;
  CFM_TRY(jacobian_fork(c, F, nt, &jac_side));
  CFM_TRY(velocities(c, F, &S));
  double bad = 0.0;
  CFM_TRY(residual(c, F, nt, F->err.p, &bad));
  if (jac_side) CFM_HIP(hipStreamWaitEvent(c->stream, F->ev_jac, 0));
  if (!std::isfinite(bad)) return make_status(CFM_ENONFINITE, "non-finite residual at step %d", nt);
  bool too_large = bad > 1.0;
  if (too_large && !jac_side) CFM_TRY(jacobian(c, F, nt));
  int it = 0;
  bool upd_ready = false;  // the update was formed with the last low-rank column
  // The low-rank pair of iteration `it` and the next update, from the new error
  // (broyden_solver.cpp:69-125).  Queued before the host reads the residual: when the
  // step has converged they are not used (the next step starts from column 0).
  const std::function<Status()> next_column = [&]() -> Status {
    double* Ci = F->C.p + (size_t)it * L;
    double* Di = F->D.p + (size_t)it * L;
    CFM_TRY(low_rank_apply(c, F, F->nerr.p, Ci, it));
    CFM_TRY(dev_axpby(c, 1.0, F->nerr.p, -1.0, F->err.p, Di, L));
    CFM_TRY(dev_norm2(c, F->partial.p, Di, L, F->scal.p + 1));
    CFM_TRY(dev_scale_by_inv(c, Ci, F->scal.p + 1, Ci, L));
    CFM_TRY(dev_scale_by_inv(c, Di, F->scal.p + 1, Di, L));
    // The next update -gamma J0^-1 e - sum_{i<=it} (e . D_i) C_i (e = the new error) is
    // C_it's unnormalised column minus its own last term: C_it (n - e . D_it) with
    // n = |e - e_old|.  Formed here, it saves a J0^-1 product and it dot products.
    CFM_TRY(dev_dots(c, F->partial.p, Di, (size_t)L, 1, F->nerr.p, L, F->scal.p + 2));
    hipLaunchKernelGGL(scale_by_diff_kernel, dim3((L + 255) / 256), dim3(256), 0, c->stream, Ci,
                       F->scal.p + 1, F->scal.p + 2, F->upd.p, L);
    CFM_CHECK_LAUNCH();
    return {};
  };
  while (too_large && it < maxit) {
    if (!upd_ready) CFM_TRY(low_rank_apply(c, F, F->err.p, F->upd.p, it));
    upd_ready = false;
    CFM_TRY(update(c, F, F->upd.p));
    CFM_TRY(fil_forces(c, F->K, nt));
    CFM_TRY(velocities(c, F, &S));
    const bool more = it < maxit - 1;  // a further iteration would be allowed
    CFM_TRY(residual(c, F, nt, F->nerr.p, &bad, more ? &next_column : nullptr));
    if (!std::isfinite(bad))
      return make_status(CFM_ENONFINITE, "Broyden's method diverged after %d iterations during step %d",
                         it + 1, nt + 1);
    too_large = bad > 1.0;
    if (too_large && more) {
      upd_ready = true;
      std::swap(F->err.p, F->nerr.p);
    }
    ++it;
  }
  CFM_HIP(hipStreamSynchronize(c->stream));
  F->stepped = true;
  S.broyden_iterations = it;
  S.converged = !too_large;
  S.final_error = bad;
  S.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (st) *st = S;
  return {};
}

// ---- equal independent systems in one context (members) ----------------------------
// Each member is a wall-bounded set of fpm filaments.  One product launch evaluates every
// active member's M_ss f (no coupling between members), the barrier skips pairs of
// different members, and each member runs its own Broyden iteration: its low-rank
// coefficients and norms are reductions over its own rows, it stops when its own
// badness meets TOL, and from then on its state is held (zero update).  Every member
// takes the arithmetic a lone system of fpm filaments would (the same product
// decomposition and reduction orders), so its states equal those of a lone run.

static Status member_velocities(cfm_ctx* c, FilState* F) {
  if (F->prm.barrier) CFM_TRY(fil_barrier(c, F->K, F->grid));
  const int saved = c->cfg.precision;
  c->cfg.precision = F->prm.pair_precision;
  const Status s = sym_members_segments(c, F->members, F->m_active.p);
  c->cfg.precision = saved;
  return s;
}

// err = the residual of every member; bad[e] read back to the host.
static Status member_residual(cfm_ctx* c, FilState* F, int nt, double* err, std::vector<double>& bad) {
  const int E = F->members, Lm = F->nbroy / E;
  CFM_TRY(fil_errors(c, F->K, nt, err, F->bad_part.p, F->scal.p, 0));
  CFM_TRY(fil_member_badness(c, F->K, err, Lm, E, F->m_bad.p));
  CFM_HIP(hipMemcpyAsync(F->m_host, F->m_bad.p, sizeof(double) * E, hipMemcpyDeviceToHost, c->stream));
  CFM_HIP(hipStreamSynchronize(c->stream));
  bad.assign(F->m_host, F->m_host + E);
  return {};
}

// The members' active flags to the device (through pinned memory, waited for: the next
// call may rewrite the staging words).
static Status member_flags(cfm_ctx* c, FilState* F, const std::vector<int>& act) {
  std::memcpy(F->m_host_flags, act.data(), sizeof(int) * act.size());
  CFM_HIP(hipMemcpyAsync(F->m_active.p, F->m_host_flags, sizeof(int) * act.size(),
                         hipMemcpyHostToDevice, c->stream));
  CFM_HIP(hipStreamSynchronize(c->stream));
  return {};
}

// out = -gamma J0^-1 e - sum_{i<it} (e . D_i) C_i, member by member (active members)
static Status member_low_rank(cfm_ctx* c, FilState* F, const double* e, double* out, int it) {
  const int E = F->members, L = F->nbroy, Lm = L / E, ldc = F->prm.max_broyden + 1;
  CFM_TRY(fil_jinv_mult(c, F->K, e, out, -F->prm.gamma));
  if (it > 0) {
    CFM_TRY(dev_member_dots(c, F->m_partial.p, F->D.p, (size_t)L, it, e, Lm, E, F->m_coef.p, ldc,
                            false, F->m_active.p));
    CFM_TRY(dev_member_gemv_sub(c, F->C.p, (size_t)L, it, F->m_coef.p, ldc, out, L, Lm));
  }
  return {};
}

static Status member_step(cfm_ctx* c, FilState* F, int nt, cfm_step_stats* st) {
  const auto t0 = std::chrono::steady_clock::now();
  F->nt = nt;
  const int E = F->members, L = F->nbroy, Lm = L / E;
  const int maxit = F->prm.max_broyden;
  std::vector<int> act(E, 1), iters(E, 0);
  std::vector<double> bad(E, 0.0), last(E, 0.0);
  CFM_TRY(initial_guess(c, F, nt));
  CFM_TRY(fil_forces(c, F->K, nt));
  bool jac_side = false;
  CFM_TRY(jacobian_fork(c, F, nt, &jac_side));
  CFM_TRY(member_flags(c, F, act));
  CFM_TRY(member_velocities(c, F));
  CFM_TRY(member_residual(c, F, nt, F->err.p, bad));
  if (jac_side) CFM_HIP(hipStreamWaitEvent(c->stream, F->ev_jac, 0));
  bool any = false;
  for (int e = 0; e < E; ++e) {
    if (!std::isfinite(bad[e]))
      return make_status(CFM_ENONFINITE, "non-finite residual of member %d at step %d", e, nt);
    act[e] = bad[e] > 1.0;
    last[e] = bad[e];
    any = any || act[e];
  }
  CFM_TRY(member_flags(c, F, act));
  if (any && !jac_side) CFM_TRY(jacobian(c, F, nt));
  int it = 0;
  bool upd_ready = false;  // the update was formed with the last low-rank column (as step())
  while (any && it < maxit) {
    if (!upd_ready) CFM_TRY(member_low_rank(c, F, F->err.p, F->upd.p, it));
    upd_ready = false;
    CFM_TRY(dev_member_mask(c, F->upd.p, L, Lm, F->m_active.p));  // held members: zero update
    CFM_TRY(update(c, F, F->upd.p));
    CFM_TRY(fil_forces(c, F->K, nt));
    CFM_TRY(member_velocities(c, F));
    CFM_TRY(member_residual(c, F, nt, F->nerr.p, bad));
    std::vector<int> next(E, 0);
    bool cont = false;
    for (int e = 0; e < E; ++e) {
      if (!act[e]) continue;
      if (!std::isfinite(bad[e]))
        return make_status(CFM_ENONFINITE,
                           "Broyden's method diverged for member %d after %d iterations during step %d",
                           e, it + 1, nt + 1);
      last[e] = bad[e];
      next[e] = bad[e] > 1.0;
      if (!next[e]) iters[e] = it + 1;
      cont = cont || next[e];
    }
    if (cont && it < maxit - 1) {
      // the low-rank pair of the members that continue (broyden_solver.cpp:140-174)
      CFM_TRY(member_flags(c, F, next));
      double* Ci = F->C.p + (size_t)it * L;
      double* Di = F->D.p + (size_t)it * L;
      CFM_TRY(member_low_rank(c, F, F->nerr.p, Ci, it));
      CFM_TRY(dev_axpby(c, 1.0, F->nerr.p, -1.0, F->err.p, Di, L));
      CFM_TRY(dev_member_dots(c, F->m_partial.p, Di, 0, 1, Di, Lm, E, F->m_norm.p, 1, true,
                              F->m_active.p));
      CFM_TRY(dev_member_scale_by_inv(c, Ci, F->m_norm.p, Ci, L, Lm));
      CFM_TRY(dev_member_scale_by_inv(c, Di, F->m_norm.p, Di, L, Lm));
      // the next update of each continuing member, C_it (n - e . D_it) (as step())
      CFM_TRY(dev_member_dots(c, F->m_partial.p, Di, 0, 1, F->nerr.p, Lm, E, F->m_dot.p, 1, false,
                              F->m_active.p));
      hipLaunchKernelGGL(member_scale_by_diff_kernel, dim3((L + 255) / 256), dim3(256), 0, c->stream,
                         Ci, F->m_norm.p, F->m_dot.p, F->upd.p, L, Lm);
      CFM_CHECK_LAUNCH();
      upd_ready = true;
      std::swap(F->err.p, F->nerr.p);
    } else if (cont) {
      CFM_TRY(member_flags(c, F, next));
    }
    act = next;
    any = cont;
    ++it;
  }
  CFM_HIP(hipStreamSynchronize(c->stream));
  F->stepped = true;
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  F->m_stats.assign(E, cfm_step_stats{});
  cfm_step_stats S{};
  S.converged = 1;
  for (int e = 0; e < E; ++e) {
    cfm_step_stats& m = F->m_stats[e];
    m.broyden_iterations = act[e] ? it : iters[e];
    m.converged = act[e] ? 0 : 1;
    m.final_error = last[e];
    m.seconds = secs;
    S.broyden_iterations = std::max(S.broyden_iterations, m.broyden_iterations);
    S.converged = S.converged && m.converged;
    S.final_error = std::max(S.final_error, m.final_error);
  }
  S.seconds = secs;
  if (st) *st = S;
  return {};
}

static FilState* fil_of(cfm_ctx* c) { return reinterpret_cast<FilState*>(c->fil); }

}  // namespace cfm

using namespace cfm;

static int fail_c(cfm_ctx* c, const Status& s) {
  c->err = s.msg;
  return s.code;
}

static bool finite_all(const double* a, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(a[i])) return false;
  return true;
}

void cfm_fil_release(cfm_ctx* c) {
  FilState* F = fil_of(c);
  if (!F) return;
  DevBuf<double>* bufs[] = {&F->q, &F->qm1, &F->u, &F->um1, &F->xm1, &F->xm2, &F->lam,
                            &F->lam_m1, &F->lam_m2, &F->clamp, &F->clamp_m1, &F->clamp_m2,
                            &F->base, &F->pert, &F->fint, &F->jinv, &F->err, &F->nerr, &F->upd,
                            &F->C, &F->D, &F->coef, &F->partial, &F->bad_part, &F->scal,
                            &F->fb_prev, &F->rhs_b, &F->sw_state, &F->sw_fref, &F->sw_fext,
                            &F->sw_mob, &F->sw_BJ, &F->sw_Sf, &F->sw_clamp, &F->sw_rep, &F->sw_tmp,
                            &F->bx_fb, &F->bx_fb_m1, &F->bx_fb_m2};
  for (auto* b : bufs) b->release();
  F->status.release();
  F->grid.release();
  F->hist.B.release();
  F->hist.X.release();
  F->hist.G.release();
  F->hist.g.release();
  F->hist.coef.release();
  F->hist.used.release();
  F->hist.partial.release();
  if (F->host) (void)hipHostFree(F->host);
  if (F->m_host) (void)hipHostFree(F->m_host);
  if (F->m_host_flags) (void)hipHostFree(F->m_host_flags);
  if (F->ev_jac) (void)hipEventDestroy(F->ev_jac);
  if (F->ev_res) (void)hipEventDestroy(F->ev_res);
  F->m_bad.release();
  F->m_norm.release();
  F->m_dot.release();
  F->m_coef.release();
  F->m_partial.release();
  F->m_active.release();
  delete F;
  c->fil = nullptr;
}

extern "C" {

int cfm_fil_init(cfm_ctx* c, const cfm_fil_params* prm, const double* q_seg, const double* pert) {
  if (!c || !prm || !q_seg) return c ? fail_c(c, make_status(CFM_EINVAL, "null argument")) : CFM_EINVAL;
  if (!c->have_positions) return fail_c(c, make_status(CFM_ESTATE, "positions not set"));
  Particles& p = c->p;
  if (p.seg_dof != 6) return fail_c(c, make_status(CFM_EINVAL, "filaments need seg_dof = 6"));
  if (prm->nseg < 2 || prm->nseg > kMaxFilSeg || p.n_seg % prm->nseg != 0)
    return fail_c(c, make_status(CFM_EINVAL, "nseg must be in [2, %d] and divide n_seg (%d)", kMaxFilSeg,
                                 p.n_seg));
  if (prm->max_broyden < 1 || !(prm->dt > 0) || !(prm->dl > 0) || !(prm->tol > 0))
    return fail_c(c, make_status(CFM_EINVAL, "need max_broyden >= 1 and dt, dl, tol > 0"));
  if (prm->pair_precision != CFM_FP64 && prm->pair_precision != CFM_FP32C)
    return fail_c(c, make_status(CFM_EINVAL, "unknown pair precision"));
  if (!finite_all(q_seg, 4 * (size_t)p.n_seg)) return fail_c(c, make_status(CFM_ENONFINITE, "non-finite q"));
  (void)hipSetDevice(c->device);
  cfm_fil_release(c);
  FilState* F = new FilState();
  c->fil = F;
  F->prm = *prm;
  FilK& K = F->K;
  K.nseg = prm->nseg;
  K.nfil = p.n_seg / prm->nseg;
  K.nseg_total = p.n_seg;
  K.num_euler = prm->num_euler_steps;
  K.dl = prm->dl;
  K.dt = prm->dt;
  K.kb = prm->kb;
  K.kt = prm->kt;
  K.mu = c->cfg.mu;
  K.a = c->cfg.a_seg;
  K.end_force = prm->end_force;
  K.rep_factor = prm->repulsive_factor;
  K.tol = prm->tol;
  const size_t ns = p.n_seg, nf = K.nfil;
  F->bx = prm->broyden_only != 0 && p.n_blob > 0;
  F->nbroy = 6 * p.n_seg + (F->bx ? 3 * p.n_blob : 0);
  auto run = [&]() -> Status {
    CFM_TRY(F->q.alloc(4 * ns)); CFM_TRY(F->qm1.alloc(4 * ns));
    CFM_TRY(F->u.alloc(3 * ns)); CFM_TRY(F->um1.alloc(3 * ns));
    CFM_TRY(F->xm1.alloc(3 * ns)); CFM_TRY(F->xm2.alloc(3 * ns));
    CFM_TRY(F->lam.alloc(3 * ns)); CFM_TRY(F->lam_m1.alloc(3 * ns)); CFM_TRY(F->lam_m2.alloc(3 * ns));
    CFM_TRY(F->clamp.alloc(3 * nf)); CFM_TRY(F->clamp_m1.alloc(3 * nf)); CFM_TRY(F->clamp_m2.alloc(3 * nf));
    CFM_TRY(F->base.alloc(3 * nf)); CFM_TRY(F->pert.alloc(3 * nf));
    CFM_TRY(F->fint.alloc(6 * ns));
    CFM_TRY(F->jinv.alloc((size_t)nf * 36 * K.nseg * K.nseg));
    const int L = F->nbroy;
    CFM_TRY(F->err.alloc(L)); CFM_TRY(F->nerr.alloc(L)); CFM_TRY(F->upd.alloc(L));
    CFM_TRY(F->C.alloc((size_t)L * prm->max_broyden));
    CFM_TRY(F->D.alloc((size_t)L * prm->max_broyden));
    CFM_TRY(F->coef.alloc(prm->max_broyden + 1));
    CFM_TRY(F->partial.alloc(dev_dots_partial_size(L, prm->max_broyden + 1)));
    CFM_TRY(F->bad_part.alloc((p.n_seg + 255) / 256 + 1 + (F->bx ? (p.n_blob + 255) / 256 : 0)));
    if (F->bx) {
      for (auto* b : {&F->bx_fb, &F->bx_fb_m1, &F->bx_fb_m2}) {
        CFM_TRY(b->alloc(3 * (size_t)p.n_blob));
        CFM_TRY(dev_fill(c, b->p, 0.0, 3 * p.n_blob));
      }
      CFM_TRY(F->sw_tmp.alloc(6 * ns));
    }
    CFM_TRY(F->scal.alloc(8));
    CFM_TRY(F->grid.ensure(p.n_seg + p.n_blob));
    CFM_TRY(F->status.alloc(4));
    if (p.n_blob) {
      CFM_TRY(F->fb_prev.alloc(3 * (size_t)p.n_blob));
      CFM_TRY(dev_fill(c, F->fb_prev.p, 0.0, 3 * p.n_blob));
    }
    DevBuf<double>* zero[] = {&F->u, &F->um1, &F->lam, &F->lam_m1, &F->lam_m2, &F->clamp,
                              &F->clamp_m1, &F->clamp_m2, &F->pert};
    for (auto* b : zero) CFM_TRY(dev_fill(c, b->p, 0.0, (int)b->n));
    CFM_TRY(upload_aos(c, q_seg, p.n_seg, 4, F->q.p, p.n_seg, 0));
    CFM_HIP(hipMemcpyAsync(F->qm1.p, F->q.p, sizeof(double) * 4 * ns, hipMemcpyDeviceToDevice, c->stream));
    for (int k = 0; k < 3; ++k) {
      const double* src = (k == 0 ? p.x.p : k == 1 ? p.y.p : p.z.p);
      CFM_HIP(hipMemcpyAsync(F->xm1.p + k * ns, src, sizeof(double) * ns, hipMemcpyDeviceToDevice, c->stream));
      CFM_HIP(hipMemcpyAsync(F->xm2.p + k * ns, src, sizeof(double) * ns, hipMemcpyDeviceToDevice, c->stream));
    }
    // base of filament f = position of its first segment
    for (int k = 0; k < 3; ++k) {
      const double* src = (k == 0 ? p.x.p : k == 1 ? p.y.p : p.z.p);
      CFM_HIP(hipMemcpy2DAsync(F->base.p + k * nf, sizeof(double), src, sizeof(double) * K.nseg,
                               sizeof(double), nf, hipMemcpyDeviceToDevice, c->stream));
    }
    if (pert) CFM_TRY(upload_aos(c, pert, (int)nf, 3, F->pert.p, (int)nf, 0));
    CFM_HIP(hipStreamSynchronize(c->stream));
    return {};
  };
  Status s = run();
  if (!s.ok()) {
    cfm_fil_release(c);
    return fail_c(c, s);
  }
  K.x = p.x.p; K.y = p.y.p; K.z = p.z.p;
  K.q = F->q.p; K.qm1 = F->qm1.p; K.u = F->u.p; K.um1 = F->um1.p;
  K.xm1 = F->xm1.p; K.xm2 = F->xm2.p;
  K.lam = F->lam.p; K.lam_m1 = F->lam_m1.p; K.lam_m2 = F->lam_m2.p;
  K.clamp = F->clamp.p; K.clamp_m1 = F->clamp_m1.p; K.clamp_m2 = F->clamp_m2.p;
  K.base = F->base.p; K.pert = F->pert.p; K.fint = F->fint.p; K.jinv = F->jinv.p;
  if (F->bx) {
    BxK& X = F->X;
    X.m = p.n_blob;
    X.num_euler = prm->num_euler_steps;
    X.dt = prm->dt;
    X.tol = prm->tol;
    X.mu = c->cfg.mu;
    X.a = c->cfg.a_blob;
    X.fb = F->bx_fb.p;
    X.fb_m1 = F->bx_fb_m1.p;
    X.fb_m2 = F->bx_fb_m2.p;
  }
  return CFM_OK;
}

#define FIL_OR_FAIL(c)                                                              \
  if (!c) return CFM_EINVAL;                                                        \
  FilState* F = fil_of(c);                                                          \
  if (!F) return fail_c(c, make_status(CFM_ESTATE, "cfm_fil_init has not been called")); \
  (void)hipSetDevice(c->device)

int cfm_fil_step(cfm_ctx* c, int32_t nt, cfm_step_stats* st) {
  FIL_OR_FAIL(c);
  if (nt < 0) return fail_c(c, make_status(CFM_EINVAL, "nt must be >= 0"));
  Status s = F->members > 1 ? member_step(c, F, nt, st) : step(c, F, nt, st);
  return s.ok() ? CFM_OK : fail_c(c, s);
}

int cfm_fil_set_members(cfm_ctx* c, int32_t members) {
  FIL_OR_FAIL(c);
  const Particles& p = c->p;
  if (members < 1 || F->K.nfil % members != 0)
    return fail_c(c, make_status(CFM_EINVAL, "members must divide the number of filaments (%d)", F->K.nfil));
  if (members > 1 && (p.n_blob > 0 || F->swim || F->bx || !c->cfg.wall))
    return fail_c(c, make_status(CFM_EINVAL, "members: wall-bounded filaments only (no blobs or bodies)"));
  const int L = F->nbroy, Lm = L / members, ldc = F->prm.max_broyden + 1;
  auto run = [&]() -> Status {
    F->members = members;
    F->K.fpm = members > 1 ? F->K.nfil / members : 0;
    if (members == 1) return {};
    CFM_TRY(F->m_bad.alloc(members));
    CFM_TRY(F->m_norm.alloc(members));
    CFM_TRY(F->m_dot.alloc(members));
    CFM_TRY(F->m_coef.alloc((size_t)members * ldc));
    CFM_TRY(dev_fill(c, F->m_coef.p, 0.0, members * ldc));
    CFM_TRY(F->m_partial.alloc(dev_member_dots_partial_size(Lm, ldc, members)));
    CFM_TRY(F->m_active.alloc(members));
    if (F->m_host) (void)hipHostFree(F->m_host);
    if (F->m_host_flags) (void)hipHostFree(F->m_host_flags);
    F->m_host = nullptr;
    F->m_host_flags = nullptr;
    CFM_HIP(hipHostMalloc((void**)&F->m_host, sizeof(double) * members, hipHostMallocDefault));
    CFM_HIP(hipHostMalloc((void**)&F->m_host_flags, sizeof(int) * members, hipHostMallocDefault));
    CFM_HIP(hipStreamSynchronize(c->stream));
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : fail_c(c, s);
}

int cfm_fil_member_stats(cfm_ctx* c, cfm_step_stats* per_member) {
  FIL_OR_FAIL(c);
  if (!per_member) return fail_c(c, make_status(CFM_EINVAL, "null argument"));
  if ((int)F->m_stats.size() != F->members || F->members < 2)
    return fail_c(c, make_status(CFM_ESTATE, "no member step has run"));
  for (int e = 0; e < F->members; ++e) per_member[e] = F->m_stats[e];
  return CFM_OK;
}

int cfm_fil_get_state(cfm_ctx* c, double* x, double* q, double* u, double* lam, double* clamp) {
  FIL_OR_FAIL(c);
  const int ns = c->p.n_seg;
  auto run = [&]() -> Status {
    if (x) {
      // positions are the context's particle rows x, y, z (separate arrays)
      std::vector<double> tmp(3 * (size_t)ns);
      CFM_HIP(hipMemcpyAsync(tmp.data(), c->p.x.p, sizeof(double) * ns, hipMemcpyDeviceToHost, c->stream));
      CFM_HIP(hipMemcpyAsync(tmp.data() + ns, c->p.y.p, sizeof(double) * ns, hipMemcpyDeviceToHost, c->stream));
      CFM_HIP(hipMemcpyAsync(tmp.data() + 2 * ns, c->p.z.p, sizeof(double) * ns, hipMemcpyDeviceToHost, c->stream));
      CFM_HIP(hipStreamSynchronize(c->stream));
      for (int i = 0; i < ns; ++i)
        for (int k = 0; k < 3; ++k) x[3 * i + k] = tmp[(size_t)k * ns + i];
    }
    if (q) CFM_TRY(download_aos(c, q, ns, 4, F->q.p, ns, 0));
    if (u) CFM_TRY(download_aos(c, u, ns, 3, F->u.p, ns, 0));
    if (lam) CFM_TRY(download_aos(c, lam, ns, 3, F->lam.p, ns, 0));
    if (clamp) CFM_TRY(download_aos(c, clamp, F->K.nfil, 3, F->clamp.p, F->K.nfil, 0));
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : fail_c(c, s);
}

int cfm_fil_set_state(cfm_ctx* c, const double* q, const double* qm1, const double* u,
                      const double* um1, const double* lam, const double* clamp,
                      const double* xm1, const double* xm2) {
  FIL_OR_FAIL(c);
  const int ns = c->p.n_seg;
  auto run = [&]() -> Status {
    if (q) CFM_TRY(upload_aos(c, q, ns, 4, F->q.p, ns, 0));
    if (qm1) CFM_TRY(upload_aos(c, qm1, ns, 4, F->qm1.p, ns, 0));
    if (u) CFM_TRY(upload_aos(c, u, ns, 3, F->u.p, ns, 0));
    if (um1) CFM_TRY(upload_aos(c, um1, ns, 3, F->um1.p, ns, 0));
    if (lam) CFM_TRY(upload_aos(c, lam, ns, 3, F->lam.p, ns, 0));
    if (clamp) CFM_TRY(upload_aos(c, clamp, F->K.nfil, 3, F->clamp.p, F->K.nfil, 0));
    if (xm1) CFM_TRY(upload_aos(c, xm1, ns, 3, F->xm1.p, ns, 0));
    if (xm2) CFM_TRY(upload_aos(c, xm2, ns, 3, F->xm2.p, ns, 0));
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : fail_c(c, s);
}

int cfm_fil_internal_forces(cfm_ctx* c, int32_t nt, double* f_seg) {
  FIL_OR_FAIL(c);
  if (!f_seg) return fail_c(c, make_status(CFM_EINVAL, "null output"));
  auto run = [&]() -> Status {
    CFM_TRY(fil_forces(c, F->K, nt));
    return download_aos(c, f_seg, c->p.n_seg, 6, F->fint.p, c->p.n_seg, 0);
  };
  Status s = run();
  return s.ok() ? CFM_OK : fail_c(c, s);
}

int cfm_barrier_forces(cfm_ctx* c, double* f_seg) {
  FIL_OR_FAIL(c);
  if (!f_seg) return fail_c(c, make_status(CFM_EINVAL, "null output"));
  auto run = [&]() -> Status {
    CFM_TRY(dev_fill(c, c->p.f.p, 0.0, 3 * c->p.n));
    CFM_TRY(fil_barrier(c, F->K, F->grid));
    return download_aos(c, f_seg, c->p.n_seg, 3, c->p.f.p, c->p.n, 0);
  };
  Status s = run();
  return s.ok() ? CFM_OK : fail_c(c, s);
}

int cfm_fil_jacobian_inverse(cfm_ctx* c, int32_t nt, double* jinv) {
  FIL_OR_FAIL(c);
  auto run = [&]() -> Status {
    CFM_TRY(fil_forces(c, F->K, nt));
    CFM_HIP(hipMemsetAsync(F->status.p, 0, sizeof(int), c->stream));
    CFM_TRY(fil_jacobian(c, F->K, nt, F->status.p));
    int stv = 0;
    CFM_HIP(hipMemcpyAsync(&stv, F->status.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CFM_HIP(hipStreamSynchronize(c->stream));
    if (stv) return make_status(CFM_ENONFINITE, "singular approximate Jacobian");
    if (jinv) {
      const int N = 6 * F->K.nseg;
      const size_t tot = (size_t)F->K.nfil * N * N;
      std::vector<double> t(tot);
      CFM_HIP(copy_to_host(c->stream, t.data(), F->jinv.p, sizeof(double) * tot));
      for (int f = 0; f < F->K.nfil; ++f)
        for (int r = 0; r < N; ++r)
          for (int cc = 0; cc < N; ++cc)
            jinv[(size_t)f * N * N + (size_t)r * N + cc] = t[(size_t)f * N * N + (size_t)cc * N + r];
    }
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : fail_c(c, s);
}

int cfm_fil_errors(cfm_ctx* c, int32_t nt, const double* v_seg, double* err, double* badness) {
  FIL_OR_FAIL(c);
  if (!v_seg || !err) return fail_c(c, make_status(CFM_EINVAL, "null argument"));
  if (F->bx) return fail_c(c, make_status(CFM_EINVAL, "with broyden_only the residual also needs the blob "
                                                      "velocities; it is formed inside cfm_fil_step"));
  auto run = [&]() -> Status {
    CFM_TRY(upload_aos(c, v_seg, c->p.n_seg, 6, c->p.v.p, c->p.n, 0));
    double bad = 0.0;
    CFM_TRY(residual(c, F, nt, F->err.p, &bad));
    if (badness) *badness = bad;
    CFM_HIP(copy_to_host(c->stream, err, F->err.p, sizeof(double) * F->nbroy));
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : fail_c(c, s);
}

int cfm_fil_update(cfm_ctx* c, const double* upd) {
  FIL_OR_FAIL(c);
  if (!upd) return fail_c(c, make_status(CFM_EINVAL, "null argument"));
  auto run = [&]() -> Status {
    CFM_HIP(copy_to_device(c->stream, F->upd.p, upd, sizeof(double) * F->nbroy));
    CFM_TRY(update(c, F, F->upd.p));
    CFM_HIP(hipStreamSynchronize(c->stream));
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : fail_c(c, s);
}

int cfm_fil_attach_bodies(cfm_ctx* c, const double* f_ext) {
  FIL_OR_FAIL(c);
  if (F->swim) return fail_c(c, make_status(CFM_ESTATE, "bodies are already attached"));
  if (F->members > 1)
    return fail_c(c, make_status(CFM_ESTATE, "members are wall-bounded filaments only (cfm_fil_set_members)"));
  BodyInfo bi;
  Status st = body_info(c, &bi);
  if (!st.ok()) return fail_c(c, st);
  const int nfil = F->K.nfil, B = bi.n_body;
  if (nfil % B != 0)
    return fail_c(c, make_status(CFM_EINVAL, "%d filaments cannot be shared equally by %d bodies", nfil, B));
  if (c->cfg.a_blob != c->cfg.a_seg)
    return fail_c(c, make_status(CFM_EINVAL, "free swimmers need equal blob and segment radii"));
  const int pc = F->prm.blob_solve.precond;
  if (pc == CFM_PC_NEAR)
    return fail_c(c, make_status(CFM_EINVAL, "free swimmers take blob preconditioners NONE, DIAG, SELF "
                                             "or DENSE"));
  if (f_ext)
    for (int i = 0; i < 6 * B; ++i)
      if (!std::isfinite(f_ext[i])) return fail_c(c, make_status(CFM_ENONFINITE, "non-finite external load"));
  const int ns = c->p.n_seg, m = c->p.n_blob;
  auto run = [&]() -> Status {
    // base offsets in the body frame from the current first segments
    std::vector<double> base(3 * (size_t)nfil), fref(3 * (size_t)nfil);
    CFM_HIP(copy_to_host(c->stream, base.data(), F->base.p, sizeof(double) * base.size()));
    for (int f = 0; f < nfil; ++f) {
      const int b = f / (nfil / B);
      const double* q = &bi.q[4 * (size_t)b];
      const double s0 = q[0], a = q[1], bb = q[2], cc = q[3];
      const double R[9] = {1 - 2 * (bb * bb + cc * cc), 2 * (a * bb - s0 * cc), 2 * (a * cc + s0 * bb),
                           2 * (a * bb + s0 * cc), 1 - 2 * (a * a + cc * cc), 2 * (bb * cc - s0 * a),
                           2 * (a * cc - s0 * bb), 2 * (bb * cc + s0 * a), 1 - 2 * (a * a + bb * bb)};
      const double d[3] = {base[f] - bi.x[3 * b], base[nfil + f] - bi.x[3 * b + 1],
                           base[2 * (size_t)nfil + f] - bi.x[3 * b + 2]};
      for (int k = 0; k < 3; ++k) fref[3 * (size_t)f + k] = R[k] * d[0] + R[3 + k] * d[1] + R[6 + k] * d[2];
    }
    // per body: x, xm1, xm2, u, um1 (3 each), q, qm1 (4 each), Qinit 9, vel 6, force 6,
    // S0 36, Sinv 36, tb 6
    const int per = 3 * 5 + 4 * 2 + 9 + 6 + 6 + 36 + 36 + 6;
    CFM_TRY(F->sw_state.alloc((size_t)per * B));
    double* d0 = F->sw_state.p;
    SwimK& S = F->S;
    size_t o = 0;
    auto take = [&](int k) { double* p = d0 + o; o += (size_t)k * B; return p; };
    S.x = take(3); S.xm1 = take(3); S.xm2 = take(3); S.u = take(3); S.um1 = take(3);
    S.q = take(4); S.qm1 = take(4); S.Qinit = take(9); S.vel = take(6); S.force = take(6);
    S.S0 = take(36); S.Sinv = take(36); S.tb = take(6);
    std::vector<double> hx(3 * (size_t)B), hq(4 * (size_t)B), hQ(9 * (size_t)B);
    for (int b = 0; b < B; ++b) {
      for (int k = 0; k < 3; ++k) hx[3 * b + k] = bi.x[3 * b + k];
      for (int k = 0; k < 4; ++k) hq[4 * b + k] = bi.q[4 * b + k];
      const double* q = &bi.q[4 * (size_t)b];
      const double s0 = q[0], a = q[1], bb = q[2], cc = q[3];
      const double R[9] = {1 - 2 * (bb * bb + cc * cc), 2 * (a * bb - s0 * cc), 2 * (a * cc + s0 * bb),
                           2 * (a * bb + s0 * cc), 1 - 2 * (a * a + cc * cc), 2 * (bb * cc - s0 * a),
                           2 * (a * cc - s0 * bb), 2 * (bb * cc + s0 * a), 1 - 2 * (a * a + bb * bb)};
      for (int k = 0; k < 9; ++k) hQ[9 * b + k] = R[k];
    }
    CFM_TRY(dev_fill(c, d0, 0.0, per * B));
    for (double* dst : {S.x, S.xm1, S.xm2})
      CFM_HIP(hipMemcpyAsync(dst, hx.data(), sizeof(double) * hx.size(), hipMemcpyHostToDevice, c->stream));
    for (double* dst : {S.q, S.qm1})
      CFM_HIP(hipMemcpyAsync(dst, hq.data(), sizeof(double) * hq.size(), hipMemcpyHostToDevice, c->stream));
    CFM_HIP(hipMemcpyAsync(S.Qinit, hQ.data(), sizeof(double) * hQ.size(), hipMemcpyHostToDevice, c->stream));
    CFM_TRY(F->sw_fref.alloc(fref.size()));
    CFM_HIP(hipMemcpyAsync(F->sw_fref.p, fref.data(), sizeof(double) * fref.size(), hipMemcpyHostToDevice, c->stream));
    CFM_TRY(F->sw_fext.alloc(6 * (size_t)B));
    CFM_TRY(dev_fill(c, F->sw_fext.p, 0.0, 6 * B));
    if (f_ext)
      CFM_HIP(hipMemcpyAsync(F->sw_fext.p, f_ext, sizeof(double) * 6 * B, hipMemcpyHostToDevice, c->stream));
    CFM_TRY(F->sw_mob.alloc(36));
    CFM_HIP(hipMemcpyAsync(F->sw_mob.p, bi.mob, sizeof(double) * 36, hipMemcpyHostToDevice, c->stream));
    const int N = 6 * F->K.nseg;
    CFM_TRY(F->sw_BJ.alloc((size_t)nfil * 6 * N));
    CFM_TRY(F->sw_Sf.alloc(36 * (size_t)nfil));
    CFM_TRY(F->sw_clamp.alloc(18 * (size_t)nfil));
    CFM_TRY(F->sw_rep.alloc(3 * (size_t)m));
    CFM_TRY(dev_fill(c, F->sw_rep.p, 0.0, 3 * m));
    CFM_TRY(F->sw_tmp.alloc(6 * (size_t)ns));
    S.nbody = B;
    S.fpb = nfil / B;
    S.nfil = nfil;
    S.nseg = F->K.nseg;
    S.nb = bi.nb;
    S.num_euler = F->K.num_euler;
    S.dl = F->K.dl;
    S.dt = F->K.dt;
    S.tol = F->K.tol;
    S.end_force = F->K.end_force;
    S.a_seg = c->cfg.a_seg;
    S.mu = c->cfg.mu;
    S.fref = F->sw_fref.p;
    S.fext = F->sw_fext.p;
    S.mob = F->sw_mob.p;
    S.BJ = F->sw_BJ.p;
    S.Sf = F->sw_Sf.p;
    // Broyden vectors grow by 6 per body (and keep the blob forces with broyden_only);
    // the blob solve by the body velocities
    const int L = 6 * ns + 6 * B + (F->bx ? 3 * m : 0);
    F->nbroy = L;
    for (auto* buf : {&F->err, &F->nerr, &F->upd}) CFM_TRY(buf->alloc(L));
    CFM_TRY(F->C.alloc((size_t)L * F->prm.max_broyden));
    CFM_TRY(F->D.alloc((size_t)L * F->prm.max_broyden));
    CFM_TRY(F->partial.alloc(dev_dots_partial_size(L, F->prm.max_broyden + 1)));
    CFM_TRY(F->bad_part.alloc((ns + 255) / 256 + 1 + B + (F->bx ? (m + 255) / 256 : 0)));
    CFM_TRY(F->fb_prev.alloc(3 * (size_t)m + 6 * B));
    CFM_TRY(dev_fill(c, F->fb_prev.p, 0.0, 3 * m + 6 * B));
    F->K.bx = S.x;
    F->K.bq = S.q;
    F->K.bu = S.u;
    F->K.fref = F->sw_fref.p;
    F->K.fpb = S.fpb;
    F->K.clampblk = F->sw_clamp.p;
    F->prm.guess = CFM_GUESS_PREVIOUS;  // the blob matrix changes as the bodies move
    body_freeze_dense(c, pc == CFM_PC_DENSE);
    F->swim = true;
    CFM_TRY(body_place_device(c, S.x, S.q));
    if (F->bx) {
      F->X.nbody = B;
      F->X.nb = bi.nb;
      F->X.ref = body_ref_device(c);
      if (!F->X.ref) return make_status(CFM_ESTATE, "no body-frame blob offsets on the device");
    }
    CFM_HIP(hipStreamSynchronize(c->stream));
    return {};
  };
  st = run();
  if (!st.ok()) {
    F->swim = false;
    return fail_c(c, st);
  }
  return CFM_OK;
}

int cfm_fil_set_body_sphere(cfm_ctx* c, double radius) {
  if (!c) return CFM_EINVAL;
  FilState* F = fil_of(c);
  if (!F || !F->swim) return fail_c(c, make_status(CFM_ESTATE, "cfm_fil_attach_bodies has not been called"));
  if (!(radius >= 0.0) || !std::isfinite(radius))
    return fail_c(c, make_status(CFM_EINVAL, "sphere radius must be finite and >= 0"));
  F->S.sphere_R = radius;
  return CFM_OK;
}

int cfm_fil_get_body_state(cfm_ctx* c, double* x, double* q, double* u, double* vel) {
  FIL_OR_FAIL(c);
  if (!F->swim) return fail_c(c, make_status(CFM_ESTATE, "no bodies attached (cfm_fil_attach_bodies)"));
  const int B = F->S.nbody;
  auto run = [&]() -> Status {
    if (x) CFM_HIP(hipMemcpyAsync(x, F->S.x, sizeof(double) * 3 * B, hipMemcpyDeviceToHost, c->stream));
    if (q) CFM_HIP(hipMemcpyAsync(q, F->S.q, sizeof(double) * 4 * B, hipMemcpyDeviceToHost, c->stream));
    if (u) CFM_HIP(hipMemcpyAsync(u, F->S.u, sizeof(double) * 3 * B, hipMemcpyDeviceToHost, c->stream));
    if (vel) CFM_HIP(hipMemcpyAsync(vel, F->S.vel, sizeof(double) * 6 * B, hipMemcpyDeviceToHost, c->stream));
    CFM_HIP(hipStreamSynchronize(c->stream));
    return {};
  };
  Status st = run();
  return st.ok() ? CFM_OK : fail_c(c, st);
}

int cfm_fil_get_blob_forces(cfm_ctx* c, double* f_blob) {
  FIL_OR_FAIL(c);
  if (!f_blob) return fail_c(c, make_status(CFM_EINVAL, "null output"));
  const int m = c->p.n_blob;
  if (m == 0) return CFM_OK;
  Status st = download_aos(c, f_blob, m, 3, F->bx ? F->bx_fb.p : F->fb_prev.p, m, 0);
  return st.ok() ? CFM_OK : fail_c(c, st);
}

int cfm_fil_initial_guess(cfm_ctx* c, int32_t nt) {
  FIL_OR_FAIL(c);
  Status s = initial_guess(c, F, nt);
  if (s.ok()) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) s = make_status(CFM_EHIP, "%s", hipGetErrorString(e));
  }
  return s.ok() ? CFM_OK : fail_c(c, s);
}

}  // extern "C"
