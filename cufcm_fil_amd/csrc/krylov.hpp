// krylov.hpp -- device workspace and entry points of the GMRES solver.
#pragma once
#include "ctx.hpp"

namespace cfm {

// Krylov workspace (device), sized for vectors of length L and m basis vectors.
struct Krylov {
  int L = 0, m = 0;
  // Q: Arnoldi basis; Z: its preconditioned images P^-1 q_j, kept so the update
  // x += P^-1 Q y is the product Z y (no further preconditioner application)
  DevBuf<double> Q, Z, H, cs, sn, beta, h, h2, y, scal, r, w, z, t, d, dx, rhs, partial, res_ring;
  double* host_res = nullptr;  // pinned scalar read-back
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  Status reserve(int L_, int m_);
  void release();
};

// One linear system over a contiguous particle range of the combined index space.
struct SystemDesc {
  int m = 0;      // particles
  int dof = 3;    // unknowns per particle
  int off = 0;    // first particle (combined id)
  int L = 0;      // m * dof
  int which = 0;  // CFM_SS or CFM_BB
  double a = 1.0;
  // Coupled prescribed-cilia system (cilia.hip): `extra` scalar unknowns (phase and
  // shape-angle speeds) follow the [dof][m] force block, L = m dof + extra.
  bool cilia = false;
  int extra = 0;
  // Free rigid bodies (body.hip): 6 unknowns per body (V, Omega) follow the [3][m]
  // blob-force block, L = 3 m + 6 n_body.
  bool body = false;
};

// Coupling rows / columns and block preconditioner of the coupled cilia system.
Status cilia_couple_apply(cfm_ctx* c, const double* x, double* y);
Status cilia_precond(cfm_ctx* c, int kind, const double* x, double* y);

// Coupling rows / columns, block preconditioners and description of the free-body system.
Status body_couple_apply(cfm_ctx* c, const double* x, double* y);
Status body_precond(cfm_ctx* c, int kind, const double* x, double* y);
Status body_system_desc(cfm_ctx* c, SystemDesc* s);

// Optional correction added to the preconditioner output (learned near-field term),
// and optional work queued once the solve has queued its first read-back (||b|| and the
// residual of the guess) and before the host waits for it: work independent of the
// solve that keeps the device busy meanwhile.  solve() calls before_first_wait at most
// once, and always when it returns CFM_OK.
struct GmresHooks {
  Status (*precond_correction)(cfm_ctx*, const SystemDesc&, const double* in, double* out,
                               void* user) = nullptr;
  void* user = nullptr;
  Status (*before_first_wait)(cfm_ctx*, void* user) = nullptr;
  void* wait_user = nullptr;
};

SystemDesc system_desc(const cfm_ctx* c, int system);
Status system_apply(cfm_ctx* c, const SystemDesc& s, int precision, const double* x, double* y);
Status system_precond(cfm_ctx* c, const SystemDesc& s, int kind, const double* x, double* y);
Status solve(cfm_ctx* c, Krylov& K, const SystemDesc& s, const cfm_solve_opts& o, const double* b,
             double* x, cfm_solve_stats* st, const GmresHooks* hooks);
Status dots(cfm_ctx* c, Krylov& K, const double* Q, size_t ldq, int k, const double* w, int L,
            double* out, bool accumulate);
Status norm2(cfm_ctx* c, Krylov& K, const double* w, int L, double* out);
Status read_scalar(cfm_ctx* c, Krylov& K, const double* dev, double* host);

}  // namespace cfm

namespace cfm {
Status dev_dots(cfm_ctx* c, double* partial, const double* Q, size_t ldq, int k, const double* w,
                int L, double* out);
// As dev_dots, also writing out2[i * ld2] = out[i].
Status dev_dots2(cfm_ctx* c, double* partial, const double* Q, size_t ldq, int k, const double* w,
                 int L, double* out, double* out2, int ld2);
size_t dev_dots_partial_size(int L, int k);
// E independent vectors of length Lm laid end to end (members): out[e ldo + i] =
// Q[i][e Lm ..]^T w[e Lm ..] (norm: the square root, k = 1), members with active[e] == 0
// skipped (active may be null); the partials hold dev_member_dots_partial_size values.
Status dev_member_dots(cfm_ctx* c, double* partial, const double* Q, size_t ldq, int k,
                       const double* w, int Lm, int E, double* out, int ldo, bool norm,
                       const int* active);
size_t dev_member_dots_partial_size(int Lm, int k, int E);
// w[r] -= sum_{i<k} h[m ldh + i] Q[i][r] with m = r / Lm
Status dev_member_gemv_sub(cfm_ctx* c, const double* Q, size_t ldq, int k, const double* h, int ldh,
                           double* w, int L, int Lm);
// out[r] = x[r] / s[r / Lm]; x[r] = 0 where active[r / Lm] == 0
Status dev_member_scale_by_inv(cfm_ctx* c, const double* x, const double* s, double* out, int L, int Lm);
Status dev_member_mask(cfm_ctx* c, double* x, int L, int Lm, const int* active);
Status dev_gemv_sub(cfm_ctx* c, const double* Q, size_t ldq, int k, const double* h, double* w, int L);
Status dev_axpby(cfm_ctx* c, double a, const double* x, double b, const double* y, double* out, int L);
Status dev_fill(cfm_ctx* c, double* x, double v, int L);
Status dev_gemv(cfm_ctx* c, const double* Q, size_t ldq, int k, const double* y, double* out, int L);
Status dev_scale_by_inv(cfm_ctx* c, const double* x, const double* s, double* out, int L);
Status dev_norm2(cfm_ctx* c, double* partial, const double* w, int L, double* out);
Status dev_copy_rows(cfm_ctx* c, const double* src, size_t src_ld, int src_off, double* dst,
                     size_t dst_ld, int dst_off, int dof, int m);
}  // namespace cfm
