// squirmer.hip -- squirmer bodies: rigid bodies of blobs driven by a prescribed surface
// slip velocity, with no cilia (the reference's CILIA_TYPE 4, NO_CILIA_SQUIRMER).
//
// Slip (mobility_solver.cpp:470-505 with util.cu:69-115): for a blob at body-frame
// offset r with polar angle theta (from the body's z axis) and azimuth phi,
//     u_r = sum_n A_n P_n(cos theta),  u_theta = sum_n B_n V_n(cos theta),
//     V_n(x) = 2 sin(theta) P_n'(x) / (n (n + 1)),
// mapped to Cartesian components in the body frame and turned by R(q_b).  The reference
// evaluates this slip and leaves its addition to the right-hand side commented out
// (:502), so its squirmer has no slip; here it is applied.  Each step solves the force-
// and torque-free body system (body.hip) M F - K U = u_s, -K^T F = 0, and from the second
// step on first advances every body explicitly with the previous step's velocities,
// x += dt V, q = exp(dt Omega) q (cilia_sim_main.cu:724-740).  The slip is constant in
// the body frame: it is formed once on the host and turned by R(q_b) on the device at
// every step.
#include <cmath>
#include <vector>

#include "body.hpp"
#include "ctx.hpp"
#include "krylov.hpp"

namespace cfm {
namespace {

struct Squirmer {
  cfm_squirmer_params prm{};
  DevBuf<double> slip0;      // body-frame slip of one body's blobs [3][nb]
  DevBuf<double> rot;        // R(q_b) of every body (9 per body, row-major)
  DevBuf<double> rhs, x;     // body-system vectors (3 n_blob + 6 n_body)
  std::vector<double> vel;   // (V, Omega) per body of the last step
  bool have_vel = false;
};

Squirmer* sq_of(cfm_ctx* c) { return reinterpret_cast<Squirmer*>(c->squirmer); }

// P_n(x) and P_n'(x) by the three-term recurrences.
void legendre(int n, double x, double* p, double* dp) {
  double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
  if (n == 0) {
    *p = 1.0;
    *dp = 0.0;
    return;
  }
  for (int k = 2; k <= n; ++k) {
    const double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
    const double d2 = d0 + (2 * k - 1) * p1;
    p0 = p1;
    p1 = p2;
    d0 = d1;
    d1 = d2;
  }
  *p = p1;
  *dp = d1;
}

void quat_rot(const double* q, double R[9]) {
  const double s = q[0], a = q[1], b = q[2], cc = q[3];
  R[0] = 1.0 - 2.0 * (b * b + cc * cc);
  R[1] = 2.0 * (a * b - s * cc);
  R[2] = 2.0 * (a * cc + s * b);
  R[3] = 2.0 * (a * b + s * cc);
  R[4] = 1.0 - 2.0 * (a * a + cc * cc);
  R[5] = 2.0 * (b * cc - s * a);
  R[6] = 2.0 * (a * cc - s * b);
  R[7] = 2.0 * (b * cc + s * a);
  R[8] = 1.0 - 2.0 * (a * a + b * b);
}

// rhs blob rows: R(q_b) slip0 of the body's blobs; body rows zero.
__global__ void squirmer_rhs_kernel(int nb, int n_body, const double* __restrict__ slip0,
                                    const double* __restrict__ rot, double* __restrict__ rhs) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = nb * n_body;
  if (g < 6 * n_body) rhs[3 * (size_t)m + g] = 0.0;
  if (g >= m) return;
  const int b = g / nb, i = g - b * nb;
  const double* R = rot + 9 * (size_t)b;
  const double u[3] = {slip0[i], slip0[nb + i], slip0[2 * (size_t)nb + i]};
  for (int k = 0; k < 3; ++k) rhs[(size_t)k * m + g] = R[3 * k] * u[0] + R[3 * k + 1] * u[1] + R[3 * k + 2] * u[2];
}

}  // namespace

void cfm_squirmer_release(cfm_ctx* c) {
  Squirmer* S = sq_of(c);
  if (!S) return;
  S->slip0.release();
  S->rot.release();
  S->rhs.release();
  S->x.release();
  delete S;
  c->squirmer = nullptr;
}

}  // namespace cfm

using namespace cfm;

static int sfail(cfm_ctx* c, const Status& s) {
  c->err = s.msg;
  return s.code;
}

extern "C" int cfm_squirmer_init(cfm_ctx* c, const cfm_squirmer_params* p) {
  if (!c || !p) return CFM_EINVAL;
  if (p->n_modes < 1 || p->n_modes > 8) return sfail(c, make_status(CFM_EINVAL, "n_modes must be in [1, 8]"));
  if (!(p->dt > 0.0) || !(p->solve.rtol > 0.0) || p->solve.restart < 0 || p->solve.max_iter < 0)
    return sfail(c, make_status(CFM_EINVAL, "need dt > 0 and valid solve options"));
  for (int n = 0; n < p->n_modes; ++n)
    if (!std::isfinite(p->A[n]) || !std::isfinite(p->B[n]))
      return sfail(c, make_status(CFM_ENONFINITE, "non-finite slip mode"));
  if (c->p.n_seg != 0) return sfail(c, make_status(CFM_EINVAL, "squirmer contexts hold blobs only"));
  BodyInfo bi;
  Status st = body_info(c, &bi);
  if (!st.ok()) return sfail(c, st);
  (void)hipSetDevice(c->device);
  cfm_squirmer_release(c);
  Squirmer* S = new Squirmer();
  c->squirmer = S;
  S->prm = *p;
  auto run = [&]() -> Status {
    const int nb = bi.nb;
    std::vector<double> u(3 * (size_t)nb);
    for (int i = 0; i < nb; ++i) {
      const double* r = &bi.ref[3 * (size_t)i];
      const double rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      if (!(rn > 0.0)) return make_status(CFM_EINVAL, "blob %d sits at the body centre", i);
      const double phi = std::atan2(r[1], r[0]);
      const double th = std::acos(std::fmax(-1.0, std::fmin(1.0, r[2] / rn)));
      const double x = std::cos(th), s = std::sin(th);
      double ur = 0.0, ut = 0.0;
      for (int n = 1; n <= p->n_modes; ++n) {
        double pn = 0.0, dpn = 0.0;
        legendre(n, x, &pn, &dpn);
        ur += p->A[n - 1] * pn;
        ut += p->B[n - 1] * 2.0 * s * dpn / (n * (n + 1));
      }
      u[i] = ur * s * std::cos(phi) + ut * x * std::cos(phi);
      u[nb + i] = ur * s * std::sin(phi) + ut * x * std::sin(phi);
      u[2 * (size_t)nb + i] = ur * x - ut * s;
    }
    CFM_TRY(S->slip0.alloc(u.size()));
    CFM_HIP(copy_to_device(c->stream, S->slip0.p, u.data(), sizeof(double) * u.size()));
    const size_t L = 3 * (size_t)c->p.n_blob + 6 * (size_t)bi.n_body;
    CFM_TRY(S->rot.alloc(9 * (size_t)bi.n_body));
    CFM_TRY(S->rhs.alloc(L));
    CFM_TRY(S->x.alloc(L));
    CFM_TRY(dev_fill(c, S->x.p, 0.0, (int)L));
    S->vel.assign(6 * (size_t)bi.n_body, 0.0);
    return {};
  };
  st = run();
  if (!st.ok()) {
    cfm_squirmer_release(c);
    return sfail(c, st);
  }
  return CFM_OK;
}

extern "C" int cfm_squirmer_step(cfm_ctx* c, int32_t nt, double* vel, cfm_solve_stats* stats) {
  if (!c) return CFM_EINVAL;
  Squirmer* S = sq_of(c);
  if (!S) return sfail(c, make_status(CFM_ESTATE, "cfm_squirmer_init has not been called"));
  if (nt < 0) return sfail(c, make_status(CFM_EINVAL, "nt must be >= 0"));
  (void)hipSetDevice(c->device);
  auto run = [&]() -> Status {
    BodyInfo bi;
    CFM_TRY(body_info(c, &bi));
    const int B = bi.n_body;
    if (nt > 0 && S->have_vel) {
      // explicit update with the previous step's velocities (cilia_sim_main.cu:724-740)
      const double dt = S->prm.dt;
      std::vector<double> x = bi.x, q = bi.q;
      for (int b = 0; b < B; ++b) {
        const double* v = &S->vel[6 * (size_t)b];
        for (int k = 0; k < 3; ++k) x[3 * b + k] += dt * v[k];
        const double u[3] = {dt * v[3], dt * v[4], dt * v[5]};
        const double th = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        double e[4] = {1.0, 0.0, 0.0, 0.0};
        if (th > 0.0) {
          const double sh = std::sin(0.5 * th) / th;
          e[0] = std::cos(0.5 * th);
          e[1] = sh * u[0];
          e[2] = sh * u[1];
          e[3] = sh * u[2];
        }
        double* qb = &q[4 * (size_t)b];
        const double w = e[0] * qb[0] - e[1] * qb[1] - e[2] * qb[2] - e[3] * qb[3];
        const double a = e[0] * qb[1] + e[1] * qb[0] + e[2] * qb[3] - e[3] * qb[2];
        const double bb = e[0] * qb[2] - e[1] * qb[3] + e[2] * qb[0] + e[3] * qb[1];
        const double cc = e[0] * qb[3] + e[1] * qb[2] - e[2] * qb[1] + e[3] * qb[0];
        const double nq = std::sqrt(w * w + a * a + bb * bb + cc * cc);
        qb[0] = w / nq;
        qb[1] = a / nq;
        qb[2] = bb / nq;
        qb[3] = cc / nq;
      }
      if (cfm_body_set_state(c, x.data(), q.data()) != CFM_OK) return make_status(CFM_EHIP, "%s", c->err.c_str());
      bi.x = x;
      bi.q = q;
    }
    std::vector<double> R(9 * (size_t)B);
    for (int b = 0; b < B; ++b) quat_rot(&bi.q[4 * (size_t)b], &R[9 * (size_t)b]);
    CFM_HIP(copy_to_device(c->stream, S->rot.p, R.data(), sizeof(double) * R.size()));
    const int m = c->p.n_blob;
    hipLaunchKernelGGL(squirmer_rhs_kernel, dim3((std::max(m, 6 * B) + 255) / 256), dim3(256), 0, c->stream,
                       bi.nb, B, S->slip0.p, S->rot.p, S->rhs.p);
    CFM_CHECK_LAUNCH();
    SystemDesc s;
    CFM_TRY(body_system_desc(c, &s));
    cfm_solve_opts o = S->prm.solve;
    o.system = CFM_SYS_BODY;
    o.use_guess = S->have_vel ? 1 : 0;
    if (!c->krylov) c->krylov = new Krylov();
    cfm_solve_stats st{};
    CFM_TRY(solve(c, *c->krylov, s, o, S->rhs.p, S->x.p, &st, nullptr));
    CFM_HIP(copy_to_host(c->stream, S->vel.data(), S->x.p + 3 * (size_t)m, sizeof(double) * 6 * B));
    S->have_vel = true;
    if (vel) std::copy(S->vel.begin(), S->vel.end(), vel);
    if (stats) *stats = st;
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : sfail(c, s);
}
