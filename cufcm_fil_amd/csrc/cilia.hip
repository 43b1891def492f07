// cilia.hip -- prescribed-shape cilia above the wall (fitted Fulford-Blake beat).
//
// Restates the reference's CILIA_TYPE 3 path with WRITE_GENERALISED_FORCES (the
// mode that wrote its input/forcing reference files): every cilium follows the
// fitted beat at a prescribed phase speed, the segment forces solve the wall-RPY
// resistance problem M F = phase_dot * vel_dir_phase (3 unknowns per segment), and
// each cilium reports its phase and angle generalised forces sum_n vel_dir . F_n.
//   beat coefficients      filament.cpp:684-758 (Ito, Omori & Ishikawa 2019, Table 1)
//   shape / derivatives    filament.cpp:964-1045
//   segment parameters     filament.cpp:1047-1160 (trapezoid arc length, bisection)
//   placement              filament.cpp:1336-1490, base frame filament.cpp:225-226
//   rhs / generalised f.   mobility_solver.cpp:301-420, cilia_sim_main.cu:405-438
// One thread per cilium evaluates the geometry (the segment parameters need a
// sequential bisection per segment); the solve is the library's segment GMRES.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "ctx.hpp"
#include "krylov.hpp"

namespace cfm {

Status upload_aos(cfm_ctx* c, const double* host, int cnt, int dof, double* dst, int ld, int off);
Status download_aos(cfm_ctx* c, double* host, int cnt, int dof, const double* src, int ld, int off);

constexpr int kMaxCiliaSeg = 64;

// Fourier modes x polynomial degree (s, s^2, s^3); x along the filament axis of the
// reference frame, y along the beat direction.  Values of Ito, Omori & Ishikawa
// (2019) Table 1 as the reference uses them (rescaled to unit mean length, zero
// mode halved).
struct BeatCoeffs {
  double ax[4][3], ay[4][3], bx[3][3], by[3][3];
};

static BeatCoeffs fulford_blake() {
  BeatCoeffs b{};
  const double ax[4][3] = {{9.7204e-01, -2.8315e-01, 4.9243e-02},
                           {-1.8466e-02, -1.2926e-01, 2.6981e-01},
                           {1.6209e-01, -3.4983e-01, 1.9082e-01},
                           {1.0259e-02, 3.5907e-02, -6.8736e-02}};
  const double ay[4][3] = {{-3.3547e-01, 4.0369e-01, 1.0362e-01},
                           {4.0318e-01, -1.5553e+00, 7.3455e-01},
                           {-9.9513e-02, 3.2829e-02, -1.2106e-01},
                           {8.1046e-02, -3.0982e-01, 1.4568e-01}};
  const double bx[3][3] = {{1.9697e-01, -5.1193e-01, 3.4778e-01},
                           {-5.1295e-02, 4.3396e-01, -3.3547e-01},
                           {1.2311e-02, 1.4157e-01, -1.1695e-01}};
  const double by[3][3] = {{2.9136e-01, 1.0721e+00, -1.0433e+00},
                           {6.1554e-03, 3.2521e-01, -2.8315e-01},
                           {-6.0528e-02, 2.3185e-01, -2.0108e-01}};
  std::memcpy(b.ax, ax, sizeof ax);
  std::memcpy(b.ay, ay, sizeof ay);
  std::memcpy(b.bx, bx, sizeof bx);
  std::memcpy(b.by, by, sizeof by);
  return b;
}

struct CiliaK {
  BeatCoeffs cf;
  int ncil, nseg, nseg_total;
  double dl, len;  // segment spacing, filament length dl (nseg - 1)
  const double* base;    // 3 x ncil (SoA)
  const double* phase;   // ncil
  const double* angle;   // ncil (rotation of the beat about the reference z axis)
  double *x, *y, *z;     // segment positions (the context's particle arrays)
  double* vphase;        // 3 x nseg_total
  double* vangle;        // 3 x nseg_total
  double* svals;         // ncil x nseg
};

// Mode vectors (1, cos psi, cos 2psi, cos 3psi) and (sin psi, sin 2psi, sin 3psi).
__device__ inline void modes(double psi, double cv[4], double sv[3]) {
  cv[0] = 1.0;
  for (int n = 1; n < 4; ++n) {
    cv[n] = cos(n * psi);
    sv[n - 1] = sin(n * psi);
  }
}

// sum_m (cv_m A_m + sv_m B_m) . p over the three polynomial degrees
__device__ inline double combine(const double A[4][3], const double B[3][3], const double cv[4],
                                 const double sv[3], const double p[3]) {
  double r = 0.0;
  for (int d = 0; d < 3; ++d) {
    double a = 0.0;
    for (int m = 0; m < 4; ++m) a += cv[m] * A[m][d];
    for (int m = 0; m < 3; ++m) a += sv[m] * B[m][d];
    r += a * p[d];
  }
  return r;
}

__device__ inline double tangent_norm(const BeatCoeffs& cf, const double cv[4], const double sv[3],
                                      double s) {
  const double p[3] = {1.0, 2.0 * s, 3.0 * s * s};
  const double tx = combine(cf.ax, cf.bx, cv, sv, p);
  const double ty = combine(cf.ay, cf.by, cv, sv, p);
  return sqrt(tx * tx + ty * ty);
}

// Trapezoid arc length on 10 nseg panels, the last panel interpolated.
__device__ double curve_length(const BeatCoeffs& cf, const double cv[4], const double sv[3],
                               double s, int nseg) {
  if (s <= 0.0) return 0.0;
  const double dl = 1.0 / (10.0 * nseg);
  double f0 = tangent_norm(cf, cv, sv, 0.0);
  const double fl = floor(s / dl);
  double length = 0.0;
  for (int n = 1; n <= (int)fl; ++n) {
    const double f1 = tangent_norm(cf, cv, sv, n * dl);
    length += 0.5 * dl * (f0 + f1);
    f0 = f1;
  }
  const double ce = ceil(s / dl);
  if (ce > fl) {
    const double f1 = tangent_norm(cf, cv, sv, ce * dl);
    length += ((s / dl - fl) / (ce - fl)) * 0.5 * dl * (f0 + f1);
  }
  return length;
}

__global__ void cilia_geometry_kernel(CiliaK k) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k.ncil) return;
  const int n = k.nseg, s0 = c * n;
  const double psi = k.phase[c];
  double cv[4], sv[3];
  modes(psi, cv, sv);
  // segment parameters: bisection on the arc-length fraction
  double s[kMaxCiliaSeg];
  s[0] = 0.0;
  s[n - 1] = 1.0;
  const double total = curve_length(k.cf, cv, sv, 1.0, n);
  for (int i = 1; i < n - 1; ++i) {
    const double target = double(i) / double(n - 1);
    double lo = s[i - 1], hi = 1.0;
    double est = 0.5 * (lo + hi);
    double frac = curve_length(k.cf, cv, sv, est, n) / total;
    for (int it = 0; it < 200 && fabs(frac - target) > 0.1 / n; ++it) {
      if (frac > target)
        hi = est;
      else
        lo = est;
      est = 0.5 * (lo + hi);
      frac = curve_length(k.cf, cv, sv, est, n) / total;
    }
    s[i] = est;
  }
  // placement: reference x -> lab z, reference y -> lab y, reference z -> lab -x
  const double bxp = k.base[c], byp = k.base[k.ncil + c], bzp = k.base[2 * k.ncil + c];
  const double ca = cos(k.angle[c]), sa = sin(k.angle[c]);
  const size_t ns = k.nseg_total;
  double dc[4], ds[3];
  dc[0] = 0.0;
  for (int m = 1; m < 4; ++m) {
    dc[m] = -m * sin(m * psi);
    ds[m - 1] = m * cos(m * psi);
  }
  for (int i = 0; i < n; ++i) {
    const int g = s0 + i;
    k.svals[(size_t)c * n + i] = s[i];
    if (i == 0) {
      k.x[g] = bxp;
      k.y[g] = byp;
      k.z[g] = bzp;
      for (int d = 0; d < 3; ++d) {
        k.vphase[d * ns + g] = 0.0;
        k.vangle[d * ns + g] = 0.0;
      }
      continue;
    }
    const double p[3] = {s[i], s[i] * s[i], s[i] * s[i] * s[i]};
    // shape in the reference frame, rotated by the beat angle about its z axis
    const double sx = combine(k.cf.ax, k.cf.bx, cv, sv, p);
    const double sy = combine(k.cf.ay, k.cf.by, cv, sv, p);
    const double rx = k.len * (ca * sx - sa * sy), ry = k.len * (sa * sx + ca * sy);
    k.x[g] = bxp;
    k.y[g] = byp + ry;
    k.z[g] = bzp + rx;
    // e_z x ref in the reference frame is (-ry, rx, 0) -> lab (0, rx, -ry)
    k.vangle[g] = 0.0;
    k.vangle[ns + g] = rx;
    k.vangle[2 * ns + g] = -ry;
    const double dx = combine(k.cf.ax, k.cf.bx, dc, ds, p);
    const double dy = combine(k.cf.ay, k.cf.by, dc, ds, p);
    const double ux = k.len * (ca * dx - sa * dy), uy = k.len * (sa * dx + ca * dy);
    k.vphase[g] = 0.0;
    k.vphase[ns + g] = uy;
    k.vphase[2 * ns + g] = ux;
  }
}

// rhs = phase_dot vel_dir_phase ([3][n_seg] system layout)
__global__ void cilia_rhs_kernel(int nseg_total, int nseg, const double* __restrict__ pdot,
                                 const double* __restrict__ vphase, double* __restrict__ rhs) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nseg_total) return;
  const double w = pdot[g / nseg];
  for (int d = 0; d < 3; ++d) rhs[(size_t)d * nseg_total + g] = w * vphase[(size_t)d * nseg_total + g];
}

// q_phase, q_angle per cilium (fixed order over segments and components)
__global__ void cilia_genforce_kernel(int ncil, int nseg, int nseg_total,
                                      const double* __restrict__ vphase,
                                      const double* __restrict__ vangle,
                                      const double* __restrict__ f, double* __restrict__ q) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncil) return;
  double qp = 0.0, qa = 0.0;
  for (int i = 0; i < nseg; ++i) {
    const size_t g = (size_t)c * nseg + i;
    for (int d = 0; d < 3; ++d) {
      const double fv = f[d * (size_t)nseg_total + g];
      qp += vphase[d * (size_t)nseg_total + g] * fv;
      qa += vangle[d * (size_t)nseg_total + g] * fv;
    }
  }
  q[c] = qp;
  q[ncil + c] = qa;
}

__global__ void cilia_advance_kernel(int ncil, double dt, const double* __restrict__ pdot,
                                     double* __restrict__ phase) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncil) phase[c] += pdot[c] * dt;
}

struct CiliaState {
  cfm_cilia_params prm{};
  DevBuf<double> base, phase, pdot, angle, vphase, vangle, svals, rhs, f, q;
  double* host = nullptr;  // pinned read-back of q
  bool have_geometry = false;
};

static CiliaState* cil_of(cfm_ctx* c) { return reinterpret_cast<CiliaState*>(c->cilia); }

static CiliaK kargs(cfm_ctx* c, CiliaState* S) {
  CiliaK k{};
  k.cf = fulford_blake();
  k.ncil = S->prm.n_cilia;
  k.nseg = S->prm.nseg;
  k.nseg_total = c->p.n_seg;
  k.dl = S->prm.dl;
  k.len = S->prm.dl * (S->prm.nseg - 1);
  k.base = S->base.p;
  k.phase = S->phase.p;
  k.angle = S->angle.p;
  k.x = c->p.x.p;
  k.y = c->p.y.p;
  k.z = c->p.z.p;
  k.vphase = S->vphase.p;
  k.vangle = S->vangle.p;
  k.svals = S->svals.p;
  return k;
}

static inline dim3 gc(int n) { return dim3((unsigned)((n + 63) / 64)); }

static Status geometry(cfm_ctx* c, CiliaState* S) {
  hipLaunchKernelGGL(cilia_geometry_kernel, gc(S->prm.n_cilia), dim3(64), 0, c->stream, kargs(c, S));
  CFM_CHECK_LAUNCH();
  c->have_positions = true;
  S->have_geometry = true;
  return {};
}

}  // namespace cfm

using namespace cfm;

void cfm_cilia_release(cfm_ctx* c) {
  CiliaState* S = cil_of(c);
  if (!S) return;
  for (auto* b : {&S->base, &S->phase, &S->pdot, &S->angle, &S->vphase, &S->vangle, &S->svals,
                  &S->rhs, &S->f, &S->q})
    b->release();
  if (S->host) (void)hipHostFree(S->host);
  delete S;
  c->cilia = nullptr;
}

static int cfail(cfm_ctx* c, const Status& s) {
  c->err = s.msg;
  return s.code;
}

extern "C" {

int cfm_cilia_init(cfm_ctx* c, const cfm_cilia_params* prm, const double* base,
                   const double* phase, const double* phase_dot) {
  if (!c) return CFM_EINVAL;
  if (!prm || !base || !phase || !phase_dot)
    return cfail(c, make_status(CFM_EINVAL, "null argument"));
  const Particles& p = c->p;
  if (prm->beat != CFM_BEAT_FULFORD_BLAKE)
    return cfail(c, make_status(CFM_EINVAL, "unknown beat %d", prm->beat));
  if (prm->nseg < 2 || prm->nseg > kMaxCiliaSeg || prm->n_cilia <= 0 ||
      (long long)prm->n_cilia * prm->nseg != p.n_seg)
    return cfail(c, make_status(CFM_EINVAL, "n_cilia x nseg must equal the context's n_seg, 2 <= nseg <= %d",
                                kMaxCiliaSeg));
  if (p.n_blob != 0 || p.seg_dof != 3 || !c->cfg.wall)
    return cfail(c, make_status(CFM_EINVAL, "prescribed cilia need a wall context with 3-DOF segments and no blobs"));
  if (!(prm->dl > 0.0) || !(prm->dt > 0.0))
    return cfail(c, make_status(CFM_EINVAL, "dl and dt must be positive"));
  if (prm->solve.precond < CFM_PC_NONE || prm->solve.precond > CFM_PC_SELF)
    return cfail(c, make_status(CFM_EINVAL, "cilia solves take CFM_PC_NONE, _DIAG or _SELF"));
  const int nc = prm->n_cilia;
  for (int i = 0; i < nc; ++i) {
    if (!std::isfinite(phase[i]) || !std::isfinite(phase_dot[i]))
      return cfail(c, make_status(CFM_EINVAL, "non-finite phase or phase speed"));
    for (int d = 0; d < 3; ++d)
      if (!std::isfinite(base[3 * i + d])) return cfail(c, make_status(CFM_EINVAL, "non-finite base"));
    if (!(base[3 * i + 2] > 0.0))
      return cfail(c, make_status(CFM_EINVAL, "cilium %d base is not above the wall", i));
  }
  (void)hipSetDevice(c->device);
  cfm_cilia_release(c);
  CiliaState* S = new (std::nothrow) CiliaState();
  if (!S) return CFM_ENOMEM;
  c->cilia = S;
  S->prm = *prm;
  auto run = [&]() -> Status {
    const size_t ns = p.n_seg;
    CFM_TRY(S->base.alloc(3 * (size_t)nc));
    CFM_TRY(S->phase.alloc(nc));
    CFM_TRY(S->pdot.alloc(nc));
    CFM_TRY(S->angle.alloc(nc));
    CFM_TRY(S->vphase.alloc(3 * ns));
    CFM_TRY(S->vangle.alloc(3 * ns));
    CFM_TRY(S->svals.alloc((size_t)nc * prm->nseg));
    CFM_TRY(S->rhs.alloc(3 * ns));
    CFM_TRY(S->f.alloc(3 * ns));
    CFM_TRY(S->q.alloc(2 * (size_t)nc));
    CFM_HIP(hipHostMalloc((void**)&S->host, sizeof(double) * 2 * nc, hipHostMallocDefault));
    CFM_TRY(upload_aos(c, base, nc, 3, S->base.p, nc, 0));
    CFM_TRY(upload_aos(c, phase, nc, 1, S->phase.p, nc, 0));
    CFM_TRY(upload_aos(c, phase_dot, nc, 1, S->pdot.p, nc, 0));
    CFM_TRY(dev_fill(c, S->angle.p, 0.0, nc));
    CFM_TRY(dev_fill(c, S->f.p, 0.0, 3 * p.n_seg));
    CFM_TRY(geometry(c, S));
    CFM_HIP(hipStreamSynchronize(c->stream));
    return {};
  };
  Status s = run();
  if (!s.ok()) {
    cfm_cilia_release(c);
    return cfail(c, s);
  }
  return CFM_OK;
}

int cfm_cilia_step(cfm_ctx* c, int32_t nt, double* q_phase, double* q_angle, cfm_solve_stats* st) {
  if (!c) return CFM_EINVAL;
  CiliaState* S = cil_of(c);
  if (!S) return cfail(c, make_status(CFM_ESTATE, "cfm_cilia_init has not been called"));
  (void)hipSetDevice(c->device);
  const int nc = S->prm.n_cilia;
  auto run = [&]() -> Status {
    if (nt > 0)  // phase += phase_dot dt before the step (filament.cpp:1340-1343)
      hipLaunchKernelGGL(cilia_advance_kernel, gc(nc), dim3(64), 0, c->stream, nc, S->prm.dt,
                         S->pdot.p, S->phase.p);
    CFM_TRY(geometry(c, S));
    const SystemDesc sd = system_desc(c, CFM_SYS_SEG);
    hipLaunchKernelGGL(cilia_rhs_kernel, dim3((c->p.n_seg + 255) / 256), dim3(256), 0, c->stream,
                       c->p.n_seg, S->prm.nseg, S->pdot.p, S->vphase.p, S->rhs.p);
    CFM_CHECK_LAUNCH();
    cfm_solve_opts o = S->prm.solve;
    o.system = CFM_SYS_SEG;
    o.use_guess = 1;  // previous step's forces
    if (!c->krylov) c->krylov = new Krylov();
    cfm_solve_stats ss{};
    CFM_TRY(solve(c, *c->krylov, sd, o, S->rhs.p, S->f.p, &ss, nullptr));
    hipLaunchKernelGGL(cilia_genforce_kernel, gc(nc), dim3(64), 0, c->stream, nc, S->prm.nseg,
                       c->p.n_seg, S->vphase.p, S->vangle.p, S->f.p, S->q.p);
    CFM_CHECK_LAUNCH();
    CFM_HIP(hipMemcpyAsync(S->host, S->q.p, sizeof(double) * 2 * nc, hipMemcpyDeviceToHost, c->stream));
    CFM_HIP(hipStreamSynchronize(c->stream));
    if (q_phase) std::memcpy(q_phase, S->host, sizeof(double) * nc);
    if (q_angle) std::memcpy(q_angle, S->host + nc, sizeof(double) * nc);
    if (st) *st = ss;
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : cfail(c, s);
}

int cfm_cilia_get_state(cfm_ctx* c, double* x, double* vel_dir_phase, double* vel_dir_angle,
                        double* s_values, double* phase, double* f_seg) {
  if (!c) return CFM_EINVAL;
  CiliaState* S = cil_of(c);
  if (!S) return cfail(c, make_status(CFM_ESTATE, "cfm_cilia_init has not been called"));
  (void)hipSetDevice(c->device);
  const int ns = c->p.n_seg, nc = S->prm.n_cilia;
  auto run = [&]() -> Status {
    if (x) {
      std::vector<double> t(3 * (size_t)ns);
      CFM_HIP(hipMemcpy(t.data(), c->p.x.p, sizeof(double) * ns, hipMemcpyDeviceToHost));
      CFM_HIP(hipMemcpy(t.data() + ns, c->p.y.p, sizeof(double) * ns, hipMemcpyDeviceToHost));
      CFM_HIP(hipMemcpy(t.data() + 2 * (size_t)ns, c->p.z.p, sizeof(double) * ns, hipMemcpyDeviceToHost));
      for (int i = 0; i < ns; ++i)
        for (int d = 0; d < 3; ++d) x[3 * i + d] = t[(size_t)d * ns + i];
    }
    if (vel_dir_phase) CFM_TRY(download_aos(c, vel_dir_phase, ns, 3, S->vphase.p, ns, 0));
    if (vel_dir_angle) CFM_TRY(download_aos(c, vel_dir_angle, ns, 3, S->vangle.p, ns, 0));
    if (f_seg) CFM_TRY(download_aos(c, f_seg, ns, 3, S->f.p, ns, 0));
    if (s_values)
      CFM_HIP(hipMemcpy(s_values, S->svals.p, sizeof(double) * nc * S->prm.nseg, hipMemcpyDeviceToHost));
    if (phase) CFM_HIP(hipMemcpy(phase, S->phase.p, sizeof(double) * nc, hipMemcpyDeviceToHost));
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : cfail(c, s);
}

}  // extern "C"
