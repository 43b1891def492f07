// cilia.hip -- prescribed-shape cilia above the wall (fitted Fulford-Blake beat).
//
// Restates the reference's CILIA_TYPE 3 path.  With WRITE_GENERALISED_FORCES (the
// mode that wrote its input/forcing reference files) every cilium follows the fitted
// beat at a prescribed phase speed, the segment forces solve the wall-RPY resistance
// problem M F = phase_dot * vel_dir_phase (3 unknowns per segment), and each cilium
// reports its phase and angle generalised forces sum_n vel_dir . F_n.  With
// DYNAMIC_PHASE_EVOLUTION (and DYNAMIC_SHAPE_ROTATION) the phase (and shape-angle)
// speeds are unknowns of a coupled system driven by recorded generalised forces
// (mobility_solver.cpp:242-356 driving, :1891-2115 operator, :1104-1244 block
// preconditioner).
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
#include "precond.hpp"

namespace cfm {

Status upload_aos(cfm_ctx* c, const double* host, int cnt, int dof, double* dst, int ld, int off);
Status download_aos(cfm_ctx* c, double* host, int cnt, int dof, const double* src, int ld, int off);

constexpr int kMaxCiliaSeg = 64;

// Fourier modes x polynomial degree (s, s^2, s^3); x along the filament axis of the
// reference frame, y along the beat direction; cos0 weights the zero cosine mode.
struct BeatCoeffs {
  double ax[4][3], ay[4][3], bx[3][3], by[3][3];
  double cos0;
};

// Ito, Omori & Ishikawa (2019) Table 1 as SHAPE_SEQUENCE 1 uses it: rescaled to unit
// mean length, zero mode halved into the table (filament.cpp:684-758).
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
  b.cos0 = 1.0;
  return b;
}

// The unaltered Table 1 values (SHAPE_SEQUENCE 4, filament.cpp:813-868) with the zero
// cosine mode weighted 0.5 at evaluation (filament.cpp:977-979, 1006-1008).
static BeatCoeffs fulford_blake_original() {
  BeatCoeffs b{};
  const double ax[4][3] = {{1.895, -0.552, 0.096}, {-0.018, -0.126, 0.263}, {0.158, -0.341, 0.186},
                           {0.010, 0.035, -0.067}};
  const double ay[4][3] = {{-0.654, 0.787, 0.202}, {0.393, -1.516, 0.716}, {-0.097, 0.032, -0.118},
                           {0.079, -0.302, 0.142}};
  const double bx[3][3] = {{0.192, -0.499, 0.339}, {-0.050, 0.423, -0.327}, {0.012, 0.138, -0.114}};
  const double by[3][3] = {{0.284, 1.045, -1.017}, {0.006, 0.317, -0.276}, {-0.059, 0.226, -0.196}};
  std::memcpy(b.ax, ax, sizeof ax);
  std::memcpy(b.ay, ay, sizeof ay);
  std::memcpy(b.bx, bx, sizeof bx);
  std::memcpy(b.by, by, sizeof by);
  b.cos0 = 0.5;
  return b;
}

struct CiliaK {
  BeatCoeffs cf;
  int ncil, nseg, nseg_total;
  double dl, len;  // segment spacing, filament length dl (nseg - 1)
  double s_tol;    // bisection tolerance on the arc-length fraction
  const double* s_table;  // recorded segment parameters (rows x cols) or null
  int s_rows, s_cols;
  const double* base;    // 3 x ncil (SoA)
  const double* phase;   // ncil
  const double* angle;   // ncil (rotation of the beat about the reference z axis)
  double *x, *y, *z;     // segment positions (the context's particle arrays)
  double* vphase;        // 3 x nseg_total
  double* vangle;        // 3 x nseg_total
  double* svals;         // ncil x nseg
};

// Mode vectors (cos0, cos psi, cos 2psi, cos 3psi) and (sin psi, sin 2psi, sin 3psi).
__device__ inline void modes(double psi, double cos0, double cv[4], double sv[3]) {
  cv[0] = cos0;
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
  modes(psi, k.cf.cos0, cv, sv);
  double s[kMaxCiliaSeg];
  s[0] = 0.0;
  s[n - 1] = 1.0;
  if (k.s_table) {
    // recorded parameters, linear in phase between the table's rows and linear in
    // segment position between its columns (filament.cpp:1190-1235)
    double x = 0.5 * psi / 3.14159265358979323846;
    x = (x - floor(x)) * k.s_rows;
    const int lo = (int)x;
    const double *a = k.s_table, *b = k.s_table;
    double w = 0.0;
    if (lo < k.s_rows) {  // lo == s_rows only when rounding lands x on s_rows: row 0
      a = k.s_table + (size_t)lo * k.s_cols;
      b = k.s_table + (size_t)(lo == k.s_rows - 1 ? 0 : lo + 1) * k.s_cols;
      w = x - lo;
    }
    for (int i = 1; i < n - 1; ++i) {
      const double si = (double)(k.s_cols - 1) * i / (double)(n - 1);
      const int j = (int)si;
      const double vj = a[j] + (b[j] - a[j]) * w;
      const double vj1 = a[j + 1] + (b[j + 1] - a[j + 1]) * w;
      s[i] = vj + (si - j) * (vj1 - vj);
    }
  } else {
    // bisection on the arc-length fraction (filament.cpp:1096-1185)
    const double total = curve_length(k.cf, cv, sv, 1.0, n);
    for (int i = 1; i < n - 1; ++i) {
      const double target = double(i) / double(n - 1);
      double lo = s[i - 1], hi = 1.0;
      double est = 0.5 * (lo + hi);
      double frac = curve_length(k.cf, cv, sv, est, n) / total;
      for (int it = 0; it < 200 && fabs(frac - target) > k.s_tol; ++it) {
        if (frac > target)
          hi = est;
        else
          lo = est;
        est = 0.5 * (lo + hi);
        frac = curve_length(k.cf, cv, sv, est, n) / total;
      }
      s[i] = est;
    }
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

// ---- coupled system: unknowns [F (3 x n_seg)] ++ [phase speeds] ++ [angle speeds] ----

struct CoupleK {
  int ncil, nseg, m;        // m = n_seg of the context
  int dyn_angle;
  const double* vphase;     // 3 x m
  const double* vangle;     // 3 x m
};

// y_F -= phi_dot K + psi_dot A (own segments); y_phi = -K . F, y_psi = -A . F
// (mobility_solver.cpp:2002-2050 with the body velocities prescribed).
__global__ void cilia_couple_kernel(CoupleK k, const double* __restrict__ x, double* __restrict__ y) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k.ncil) return;
  const size_t m = k.m;
  const double wp = x[3 * m + c];
  const double wa = k.dyn_angle ? x[3 * m + k.ncil + c] : 0.0;
  double dp = 0.0, da = 0.0;
  for (int i = 0; i < k.nseg; ++i) {
    const size_t g = (size_t)c * k.nseg + i;
    for (int d = 0; d < 3; ++d) {
      const double vp = k.vphase[d * m + g], va = k.vangle[d * m + g];
      const double fv = x[d * m + g];
      dp += vp * fv;
      da += va * fv;
      y[d * m + g] -= wp * vp + (k.dyn_angle ? wa * va : 0.0);
    }
  }
  y[3 * m + c] = -dp;
  if (k.dyn_angle) y[3 * m + k.ncil + c] = -da;
}

// Block preconditioner: F = P (r1 + K w), (K^T P K) w = -r2 - K^T P r1 with P the
// particle preconditioner (6 pi mu a: exactly mobility_solver.cpp:1104-1200; or the
// wall self-block inverse), one 1x1 or 2x2 Schur block per cilium.
__global__ void cilia_precond_kernel(CoupleK k, int kind, double mu, const double* __restrict__ z,
                                     const double* __restrict__ r, double* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k.ncil) return;
  const size_t m = k.m;
  double s11 = 0.0, s12 = 0.0, s22 = 0.0, dp = 0.0, da = 0.0;
  for (int i = 0; i < k.nseg; ++i) {
    const size_t g = (size_t)c * k.nseg + i;
    double rv[3], vp[3], va[3], pr[3], pp[3], pa[3];
    for (int d = 0; d < 3; ++d) {
      rv[d] = r[d * m + g];
      vp[d] = k.vphase[d * m + g];
      va[d] = k.vangle[d * m + g];
    }
    particle_precond(kind, 1, 3, z[g], 1.0, mu, rv, pr);
    particle_precond(kind, 1, 3, z[g], 1.0, mu, vp, pp);
    particle_precond(kind, 1, 3, z[g], 1.0, mu, va, pa);
    for (int d = 0; d < 3; ++d) {
      s11 += vp[d] * pp[d];
      s12 += vp[d] * pa[d];
      s22 += va[d] * pa[d];
      dp += vp[d] * pr[d];
      da += va[d] * pr[d];
      out[d * m + g] = pr[d];
    }
  }
  const double tp = -r[3 * m + c] - dp;
  double wp, wa = 0.0;
  if (k.dyn_angle) {
    const double ta = -r[3 * m + k.ncil + c] - da;
    const double det = s11 * s22 - s12 * s12;
    wp = (s22 * tp - s12 * ta) / det;
    wa = (s11 * ta - s12 * tp) / det;
    out[3 * m + k.ncil + c] = wa;
  } else {
    wp = tp / s11;
  }
  out[3 * m + c] = wp;
  for (int i = 0; i < k.nseg; ++i) {
    const size_t g = (size_t)c * k.nseg + i;
    double vp[3], va[3], pp[3], pa[3];
    for (int d = 0; d < 3; ++d) {
      vp[d] = k.vphase[d * m + g];
      va[d] = k.vangle[d * m + g];
    }
    particle_precond(kind, 1, 3, z[g], 1.0, mu, vp, pp);
    particle_precond(kind, 1, 3, z[g], 1.0, mu, va, pa);
    for (int d = 0; d < 3; ++d) out[d * m + g] += wp * pp[d] + (k.dyn_angle ? wa * pa[d] : 0.0);
  }
}

// Periodic cubic Hermite interpolation of a table sampled at 2 pi k / n with
// central-difference slopes (mobility_solver.cpp:242-281).
__device__ double cyclic_hermite(const double* tab, int n, double phase) {
  double x = 0.5 * phase / 3.14159265358979323846;
  x = (x - floor(x)) * n;
  const int lo = (int)x;
  if (lo == n) return tab[0];
  const int hi = (lo == n - 1) ? 0 : lo + 1;
  const double d_lo = (lo == 0) ? (tab[1] - tab[0]) : 0.5 * (tab[hi] - tab[lo - 1]);
  const double d_hi = (hi == 0) ? (tab[0] - tab[n - 1])
                                : 0.5 * (tab[(hi == n - 1) ? 0 : hi + 1] - tab[lo]);
  const double t = x - lo, t2 = t * t, t3 = t2 * t;
  return tab[lo] * (2.0 * t3 - 3.0 * t2 + 1.0) + d_lo * (t3 - 2.0 * t2 + t) +
         tab[hi] * (3.0 * t2 - 2.0 * t3) + d_hi * (t3 - t2);
}

// rhs = [0 ; -q_phase ; -q_angle] with the recorded driving forces scaled by the
// cilium's natural frequency and the linear torsional spring on the shape angle
// (mobility_solver.cpp:301-356, FORCE_NOISE_MAG = 0, FENE_MODEL = 0).
__global__ void cilia_drive_kernel(int ncil, int m, int dyn_angle, double len, double spring,
                                   const double* __restrict__ phase, const double* __restrict__ angle,
                                   const double* __restrict__ omega0,
                                   const double* __restrict__ tab_p, int np,
                                   const double* __restrict__ tab_a, int na,
                                   double* __restrict__ rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int L = 3 * m + ncil * (dyn_angle ? 2 : 1);
  if (i >= L) return;
  if (i < 3 * m) {
    rhs[i] = 0.0;
    return;
  }
  const int c = (i - 3 * m) % ncil;
  const double sc = 0.5 * omega0[c] / 3.14159265358979323846;
  if (i < 3 * m + ncil) {
    rhs[i] = -cyclic_hermite(tab_p, np, phase[c]) * sc;
  } else {
    const double q = cyclic_hermite(tab_a, na, phase[c]) * sc - spring * omega0[c] * len * len * len * angle[c];
    rhs[i] = -q;
  }
}

// speeds from the solution; F copied to the force array
__global__ void cilia_unpack_kernel(int ncil, int m, int dyn_angle, const double* __restrict__ z,
                                    double* __restrict__ pdot, double* __restrict__ adot,
                                    double* __restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * m) f[i] = z[i];
  if (i < ncil) {
    pdot[i] = z[3 * m + i];
    adot[i] = dyn_angle ? z[3 * m + ncil + i] : 0.0;
  }
}

__global__ void cilia_advance_angle_kernel(int ncil, double dt, const double* __restrict__ adot,
                                           double* __restrict__ angle) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncil) angle[c] += adot[c] * dt;
}

struct CiliaState {
  cfm_cilia_params prm{};
  DevBuf<double> base, phase, pdot, angle, vphase, vangle, svals, rhs, f, q;
  double* host = nullptr;  // pinned read-back of q
  bool have_geometry = false;
  // dynamic evolution (cfm_cilia_set_driving): phase (and shape-angle) speeds are
  // unknowns of the coupled system, driven by recorded generalised forces
  bool dynamic = false, dyn_angle = false;
  double spring = 0.0;
  DevBuf<double> omega0, adot, tab_phase, tab_angle, z, zrhs;
  int n_tab_phase = 0, n_tab_angle = 0;
  // recorded segment parameters (cfm_cilia_set_s_table)
  DevBuf<double> s_table;
  int s_rows = 0, s_cols = 0;
};

static CiliaState* cil_of(cfm_ctx* c) { return reinterpret_cast<CiliaState*>(c->cilia); }

static CiliaK kargs(cfm_ctx* c, CiliaState* S) {
  CiliaK k{};
  k.cf = S->prm.beat == CFM_BEAT_FULFORD_BLAKE_ORIGINAL ? fulford_blake_original() : fulford_blake();
  k.s_tol = S->prm.s_tol > 0.0 ? S->prm.s_tol : 0.1 / S->prm.nseg;
  k.s_table = S->s_rows > 0 ? S->s_table.p : nullptr;
  k.s_rows = S->s_rows;
  k.s_cols = S->s_cols;
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

static CoupleK couple_args(cfm_ctx* c, CiliaState* S) {
  CoupleK k{};
  k.ncil = S->prm.n_cilia;
  k.nseg = S->prm.nseg;
  k.m = c->p.n_seg;
  k.dyn_angle = S->dyn_angle ? 1 : 0;
  k.vphase = S->vphase.p;
  k.vangle = S->vangle.p;
  return k;
}

Status cilia_couple_apply(cfm_ctx* c, const double* x, double* y) {
  CiliaState* S = cil_of(c);
  if (!S || !S->dynamic) return make_status(CFM_ESTATE, "no coupled cilia system");
  hipLaunchKernelGGL(cilia_couple_kernel, gc(S->prm.n_cilia), dim3(64), 0, c->stream,
                     couple_args(c, S), x, y);
  CFM_CHECK_LAUNCH();
  return {};
}

Status cilia_precond(cfm_ctx* c, int kind, const double* x, double* y) {
  CiliaState* S = cil_of(c);
  if (!S || !S->dynamic) return make_status(CFM_ESTATE, "no coupled cilia system");
  if (kind != CFM_PC_DIAG && kind != CFM_PC_SELF)
    return make_status(CFM_EINVAL, "the coupled cilia system takes CFM_PC_DIAG or CFM_PC_SELF");
  hipLaunchKernelGGL(cilia_precond_kernel, gc(S->prm.n_cilia), dim3(64), 0, c->stream,
                     couple_args(c, S), kind, c->cfg.mu, c->p.z.p, x, y);
  CFM_CHECK_LAUNCH();
  return {};
}

}  // namespace cfm

using namespace cfm;

void cfm_cilia_release(cfm_ctx* c) {
  CiliaState* S = cil_of(c);
  if (!S) return;
  for (auto* b : {&S->base, &S->phase, &S->pdot, &S->angle, &S->vphase, &S->vangle, &S->svals,
                  &S->rhs, &S->f, &S->q, &S->omega0, &S->adot, &S->tab_phase, &S->tab_angle,
                  &S->z, &S->zrhs, &S->s_table})
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
  if (prm->beat != CFM_BEAT_FULFORD_BLAKE && prm->beat != CFM_BEAT_FULFORD_BLAKE_ORIGINAL)
    return cfail(c, make_status(CFM_EINVAL, "unknown beat %d", prm->beat));
  if (!(prm->s_tol >= 0.0) || prm->s_tol >= 0.5)
    return cfail(c, make_status(CFM_EINVAL, "s_tol must lie in [0, 0.5) (0: 0.1 / nseg)"));
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
    CFM_TRY(S->omega0.alloc(nc));
    CFM_TRY(S->adot.alloc(nc));
    CFM_HIP(hipMemcpyAsync(S->omega0.p, S->pdot.p, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
    CFM_TRY(dev_fill(c, S->adot.p, 0.0, nc));
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
    if (nt > 0) {  // phase += phase_dot dt before the step (filament.cpp:1340-1356)
      hipLaunchKernelGGL(cilia_advance_kernel, gc(nc), dim3(64), 0, c->stream, nc, S->prm.dt,
                         S->pdot.p, S->phase.p);
      if (S->dyn_angle)
        hipLaunchKernelGGL(cilia_advance_angle_kernel, gc(nc), dim3(64), 0, c->stream, nc,
                           S->prm.dt, S->adot.p, S->angle.p);
    }
    CFM_TRY(geometry(c, S));
    SystemDesc sd = system_desc(c, CFM_SYS_SEG);
    if (S->dynamic) {
      // coupled system: forces and speeds, driven by the recorded generalised forces
      sd.cilia = true;
      sd.extra = nc * (S->dyn_angle ? 2 : 1);
      sd.L = 3 * c->p.n_seg + sd.extra;
      hipLaunchKernelGGL(cilia_drive_kernel, dim3((sd.L + 255) / 256), dim3(256), 0, c->stream, nc,
                         c->p.n_seg, S->dyn_angle ? 1 : 0, S->prm.dl * (S->prm.nseg - 1), S->spring,
                         S->phase.p, S->angle.p, S->omega0.p, S->tab_phase.p, S->n_tab_phase,
                         S->tab_angle.p, S->n_tab_angle, S->zrhs.p);
      CFM_CHECK_LAUNCH();
      cfm_solve_opts o = S->prm.solve;
      o.system = CFM_SYS_SEG;
      o.use_guess = 1;  // previous step's forces and speeds
      if (!c->krylov) c->krylov = new Krylov();
      cfm_solve_stats ss{};
      CFM_TRY(solve(c, *c->krylov, sd, o, S->zrhs.p, S->z.p, &ss, nullptr));
      hipLaunchKernelGGL(cilia_unpack_kernel, dim3((3 * c->p.n_seg + 255) / 256), dim3(256), 0,
                         c->stream, nc, c->p.n_seg, S->dyn_angle ? 1 : 0, S->z.p, S->pdot.p,
                         S->adot.p, S->f.p);
      CFM_CHECK_LAUNCH();
      hipLaunchKernelGGL(cilia_genforce_kernel, gc(nc), dim3(64), 0, c->stream, nc, S->prm.nseg,
                         c->p.n_seg, S->vphase.p, S->vangle.p, S->f.p, S->q.p);
      CFM_CHECK_LAUNCH();
      CFM_HIP(hipMemcpyAsync(S->host, S->q.p, sizeof(double) * 2 * nc, hipMemcpyDeviceToHost, c->stream));
      CFM_HIP(hipStreamSynchronize(c->stream));
      if (q_phase) std::memcpy(q_phase, S->host, sizeof(double) * nc);
      if (q_angle) std::memcpy(q_angle, S->host + nc, sizeof(double) * nc);
      if (st) *st = ss;
      return {};
    }
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

int cfm_cilia_set_driving(cfm_ctx* c, const double* q_phase_table, int32_t n_phase,
                          const double* q_angle_table, int32_t n_angle, double spring) {
  if (!c) return CFM_EINVAL;
  CiliaState* S = cil_of(c);
  if (!S) return cfail(c, make_status(CFM_ESTATE, "cfm_cilia_init has not been called"));
  if (!q_phase_table || n_phase < 2 || (q_angle_table && n_angle < 2) || !std::isfinite(spring))
    return cfail(c, make_status(CFM_EINVAL, "driving tables need at least 2 values"));
  for (int i = 0; i < n_phase; ++i)
    if (!std::isfinite(q_phase_table[i])) return cfail(c, make_status(CFM_EINVAL, "non-finite table"));
  if (q_angle_table)
    for (int i = 0; i < n_angle; ++i)
      if (!std::isfinite(q_angle_table[i])) return cfail(c, make_status(CFM_EINVAL, "non-finite table"));
  (void)hipSetDevice(c->device);
  const int nc = S->prm.n_cilia;
  auto run = [&]() -> Status {
    CFM_TRY(S->tab_phase.alloc(n_phase));
    CFM_HIP(hipMemcpy(S->tab_phase.p, q_phase_table, sizeof(double) * n_phase, hipMemcpyHostToDevice));
    S->n_tab_phase = n_phase;
    S->dyn_angle = q_angle_table != nullptr;
    if (S->dyn_angle) {
      CFM_TRY(S->tab_angle.alloc(n_angle));
      CFM_HIP(hipMemcpy(S->tab_angle.p, q_angle_table, sizeof(double) * n_angle, hipMemcpyHostToDevice));
      S->n_tab_angle = n_angle;
    }
    S->spring = spring;
    const size_t L = 3 * (size_t)c->p.n_seg + (size_t)nc * (S->dyn_angle ? 2 : 1);
    CFM_TRY(S->z.alloc(L));
    CFM_TRY(S->zrhs.alloc(L));
    CFM_TRY(dev_fill(c, S->z.p, 0.0, (int)L));
    // initial guess of the speeds: the natural frequencies, no shape rotation
    CFM_HIP(hipMemcpyAsync(S->z.p + 3 * (size_t)c->p.n_seg, S->omega0.p, sizeof(double) * nc,
                           hipMemcpyDeviceToDevice, c->stream));
    CFM_HIP(hipStreamSynchronize(c->stream));
    S->dynamic = true;
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : cfail(c, s);
}

int cfm_cilia_set_s_table(cfm_ctx* c, const double* table, int32_t rows, int32_t cols) {
  if (!c) return CFM_EINVAL;
  CiliaState* S = cil_of(c);
  if (!S) return cfail(c, make_status(CFM_ESTATE, "cfm_cilia_init has not been called"));
  (void)hipSetDevice(c->device);
  if (!table) {  // back to the bisection
    S->s_rows = S->s_cols = 0;
  } else {
    if (rows <= 0 || cols < 2)
      return cfail(c, make_status(CFM_EINVAL, "s table needs rows >= 1 and cols >= 2"));
    for (long long i = 0; i < (long long)rows * cols; ++i)
      if (!std::isfinite(table[i])) return cfail(c, make_status(CFM_ENONFINITE, "non-finite s table"));
    auto run = [&]() -> Status {
      CFM_TRY(S->s_table.alloc((size_t)rows * cols));
      CFM_HIP(hipMemcpyAsync(S->s_table.p, table, sizeof(double) * rows * cols, hipMemcpyHostToDevice,
                             c->stream));
      CFM_HIP(hipStreamSynchronize(c->stream));
      return {};
    };
    Status st = run();
    if (!st.ok()) return cfail(c, st);
    S->s_rows = rows;
    S->s_cols = cols;
  }
  Status st = geometry(c, S);  // re-place the segments with the new parameters
  return st.ok() ? CFM_OK : cfail(c, st);
}

int cfm_cilia_get_speeds(cfm_ctx* c, double* phase, double* phase_dot, double* angle,
                         double* angle_dot) {
  if (!c) return CFM_EINVAL;
  CiliaState* S = cil_of(c);
  if (!S) return cfail(c, make_status(CFM_ESTATE, "cfm_cilia_init has not been called"));
  (void)hipSetDevice(c->device);
  const int nc = S->prm.n_cilia;
  auto run = [&]() -> Status {
    CFM_HIP(hipStreamSynchronize(c->stream));
    if (phase) CFM_HIP(hipMemcpy(phase, S->phase.p, sizeof(double) * nc, hipMemcpyDeviceToHost));
    if (phase_dot) CFM_HIP(hipMemcpy(phase_dot, S->pdot.p, sizeof(double) * nc, hipMemcpyDeviceToHost));
    if (angle) CFM_HIP(hipMemcpy(angle, S->angle.p, sizeof(double) * nc, hipMemcpyDeviceToHost));
    if (angle_dot) CFM_HIP(hipMemcpy(angle_dot, S->adot.p, sizeof(double) * nc, hipMemcpyDeviceToHost));
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : cfail(c, s);
}

int cfm_cilia_set_phases(cfm_ctx* c, const double* phase, const double* angle) {
  if (!c) return CFM_EINVAL;
  CiliaState* S = cil_of(c);
  if (!S) return cfail(c, make_status(CFM_ESTATE, "cfm_cilia_init has not been called"));
  const int nc = S->prm.n_cilia;
  for (int i = 0; i < nc; ++i)
    if ((phase && !std::isfinite(phase[i])) || (angle && !std::isfinite(angle[i])))
      return cfail(c, make_status(CFM_ENONFINITE, "non-finite phase or angle"));
  (void)hipSetDevice(c->device);
  auto run = [&]() -> Status {
    if (phase) CFM_HIP(hipMemcpyAsync(S->phase.p, phase, sizeof(double) * nc, hipMemcpyHostToDevice, c->stream));
    if (angle) CFM_HIP(hipMemcpyAsync(S->angle.p, angle, sizeof(double) * nc, hipMemcpyHostToDevice, c->stream));
    CFM_TRY(geometry(c, S));
    CFM_HIP(hipStreamSynchronize(c->stream));
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
