// cilia.hip -- prescribed-shape cilia above the wall or on free bodies (fitted
// Fulford-Blake, volvox and coral-larvae beats; the build-a-beat tangent-angle model).
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
//   beat coefficients      filament.cpp:684-758 (Ito, Omori & Ishikawa 2019, Table 1),
//                          :748-812 volvox, :813-868 unaltered, :869-893 coral larvae
//   build-a-beat           filament.cpp:1238-1334 angle law, :1401-1546 placement
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
#include "body.hpp"
#include "rpy_pair.hpp"
#include "fil_device.hpp"
#include "precond.hpp"

namespace cfm {

Status upload_aos(cfm_ctx* c, const double* host, int cnt, int dof, double* dst, int ld, int off);
Status download_aos(cfm_ctx* c, double* host, int cnt, int dof, const double* src, int ld, int off);

constexpr int kMaxCiliaSeg = 64;

// Fourier modes x polynomial degree (s, s^2, ... s^ndeg, ndeg <= 4); x along the
// filament axis of the reference frame, y along the beat direction; cos0 weights the
// zero cosine mode.  build: the build-a-beat tangent-angle model instead.
constexpr int kMaxDeg = 4;
struct BeatCoeffs {
  double ax[4][kMaxDeg], ay[4][kMaxDeg], bx[3][kMaxDeg], by[3][kMaxDeg];
  double cos0;
  int ndeg;
  int build;
};

template <int D>
static void set_table(BeatCoeffs& b, const double (&ax)[4][D], const double (&ay)[4][D],
                      const double (&bx)[3][D], const double (&by)[3][D]) {
  b = BeatCoeffs{};
  for (int m = 0; m < 4; ++m)
    for (int d = 0; d < D; ++d) {
      b.ax[m][d] = ax[m][d];
      b.ay[m][d] = ay[m][d];
      if (m < 3) {
        b.bx[m][d] = bx[m][d];
        b.by[m][d] = by[m][d];
      }
    }
  b.ndeg = D;
  b.cos0 = 1.0;
}

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
  set_table(b, ax, ay, bx, by);
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
  set_table(b, ax, ay, bx, by);
  b.cos0 = 0.5;
  return b;
}

// The reference's volvox fit (SHAPE_SEQUENCE 3, filament.cpp:748-812): four polynomial
// degrees, zero cosine mode at weight 1.
static BeatCoeffs volvox() {
  BeatCoeffs b{};
  const double ax[4][4] = {{1.00490080, -0.19958656, -0.20422343, 0.10887946},
                           {0.00314934, -0.22239636, 1.18659976, -0.87409016},
                           {0.18942274, -1.24525818, 1.94063304, -0.78632572},
                           {-0.11162123, 1.07262966, -2.15142914, 1.20544741}};
  const double ay[4][4] = {{-0.01665535, -0.27457190, 1.08406999, -0.56102820},
                           {0.17505739, -2.09412076, 1.99556763, -0.50179439},
                           {-0.12408600, 1.62438155, -3.21015214, 1.70178281},
                           {-0.16672661, 0.90896038, -1.50867276, 0.74387795}};
  const double bx[3][4] = {{0.09140608, -0.54804089, 1.40643137, -0.60109684},
                           {-0.09670739, 1.14634486, -2.60675971, 1.55525152},
                           {-0.16451150, 0.98846449, -1.49147276, 0.64896090}};
  const double by[3][4] = {{-0.06343445, 2.15509751, -4.19609063, 2.05489100},
                           {-0.28828336, 1.45223183, -1.56580487, 0.44541361},
                           {0.02187088, -0.33070891, 0.94906955, -0.66157093}};
  set_table(b, ax, ay, bx, by);
  return b;
}

// The reference's coral-larvae fit (SHAPE_SEQUENCE 2, filament.cpp:869-893).
static BeatCoeffs coral_larvae() {
  BeatCoeffs b{};
  const double ax[4][3] = {{9.034276e-01, -6.634681e-01, 2.016123e-01},
                           {-2.468656e-01, 1.215869e+00, -5.867236e-01},
                           {3.075836e-01, -6.966322e-01, 4.554729e-01},
                           {2.378489e-02, -1.355582e-01, 7.315831e-02}};
  const double ay[4][3] = {{-2.766076e-01, 6.368662e-01, -8.513947e-02},
                           {3.091205e-01, -1.993712e+00, 1.091847e+00},
                           {-2.952120e-01, 7.189021e-01, -6.328589e-01},
                           {-9.868018e-02, 3.312812e-01, -2.578206e-01}};
  const double bx[3][3] = {{1.336807e-01, -6.807852e-01, 6.008592e-01},
                           {8.146824e-02, 3.472676e-01, -2.744220e-01},
                           {3.615272e-02, 8.619119e-02, -6.122992e-02}};
  const double by[3][3] = {{9.165546e-01, 8.848315e-02, -4.838832e-01},
                           {-2.946211e-01, 1.057964e+00, -6.335721e-01},
                           {6.122744e-02, -2.319094e-01, 1.906292e-01}};
  set_table(b, ax, ay, bx, by);
  return b;
}

// Build-a-beat (SHAPE_SEQUENCE 0, filament.cpp:1238-1334) with the reference's
// configuration (config.hpp:133-139): amplitude 1.99 filament lengths, recovery-stroke
// window 1/12 of the cycle, effective stroke 0.75 of it, zero-velocity avoidance 0.05.
constexpr double kBuildAmplitude = 1.99, kBuildRecovery = 1.0 / 12.0, kBuildEffective = 0.75,
                 kBuildShift = 0.05;
static BeatCoeffs build_a_beat() {
  BeatCoeffs b{};
  b.build = 1;
  b.ndeg = 0;
  return b;
}

// Segments of one cilium: half the unit for the paired beats.
static int cilium_segments(const cfm_cilia_params& prm) {
  const bool pair = prm.beat == CFM_BEAT_BICILIA || prm.beat == CFM_BEAT_BICILIA_LONGT;
  return pair ? prm.nseg / 2 : prm.nseg;
}
// The reference's FIL_LENGTH = DL (NSEG_PER_CILIA - 1) (config.hpp:579).
static double cilium_length(const cfm_cilia_params& prm) { return prm.dl * (cilium_segments(prm) - 1); }

// Tangent angle of the build-a-beat model at arc fraction s and phase psi, and its
// phase derivative.
__device__ inline void build_angle(double s, double psi, double* theta, double* dtheta) {
  const double pi = 3.14159265358979323846;
  const double th_s = acos(0.5 * kBuildAmplitude);
  const double wR = kBuildRecovery * 2.0 * pi, wE = kBuildEffective * 2.0 * pi;
  const double trans = wE + s * (2.0 * pi - wE - wR);
  double sp = psi - s * 2.0 * pi * kBuildShift;
  sp -= 2.0 * pi * floor(0.5 * sp / pi);
  double delta, dd;
  if (sp < wE) {
    const double t = pi * (wE - sp) / wE;
    delta = t - sin(t) * cos(t);
    dd = -sin(t) * sin(t) * 2.0 * pi / wE / pi;  // d delta / d psi
  } else if (sp < trans) {
    delta = 0.0;
    dd = 0.0;
  } else if (sp < trans + wR) {
    const double t = pi * (sp - trans) / wR;
    delta = t - sin(t) * cos(t);
    dd = sin(t) * sin(t) * 2.0 * pi / wR / pi;
  } else {
    delta = pi;
    dd = 0.0;
  }
  *theta = th_s + (pi - 2.0 * th_s) * delta / pi;
  *dtheta = (pi - 2.0 * th_s) * dd;
}

struct CiliaK {
  BeatCoeffs cf;
  int ncil, nseg, nseg_total;
  double dl, len;  // segment spacing, length dl (m - 1) of one cilium of m segments
  // paired cilia (CFM_BEAT_BICILIA / _LONGT): each unit of nseg segments is two cilia of
  // m = nseg / 2, the second at phase psi + 2 pi pair_dp (pair 1) or pair_dp psi (pair 2),
  // its base moved by zdis along the frame's third axis (filament.cpp:371-375,1429-1466)
  int pair;
  double pair_dp, zdis;
  double s_tol;    // bisection tolerance on the arc-length fraction
  const double* s_table;  // recorded segment parameters (rows x cols) or null
  int s_rows, s_cols;
  const double* base;    // 3 x ncil (SoA); unused for cilia on bodies
  // cilia carried by free bodies (null bx: cilia on the wall in the fixed frame)
  const double* bx;      // body centres, 3 per body
  const double* bq;      // body quaternions, 4 per body
  const double* fref;    // base offsets of one body's cilia (body frame), 3 each
  const double* fq;      // frame quaternions of one body's cilia (body frame), 4 each
  int per_body;          // cilia per body
  const double* phase;   // ncil
  const double* angle;   // ncil (rotation of the beat about the reference z axis)
  double *x, *y, *z;     // segment positions (the context's particle arrays)
  double* vphase;        // 3 x nseg_total
  double* vangle;        // 3 x nseg_total
  double* svals;         // ncil x nseg
};

// Beat table, pairing, filament length and bisection tolerance of a parameter set; the
// paired beats use the volvox fit (filament.cpp:895-960 holds the same values) and put
// the second cilium 2.2 segment radii from the first (filament.cpp:1429-1437).
static void beat_args(const cfm_cilia_params& prm, double pair_dp, double a_seg, CiliaK& k) {
  switch (prm.beat) {
    case CFM_BEAT_FULFORD_BLAKE_ORIGINAL: k.cf = fulford_blake_original(); break;
    case CFM_BEAT_VOLVOX:
    case CFM_BEAT_BICILIA:
    case CFM_BEAT_BICILIA_LONGT: k.cf = volvox(); break;
    case CFM_BEAT_CORAL_LARVAE: k.cf = coral_larvae(); break;
    case CFM_BEAT_BUILD_A_BEAT: k.cf = build_a_beat(); break;
    default: k.cf = fulford_blake();
  }
  k.pair = prm.beat == CFM_BEAT_BICILIA ? 1 : prm.beat == CFM_BEAT_BICILIA_LONGT ? 2 : 0;
  k.pair_dp = pair_dp;
  k.zdis = 2.2 * a_seg;
  k.s_tol = prm.s_tol > 0.0 ? prm.s_tol : 0.1 / cilium_segments(prm);
  k.len = cilium_length(prm);
}

// Mode vectors (cos0, cos psi, cos 2psi, cos 3psi) and (sin psi, sin 2psi, sin 3psi).
__device__ inline void modes(double psi, double cos0, double cv[4], double sv[3]) {
  cv[0] = cos0;
  for (int n = 1; n < 4; ++n) {
    cv[n] = cos(n * psi);
    sv[n - 1] = sin(n * psi);
  }
}

// sum_m (cv_m A_m + sv_m B_m) . p over the polynomial degrees
__device__ inline double combine(const double A[4][kMaxDeg], const double B[3][kMaxDeg], const double cv[4],
                                 const double sv[3], const double p[kMaxDeg], int ndeg) {
  double r = 0.0;
  for (int d = 0; d < ndeg; ++d) {
    double a = 0.0;
    for (int m = 0; m < 4; ++m) a += cv[m] * A[m][d];
    for (int m = 0; m < 3; ++m) a += sv[m] * B[m][d];
    r += a * p[d];
  }
  return r;
}

__device__ inline double tangent_norm(const BeatCoeffs& cf, const double cv[4], const double sv[3],
                                      double s) {
  const double p[kMaxDeg] = {1.0, 2.0 * s, 3.0 * s * s, 4.0 * s * s * s};
  const double tx = combine(cf.ax, cf.bx, cv, sv, p, cf.ndeg);
  const double ty = combine(cf.ay, cf.by, cv, sv, p, cf.ndeg);
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

// Running trapezoid sums P[j] (j = 0 .. 10 nseg) of curve_length's panels at one phase,
// accumulated in the order of curve_length's loop (same bits); P lives in LDS.
__device__ void curve_table(const BeatCoeffs& cf, const double cv[4], const double sv[3],
                            int nseg, double* P) {
  const double dl = 1.0 / (10.0 * nseg);
  double f0 = tangent_norm(cf, cv, sv, 0.0);
  double length = 0.0;
  P[0] = 0.0;
  for (int n = 1; n <= 10 * nseg; ++n) {
    const double f1 = tangent_norm(cf, cv, sv, n * dl);
    length += 0.5 * dl * (f0 + f1);
    P[n] = length;
    f0 = f1;
  }
}

// curve_length from the table: the whole panels read from P, the last panel evaluated
// as curve_length evaluates it.
__device__ double curve_length_tab(const BeatCoeffs& cf, const double cv[4], const double sv[3],
                                   double s, int nseg, const double* P) {
  if (s <= 0.0) return 0.0;
  const double dl = 1.0 / (10.0 * nseg);
  const double fl = floor(s / dl);
  if (fl > 10.0 * nseg) return curve_length(cf, cv, sv, s, nseg);  // s > 1: not tabulated
  double length = P[(int)fl];
  const double ce = ceil(s / dl);
  if (ce > fl) {
    const double f0 = tangent_norm(cf, cv, sv, (int)fl * dl);
    const double f1 = tangent_norm(cf, cv, sv, ce * dl);
    length += ((s / dl - fl) / (ce - fl)) * 0.5 * dl * (f0 + f1);
  }
  return length;
}

// Rotation matrix (row-major) of a unit quaternion (w, x, y, z) (quaternion.cpp:305-347).
__device__ inline void quat_to_mat(const double* q, double R[9]) {
  const double s = q[0], a = q[1], b = q[2], c = q[3];
  R[0] = 1.0 - 2.0 * (b * b + c * c);
  R[1] = 2.0 * (a * b - s * c);
  R[2] = 2.0 * (a * c + s * b);
  R[3] = 2.0 * (a * b + s * c);
  R[4] = 1.0 - 2.0 * (a * a + c * c);
  R[5] = 2.0 * (b * c - s * a);
  R[6] = 2.0 * (a * c - s * b);
  R[7] = 2.0 * (b * c + s * a);
  R[8] = 1.0 - 2.0 * (a * a + b * b);
}

// Places one cilium of m segments: unit c's segments [first, first + m) (its segment
// parameters at svals[c nseg + first ...]) at phase psi, base (bxp, byp, bzp) and frame
// axes e1 (reference x) and e2 (reference y); `wall`: the wall frame e1 = e_z, e2 = e_y.
__device__ __forceinline__ void place_cilium(const CiliaK& k, int c, double psi, int first, int m, double bxp,
                             double byp, double bzp, const double e1[3], const double e2[3],
                             bool wall, double* P) {
  const int n = m, s0 = c * k.nseg + first;
  double cv[4], sv[3];
  modes(psi, k.cf.cos0, cv, sv);
  double s[kMaxCiliaSeg];
  s[0] = 0.0;
  s[n - 1] = 1.0;
  if (k.cf.build) {
    // build-a-beat: segments at uniform arc fraction (filament.cpp:1489-1493)
    for (int i = 1; i < n - 1; ++i) s[i] = double(i) / double(n - 1);
  } else if (k.s_table) {
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
    // bisection on the arc-length fraction (filament.cpp:1096-1185), the panel sums
    // tabulated once for this phase
    curve_table(k.cf, cv, sv, n, P);
    const double total = curve_length_tab(k.cf, cv, sv, 1.0, n, P);
    for (int i = 1; i < n - 1; ++i) {
      const double target = double(i) / double(n - 1);
      double lo = s[i - 1], hi = 1.0;
      double est = 0.5 * (lo + hi);
      double frac = curve_length_tab(k.cf, cv, sv, est, n, P) / total;
      for (int it = 0; it < 200 && fabs(frac - target) > k.s_tol; ++it) {
        if (frac > target)
          hi = est;
        else
          lo = est;
        est = 0.5 * (lo + hi);
        frac = curve_length_tab(k.cf, cv, sv, est, n, P) / total;
      }
      s[i] = est;
    }
  }
  const double ca = cos(k.angle[c]), sa = sin(k.angle[c]);
  const size_t ns = k.nseg_total;
  double dc[4], ds[3];
  dc[0] = 0.0;
  for (int mo = 1; mo < 4; ++mo) {
    dc[mo] = -mo * sin(mo * psi);
    ds[mo - 1] = mo * cos(mo * psi);
  }
  // build-a-beat: trapezoid sums of the unit tangent (sin th, cos th) and of its phase
  // derivative over the segments (filament.cpp:1401-1546), in filament lengths
  double bt[2] = {0.0, 0.0}, bk[2] = {0.0, 0.0}, sum_t[2] = {0.0, 0.0}, sum_k[2] = {0.0, 0.0};
  if (k.cf.build) {
    double th, dth;
    build_angle(0.0, psi, &th, &dth);
    bt[0] = sin(th);
    bt[1] = cos(th);
    bk[0] = dth * cos(th);
    bk[1] = -dth * sin(th);
  }
  for (int i = 0; i < n; ++i) {
    const int g = s0 + i;
    k.svals[(size_t)c * k.nseg + first + i] = s[i];
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
    double sx, sy, dx, dy;  // shape and its phase derivative in filament lengths
    if (k.cf.build) {
      double th, dth;
      build_angle(s[i], psi, &th, &dth);
      const double t2[2] = {sin(th), cos(th)}, k2[2] = {dth * cos(th), -dth * sin(th)};
      const double h = 0.5 / double(n - 1);
      for (int d = 0; d < 2; ++d) {
        sum_t[d] += h * (bt[d] + t2[d]);
        sum_k[d] += h * (bk[d] + k2[d]);
        bt[d] = t2[d];
        bk[d] = k2[d];
      }
      sx = sum_t[0];
      sy = sum_t[1];
      dx = sum_k[0];
      dy = sum_k[1];
    } else {
      const double p[kMaxDeg] = {s[i], s[i] * s[i], s[i] * s[i] * s[i], s[i] * s[i] * s[i] * s[i]};
      sx = combine(k.cf.ax, k.cf.bx, cv, sv, p, k.cf.ndeg);
      sy = combine(k.cf.ay, k.cf.by, cv, sv, p, k.cf.ndeg);
      dx = combine(k.cf.ax, k.cf.bx, dc, ds, p, k.cf.ndeg);
      dy = combine(k.cf.ay, k.cf.by, dc, ds, p, k.cf.ndeg);
    }
    // shape in the reference frame, rotated by the beat angle about its z axis
    const double rx = k.len * (ca * sx - sa * sy), ry = k.len * (sa * sx + ca * sy);
    const double ux = k.len * (ca * dx - sa * dy), uy = k.len * (sa * dx + ca * dy);
    if (wall) {  // wall frame written out: lab (0, ry, rx)
      k.x[g] = bxp;
      k.y[g] = byp + ry;
      k.z[g] = bzp + rx;
      // e_z x ref in the reference frame is (-ry, rx, 0) -> lab (0, rx, -ry)
      k.vangle[g] = 0.0;
      k.vangle[ns + g] = rx;
      k.vangle[2 * ns + g] = -ry;
      k.vphase[g] = 0.0;
      k.vphase[ns + g] = uy;
      k.vphase[2 * ns + g] = ux;
      continue;
    }
    k.x[g] = bxp + e1[0] * rx + e2[0] * ry;
    k.y[g] = byp + e1[1] * rx + e2[1] * ry;
    k.z[g] = bzp + e1[2] * rx + e2[2] * ry;
    for (int d = 0; d < 3; ++d) {
      k.vangle[d * ns + g] = e2[d] * rx - e1[d] * ry;
      k.vphase[d * ns + g] = e1[d] * ux + e2[d] * uy;
    }
  }
}

// One thread per cilium: per unit, or per half of a pair (the second half of unit c is
// thread 2 c + 1).  Dynamic LDS: 10 m + 1 doubles per thread (the arc-length table).
__global__ void cilia_geometry_kernel(CiliaK k) {
  extern __shared__ double cilia_lds[];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int halves = k.pair ? 2 : 1;
  const int c = t / halves, h = t - c * halves;
  if (c >= k.ncil) return;
  // frame of the cilium: its first two columns carry the reference x (filament axis)
  // and y (beat direction) axes.  On the wall: reference x -> lab z, y -> lab y
  // (filament.cpp:225-226 for dir = e_z).  On a body: R(q_body) R(q_ref) with the base
  // at x_body + R(q_body) ref (swimmer.cpp:448-462, filament.cpp:653-663).
  double e1[3] = {0.0, 0.0, 1.0}, e2[3] = {0.0, 1.0, 0.0};
  double bxp, byp, bzp;
  if (k.bx) {
    const int b = c / k.per_body, l = c - b * k.per_body;
    double Rb[9], Rf[9];
    quat_to_mat(k.bq + 4 * b, Rb);
    quat_to_mat(k.fq + 4 * l, Rf);
    for (int r = 0; r < 3; ++r) {
      e1[r] = Rb[3 * r] * Rf[0] + Rb[3 * r + 1] * Rf[3] + Rb[3 * r + 2] * Rf[6];
      e2[r] = Rb[3 * r] * Rf[1] + Rb[3 * r + 1] * Rf[4] + Rb[3 * r + 2] * Rf[7];
    }
    const double* o = k.fref + 3 * l;
    bxp = k.bx[3 * b] + Rb[0] * o[0] + Rb[1] * o[1] + Rb[2] * o[2];
    byp = k.bx[3 * b + 1] + Rb[3] * o[0] + Rb[4] * o[1] + Rb[5] * o[2];
    bzp = k.bx[3 * b + 2] + Rb[6] * o[0] + Rb[7] * o[1] + Rb[8] * o[2];
  } else {
    bxp = k.base[c];
    byp = k.base[k.ncil + c];
    bzp = k.base[2 * k.ncil + c];
  }
  const bool wall = !k.bx;
  double psi = k.phase[c];
  if (h == 1) {
    // the second cilium of a pair at its own phase, its base zdis along the frame's third
    // axis e1 x e2 (the reference's z axis: lab -x on the wall)
    psi = k.pair == 1 ? psi + k.pair_dp * 2.0 * 3.14159265358979323846 : k.pair_dp * psi;
    bxp += k.zdis * (e1[1] * e2[2] - e1[2] * e2[1]);
    byp += k.zdis * (e1[2] * e2[0] - e1[0] * e2[2]);
    bzp += k.zdis * (e1[0] * e2[1] - e1[1] * e2[0]);
  }
  const int m = k.nseg / halves;
  place_cilium(k, c, psi, h * m, m, bxp, byp, bzp, e1, e2, wall,
               cilia_lds + (size_t)threadIdx.x * (10 * m + 1));
}

// Launch shape of the geometry kernel: threads, threads per workgroup (the arc-length
// tables within 64 KiB of LDS) and the LDS bytes per workgroup.
struct GeoLaunch {
  dim3 grid, block;
  size_t lds;
};
static GeoLaunch geometry_launch(const CiliaK& k) {
  const int threads = k.ncil * (k.pair ? 2 : 1);
  const int m = k.pair ? k.nseg / 2 : k.nseg;
  const size_t per = sizeof(double) * (10 * (size_t)m + 1);
  const int tpb = (int)std::max<size_t>(1, std::min<size_t>(64, 65536 / per));
  return {dim3((unsigned)((threads + tpb - 1) / tpb)), dim3((unsigned)tpb), per * tpb};
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
// Normal deviate of (seed, step, cilium): SplitMix64 mixing of the key into two
// 53-bit uniforms on (0, 1), then Box-Muller.
__host__ __device__ inline unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ inline double normal_deviate(unsigned long long seed, int nt, int c) {
  const unsigned long long key = mix64(seed ^ mix64(((unsigned long long)(unsigned)nt << 32) | (unsigned)c));
  const unsigned long long h1 = mix64(key), h2 = mix64(key + 1ull);
  const double u1 = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53;
  const double u2 = ((double)(h2 >> 11) + 0.5) * 0x1.0p-53;
  return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
}

struct DriveOpts {
  int fene;                 // 1: FENE form of the angle spring
  double noise;             // FORCE_NOISE_MAG
  unsigned long long seed;  // of the noise
  int nt;                   // step
};

// rhs = [0 ; -q_phase ; -q_angle] with the recorded driving forces scaled by the
// cilium's natural frequency, the torsional spring on the shape angle (linear, or the
// FENE form beyond |angle| = 0.78) and optional noise on the phase drive
// (mobility_solver.cpp:301-356).  The phase slots start at `off` (after the forces,
// and the body velocities when the cilia ride on free bodies); everything before them
// is zero.
__global__ void cilia_drive_kernel(int ncil, int off, int dyn_angle, double len, double spring,
                                   const double* __restrict__ phase, const double* __restrict__ angle,
                                   const double* __restrict__ omega0,
                                   const double* __restrict__ tab_p, int np,
                                   const double* __restrict__ tab_a, int na, DriveOpts dopt,
                                   double* __restrict__ rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int L = off + ncil * (dyn_angle ? 2 : 1);
  if (i >= L) return;
  if (i < off) {
    rhs[i] = 0.0;
    return;
  }
  const int c = (i - off) % ncil;
  const double sc = 0.5 * omega0[c] / 3.14159265358979323846;
  if (i < off + ncil) {
    double q = cyclic_hermite(tab_p, np, phase[c]) * sc;
    if (dopt.noise != 0.0) q += dopt.noise * normal_deviate(dopt.seed, dopt.nt, c);
    rhs[i] = -q;
  } else {
    const double a = angle[c], amax = 0.78;
    double fac = a;
    if (dopt.fene && fabs(a) > amax) fac = 0.5 * a * (1.0 + a * a / (amax * amax) * exp(fabs(a) - amax));
    const double q = cyclic_hermite(tab_a, na, phase[c]) * sc - spring * omega0[c] * len * len * len * fac;
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

// ---- cilia carried by free rigid bodies ---------------------------------------------
// System of the reference's PRESCRIBED_CILIA swimmers with free bodies
// (mobility_solver.cpp:1891-2115): unknowns [F: 3 x M by component, M = n_seg + n_blob]
// ++ [V, Omega per body] ++ [phase speeds] ++ [angle speeds];
//   rows of particle i of body b:  (M F)_i - (V_b + Omega_b x r_i) - phi_dot k_phi,i - theta_dot k_theta,i
//   rows of body b:                 -sum_i F_i, -sum_i r_i x F_i
//   rows of cilium c:               -k_phi . F, -k_theta . F
// with r_i = x_i - x_b.  The block preconditioner inverts this saddle matrix exactly
// with M^-1 replaced by the particle preconditioner P (6 pi mu a: the reference's
// mobility_solver.cpp:1250-1515 with make_precon_mat, swimmer.cpp:1175-1347).
// With the bodies' rotation prescribed (cfm_body_set_motion, BODY_VELOCITY_TYPE 2) the
// Omega_b columns leave the particle rows and the torque rows read Omega_b
// (mobility_solver.cpp:1947-1994, 2177-2181); the preconditioner keeps the translation
// block of each body's Schur complement and returns Omega_b = r_Omega.

struct BodyCiliaK {
  int ncil, nseg, nseg_total, nb, nbody, per_body;
  int M;  // particles of the system: segments, then blobs
  int dynamic, dyn_angle;
  int rot;  // prescribed rotation: Omega_b held at the torque rows' right-hand side (zero)
  const double *px, *py, *pz;  // particle positions
  const double* bx;            // body centres, 3 per body
  const double* vphase;        // [3][n_seg]
  const double* vangle;        // [3][n_seg]
  int pc_kind, wall;
  double mu, a_seg, a_blob;
  double* pc_c;  // per cilium (15): S_cc^-1 (i11, i12, i22), B_phi (6), B_theta (6)
  double* pc_b;  // per body (36): Schur complement inverse, row-major
  // CFM_PC_DENSE: the blob block is the resident inverse of M_bb (applied by the body
  // module); the particle kernels skip blobs, pc_b holds the segment part of the
  // Schur complement (not inverted) and tU the body right-hand sides
  int dense;
  double* tU;
  // CFM_PC_NEAR / CFM_PC_DENSE: each cilium's own mobility block inverted (P_c,
  // 3 nseg x 3 nseg), its images P_c (K columns, k_phi, k_theta) and K^T P_c K
  int blk;
  double* pblk;  // per cilium (3 nseg)^2
  double* pkc;   // per cilium 8 x 3 nseg: P_c times (e_x, e_y, e_z, e_x x r, e_y x r, e_z x r, k_phi, k_theta)
  double* dcil;  // per cilium 36: K^T P_c K
};

__device__ inline int cb_body(const BodyCiliaK& k, int i) {
  return i < k.nseg_total ? (i / k.nseg) / k.per_body : (i - k.nseg_total) / k.nb;
}

// Diagonal of the 3-DOF particle preconditioner (6 pi mu a, or the wall self-block).
__device__ inline void cb_pdiag(const BodyCiliaK& k, int i, double p[3]) {
  const double a = i < k.nseg_total ? k.a_seg : k.a_blob;
  for (int d = 0; d < 3; ++d) {
    double e[3] = {0.0, 0.0, 0.0}, o[3];
    e[d] = 1.0;
    particle_precond(k.pc_kind, k.wall, 3, k.pz[i], a, k.mu, e, o);
    p[d] = o[d];
  }
}

// Particle i of body b: the body's particles are its cilia's segments, then its blobs.
__device__ inline int cb_particle(const BodyCiliaK& k, int b, int t) {
  const int ns = k.per_body * k.nseg;
  return t < ns ? b * ns + t : k.nseg_total + b * k.nb + (t - ns);
}

// Sum over a 256-thread block; the result is valid in thread 0.
template <int NV>
__device__ inline void cb_block_sum(double (&v)[NV]) {
  __shared__ double part[4][NV];
  for (int j = 0; j < NV; ++j)
    for (int o = 32; o > 0; o >>= 1) v[j] += __shfl_down(v[j], o);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int j = 0; j < NV; ++j) part[w][j] = v[j];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int j = 0; j < NV; ++j) v[j] = (part[0][j] + part[1][j]) + (part[2][j] + part[3][j]);
  __syncthreads();
}

// y_i -= V_b + Omega_b x r_i (+ phi_dot k_phi + theta_dot k_theta on segments)
__global__ __launch_bounds__(256) void cb_couple_particles_kernel(BodyCiliaK k, const double* __restrict__ x,
                                                                  double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k.M) return;
  const size_t M = k.M;
  const int b = cb_body(k, i);
  const double* U = x + 3 * M + 6 * (size_t)b;
  const double r0 = k.px[i] - k.bx[3 * b], r1 = k.py[i] - k.bx[3 * b + 1], r2 = k.pz[i] - k.bx[3 * b + 2];
  const double w0 = k.rot ? 0.0 : U[3], w1 = k.rot ? 0.0 : U[4], w2 = k.rot ? 0.0 : U[5];
  double u0 = U[0] + (w1 * r2 - w2 * r1);
  double u1 = U[1] + (w2 * r0 - w0 * r2);
  double u2 = U[2] + (w0 * r1 - w1 * r0);
  if (k.dynamic && i < k.nseg_total) {
    const size_t ph = 3 * M + 6 * (size_t)k.nbody, ns = k.nseg_total;
    const int c = i / k.nseg;
    const double wp = x[ph + c];
    u0 += wp * k.vphase[i];
    u1 += wp * k.vphase[ns + i];
    u2 += wp * k.vphase[2 * ns + i];
    if (k.dyn_angle) {
      const double wa = x[ph + k.ncil + c];
      u0 += wa * k.vangle[i];
      u1 += wa * k.vangle[ns + i];
      u2 += wa * k.vangle[2 * ns + i];
    }
  }
  y[i] -= u0;
  y[M + i] -= u1;
  y[2 * M + i] -= u2;
}

// y_phi_c = -k_phi . F, y_theta_c = -k_theta . F
__global__ void cb_couple_cilia_kernel(BodyCiliaK k, const double* __restrict__ x, double* __restrict__ y) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k.ncil) return;
  const size_t M = k.M, ns = k.nseg_total, ph = 3 * M + 6 * (size_t)k.nbody;
  double dp = 0.0, da = 0.0;
  for (int i = 0; i < k.nseg; ++i) {
    const size_t g = (size_t)c * k.nseg + i;
    for (int d = 0; d < 3; ++d) {
      const double f = x[d * M + g];
      dp += k.vphase[d * ns + g] * f;
      da += k.vangle[d * ns + g] * f;
    }
  }
  y[ph + c] = -dp;
  if (k.dyn_angle) y[ph + k.ncil + c] = -da;
}

// y_body = -(sum F_i, sum r_i x F_i) over the body's particles (one block per body)
__global__ __launch_bounds__(256) void cb_couple_body_kernel(BodyCiliaK k, const double* __restrict__ x,
                                                             double* __restrict__ y) {
  const int b = blockIdx.x;
  const size_t M = k.M;
  const int cnt = k.per_body * k.nseg + k.nb;
  const double c0 = k.bx[3 * b], c1 = k.bx[3 * b + 1], c2 = k.bx[3 * b + 2];
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
    const int i = cb_particle(k, b, t);
    const double f0 = x[i], f1 = x[M + i], f2 = x[2 * M + i];
    const double r0 = k.px[i] - c0, r1 = k.py[i] - c1, r2 = k.pz[i] - c2;
    acc[0] += f0;
    acc[1] += f1;
    acc[2] += f2;
    acc[3] += r1 * f2 - r2 * f1;
    acc[4] += r2 * f0 - r0 * f2;
    acc[5] += r0 * f1 - r1 * f0;
  }
  cb_block_sum<6>(acc);
  if (threadIdx.x == 0)
    for (int d = 0; d < 6; ++d) {
      const size_t o = 3 * M + 6 * (size_t)b + d;
      y[o] = (k.rot && d >= 3) ? x[o] : -acc[d];
    }
}

// Per cilium: S_cc = (k_phi, k_theta)^T P (k_phi, k_theta) inverted, and
// B = K_b^T P (k_phi, k_theta), K_b^T f = (f, r x f) (make_precon_mat's C^-1, v and u sums).
__global__ void cb_pc_cilia_prep_kernel(BodyCiliaK k) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k.ncil) return;
  const int b = c / k.per_body;
  const size_t ns = k.nseg_total;
  double c11 = 0.0, c12 = 0.0, c22 = 0.0, Bp[6] = {0, 0, 0, 0, 0, 0}, Ba[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < k.nseg; ++i) {
    const int g = c * k.nseg + i;
    double p[3];
    cb_pdiag(k, g, p);
    const double r[3] = {k.px[g] - k.bx[3 * b], k.py[g] - k.bx[3 * b + 1], k.pz[g] - k.bx[3 * b + 2]};
    double kp[3], ka[3], pp[3], pa[3];
    for (int d = 0; d < 3; ++d) {
      kp[d] = k.vphase[d * ns + g];
      ka[d] = k.vangle[d * ns + g];
      pp[d] = p[d] * kp[d];
      pa[d] = p[d] * ka[d];
      c11 += kp[d] * pp[d];
      c12 += kp[d] * pa[d];
      c22 += ka[d] * pa[d];
      Bp[d] += pp[d];
      Ba[d] += pa[d];
    }
    Bp[3] += r[1] * pp[2] - r[2] * pp[1];
    Bp[4] += r[2] * pp[0] - r[0] * pp[2];
    Bp[5] += r[0] * pp[1] - r[1] * pp[0];
    Ba[3] += r[1] * pa[2] - r[2] * pa[1];
    Ba[4] += r[2] * pa[0] - r[0] * pa[2];
    Ba[5] += r[0] * pa[1] - r[1] * pa[0];
  }
  double* o = k.pc_c + 15 * (size_t)c;
  if (k.dyn_angle) {
    const double det = c11 * c22 - c12 * c12;
    o[0] = c22 / det;
    o[1] = -c12 / det;
    o[2] = c11 / det;
  } else {
    o[0] = 1.0 / c11;
    o[1] = 0.0;
    o[2] = 0.0;
  }
  for (int d = 0; d < 6; ++d) {
    o[3 + d] = Bp[d];
    o[9 + d] = Ba[d];
  }
}

// In-place inverse of a 6 x 6 row-major matrix (Gauss-Jordan, partial pivoting).
__device__ inline void cb_invert6(double* A) {
  double I[36];
  for (int i = 0; i < 36; ++i) I[i] = (i % 7 == 0) ? 1.0 : 0.0;
  for (int col = 0; col < 6; ++col) {
    int piv = col;
    for (int r = col + 1; r < 6; ++r)
      if (fabs(A[6 * r + col]) > fabs(A[6 * piv + col])) piv = r;
    if (piv != col)
      for (int j = 0; j < 6; ++j) {
        double t = A[6 * col + j];
        A[6 * col + j] = A[6 * piv + j];
        A[6 * piv + j] = t;
        t = I[6 * col + j];
        I[6 * col + j] = I[6 * piv + j];
        I[6 * piv + j] = t;
      }
    const double inv = 1.0 / A[6 * col + col];
    for (int j = 0; j < 6; ++j) {
      A[6 * col + j] *= inv;
      I[6 * col + j] *= inv;
    }
    for (int r = 0; r < 6; ++r) {
      if (r == col) continue;
      const double f = A[6 * r + col];
      for (int j = 0; j < 6; ++j) {
        A[6 * r + j] -= f * A[6 * col + j];
        I[6 * r + j] -= f * I[6 * col + j];
      }
    }
  }
  for (int i = 0; i < 36; ++i) A[i] = I[i];
}

// Per body: (K_b^T P K_b - sum_c B_c S_cc^-1 B_c^T)^-1 (one block per body).
__global__ __launch_bounds__(256) void cb_pc_body_prep_kernel(BodyCiliaK k) {
  const int b = blockIdx.x;
  const int nsb = k.blk ? 0 : k.per_body * k.nseg;  // segments handled here
  const int cnt = nsb + (k.dense ? 0 : k.nb);
  const double c0 = k.bx[3 * b], c1 = k.bx[3 * b + 1], c2 = k.bx[3 * b + 2];
  double A[36];
  for (int j = 0; j < 36; ++j) A[j] = 0.0;
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
    const int i = t < nsb ? cb_particle(k, b, t) : k.nseg_total + b * k.nb + (t - nsb);
    double p[3];
    cb_pdiag(k, i, p);
    const double r0 = k.px[i] - c0, r1 = k.py[i] - c1, r2 = k.pz[i] - c2;
    // K_i = [I | -[r]x]: row d holds e_d and -([r]x)_d
    const double K[3][6] = {{1.0, 0.0, 0.0, 0.0, r2, -r1},
                            {0.0, 1.0, 0.0, -r2, 0.0, r0},
                            {0.0, 0.0, 1.0, r1, -r0, 0.0}};
    for (int a = 0; a < 6; ++a)
      for (int e = 0; e < 6; ++e)
        A[6 * a + e] += K[0][a] * p[0] * K[0][e] + K[1][a] * p[1] * K[1][e] + K[2][a] * p[2] * K[2][e];
  }
  cb_block_sum<36>(A);
  if (threadIdx.x != 0) return;
  if (k.blk)
    for (int l = 0; l < k.per_body; ++l)
      for (int j = 0; j < 36; ++j) A[j] += k.dcil[36 * (size_t)(b * k.per_body + l) + j];
  if (k.dynamic) {
    for (int l = 0; l < k.per_body; ++l) {
      const double* o = k.pc_c + 15 * (size_t)(b * k.per_body + l);
      const double *Bp = o + 3, *Ba = o + 9;
      for (int a = 0; a < 6; ++a)
        for (int e = 0; e < 6; ++e)
          A[6 * a + e] -= o[0] * Bp[a] * Bp[e] + o[1] * (Bp[a] * Ba[e] + Ba[a] * Bp[e]) + o[2] * Ba[a] * Ba[e];
    }
  }
  if (k.rot && !k.dense)  // translation block; Omega rows -I against t_Omega = -r_Omega
    for (int a = 0; a < 6; ++a)
      for (int e = 0; e < 6; ++e)
        if (a >= 3 || e >= 3) A[6 * a + e] = (a == e) ? -1.0 : 0.0;
  if (!k.dense) cb_invert6(A);
  for (int j = 0; j < 36; ++j) k.pc_b[36 * (size_t)b + j] = A[j];
}

// mode 0: out_i = P r_i;  mode 1: out_i += P (V_b + Omega_b x r_i + phi k_phi + theta k_theta)
__global__ __launch_bounds__(256) void cb_pc_particles_kernel(BodyCiliaK k, int mode, const double* __restrict__ r,
                                                              double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k.M || (k.dense && i >= k.nseg_total) || (k.blk && i < k.nseg_total)) return;
  const size_t M = k.M;
  double p[3];
  cb_pdiag(k, i, p);
  if (mode == 0) {
    for (int d = 0; d < 3; ++d) out[d * M + i] = p[d] * r[d * M + i];
    return;
  }
  const int b = cb_body(k, i);
  const double* U = out + 3 * M + 6 * (size_t)b;
  const double r0 = k.px[i] - k.bx[3 * b], r1 = k.py[i] - k.bx[3 * b + 1], r2 = k.pz[i] - k.bx[3 * b + 2];
  const double w0 = k.rot ? 0.0 : U[3], w1 = k.rot ? 0.0 : U[4], w2 = k.rot ? 0.0 : U[5];
  double u[3] = {U[0] + (w1 * r2 - w2 * r1), U[1] + (w2 * r0 - w0 * r2), U[2] + (w0 * r1 - w1 * r0)};
  if (k.dynamic && i < k.nseg_total) {
    const size_t ph = 3 * M + 6 * (size_t)k.nbody, ns = k.nseg_total;
    const int c = i / k.nseg;
    const double wp = out[ph + c], wa = k.dyn_angle ? out[ph + k.ncil + c] : 0.0;
    for (int d = 0; d < 3; ++d) u[d] += wp * k.vphase[d * ns + i] + (k.dyn_angle ? wa * k.vangle[d * ns + i] : 0.0);
  }
  for (int d = 0; d < 3; ++d) out[d * M + i] += p[d] * u[d];
}

// mode 0: (phi, theta)_C = S_cc^-1 t with t = -r_c - (k_phi, k_theta) . (P r);
// mode 1: (phi, theta) = (phi, theta)_C - S_cc^-1 B_c^T U_b
__global__ void cb_pc_cilia_kernel(BodyCiliaK k, int mode, const double* __restrict__ r, double* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k.ncil) return;
  const size_t M = k.M, ns = k.nseg_total, ph = 3 * M + 6 * (size_t)k.nbody;
  const double* o = k.pc_c + 15 * (size_t)c;
  if (mode == 0) {
    double tp = -r[ph + c], ta = k.dyn_angle ? -r[ph + k.ncil + c] : 0.0;
    for (int i = 0; i < k.nseg; ++i) {
      const size_t g = (size_t)c * k.nseg + i;
      for (int d = 0; d < 3; ++d) {
        tp -= k.vphase[d * ns + g] * out[d * M + g];
        if (k.dyn_angle) ta -= k.vangle[d * ns + g] * out[d * M + g];
      }
    }
    out[ph + c] = o[0] * tp + o[1] * ta;
    if (k.dyn_angle) out[ph + k.ncil + c] = o[1] * tp + o[2] * ta;
    return;
  }
  const int b = c / k.per_body;
  const double* U = out + 3 * M + 6 * (size_t)b;
  double bp = 0.0, ba = 0.0;
  for (int d = 0; d < (k.rot ? 3 : 6); ++d) {
    bp += o[3 + d] * U[d];
    ba += o[9 + d] * U[d];
  }
  out[ph + c] -= o[0] * bp + o[1] * ba;
  if (k.dyn_angle) out[ph + k.ncil + c] -= o[1] * bp + o[2] * ba;
}

// U_b = Sinv_b (-r_b - sum_i K_i^T (P r)_i - sum_c B_c (phi, theta)_C) (one block per body)
__global__ __launch_bounds__(256) void cb_pc_body_kernel(BodyCiliaK k, const double* __restrict__ r,
                                                         double* __restrict__ out) {
  const int b = blockIdx.x;
  const size_t M = k.M;
  const int cnt = k.per_body * k.nseg + k.nb;
  const double c0 = k.bx[3 * b], c1 = k.bx[3 * b + 1], c2 = k.bx[3 * b + 2];
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
    const int i = cb_particle(k, b, t);
    const double f0 = out[i], f1 = out[M + i], f2 = out[2 * M + i];
    const double r0 = k.px[i] - c0, r1 = k.py[i] - c1, r2 = k.pz[i] - c2;
    acc[0] += f0;
    acc[1] += f1;
    acc[2] += f2;
    acc[3] += r1 * f2 - r2 * f1;
    acc[4] += r2 * f0 - r0 * f2;
    acc[5] += r0 * f1 - r1 * f0;
  }
  cb_block_sum<6>(acc);
  if (threadIdx.x != 0) return;
  const size_t ub = 3 * M + 6 * (size_t)b, ph = 3 * M + 6 * (size_t)k.nbody;
  double t[6];
  for (int d = 0; d < 6; ++d) t[d] = -r[ub + d] - acc[d];
  if (k.dynamic) {
    for (int l = 0; l < k.per_body; ++l) {
      const int c = b * k.per_body + l;
      const double* o = k.pc_c + 15 * (size_t)c;
      const double wp = out[ph + c], wa = k.dyn_angle ? out[ph + k.ncil + c] : 0.0;
      for (int d = 0; d < 6; ++d) t[d] -= o[3 + d] * wp + o[9 + d] * wa;
    }
  }
  if (k.rot)
    for (int d = 3; d < 6; ++d) t[d] = -r[ub + d];
  if (k.dense) {  // the bodies couple through P: one full solve (cb_full_gemv_kernel)
    for (int d = 0; d < 6; ++d) k.tU[6 * (size_t)b + d] = t[d];
    return;
  }
  const double* Si = k.pc_b + 36 * (size_t)b;
  for (int a = 0; a < 6; ++a) {
    double v = 0.0;
    for (int e = 0; e < 6; ++e) v += Si[6 * a + e] * t[e];
    out[ub + a] = v;
  }
}

// U = S^-1 t with the full (6 n_body)^2 inverse (row-major)
__global__ void cb_full_gemv_kernel(int n, const double* __restrict__ Sinv, const double* __restrict__ t,
                                    double* __restrict__ U) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double v = 0.0;
  for (int j = 0; j < n; ++j) v += Sinv[(size_t)r * n + j] * t[j];
  U[r] = v;
}

// rhs with prescribed phase speeds: phase_dot k_phi on the segment rows, zero elsewhere
__global__ void cb_rhs_kernel(int L, int M, int nseg_total, int nseg, const double* __restrict__ pdot,
                              const double* __restrict__ vphase, double* __restrict__ rhs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  double v = 0.0;
  if (i < 3 * M) {
    const int d = i / M, j = i - d * M;
    if (j < nseg_total) v = pdot[j / nseg] * vphase[(size_t)d * nseg_total + j];
  }
  rhs[i] = v;
}

// solution -> segment forces, blob forces, body velocities, phase and angle speeds
__global__ void cb_unpack_kernel(BodyCiliaK k, const double* __restrict__ z, double* __restrict__ f,
                                 double* __restrict__ fb, double* __restrict__ vb, double* __restrict__ pdot,
                                 double* __restrict__ adot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t M = k.M;
  if (i < 3 * k.M) {
    const int d = i / k.M, j = i - d * k.M;
    if (j < k.nseg_total)
      f[(size_t)d * k.nseg_total + j] = z[i];
    else
      fb[(size_t)d * (k.M - k.nseg_total) + (j - k.nseg_total)] = z[i];
  }
  if (i < 6 * k.nbody) vb[i] = z[3 * M + i];
  if (k.dynamic && i < k.ncil) {
    const size_t ph = 3 * M + 6 * (size_t)k.nbody;
    pdot[i] = z[ph + i];
    adot[i] = k.dyn_angle ? z[ph + k.ncil + i] : 0.0;
  }
}

// Explicit body update after a step (cilia_sim_main.cu:724-740):
// x += dt V, q = exp(dt Omega) q.
__global__ void cb_advance_bodies_kernel(int nbody, double dt, const double* __restrict__ vb,
                                         double* __restrict__ bx, double* __restrict__ bq) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbody) return;
  for (int d = 0; d < 3; ++d) bx[3 * b + d] += dt * vb[6 * b + d];
  const fil::Q4 e = fil::lie_exp(fil::v3(dt * vb[6 * b + 3], dt * vb[6 * b + 4], dt * vb[6 * b + 5]));
  const fil::Q4 q = {bq[4 * b], bq[4 * b + 1], bq[4 * b + 2], bq[4 * b + 3]};
  const fil::Q4 r = fil::qmul(e, q);
  bq[4 * b] = r.s;
  bq[4 * b + 1] = r.x;
  bq[4 * b + 2] = r.y;
  bq[4 * b + 3] = r.z;
}

// ---- per-cilium block preconditioner (CFM_PC_NEAR, CFM_PC_DENSE) ----------------------

// Unit-load response of segment i to a load on segment j, 3 x 3 (row p, column q).
__device__ inline void cb_seg_block(const BodyCiliaK& k, int i, int j, int q, double V[3]) {
  double F[3] = {0.0, 0.0, 0.0};
  F[q] = 1.0;
  V[0] = V[1] = V[2] = 0.0;
  if (k.wall) {
    if (i == j)
      wall_self<3, 3>(k.pz[i], F, V);
    else
      wall_pair<double, 3, 3>(k.px[i] - k.px[j], k.py[i] - k.py[j], k.pz[i] - k.pz[j], k.pz[i] + k.pz[j],
                              k.pz[j], F, V);
    for (int p = 0; p < 3; ++p) V[p] /= k.mu;
  } else {
    free_pair<double, 3, 3>(k.px[i] - k.px[j], k.py[i] - k.py[j], k.pz[i] - k.pz[j], k.a_seg, k.a_seg, k.mu,
                            F, V);
  }
}

// One workgroup per cilium: P_c = M_cc^-1 by in-place Gauss-Jordan (M_cc is symmetric
// positive definite: no pivoting), then P_c times the cilium's K columns, k_phi and
// k_theta, and its contributions S_cc^-1, B and K^T P_c K.
__global__ __launch_bounds__(256) void cb_block_prep_kernel(BodyCiliaK k) {
  const int c = blockIdx.x, n = 3 * k.nseg, s0 = c * k.nseg, b = c / k.per_body;
  double* A = k.pblk + (size_t)c * n * n;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int r = e / n, col = e - r * n;
    double V[3];
    cb_seg_block(k, s0 + r / 3, s0 + col / 3, col % 3, V);
    A[e] = V[r % 3];
  }
  __syncthreads();
  __shared__ double rowk[3 * kMaxCiliaSeg], colk[3 * kMaxCiliaSeg];
  for (int kk = 0; kk < n; ++kk) {
    const double piv = 1.0 / A[kk * n + kk];
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      rowk[j] = (j == kk) ? 1.0 : A[kk * n + j] * piv;
      colk[j] = A[j * n + kk];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
      const int r = e / n, j = e - r * n;
      if (r == kk)
        A[e] = (j == kk) ? piv : rowk[j];
      else if (j == kk)
        A[e] = -colk[r] * piv;
      else
        A[e] -= colk[r] * rowk[j];
    }
    __syncthreads();
  }
  // columns of the cilium: K (6) and the velocity directions (2), then their images
  const double c0 = k.bx[3 * b], c1 = k.bx[3 * b + 1], c2 = k.bx[3 * b + 2];
  const size_t ns = k.nseg_total;
  double* W = k.pkc + (size_t)c * 8 * n;
  for (int e = threadIdx.x; e < 8 * n; e += blockDim.x) {
    const int col = e / n, r = e - col * n;
    double acc = 0.0;
    for (int j = 0; j < n; ++j) {
      const int g = s0 + j / 3, d = j % 3;
      double v;
      if (col < 3) {
        v = (d == col) ? 1.0 : 0.0;
      } else if (col < 6) {  // e_a x r_g
        const double rr[3] = {k.px[g] - c0, k.py[g] - c1, k.pz[g] - c2};
        const int a = col - 3, a1 = (a + 1) % 3, a2 = (a + 2) % 3;
        v = (d == a2) ? rr[a1] : (d == a1) ? -rr[a2] : 0.0;
      } else {
        v = (col == 6 ? k.vphase : k.vangle)[(size_t)d * ns + g];
      }
      acc += A[r * n + j] * v;
    }
    W[e] = acc;
  }
  __syncthreads();
  if (threadIdx.x >= 64) return;
  // K^T P_c K (36 entries), B = K^T P_c (k_phi, k_theta), S_cc (thread 0)
  const int t = threadIdx.x;
  auto colv = [&](int col, int j) -> double {
    const int g = s0 + j / 3, d = j % 3;
    if (col < 3) return (d == col) ? 1.0 : 0.0;
    if (col < 6) {
      const double rr[3] = {k.px[g] - c0, k.py[g] - c1, k.pz[g] - c2};
      const int a = col - 3, a1 = (a + 1) % 3, a2 = (a + 2) % 3;
      return (d == a2) ? rr[a1] : (d == a1) ? -rr[a2] : 0.0;
    }
    return (col == 6 ? k.vphase : k.vangle)[(size_t)d * ns + g];
  };
  if (t < 36) {
    const int a = t / 6, e = t - 6 * (t / 6);
    double v = 0.0;
    for (int j = 0; j < n; ++j) v += colv(a, j) * W[(size_t)e * n + j];
    k.dcil[36 * (size_t)c + t] = v;
  } else if (t < 48) {
    const int q = t - 36, a = q % 6, which = q / 6;  // B_phi (a) or B_theta (a)
    double v = 0.0;
    for (int j = 0; j < n; ++j) v += colv(a, j) * W[(size_t)(6 + which) * n + j];
    k.pc_c[15 * (size_t)c + 3 + 6 * which + a] = v;
  } else if (t == 48) {
    double c11 = 0.0, c12 = 0.0, c22 = 0.0;
    for (int j = 0; j < n; ++j) {
      c11 += colv(6, j) * W[(size_t)6 * n + j];
      c12 += colv(6, j) * W[(size_t)7 * n + j];
      c22 += colv(7, j) * W[(size_t)7 * n + j];
    }
    double* o = k.pc_c + 15 * (size_t)c;
    if (k.dyn_angle) {
      const double det = c11 * c22 - c12 * c12;
      o[0] = c22 / det;
      o[1] = -c12 / det;
      o[2] = c11 / det;
    } else {
      o[0] = 1.0 / c11;
      o[1] = 0.0;
      o[2] = 0.0;
    }
  }
}

// mode 0: out_c = P_c r_c;  mode 1: out_c += P_c (K U + phi k_phi + theta k_theta)
// (one workgroup per cilium, one thread per row)
__global__ __launch_bounds__(256) void cb_block_apply_kernel(BodyCiliaK k, int mode, const double* __restrict__ r,
                                                             double* __restrict__ out) {
  const int c = blockIdx.x, n = 3 * k.nseg, s0 = c * k.nseg;
  const size_t M = k.M;
  const int row = threadIdx.x;
  if (row >= n) return;
  const int g = s0 + row / 3, d = row % 3;
  if (mode == 0) {
    const double* A = k.pblk + (size_t)c * n * n + (size_t)row * n;
    double v = 0.0;
    for (int j = 0; j < n; ++j) v += A[j] * r[(size_t)(j % 3) * M + s0 + j / 3];
    out[(size_t)d * M + g] = v;
    return;
  }
  const int b = c / k.per_body;
  const double* U = out + 3 * M + 6 * (size_t)b;
  const double* W = k.pkc + (size_t)c * 8 * n;
  double v = 0.0;
  for (int a = 0; a < (k.rot ? 3 : 6); ++a) v += W[(size_t)a * n + row] * U[a];
  if (k.dynamic) {
    const size_t ph = 3 * M + 6 * (size_t)k.nbody;
    v += W[(size_t)6 * n + row] * out[ph + c];
    if (k.dyn_angle) v += W[(size_t)7 * n + row] * out[ph + k.ncil + c];
  }
  out[(size_t)d * M + g] += v;
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
  DriveOpts dopt{0, 0.0, 0ull, 0};  // cfm_cilia_set_driving_options
  DevBuf<double> omega0, adot, tab_phase, tab_angle, z, zrhs;
  int n_tab_phase = 0, n_tab_angle = 0;
  // recorded segment parameters (cfm_cilia_set_s_table)
  DevBuf<double> s_table;
  int s_rows = 0, s_cols = 0;
  double pair_dp = 0.0;  // phase offset of the second cilium of a pair (cfm_cilia_set_pair_dp)
  // cilia carried by free bodies (cfm_cilia_init_bodies): body centres, quaternions and
  // velocities, the cilia's body-frame bases and frames, blob forces, and the block
  // preconditioner's per-cilium and per-body matrices
  int n_body = 0, per_body = 0;
  DevBuf<double> bx, bq, vbody, fref, fq, fblob, pc_c, pc_b;
  DevBuf<double> tU, pc_full;  // CFM_PC_DENSE: body right-hand sides, full Schur inverse
  DevBuf<double> pblk, pkc, dcil;  // CFM_PC_NEAR / _DENSE: per-cilium blocks
  bool pc_ready = false;
  int pc_kind = -1, pc_rot = -1;
};

// First phase-speed slot of the system vector, and the system length.
static size_t phase_slot(const cfm_ctx* c, const CiliaState* S) {
  return S->n_body > 0 ? 3 * (size_t)(c->p.n_seg + c->p.n_blob) + 6 * (size_t)S->n_body
                       : 3 * (size_t)c->p.n_seg;
}
static size_t system_len(const cfm_ctx* c, const CiliaState* S) {
  return phase_slot(c, S) + (S->dynamic ? (size_t)S->prm.n_cilia * (S->dyn_angle ? 2 : 1) : 0);
}

static CiliaState* cil_of(cfm_ctx* c) { return reinterpret_cast<CiliaState*>(c->cilia); }

static DriveOpts drive_opts(const CiliaState* S, int nt) {
  DriveOpts d = S->dopt;
  d.nt = nt;
  return d;
}

static CiliaK kargs(cfm_ctx* c, CiliaState* S) {
  CiliaK k{};
  beat_args(S->prm, S->pair_dp, c->cfg.a_seg, k);
  k.s_table = S->s_rows > 0 ? S->s_table.p : nullptr;
  k.s_rows = S->s_rows;
  k.s_cols = S->s_cols;
  k.ncil = S->prm.n_cilia;
  k.nseg = S->prm.nseg;
  k.nseg_total = c->p.n_seg;
  k.dl = S->prm.dl;
  k.base = S->base.p;
  if (S->n_body > 0) {
    k.bx = S->bx.p;
    k.bq = S->bq.p;
    k.fref = S->fref.p;
    k.fq = S->fq.p;
    k.per_body = S->per_body;
  }
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
  const CiliaK gk = kargs(c, S);
  const GeoLaunch gl = geometry_launch(gk);
  hipLaunchKernelGGL(cilia_geometry_kernel, gl.grid, gl.block, gl.lds, c->stream, gk);
  CFM_CHECK_LAUNCH();
  c->have_positions = true;
  S->have_geometry = true;
  S->pc_ready = false;
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

static BodyCiliaK body_cilia_args(cfm_ctx* c, CiliaState* S, int pc_kind) {
  BodyCiliaK k{};
  k.ncil = S->prm.n_cilia;
  k.nseg = S->prm.nseg;
  k.nseg_total = c->p.n_seg;
  k.nbody = S->n_body;
  k.per_body = S->per_body;
  k.nb = c->p.n_blob / S->n_body;
  k.M = c->p.n_seg + c->p.n_blob;
  k.dynamic = S->dynamic ? 1 : 0;
  k.dyn_angle = S->dyn_angle ? 1 : 0;
  k.rot = body_rotation_prescribed(c) ? 1 : 0;
  k.px = c->p.x.p;
  k.py = c->p.y.p;
  k.pz = c->p.z.p;
  k.bx = S->bx.p;
  k.vphase = S->vphase.p;
  k.vangle = S->vangle.p;
  k.pc_kind = pc_kind;
  k.wall = c->cfg.wall;
  k.mu = c->cfg.mu;
  k.a_seg = c->cfg.a_seg;
  k.a_blob = c->cfg.a_blob;
  k.pc_c = S->pc_c.p;
  k.pc_b = S->pc_b.p;
  k.dense = pc_kind == CFM_PC_DENSE ? 1 : 0;
  k.tU = S->tU.p;
  k.blk = (pc_kind == CFM_PC_NEAR || pc_kind == CFM_PC_DENSE) ? 1 : 0;
  k.pblk = S->pblk.p;
  k.pkc = S->pkc.p;
  k.dcil = S->dcil.p;
  if (k.blk) k.pc_kind = CFM_PC_DIAG;  // blobs outside the dense block: 6 pi mu a
  return k;
}

// The system of cilia on free bodies: forces of every particle, body velocities, speeds.
static SystemDesc body_cilia_desc(cfm_ctx* c, CiliaState* S) {
  SystemDesc sd;
  sd.m = c->p.n_seg + c->p.n_blob;
  sd.dof = 3;
  sd.off = 0;
  sd.which = CFM_FULL;
  sd.a = c->cfg.a_seg;
  sd.cilia = true;
  sd.extra = 6 * S->n_body + (S->dynamic ? S->prm.n_cilia * (S->dyn_angle ? 2 : 1) : 0);
  sd.L = 3 * sd.m + sd.extra;
  return sd;
}

// make_precon_mat (swimmer.cpp:1175-1347) and the per-cilium blocks, for the current
// geometry and the preconditioner kind of the step's solve.
static Status body_cilia_prep(cfm_ctx* c, CiliaState* S, int kind) {
  if (kind == CFM_PC_NEAR || kind == CFM_PC_DENSE) {
    const size_t n = 3 * (size_t)S->prm.nseg, nc = S->prm.n_cilia;
    CFM_TRY(S->pblk.alloc(nc * n * n));
    CFM_TRY(S->pkc.alloc(nc * 8 * n));
    CFM_TRY(S->dcil.alloc(nc * 36));
  }
  const BodyCiliaK k = body_cilia_args(c, S, kind);
  if (k.blk)
    hipLaunchKernelGGL(cb_block_prep_kernel, dim3(k.ncil), dim3(256), 0, c->stream, k);
  else
    hipLaunchKernelGGL(cb_pc_cilia_prep_kernel, gc(k.ncil), dim3(64), 0, c->stream, k);
  hipLaunchKernelGGL(cb_pc_body_prep_kernel, dim3(k.nbody), dim3(256), 0, c->stream, k);
  CFM_CHECK_LAUNCH();
  if (k.dense) {
    // S = K_b^T P K_b (resident inverse, all bodies) + per body the segment part less
    // the cilia's Schur terms; inverted once per geometry on the host
    const int n = 6 * k.nbody;
    CFM_TRY(body_dense_ready(c));
    std::vector<double> R, D(36 * (size_t)k.nbody);
    CFM_TRY(body_dense_resistance(c, &R));
    CFM_HIP(hipMemcpyAsync(D.data(), S->pc_b.p, sizeof(double) * D.size(), hipMemcpyDeviceToHost, c->stream));
    CFM_HIP(hipStreamSynchronize(c->stream));
    for (int b = 0; b < k.nbody; ++b)
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) R[(size_t)(6 * b + i) * n + 6 * b + j] += D[36 * (size_t)b + 6 * i + j];
    if (k.rot)  // translation block; Omega rows -I against t_Omega = -r_Omega
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          if (i % 6 >= 3 || j % 6 >= 3) R[(size_t)i * n + j] = (i == j) ? -1.0 : 0.0;
    // symmetric positive definite: Gauss-Jordan with partial pivoting
    std::vector<double> I((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) I[(size_t)i * n + i] = 1.0;
    for (int col = 0; col < n; ++col) {
      int piv = col;
      for (int r = col + 1; r < n; ++r)
        if (std::fabs(R[(size_t)r * n + col]) > std::fabs(R[(size_t)piv * n + col])) piv = r;
      if (!(std::fabs(R[(size_t)piv * n + col]) > 0.0))
        return make_status(CFM_ENONFINITE, "singular body Schur complement");
      for (int j = 0; j < n; ++j) {
        std::swap(R[(size_t)col * n + j], R[(size_t)piv * n + j]);
        std::swap(I[(size_t)col * n + j], I[(size_t)piv * n + j]);
      }
      const double inv = 1.0 / R[(size_t)col * n + col];
      for (int j = 0; j < n; ++j) {
        R[(size_t)col * n + j] *= inv;
        I[(size_t)col * n + j] *= inv;
      }
      for (int r = 0; r < n; ++r) {
        if (r == col) continue;
        const double f = R[(size_t)r * n + col];
        if (f == 0.0) continue;
        for (int j = 0; j < n; ++j) {
          R[(size_t)r * n + j] -= f * R[(size_t)col * n + j];
          I[(size_t)r * n + j] -= f * I[(size_t)col * n + j];
        }
      }
    }
    CFM_TRY(S->pc_full.alloc((size_t)n * n));
    CFM_TRY(S->tU.alloc(n));
    CFM_HIP(hipMemcpyAsync(S->pc_full.p, I.data(), sizeof(double) * I.size(), hipMemcpyHostToDevice, c->stream));
    CFM_HIP(hipStreamSynchronize(c->stream));
  }
  S->pc_ready = true;
  S->pc_kind = kind;
  S->pc_rot = k.rot;
  return {};
}

Status cilia_couple_apply(cfm_ctx* c, const double* x, double* y) {
  CiliaState* S = cil_of(c);
  if (S && S->n_body > 0) {
    const BodyCiliaK k = body_cilia_args(c, S, CFM_PC_DIAG);
    hipLaunchKernelGGL(cb_couple_particles_kernel, dim3((k.M + 255) / 256), dim3(256), 0, c->stream, k, x, y);
    if (k.dynamic) hipLaunchKernelGGL(cb_couple_cilia_kernel, gc(k.ncil), dim3(64), 0, c->stream, k, x, y);
    hipLaunchKernelGGL(cb_couple_body_kernel, dim3(k.nbody), dim3(256), 0, c->stream, k, x, y);
    CFM_CHECK_LAUNCH();
    return {};
  }
  if (!S || !S->dynamic) return make_status(CFM_ESTATE, "no coupled cilia system");
  hipLaunchKernelGGL(cilia_couple_kernel, gc(S->prm.n_cilia), dim3(64), 0, c->stream,
                     couple_args(c, S), x, y);
  CFM_CHECK_LAUNCH();
  return {};
}

Status cilia_precond(cfm_ctx* c, int kind, const double* x, double* y) {
  CiliaState* S = cil_of(c);
  if (S && S->n_body > 0) {
    if (kind == CFM_PC_NONE) {
      CFM_HIP(hipMemcpyAsync(y, x, sizeof(double) * system_len(c, S), hipMemcpyDeviceToDevice, c->stream));
      return {};
    }
    if (kind != CFM_PC_DIAG && kind != CFM_PC_SELF && kind != CFM_PC_NEAR && kind != CFM_PC_DENSE)
      return make_status(CFM_EINVAL, "cilia on free bodies take CFM_PC_NONE, _DIAG, _SELF, _NEAR or _DENSE");
    if (!S->pc_ready || S->pc_kind != kind || S->pc_rot != (body_rotation_prescribed(c) ? 1 : 0))
      CFM_TRY(body_cilia_prep(c, S, kind));
    const BodyCiliaK k = body_cilia_args(c, S, kind);
    const dim3 gp((k.M + 255) / 256);
    const size_t M = k.M, ns = k.nseg_total;
    hipLaunchKernelGGL(cb_pc_particles_kernel, gp, dim3(256), 0, c->stream, k, 0, x, y);
    if (k.blk) hipLaunchKernelGGL(cb_block_apply_kernel, dim3(k.ncil), dim3(256), 0, c->stream, k, 0, x, y);
    if (k.dense) CFM_TRY(body_dense_blob_apply(c, x + ns, M, y + ns, M));
    if (k.dynamic) hipLaunchKernelGGL(cb_pc_cilia_kernel, gc(k.ncil), dim3(64), 0, c->stream, k, 0, x, y);
    hipLaunchKernelGGL(cb_pc_body_kernel, dim3(k.nbody), dim3(256), 0, c->stream, k, x, y);
    if (k.dense)
      hipLaunchKernelGGL(cb_full_gemv_kernel, gc(6 * k.nbody), dim3(64), 0, c->stream, 6 * k.nbody,
                         S->pc_full.p, S->tU.p, y + 3 * M);
    if (k.dynamic) hipLaunchKernelGGL(cb_pc_cilia_kernel, gc(k.ncil), dim3(64), 0, c->stream, k, 1, x, y);
    hipLaunchKernelGGL(cb_pc_particles_kernel, gp, dim3(256), 0, c->stream, k, 1, x, y);
    if (k.blk) hipLaunchKernelGGL(cb_block_apply_kernel, dim3(k.ncil), dim3(256), 0, c->stream, k, 1, x, y);
    if (k.dense) CFM_TRY(body_dense_rigid_add(c, y + 3 * M, y + ns, M));
    CFM_CHECK_LAUNCH();
    return {};
  }
  if (!S || !S->dynamic) return make_status(CFM_ESTATE, "no coupled cilia system");
  if (kind != CFM_PC_DIAG && kind != CFM_PC_SELF)
    return make_status(CFM_EINVAL, "the coupled cilia system takes CFM_PC_DIAG or CFM_PC_SELF");
  hipLaunchKernelGGL(cilia_precond_kernel, gc(S->prm.n_cilia), dim3(64), 0, c->stream,
                     couple_args(c, S), kind, c->cfg.mu, c->p.z.p, x, y);
  CFM_CHECK_LAUNCH();
  return {};
}

// One step of cilia on free bodies: explicit update of phases, shape angles and bodies
// with the previous step's speeds, geometry, block preconditioner, one solve.
static Status step_bodies(cfm_ctx* c, CiliaState* S, int nt, double* q_phase, double* q_angle,
                          cfm_solve_stats* st) {
  const int nc = S->prm.n_cilia;
  if (nt > 0) {
    // phases before the step (filament.cpp:1340-1356); bodies after the previous step
    // (cilia_sim_main.cu:724-740) -- the same state at this step's solve
    hipLaunchKernelGGL(cilia_advance_kernel, gc(nc), dim3(64), 0, c->stream, nc, S->prm.dt, S->pdot.p,
                       S->phase.p);
    if (S->dyn_angle)
      hipLaunchKernelGGL(cilia_advance_angle_kernel, gc(nc), dim3(64), 0, c->stream, nc, S->prm.dt,
                         S->adot.p, S->angle.p);
    hipLaunchKernelGGL(cb_advance_bodies_kernel, gc(S->n_body), dim3(64), 0, c->stream, S->n_body,
                       S->prm.dt, S->vbody.p, S->bx.p, S->bq.p);
    CFM_CHECK_LAUNCH();
    CFM_TRY(body_place_device(c, S->bx.p, S->bq.p));
  }
  CFM_TRY(geometry(c, S));
  const SystemDesc sd = body_cilia_desc(c, S);
  cfm_solve_opts o = S->prm.solve;
  o.use_guess = 1;  // previous step's solution
  if (o.precond != CFM_PC_NONE) CFM_TRY(body_cilia_prep(c, S, o.precond));
  if (S->dynamic) {
    hipLaunchKernelGGL(cilia_drive_kernel, dim3((sd.L + 255) / 256), dim3(256), 0, c->stream, nc,
                       (int)phase_slot(c, S), S->dyn_angle ? 1 : 0, cilium_length(S->prm),
                       S->spring, S->phase.p, S->angle.p, S->omega0.p, S->tab_phase.p, S->n_tab_phase,
                       S->tab_angle.p, S->n_tab_angle, drive_opts(S, nt), S->zrhs.p);
  } else {
    hipLaunchKernelGGL(cb_rhs_kernel, dim3((sd.L + 255) / 256), dim3(256), 0, c->stream, sd.L, sd.m,
                       c->p.n_seg, S->prm.nseg, S->pdot.p, S->vphase.p, S->zrhs.p);
  }
  CFM_CHECK_LAUNCH();
  if (!c->krylov) c->krylov = new Krylov();
  cfm_solve_stats ss{};
  CFM_TRY(solve(c, *c->krylov, sd, o, S->zrhs.p, S->z.p, &ss, nullptr));
  const BodyCiliaK k = body_cilia_args(c, S, o.precond);
  const int nu = std::max(3 * k.M, std::max(6 * k.nbody, nc));
  hipLaunchKernelGGL(cb_unpack_kernel, dim3((nu + 255) / 256), dim3(256), 0, c->stream, k, S->z.p, S->f.p,
                     S->fblob.p, S->vbody.p, S->pdot.p, S->adot.p);
  hipLaunchKernelGGL(cilia_genforce_kernel, gc(nc), dim3(64), 0, c->stream, nc, S->prm.nseg, c->p.n_seg,
                     S->vphase.p, S->vangle.p, S->f.p, S->q.p);
  CFM_CHECK_LAUNCH();
  CFM_HIP(hipMemcpyAsync(S->host, S->q.p, sizeof(double) * 2 * nc, hipMemcpyDeviceToHost, c->stream));
  CFM_HIP(hipStreamSynchronize(c->stream));
  if (q_phase) std::memcpy(q_phase, S->host, sizeof(double) * nc);
  if (q_angle) std::memcpy(q_angle, S->host + nc, sizeof(double) * nc);
  if (st) *st = ss;
  return {};
}

}  // namespace cfm

using namespace cfm;

void cfm_cilia_release(cfm_ctx* c) {
  CiliaState* S = cil_of(c);
  if (!S) return;
  for (auto* b : {&S->base, &S->phase, &S->pdot, &S->angle, &S->vphase, &S->vangle, &S->svals,
                  &S->rhs, &S->f, &S->q, &S->omega0, &S->adot, &S->tab_phase, &S->tab_angle,
                  &S->z, &S->zrhs, &S->s_table, &S->bx, &S->bq, &S->vbody, &S->fref, &S->fq,
                  &S->fblob, &S->pc_c, &S->pc_b, &S->tU, &S->pc_full, &S->pblk, &S->pkc, &S->dcil})
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

// Parameter checks shared by both initialisations.
static Status check_params(const cfm_ctx* c, const cfm_cilia_params* prm, const double* phase,
                           const double* phase_dot) {
  const Particles& p = c->p;
  if (prm->beat < CFM_BEAT_FULFORD_BLAKE || prm->beat > CFM_BEAT_BICILIA_LONGT)
    return make_status(CFM_EINVAL, "unknown beat %d", prm->beat);
  if (cilium_segments(*prm) != prm->nseg && (prm->nseg % 2 != 0 || prm->nseg < 4))
    return make_status(CFM_EINVAL, "paired beats take an even nseg >= 4 (two cilia of nseg / 2)");
  if (!(prm->s_tol >= 0.0) || prm->s_tol >= 0.5)
    return make_status(CFM_EINVAL, "s_tol must lie in [0, 0.5) (0: 0.1 / nseg)");
  if (prm->nseg < 2 || prm->nseg > kMaxCiliaSeg || prm->n_cilia <= 0 ||
      (long long)prm->n_cilia * prm->nseg != p.n_seg)
    return make_status(CFM_EINVAL, "n_cilia x nseg must equal the context's n_seg, 2 <= nseg <= %d",
                       kMaxCiliaSeg);
  if (p.seg_dof != 3) return make_status(CFM_EINVAL, "prescribed cilia need 3-DOF segments");
  if (!(prm->dl > 0.0) || !(prm->dt > 0.0)) return make_status(CFM_EINVAL, "dl and dt must be positive");
  if (prm->solve.precond < CFM_PC_NONE || prm->solve.precond > CFM_PC_SELF)
    return make_status(CFM_EINVAL, "cilia solves take CFM_PC_NONE, _DIAG or _SELF");
  for (int i = 0; i < prm->n_cilia; ++i)
    if (!std::isfinite(phase[i]) || !std::isfinite(phase_dot[i]))
      return make_status(CFM_EINVAL, "non-finite phase or phase speed");
  return {};
}

int cfm_cilia_init(cfm_ctx* c, const cfm_cilia_params* prm, const double* base,
                   const double* phase, const double* phase_dot) {
  if (!c) return CFM_EINVAL;
  if (!prm || !base || !phase || !phase_dot)
    return cfail(c, make_status(CFM_EINVAL, "null argument"));
  const Particles& p = c->p;
  const Status chk = check_params(c, prm, phase, phase_dot);
  if (!chk.ok()) return cfail(c, chk);
  if (p.n_blob != 0 || !c->cfg.wall)
    return cfail(c, make_status(CFM_EINVAL, "prescribed cilia need a wall context with 3-DOF segments and no blobs"));
  const int nc = prm->n_cilia;
  for (int i = 0; i < nc; ++i) {
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

int cfm_cilia_init_bodies(cfm_ctx* c, const cfm_cilia_params* prm, const double* fil_ref,
                          const double* fil_q, const double* phase, const double* phase_dot) {
  if (!c) return CFM_EINVAL;
  if (!prm || !fil_ref || !fil_q || !phase || !phase_dot)
    return cfail(c, make_status(CFM_EINVAL, "null argument"));
  cfm_cilia_params pr = *prm;
  const bool dense = pr.solve.precond == CFM_PC_DENSE;
  if (dense || pr.solve.precond == CFM_PC_NEAR) pr.solve.precond = CFM_PC_DIAG;  // checked as the others
  const Status chk = check_params(c, &pr, phase, phase_dot);
  if (!chk.ok()) return cfail(c, chk);
  BodyInfo bi;
  const Status bs = body_info(c, &bi);
  if (!bs.ok()) return cfail(c, bs);
  const int nc = prm->n_cilia;
  if (nc % bi.n_body != 0)
    return cfail(c, make_status(CFM_EINVAL, "n_cilia (%d) must be a multiple of the number of bodies (%d)",
                                nc, bi.n_body));
  const int per = nc / bi.n_body;
  for (int l = 0; l < per; ++l) {
    double qq = 0.0;
    for (int d = 0; d < 4; ++d) {
      if (!std::isfinite(fil_q[4 * l + d]) || (d < 3 && !std::isfinite(fil_ref[3 * l + d])))
        return cfail(c, make_status(CFM_ENONFINITE, "non-finite cilium base or frame"));
      qq += fil_q[4 * l + d] * fil_q[4 * l + d];
    }
    if (std::fabs(qq - 1.0) > 1e-8)
      return cfail(c, make_status(CFM_EINVAL, "cilium frame quaternion %d is not a unit quaternion", l));
  }
  (void)hipSetDevice(c->device);
  cfm_cilia_release(c);
  CiliaState* S = new (std::nothrow) CiliaState();
  if (!S) return CFM_ENOMEM;
  c->cilia = S;
  S->prm = *prm;
  S->n_body = bi.n_body;
  S->per_body = per;
  // the resident inverse of M_bb is built once and applied in the bodies' frame
  body_freeze_dense(c, dense);
  auto run = [&]() -> Status {
    const size_t ns = c->p.n_seg, nbody = bi.n_body;
    CFM_TRY(S->base.alloc(3 * (size_t)nc));
    CFM_TRY(S->phase.alloc(nc));
    CFM_TRY(S->pdot.alloc(nc));
    CFM_TRY(S->angle.alloc(nc));
    CFM_TRY(S->vphase.alloc(3 * ns));
    CFM_TRY(S->vangle.alloc(3 * ns));
    CFM_TRY(S->svals.alloc((size_t)nc * prm->nseg));
    CFM_TRY(S->f.alloc(3 * ns));
    CFM_TRY(S->q.alloc(2 * (size_t)nc));
    CFM_TRY(S->omega0.alloc(nc));
    CFM_TRY(S->adot.alloc(nc));
    CFM_TRY(S->bx.alloc(3 * nbody));
    CFM_TRY(S->bq.alloc(4 * nbody));
    CFM_TRY(S->vbody.alloc(6 * nbody));
    CFM_TRY(S->fref.alloc(3 * (size_t)per));
    CFM_TRY(S->fq.alloc(4 * (size_t)per));
    CFM_TRY(S->fblob.alloc(3 * (size_t)c->p.n_blob));
    CFM_TRY(S->pc_c.alloc(15 * (size_t)nc));
    CFM_TRY(S->pc_b.alloc(36 * nbody));
    const size_t L = system_len(c, S);
    CFM_TRY(S->z.alloc(L));
    CFM_TRY(S->zrhs.alloc(L));
    CFM_HIP(hipHostMalloc((void**)&S->host, sizeof(double) * 2 * nc, hipHostMallocDefault));
    CFM_TRY(upload_aos(c, phase, nc, 1, S->phase.p, nc, 0));
    CFM_TRY(upload_aos(c, phase_dot, nc, 1, S->pdot.p, nc, 0));
    CFM_HIP(hipMemcpyAsync(S->omega0.p, S->pdot.p, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
    CFM_HIP(hipMemcpyAsync(S->bx.p, bi.x.data(), sizeof(double) * 3 * nbody, hipMemcpyHostToDevice, c->stream));
    CFM_HIP(hipMemcpyAsync(S->bq.p, bi.q.data(), sizeof(double) * 4 * nbody, hipMemcpyHostToDevice, c->stream));
    CFM_HIP(hipMemcpyAsync(S->fref.p, fil_ref, sizeof(double) * 3 * per, hipMemcpyHostToDevice, c->stream));
    CFM_HIP(hipMemcpyAsync(S->fq.p, fil_q, sizeof(double) * 4 * per, hipMemcpyHostToDevice, c->stream));
    CFM_TRY(dev_fill(c, S->base.p, 0.0, 3 * nc));
    CFM_TRY(dev_fill(c, S->adot.p, 0.0, nc));
    CFM_TRY(dev_fill(c, S->angle.p, 0.0, nc));
    CFM_TRY(dev_fill(c, S->f.p, 0.0, (int)(3 * ns)));
    CFM_TRY(dev_fill(c, S->vbody.p, 0.0, (int)(6 * nbody)));
    CFM_TRY(dev_fill(c, S->fblob.p, 0.0, 3 * c->p.n_blob));
    CFM_TRY(dev_fill(c, S->z.p, 0.0, (int)L));
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

int cfm_cilia_get_bodies(cfm_ctx* c, double* x, double* q, double* vel, double* f_blob) {
  if (!c) return CFM_EINVAL;
  CiliaState* S = cil_of(c);
  if (!S || S->n_body == 0)
    return cfail(c, make_status(CFM_ESTATE, "cfm_cilia_init_bodies has not been called"));
  (void)hipSetDevice(c->device);
  const size_t nb = S->n_body;
  auto run = [&]() -> Status {
    CFM_HIP(hipStreamSynchronize(c->stream));
    if (x) CFM_HIP(copy_to_host(c->stream, x, S->bx.p, sizeof(double) * 3 * nb));
    if (q) CFM_HIP(copy_to_host(c->stream, q, S->bq.p, sizeof(double) * 4 * nb));
    if (vel) CFM_HIP(copy_to_host(c->stream, vel, S->vbody.p, sizeof(double) * 6 * nb));
    if (f_blob) CFM_TRY(download_aos(c, f_blob, c->p.n_blob, 3, S->fblob.p, c->p.n_blob, 0));
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : cfail(c, s);
}

int cfm_cilia_step(cfm_ctx* c, int32_t nt, double* q_phase, double* q_angle, cfm_solve_stats* st) {
  if (!c) return CFM_EINVAL;
  CiliaState* S = cil_of(c);
  if (!S) return cfail(c, make_status(CFM_ESTATE, "cfm_cilia_init has not been called"));
  (void)hipSetDevice(c->device);
  const int nc = S->prm.n_cilia;
  auto run = [&]() -> Status {
    if (S->n_body > 0) return step_bodies(c, S, nt, q_phase, q_angle, st);
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
                         3 * c->p.n_seg, S->dyn_angle ? 1 : 0, cilium_length(S->prm), S->spring,
                         S->phase.p, S->angle.p, S->omega0.p, S->tab_phase.p, S->n_tab_phase,
                         S->tab_angle.p, S->n_tab_angle, drive_opts(S, nt), S->zrhs.p);
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
    CFM_HIP(copy_to_device(c->stream, S->tab_phase.p, q_phase_table, sizeof(double) * n_phase));
    S->n_tab_phase = n_phase;
    S->dyn_angle = q_angle_table != nullptr;
    if (S->dyn_angle) {
      CFM_TRY(S->tab_angle.alloc(n_angle));
      CFM_HIP(copy_to_device(c->stream, S->tab_angle.p, q_angle_table, sizeof(double) * n_angle));
      S->n_tab_angle = n_angle;
    }
    S->spring = spring;
    S->dynamic = true;
    const size_t L = system_len(c, S);
    CFM_TRY(S->z.alloc(L));
    CFM_TRY(S->zrhs.alloc(L));
    CFM_TRY(dev_fill(c, S->z.p, 0.0, (int)L));
    // initial guess of the speeds: the natural frequencies, no shape rotation
    CFM_HIP(hipMemcpyAsync(S->z.p + phase_slot(c, S), S->omega0.p, sizeof(double) * nc,
                           hipMemcpyDeviceToDevice, c->stream));
    CFM_HIP(hipStreamSynchronize(c->stream));
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : cfail(c, s);
}

int cfm_cilia_set_driving_options(cfm_ctx* c, int32_t spring_model, double noise, uint64_t seed) {
  if (!c) return CFM_EINVAL;
  CiliaState* S = cil_of(c);
  if (!S) return cfail(c, make_status(CFM_ESTATE, "cfm_cilia_init has not been called"));
  if ((spring_model != 0 && spring_model != 1) || !std::isfinite(noise))
    return cfail(c, make_status(CFM_EINVAL, "spring_model is 0 (linear) or 1 (FENE); noise must be finite"));
  S->dopt.fene = spring_model;
  S->dopt.noise = noise;
  S->dopt.seed = seed;
  return CFM_OK;
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
    if (phase) CFM_HIP(copy_to_host(c->stream, phase, S->phase.p, sizeof(double) * nc));
    if (phase_dot) CFM_HIP(copy_to_host(c->stream, phase_dot, S->pdot.p, sizeof(double) * nc));
    if (angle) CFM_HIP(copy_to_host(c->stream, angle, S->angle.p, sizeof(double) * nc));
    if (angle_dot) CFM_HIP(copy_to_host(c->stream, angle_dot, S->adot.p, sizeof(double) * nc));
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

int cfm_cilia_set_pair_dp(cfm_ctx* c, double pair_dp) {
  if (!c) return CFM_EINVAL;
  CiliaState* S = cil_of(c);
  if (!S) return cfail(c, make_status(CFM_ESTATE, "cfm_cilia_init has not been called"));
  if (!std::isfinite(pair_dp)) return cfail(c, make_status(CFM_EINVAL, "non-finite pair_dp"));
  (void)hipSetDevice(c->device);
  auto run = [&]() -> Status {
    S->pair_dp = pair_dp;
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
      CFM_HIP(copy_to_host(c->stream, t.data(), c->p.x.p, sizeof(double) * ns));
      CFM_HIP(copy_to_host(c->stream, t.data() + ns, c->p.y.p, sizeof(double) * ns));
      CFM_HIP(copy_to_host(c->stream, t.data() + 2 * (size_t)ns, c->p.z.p, sizeof(double) * ns));
      for (int i = 0; i < ns; ++i)
        for (int d = 0; d < 3; ++d) x[3 * i + d] = t[(size_t)d * ns + i];
    }
    if (vel_dir_phase) CFM_TRY(download_aos(c, vel_dir_phase, ns, 3, S->vphase.p, ns, 0));
    if (vel_dir_angle) CFM_TRY(download_aos(c, vel_dir_angle, ns, 3, S->vangle.p, ns, 0));
    if (f_seg) CFM_TRY(download_aos(c, f_seg, ns, 3, S->f.p, ns, 0));
    if (s_values)
      CFM_HIP(copy_to_host(c->stream, s_values, S->svals.p, sizeof(double) * nc * S->prm.nseg));
    if (phase) CFM_HIP(copy_to_host(c->stream, phase, S->phase.p, sizeof(double) * nc));
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : cfail(c, s);
}

}  // extern "C"

// ---- ensembles of dynamic-phase cilia lattices -----------------------------------------
// The reference's outer loops (pyfile/find_periodic/newton_solver.py:49-74,
// pyfile/bisection/bisection.py:150-162) run each flow map as its own process.  Here the
// members of an ensemble -- the same lattice of cilia above the wall, the same beat and
// driving tables, each member with its own phases, shape angles, natural frequencies and
// time step -- step together: one geometry launch for every cilium of every member, the
// wall-RPY product of every member by the symmetric kernel with one grid slice per
// member, and one GMRES in lockstep with per-member Krylov bases, Hessenbergs, Givens
// rotations and convergence.  One Arnoldi step of the ensemble is a fixed kernel
// sequence whose step index lives in device memory, captured once into a hipGraph and
// replayed.
//
// Member e's system vector is the slice [e Lm, (e+1) Lm) of the ensemble vector with
// the single-system layout: forces [3][ns] (component-major), then n_cilia phase
// speeds and, with shape rotation, n_cilia angle speeds.  Geometry arrays (positions,
// vel_dir_phase / vel_dir_angle) are SoA over all members' segments.

namespace cfm {

Status sym_members_apply(hipStream_t stream, const double* x, const double* y, const double* z, int ns,
                         int members, const double* f, size_t mvec, double* v, double mu,
                         const int* status, int sstride, int want, DevBuf<double>& own,
                         DevBuf<double>& src);

namespace {

// member status (slot 0 of iinfo)
constexpr int CE_DONE = 0, CE_RUN = 1, CE_CHECK = 2;
constexpr int CE_ST = 0, CE_K = 1, CE_ITERS = 2, CE_UPD = 3, CE_DONE_AT = 4, CE_REJ = 5;
constexpr int CE_TARGET = 0, CE_EST = 1, CE_INV = 2, CE_NR = 3, CE_NB = 4;

struct CEK {
  int E, ncil, nseg, ns, ne, Lm, dyn_angle, m;
  size_t L, Nseg;         // E Lm; E ns (stride of the SoA geometry arrays)
  const double* vphase;   // 3 x Nseg
  const double* vangle;   // 3 x Nseg
  const double* z;        // segment heights (Nseg)
  double* Q;              // (m + 1) x L
  double* H;              // E x (m + 1) x (m + 1), column j at H[e][j][0..m]
  double *cs, *sn, *beta, *h, *h2, *y;  // E x (m + 2)
  double* part;           // (m + 1) x E
  double* sinfo;          // E x 8
  int* iinfo;             // E x 8
  int* jdev;              // [0] step j, [1] steps taken, [2] members running
};

__device__ __forceinline__ int ce_status(const CEK& k, int e) { return k.iinfo[e * 8 + CE_ST]; }

// Block preconditioner of one cilium (cilia_precond_kernel for member slices): mode 0
// on running members from Q_j, mode 1 on members whose cycle ended from `in`.
__global__ void ce_precond_kernel(CEK k, int kind, double mu, int mode, const double* __restrict__ in,
                                  double* __restrict__ out) {
  const int gcil = blockIdx.x * blockDim.x + threadIdx.x;
  if (gcil >= k.E * k.ncil) return;
  const int e = gcil / k.ncil, c = gcil - e * k.ncil;
  if (mode == 0 ? ce_status(k, e) != CE_RUN : !k.iinfo[e * 8 + CE_UPD]) return;
  const double* r = (mode == 0 ? k.Q + (size_t)k.jdev[0] * k.L : in) + (size_t)e * k.Lm;
  double* o = out + (size_t)e * k.Lm;
  const size_t ns = k.ns, N = k.Nseg;
  double s11 = 0.0, s12 = 0.0, s22 = 0.0, dp = 0.0, da = 0.0;
  for (int i = 0; i < k.nseg; ++i) {
    const size_t g = (size_t)c * k.nseg + i, gg = (size_t)e * ns + g;
    double rv[3], vp[3], va[3], pr[3], pp[3], pa[3];
    for (int d = 0; d < 3; ++d) {
      rv[d] = r[d * ns + g];
      vp[d] = k.vphase[d * N + gg];
      va[d] = k.vangle[d * N + gg];
    }
    particle_precond(kind, 1, 3, k.z[gg], 1.0, mu, rv, pr);
    particle_precond(kind, 1, 3, k.z[gg], 1.0, mu, vp, pp);
    particle_precond(kind, 1, 3, k.z[gg], 1.0, mu, va, pa);
    for (int d = 0; d < 3; ++d) {
      s11 += vp[d] * pp[d];
      s12 += vp[d] * pa[d];
      s22 += va[d] * pa[d];
      dp += vp[d] * pr[d];
      da += va[d] * pr[d];
      o[d * ns + g] = pr[d];
    }
  }
  const double tp = -r[3 * ns + c] - dp;
  double wp, wa = 0.0;
  if (k.dyn_angle) {
    const double ta = -r[3 * ns + k.ncil + c] - da;
    const double det = s11 * s22 - s12 * s12;
    wp = (s22 * tp - s12 * ta) / det;
    wa = (s11 * ta - s12 * tp) / det;
    o[3 * ns + k.ncil + c] = wa;
  } else {
    wp = tp / s11;
  }
  o[3 * ns + c] = wp;
  for (int i = 0; i < k.nseg; ++i) {
    const size_t g = (size_t)c * k.nseg + i, gg = (size_t)e * ns + g;
    double vp[3], va[3], pp[3], pa[3];
    for (int d = 0; d < 3; ++d) {
      vp[d] = k.vphase[d * N + gg];
      va[d] = k.vangle[d * N + gg];
    }
    particle_precond(kind, 1, 3, k.z[gg], 1.0, mu, vp, pp);
    particle_precond(kind, 1, 3, k.z[gg], 1.0, mu, va, pa);
    for (int d = 0; d < 3; ++d) o[d * ns + g] += wp * pp[d] + (k.dyn_angle ? wa * pa[d] : 0.0);
  }
}

// Coupling rows of the members with status `want` (after y_F = M x_F):
// y_F -= phi_dot K + psi_dot A, y_phi = -K . F, y_psi = -A . F (cilia_couple_kernel).
__global__ void ce_couple_kernel(CEK k, int want, const double* __restrict__ x, double* __restrict__ y) {
  const int gcil = blockIdx.x * blockDim.x + threadIdx.x;
  if (gcil >= k.E * k.ncil) return;
  const int e = gcil / k.ncil, c = gcil - e * k.ncil;
  if (ce_status(k, e) != want) return;
  const double* xe = x + (size_t)e * k.Lm;
  double* ye = y + (size_t)e * k.Lm;
  const size_t ns = k.ns, N = k.Nseg;
  const double wp = xe[3 * ns + c];
  const double wa = k.dyn_angle ? xe[3 * ns + k.ncil + c] : 0.0;
  double dp = 0.0, da = 0.0;
  for (int i = 0; i < k.nseg; ++i) {
    const size_t g = (size_t)c * k.nseg + i, gg = (size_t)e * ns + g;
    for (int d = 0; d < 3; ++d) {
      const double vp = k.vphase[d * N + gg], va = k.vangle[d * N + gg];
      const double fv = xe[d * ns + g];
      dp += vp * fv;
      da += va * fv;
      ye[d * ns + g] -= wp * vp + (k.dyn_angle ? wa * va : 0.0);
    }
  }
  ye[3 * ns + c] = -dp;
  if (k.dyn_angle) ye[3 * ns + k.ncil + c] = -da;
}

// part[i][e] = <Q_i, w> over member e (i <= j; or <w, w> with norm), one workgroup per
// (member, vector), summed in a fixed order.
__global__ __launch_bounds__(256) void ce_dots_kernel(CEK k, const double* __restrict__ w, int norm, int want) {
  __shared__ double red[4];
  const int e = blockIdx.x, i = blockIdx.y;
  if (ce_status(k, e) != want) return;
  if (norm ? i > 0 : i > k.jdev[0]) return;
  const double* q = (norm ? w : k.Q + (size_t)i * k.L) + (size_t)e * k.Lm;
  const double* we = w + (size_t)e * k.Lm;
  double s = 0.0;
  for (int t = threadIdx.x; t < k.Lm; t += 256) s += q[t] * we[t];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) k.part[(size_t)i * k.E + e] = (red[0] + red[1]) + (red[2] + red[3]);
}

// pass 0: h = part; pass 1: h2 = part, H[:, j] = h + h2.
__global__ void ce_finish_kernel(CEK k, int pass) {
  const int e = blockIdx.x;
  if (ce_status(k, e) != CE_RUN) return;
  const int j = k.jdev[0];
  const int ld = k.m + 2;
  for (int i = threadIdx.x; i <= j; i += blockDim.x) {
    const double v = k.part[(size_t)i * k.E + e];
    if (pass == 0) {
      k.h[(size_t)e * ld + i] = v;
    } else {
      k.h2[(size_t)e * ld + i] = v;
      k.H[((size_t)e * (k.m + 1) + j) * (k.m + 1) + i] = k.h[(size_t)e * ld + i] + v;
    }
  }
}

// w -= sum_{i <= j} h[e][i] Q_i over running members.
__global__ void ce_sub_kernel(CEK k, double* __restrict__ w, int pass) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k.L) return;
  const int e = (int)(t / k.Lm);
  if (ce_status(k, e) != CE_RUN) return;
  const int j = k.jdev[0];
  const double* hh = (pass == 0 ? k.h : k.h2) + (size_t)e * (k.m + 2);
  double acc = 0.0;
  for (int i = 0; i <= j; ++i) acc += hh[i] * k.Q[(size_t)i * k.L + t];
  w[t] -= acc;
}

__device__ void ce_backsub(const CEK& k, int e, int kk) {
  const int ldh = k.m + 1;
  const double* H = k.H + (size_t)e * ldh * ldh;
  const double* beta = k.beta + (size_t)e * (k.m + 2);
  double* y = k.y + (size_t)e * (k.m + 2);
  for (int i = kk - 1; i >= 0; --i) {
    double acc = beta[i];
    for (int c = i + 1; c < kk; ++c) acc -= H[(size_t)c * ldh + i] * y[c];
    y[i] = acc / H[(size_t)i * ldh + i];
  }
}

// Norm of w, Givens rotations of column j, residual estimate and the per-member
// decisions of this step (mobility_solver.cpp:2467-2510 rotation formulas); thread 0
// then advances the step counters.
__global__ __launch_bounds__(256) void ce_givens_kernel(CEK k) {
  __shared__ int run_count;
  if (threadIdx.x == 0) run_count = 0;
  __syncthreads();
  const int j = k.jdev[0];
  const int ldh = k.m + 1, ld = k.m + 2;
  for (int e = threadIdx.x; e < k.E; e += blockDim.x) {
    int* ii = k.iinfo + e * 8;
    ii[CE_UPD] = 0;
    if (ii[CE_ST] != CE_RUN) continue;
    const double nrm = sqrt(k.part[e]);
    double* col = k.H + ((size_t)e * ldh + j) * ldh;
    col[j + 1] = nrm;
    double* cs = k.cs + (size_t)e * ld;
    double* sn = k.sn + (size_t)e * ld;
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
    double* beta = k.beta + (size_t)e * ld;
    beta[j + 1] = -s_ * beta[j];
    beta[j] = c_ * beta[j];
    const double est = fabs(beta[j + 1]);
    double* si = k.sinfo + e * 8;
    si[CE_EST] = est;
    si[CE_INV] = nrm > 0.0 ? 1.0 / nrm : 0.0;
    ii[CE_K] = j + 1;
    ii[CE_ITERS] += 1;
    const bool conv = est <= si[CE_TARGET];
    if (conv) ii[CE_DONE_AT] = k.jdev[1];
    if (conv || j + 1 >= k.m || !(nrm > 0.0) || !isfinite(est)) {
      ce_backsub(k, e, j + 1);
      ii[CE_UPD] = 1;
      ii[CE_ST] = CE_CHECK;
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

// Q_{j+1} = w / ||w|| for members still running (j already advanced).
__global__ void ce_scale_kernel(CEK k, const double* __restrict__ w) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k.L) return;
  const int e = (int)(t / k.Lm);
  if (ce_status(k, e) != CE_RUN) return;
  k.Q[(size_t)k.jdev[0] * k.L + t] = w[t] * k.sinfo[e * 8 + CE_INV];
}

// t = Q y over the members whose cycle ended at this step.
__global__ void ce_qy_kernel(CEK k, double* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k.L) return;
  const int e = (int)(t / k.Lm);
  const int* ii = k.iinfo + e * 8;
  if (!ii[CE_UPD]) return;
  const double* y = k.y + (size_t)e * (k.m + 2);
  double acc = 0.0;
  for (int i = 0; i < ii[CE_K]; ++i) acc += y[i] * k.Q[(size_t)i * k.L + t];
  out[t] = acc;
}

// x += d over the members whose cycle ended at this step.
__global__ void ce_add_kernel(CEK k, double* __restrict__ x, const double* __restrict__ d) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k.L) return;
  const int e = (int)(t / k.Lm);
  if (!k.iinfo[e * 8 + CE_UPD]) return;
  x[t] += d[t];
}

// r = b - w over members with status `want`.
__global__ void ce_resid_kernel(CEK k, int want, const double* __restrict__ b, const double* __restrict__ w,
                                double* __restrict__ r) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k.L) return;
  const int e = (int)(t / k.Lm);
  if (ce_status(k, e) != want) return;
  r[t] = b[t] - w[t];
}

// ||b|| per member (part[0][e] holds <b, b>); every member enters CE_CHECK with target
// rtol ||b||; a zero right-hand side gives the zero solution at once
// (mobility_solver.cpp:2221-2294).
__global__ void ce_setup_kernel(CEK k, double rtol) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= k.E) return;
  double* si = k.sinfo + e * 8;
  int* ii = k.iinfo + e * 8;
  si[CE_NB] = sqrt(k.part[e]);
  si[CE_TARGET] = rtol * si[CE_NB];
  si[CE_NR] = 0.0;
  ii[CE_ST] = si[CE_NB] > 0.0 ? CE_CHECK : CE_DONE;
  ii[CE_ITERS] = 0;
  ii[CE_UPD] = 0;
  ii[CE_K] = 0;
  ii[CE_DONE_AT] = -1;
  ii[CE_REJ] = si[CE_NB] > 0.0 ? 0 : 1;
}

// The guess is rejected when its residual exceeds ||b|| (mobility_solver.cpp:2296-2388):
// part[0][e] holds ||b - A x0||^2.
__global__ void ce_guess_kernel(CEK k) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= k.E) return;
  if (ce_status(k, e) != CE_CHECK) return;
  const double* si = k.sinfo + e * 8;
  k.iinfo[e * 8 + CE_REJ] = sqrt(k.part[e]) > si[CE_NB] ? 1 : 0;
}

// Rejected guesses (and zero right-hand sides): x = 0, r = b.
__global__ void ce_reject_kernel(CEK k, const double* __restrict__ b, double* __restrict__ x,
                                 double* __restrict__ r) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k.L) return;
  const int e = (int)(t / k.Lm);
  if (!k.iinfo[e * 8 + CE_REJ]) return;
  x[t] = 0.0;
  r[t] = b[t];
}

// After the residual norms of CE_CHECK members (part[0][e]): converged -> CE_DONE, else
// restart with beta = (||r||, 0, ...).
__global__ void ce_restart_kernel(CEK k, int max_iter) {
  __shared__ int run_count;
  if (threadIdx.x == 0) run_count = 0;
  __syncthreads();
  for (int e = threadIdx.x; e < k.E; e += blockDim.x) {
    int* ii = k.iinfo + e * 8;
    if (ii[CE_ST] != CE_CHECK) continue;
    const double nr = sqrt(k.part[e]);
    double* si = k.sinfo + e * 8;
    si[CE_NR] = nr;
    if (nr <= si[CE_TARGET] || ii[CE_ITERS] >= max_iter || !isfinite(nr)) {
      ii[CE_ST] = CE_DONE;
      continue;
    }
    si[CE_INV] = 1.0 / nr;
    double* beta = k.beta + (size_t)e * (k.m + 2);
    beta[0] = nr;
    for (int i = 1; i < k.m + 2; ++i) beta[i] = 0.0;
    ii[CE_K] = 0;
    ii[CE_ST] = CE_RUN;
    atomicAdd(&run_count, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    k.jdev[0] = 0;
    k.jdev[2] = run_count;
  }
}

// Q_0 = r / ||r|| for members that restarted.
__global__ void ce_q0_kernel(CEK k, const double* __restrict__ r) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k.L) return;
  const int e = (int)(t / k.Lm);
  if (ce_status(k, e) != CE_RUN) return;
  k.Q[t] = r[t] * k.sinfo[e * 8 + CE_INV];
}

// rhs of every member: [0 ; -q_phase ; -q_angle] (cilia_drive_kernel per member slice;
// the noise deviate is keyed on the cilium's index within its member).
__global__ void ce_drive_kernel(CEK k, double len, double spring, const double* __restrict__ phase,
                                const double* __restrict__ angle, const double* __restrict__ omega0,
                                const double* __restrict__ tab_p, int np, const double* __restrict__ tab_a,
                                int na, DriveOpts dopt, double* __restrict__ rhs) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k.L) return;
  const int e = (int)(t / k.Lm);
  const int loc = (int)(t - (size_t)e * k.Lm);
  const int off = 3 * k.ns;
  if (loc < off) {
    rhs[t] = 0.0;
    return;
  }
  const int c = (loc - off) % k.ncil;
  const int gc_ = e * k.ncil + c;
  const double sc = 0.5 * omega0[gc_] / 3.14159265358979323846;
  if (loc < off + k.ncil) {
    double q = cyclic_hermite(tab_p, np, phase[gc_]) * sc;
    if (dopt.noise != 0.0) q += dopt.noise * normal_deviate(dopt.seed, dopt.nt, c);
    rhs[t] = -q;
  } else {
    const double a = angle[gc_], amax = 0.78;
    double fac = a;
    if (dopt.fene && fabs(a) > amax) fac = 0.5 * a * (1.0 + a * a / (amax * amax) * exp(fabs(a) - amax));
    const double q = cyclic_hermite(tab_a, na, phase[gc_]) * sc - spring * omega0[gc_] * len * len * len * fac;
    rhs[t] = -q;
  }
}

// Speeds and generalised forces of every cilium from the members' solutions.
__global__ void ce_unpack_kernel(CEK k, const double* __restrict__ sol, double* __restrict__ pdot,
                                 double* __restrict__ adot, double* __restrict__ q) {
  const int gcil = blockIdx.x * blockDim.x + threadIdx.x;
  if (gcil >= k.E * k.ncil) return;
  const int e = gcil / k.ncil, c = gcil - e * k.ncil;
  const double* xe = sol + (size_t)e * k.Lm;
  const size_t ns = k.ns, N = k.Nseg;
  pdot[gcil] = xe[3 * ns + c];
  adot[gcil] = k.dyn_angle ? xe[3 * ns + k.ncil + c] : 0.0;
  double qp = 0.0, qa = 0.0;
  for (int i = 0; i < k.nseg; ++i) {
    const size_t g = (size_t)c * k.nseg + i, gg = (size_t)e * ns + g;
    for (int d = 0; d < 3; ++d) {
      const double fv = xe[d * ns + g];
      qp += k.vphase[d * N + gg] * fv;
      qa += k.vangle[d * N + gg] * fv;
    }
  }
  q[gcil] = qp;
  q[(size_t)k.E * k.ncil + gcil] = qa;
}

// phase += dt_e phase_dot (and the shape angle) before a step, per member time step.
__global__ void ce_advance_kernel(int n, int ncil, int dyn_angle, const double* __restrict__ dt,
                                  const double* __restrict__ pdot, const double* __restrict__ adot,
                                  double* __restrict__ phase, double* __restrict__ angle) {
  const int gcil = blockIdx.x * blockDim.x + threadIdx.x;
  if (gcil >= n) return;
  const double h = dt[gcil / ncil];
  phase[gcil] += pdot[gcil] * h;
  if (dyn_angle) angle[gcil] += adot[gcil] * h;
}

// the initial guess of every member: forces zero, phase speeds = natural frequencies,
// shape-angle speeds zero
__global__ void ce_init_guess_kernel(CEK k, const double* __restrict__ omega0, double* __restrict__ x) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k.L) return;
  const int e = (int)(t / k.Lm);
  const int loc = (int)(t - (size_t)e * k.Lm), off = 3 * k.ns;
  x[t] = (loc >= off && loc < off + k.ncil) ? omega0[e * k.ncil + (loc - off)] : 0.0;
}

__global__ void bfill_cilia_kernel(double* __restrict__ x, double v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ void ce_zero_int_kernel(int* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void ce_status_fill_kernel(int* __restrict__ iinfo, int E, int st) {
  for (int e = threadIdx.x; e < E; e += blockDim.x) iinfo[e * 8 + CE_ST] = st;
}

inline dim3 g256(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace
}  // namespace cfm

using namespace cfm;

struct cfm_cilia_batch {
  cfm_cilia_batch_config cfg{};
  hipStream_t stream = nullptr;
  int E = 0, nc = 0, nseg = 0, ns = 0, ne = 0, Lm = 0, dyn_angle = 0, m = 0;
  size_t L = 0, N = 0;
  CiliaK geo{};
  DevBuf<double> base, phase, angle, pdot, adot, omega0, dt, x, y, z, vphase, vangle, svals, q;
  DevBuf<double> tab_phase, tab_angle;
  int n_tab_phase = 0, n_tab_angle = 0;
  DevBuf<double> sol, rhs, r, w, zt, t;
  DevBuf<double> Q, H, cs, sn, beta, h, h2, yv, part, sinfo;
  DevBuf<int> iinfo, jdev;
  DevBuf<double> own, src;  // partials of the symmetric product
  int* host_flags = nullptr;  // pinned: [0, 4) counters after a residual pass, [4, 8) ring
  double* host = nullptr;     // pinned read-back of q
  hipGraphExec_t graph = nullptr;
  std::vector<hipEvent_t> ev;
  bool have_state = false;
  std::string err;
};

namespace {

int cefail(cfm_cilia_batch* b, const Status& s) {
  if (b) b->err = s.msg;
  return s.code;
}

CEK cek_of(cfm_cilia_batch* b) {
  CEK k{};
  k.E = b->E;
  k.ncil = b->nc;
  k.nseg = b->nseg;
  k.ns = b->ns;
  k.ne = b->ne;
  k.Lm = b->Lm;
  k.dyn_angle = b->dyn_angle;
  k.m = b->m;
  k.L = b->L;
  k.Nseg = b->N;
  k.vphase = b->vphase.p;
  k.vangle = b->vangle.p;
  k.z = b->z.p;
  k.Q = b->Q.p;
  k.H = b->H.p;
  k.cs = b->cs.p;
  k.sn = b->sn.p;
  k.beta = b->beta.p;
  k.h = b->h.p;
  k.h2 = b->h2.p;
  k.y = b->yv.p;
  k.part = b->part.p;
  k.sinfo = b->sinfo.p;
  k.iinfo = b->iinfo.p;
  k.jdev = b->jdev.p;
  return k;
}

// out_F = M x_F (members with status want, FP64 symmetric product), then the coupling
// rows of those members.
Status ce_product(cfm_cilia_batch* b, int want, const double* x, double* out) {
  const CEK k = cek_of(b);
  CFM_TRY(sym_members_apply(b->stream, b->x.p, b->y.p, b->z.p, b->ns, b->E, x, (size_t)b->Lm, out, b->cfg.mu,
                            b->iinfo.p, 8, want, b->own, b->src));
  hipLaunchKernelGGL(ce_couple_kernel, gc(b->E * b->nc), dim3(64), 0, b->stream, k, want, x, out);
  CFM_CHECK_LAUNCH();
  return {};
}

// One Arnoldi step of every running member (captured into the graph).
Status ce_enqueue_step(cfm_cilia_batch* b) {
  const CEK k = cek_of(b);
  const int pc = b->cfg.prm.solve.precond;
  const double mu = b->cfg.mu;
  const int ncl = b->E * b->nc;
  hipLaunchKernelGGL(ce_precond_kernel, gc(ncl), dim3(64), 0, b->stream, k, pc, mu, 0, (const double*)nullptr,
                     b->zt.p);
  CFM_CHECK_LAUNCH();
  CFM_TRY(ce_product(b, CE_RUN, b->zt.p, b->w.p));
  for (int pass = 0; pass < 2; ++pass) {
    hipLaunchKernelGGL(ce_dots_kernel, dim3(b->E, b->m + 1), dim3(256), 0, b->stream, k, b->w.p, 0, CE_RUN);
    hipLaunchKernelGGL(ce_finish_kernel, dim3(b->E), dim3(64), 0, b->stream, k, pass);
    hipLaunchKernelGGL(ce_sub_kernel, g256(b->L), dim3(256), 0, b->stream, k, b->w.p, pass);
  }
  hipLaunchKernelGGL(ce_dots_kernel, dim3(b->E, 1), dim3(256), 0, b->stream, k, b->w.p, 1, CE_RUN);
  hipLaunchKernelGGL(ce_givens_kernel, dim3(1), dim3(256), 0, b->stream, k);
  hipLaunchKernelGGL(ce_scale_kernel, g256(b->L), dim3(256), 0, b->stream, k, b->w.p);
  // x += P^-1 (Q y) for members whose cycle ended
  hipLaunchKernelGGL(ce_qy_kernel, g256(b->L), dim3(256), 0, b->stream, k, b->t.p);
  hipLaunchKernelGGL(ce_precond_kernel, gc(ncl), dim3(64), 0, b->stream, k, pc, mu, 1, (const double*)b->t.p,
                     b->zt.p);
  hipLaunchKernelGGL(ce_add_kernel, g256(b->L), dim3(256), 0, b->stream, k, b->sol.p, (const double*)b->zt.p);
  CFM_CHECK_LAUNCH();
  return {};
}

Status ce_event(cfm_cilia_batch* b, size_t i, hipEvent_t* e) {
  while (b->ev.size() <= i) {
    hipEvent_t t;
    CFM_HIP(hipEventCreate(&t));
    b->ev.push_back(t);
  }
  *e = b->ev[i];
  return {};
}

Status ce_graph(cfm_cilia_batch* b, hipGraphExec_t* out) {
  if (!b->graph) {
    hipGraph_t g = nullptr;
    CFM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    Status s = ce_enqueue_step(b);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (!s.ok()) {
      if (g) (void)hipGraphDestroy(g);
      return s;
    }
    if (e != hipSuccess) return make_status(CFM_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return make_status(CFM_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  *out = b->graph;
  return {};
}

// Graph replays until no member is running or the cycle is full; the running count of
// step j is copied to ring slot 4 + j % 4 after the replay and read one step late.
Status ce_cycle(cfm_cilia_batch* b, hipGraphExec_t g, size_t* nev) {
  int pending = -1, pending_slot = 0;
  for (int j = 0; j < b->m; ++j) {
    CFM_HIP(hipGraphLaunch(g, b->stream));
    const int slot = 4 + (j % 4);
    CFM_HIP(hipMemcpyAsync(b->host_flags + slot, b->jdev.p + 2, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    hipEvent_t e;
    CFM_TRY(ce_event(b, *nev, &e));
    CFM_HIP(hipEventRecord(e, b->stream));
    const size_t mine = (*nev)++;
    if (pending >= 0) {
      hipEvent_t pe;
      CFM_TRY(ce_event(b, (size_t)pending, &pe));
      CFM_HIP(hipEventSynchronize(pe));
      if (b->host_flags[pending_slot] == 0) break;
    }
    pending = (int)mine;
    pending_slot = slot;
  }
  CFM_HIP(hipStreamSynchronize(b->stream));
  return {};
}

// Every member's solve of M x = rhs from the guess in b->sol (the single-system solve's
// rule: restarted GMRES(m) with the true residual at each cycle's end, the guess
// rejected when worse than zero).
Status ce_solve(cfm_cilia_batch* b, std::vector<double>* si, std::vector<int>* ii) {
  const CEK k = cek_of(b);
  const cfm_solve_opts& o = b->cfg.prm.solve;
  const int max_iter = o.max_iter > 0 ? o.max_iter : 1000;
  hipGraphExec_t g = nullptr;  // captured before the first cycle, after the first product
                               // has sized the product's partial buffers
  hipLaunchKernelGGL(ce_zero_int_kernel, dim3(1), dim3(64), 0, b->stream, b->jdev.p, 4);
  CFM_CHECK_LAUNCH();
  // ||b||: every member in CE_DONE for the norm pass, then targets
  hipLaunchKernelGGL(ce_status_fill_kernel, dim3(1), dim3(256), 0, b->stream, b->iinfo.p, b->E, CE_DONE);
  hipLaunchKernelGGL(ce_dots_kernel, dim3(b->E, 1), dim3(256), 0, b->stream, k, b->rhs.p, 1, CE_DONE);
  hipLaunchKernelGGL(ce_setup_kernel, gc(b->E), dim3(64), 0, b->stream, k, o.rtol);
  CFM_CHECK_LAUNCH();
  size_t nev = 0;
  for (int pass = 0; pass < 4 * (max_iter / std::max(1, b->m) + 2); ++pass) {
    // FP64 residual of the members to check: r = b - A x
    CFM_TRY(ce_product(b, CE_CHECK, b->sol.p, b->w.p));
    hipLaunchKernelGGL(ce_resid_kernel, g256(b->L), dim3(256), 0, b->stream, k, CE_CHECK, b->rhs.p, b->w.p, b->r.p);
    hipLaunchKernelGGL(ce_dots_kernel, dim3(b->E, 1), dim3(256), 0, b->stream, k, b->r.p, 1, CE_CHECK);
    if (pass == 0) {  // reject guesses worse than zero
      hipLaunchKernelGGL(ce_guess_kernel, gc(b->E), dim3(64), 0, b->stream, k);
      hipLaunchKernelGGL(ce_reject_kernel, g256(b->L), dim3(256), 0, b->stream, k, b->rhs.p, b->sol.p, b->r.p);
      hipLaunchKernelGGL(ce_dots_kernel, dim3(b->E, 1), dim3(256), 0, b->stream, k, b->r.p, 1, CE_CHECK);
    }
    hipLaunchKernelGGL(ce_restart_kernel, dim3(1), dim3(256), 0, b->stream, k, max_iter);
    hipLaunchKernelGGL(ce_q0_kernel, g256(b->L), dim3(256), 0, b->stream, k, b->r.p);
    CFM_CHECK_LAUNCH();
    CFM_HIP(hipMemcpyAsync(b->host_flags, b->jdev.p, sizeof(int) * 4, hipMemcpyDeviceToHost, b->stream));
    CFM_HIP(hipStreamSynchronize(b->stream));
    if (b->host_flags[2] == 0) break;  // nobody left running
    if (!g) CFM_TRY(ce_graph(b, &g));
    CFM_TRY(ce_cycle(b, g, &nev));
  }
  si->resize((size_t)8 * b->E);
  ii->resize((size_t)8 * b->E);
  CFM_HIP(copy_to_host(b->stream, si->data(), b->sinfo.p, sizeof(double) * si->size()));
  CFM_HIP(copy_to_host(b->stream, ii->data(), b->iinfo.p, sizeof(int) * ii->size()));
  return {};
}

}  // namespace

extern "C" {

static thread_local std::string g_ce_create_err;  // message of the last failed create

const char* cfm_cilia_batch_last_error(const cfm_cilia_batch* b) {
  return b ? b->err.c_str() : g_ce_create_err.c_str();
}

int cfm_cilia_batch_destroy(cfm_cilia_batch* b) {
  if (!b) return CFM_EINVAL;
  (void)hipSetDevice(b->cfg.device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
  for (auto* v : {&b->base, &b->phase, &b->angle, &b->pdot, &b->adot, &b->omega0, &b->dt, &b->x, &b->y, &b->z,
                  &b->vphase, &b->vangle, &b->svals, &b->q, &b->tab_phase, &b->tab_angle, &b->sol, &b->rhs, &b->r,
                  &b->w, &b->zt, &b->t, &b->Q, &b->H, &b->cs, &b->sn, &b->beta, &b->h, &b->h2, &b->yv, &b->part,
                  &b->sinfo, &b->own, &b->src})
    v->release();
  b->iinfo.release();
  b->jdev.release();
  if (b->host_flags) (void)hipHostFree(b->host_flags);
  if (b->host) (void)hipHostFree(b->host);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return CFM_OK;
}

int cfm_cilia_batch_create(const cfm_cilia_batch_config* cfg, const double* base, const double* q_phase_table,
                           int32_t n_phase, const double* q_angle_table, int32_t n_angle, cfm_cilia_batch** out) {
  if (!out) return CFM_EINVAL;
  *out = nullptr;
  auto fail = [&](int code, const char* msg) {
    g_ce_create_err = msg;
    return code;
  };
  if (!cfg || !base || !q_phase_table) return fail(CFM_EINVAL, "null argument");
  const cfm_cilia_params& prm = cfg->prm;
  if (cfg->n_members <= 0 || prm.n_cilia <= 0) return fail(CFM_EINVAL, "n_members and n_cilia must be positive");
  if (prm.nseg < 2 || prm.nseg > kMaxCiliaSeg) return fail(CFM_EINVAL, "nseg must be in 2..64");
  if (prm.beat < CFM_BEAT_FULFORD_BLAKE || prm.beat > CFM_BEAT_BICILIA_LONGT) return fail(CFM_EINVAL, "unknown beat");
  if (cilium_segments(prm) != prm.nseg && (prm.nseg % 2 != 0 || prm.nseg < 4))
    return fail(CFM_EINVAL, "paired beats take an even nseg >= 4 (two cilia of nseg / 2)");
  if (!(prm.dl > 0.0) || !(cfg->mu > 0.0)) return fail(CFM_EINVAL, "dl and mu must be positive");
  if (prm.solve.precond != CFM_PC_DIAG && prm.solve.precond != CFM_PC_SELF)
    return fail(CFM_EINVAL, "the coupled cilia system takes CFM_PC_DIAG or CFM_PC_SELF");
  if (prm.solve.mode != CFM_SOLVE_FP64) return fail(CFM_EINVAL, "ensembles solve in FP64 (CFM_SOLVE_FP64)");
  if (!(prm.solve.rtol > 0.0)) return fail(CFM_EINVAL, "rtol must be positive");
  if (prm.solve.restart < 0 || prm.solve.max_iter < 0) return fail(CFM_EINVAL, "restart and max_iter must be >= 0");
  if (n_phase < 2 || (q_angle_table && n_angle < 2)) return fail(CFM_EINVAL, "driving tables need at least 2 values");
  if (cfg->spring_model != 0 && cfg->spring_model != 1) return fail(CFM_EINVAL, "spring_model is 0 or 1");
  if (!std::isfinite(cfg->spring) || !std::isfinite(cfg->noise)) return fail(CFM_EINVAL, "non-finite spring or noise");
  for (int i = 0; i < n_phase; ++i)
    if (!std::isfinite(q_phase_table[i])) return fail(CFM_EINVAL, "non-finite table");
  if (q_angle_table)
    for (int i = 0; i < n_angle; ++i)
      if (!std::isfinite(q_angle_table[i])) return fail(CFM_EINVAL, "non-finite table");
  const int nc = prm.n_cilia;
  for (int i = 0; i < nc; ++i) {
    for (int d = 0; d < 3; ++d)
      if (!std::isfinite(base[3 * i + d])) return fail(CFM_EINVAL, "non-finite base");
    if (!(base[3 * i + 2] > 0.0)) return fail(CFM_EINVAL, "a cilium base is not above the wall");
  }
  if (hipSetDevice(cfg->device) != hipSuccess) return fail(CFM_EHIP, "hipSetDevice failed");
  cfm_cilia_batch* b = new (std::nothrow) cfm_cilia_batch();
  if (!b) return CFM_ENOMEM;
  b->cfg = *cfg;
  b->E = cfg->n_members;
  b->nc = nc;
  b->nseg = prm.nseg;
  b->ns = nc * prm.nseg;
  b->dyn_angle = q_angle_table ? 1 : 0;
  b->ne = nc * (b->dyn_angle ? 2 : 1);
  b->Lm = 3 * b->ns + b->ne;
  b->L = (size_t)b->E * b->Lm;
  b->N = (size_t)b->E * b->ns;
  b->m = prm.solve.restart > 0 ? prm.solve.restart : 50;
  auto run = [&]() -> Status {
    CFM_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    const size_t ncl = (size_t)b->E * nc;
    // bases replicated for every member, SoA over all cilia
    std::vector<double> hb(3 * ncl);
    for (int e = 0; e < b->E; ++e)
      for (int c = 0; c < nc; ++c)
        for (int d = 0; d < 3; ++d) hb[d * ncl + (size_t)e * nc + c] = base[3 * c + d];
    CFM_TRY(b->base.alloc(hb.size()));
    CFM_HIP(copy_to_device(b->stream, b->base.p, hb.data(), sizeof(double) * hb.size()));
    for (auto* v : {&b->phase, &b->angle, &b->pdot, &b->adot, &b->omega0}) CFM_TRY(v->alloc(ncl));
    CFM_TRY(b->q.alloc(2 * ncl));
    CFM_TRY(b->dt.alloc(b->E));
    for (auto* v : {&b->x, &b->y, &b->z}) CFM_TRY(v->alloc(b->N));
    CFM_TRY(b->vphase.alloc(3 * b->N));
    CFM_TRY(b->vangle.alloc(3 * b->N));
    CFM_TRY(b->svals.alloc(ncl * prm.nseg));
    CFM_TRY(b->tab_phase.alloc(n_phase));
    CFM_HIP(copy_to_device(b->stream, b->tab_phase.p, q_phase_table, sizeof(double) * n_phase));
    b->n_tab_phase = n_phase;
    if (q_angle_table) {
      CFM_TRY(b->tab_angle.alloc(n_angle));
      CFM_HIP(copy_to_device(b->stream, b->tab_angle.p, q_angle_table, sizeof(double) * n_angle));
      b->n_tab_angle = n_angle;
    }
    for (auto* v : {&b->sol, &b->rhs, &b->r, &b->w, &b->zt, &b->t}) CFM_TRY(v->alloc(b->L));
    const int m = b->m, E = b->E;
    CFM_TRY(b->Q.alloc((size_t)(m + 1) * b->L));
    CFM_TRY(b->H.alloc((size_t)E * (m + 1) * (m + 1)));
    for (auto* v : {&b->cs, &b->sn, &b->beta, &b->h, &b->h2, &b->yv}) CFM_TRY(v->alloc((size_t)E * (m + 2)));
    CFM_TRY(b->part.alloc((size_t)(m + 1) * E));
    CFM_TRY(b->sinfo.alloc((size_t)8 * E));
    CFM_TRY(b->iinfo.alloc((size_t)8 * E));
    CFM_TRY(b->jdev.alloc(4));
    CFM_HIP(hipHostMalloc((void**)&b->host_flags, sizeof(int) * 8, hipHostMallocDefault));
    CFM_HIP(hipHostMalloc((void**)&b->host, sizeof(double) * 2 * ncl, hipHostMallocDefault));
    // geometry of every cilium of every member (the single system's kernel over the
    // concatenated cilia)
    CiliaK& k = b->geo;
    beat_args(prm, 0.0, 1.0, k);
    k.ncil = (int)ncl;
    k.nseg = prm.nseg;
    k.nseg_total = (int)b->N;
    k.dl = prm.dl;
    k.base = b->base.p;
    k.phase = b->phase.p;
    k.angle = b->angle.p;
    k.x = b->x.p;
    k.y = b->y.p;
    k.z = b->z.p;
    k.vphase = b->vphase.p;
    k.vangle = b->vangle.p;
    k.svals = b->svals.p;
    CFM_HIP(hipStreamSynchronize(b->stream));
    return {};
  };
  Status s = run();
  if (!s.ok()) {
    g_ce_create_err = s.msg;
    const int code = s.code;
    cfm_cilia_batch_destroy(b);
    return code;
  }
  *out = b;
  return CFM_OK;
}

int cfm_cilia_batch_set_state(cfm_cilia_batch* b, const double* phase, const double* omega0, const double* angle,
                              const double* dt) {
  if (!b) return CFM_EINVAL;
  if (!phase || !omega0 || !dt) return cefail(b, make_status(CFM_EINVAL, "null argument"));
  const size_t ncl = (size_t)b->E * b->nc;
  for (size_t i = 0; i < ncl; ++i)
    if (!std::isfinite(phase[i]) || !std::isfinite(omega0[i]) || (angle && !std::isfinite(angle[i])))
      return cefail(b, make_status(CFM_EINVAL, "non-finite phase, natural frequency or angle"));
  for (int e = 0; e < b->E; ++e)
    if (!std::isfinite(dt[e])) return cefail(b, make_status(CFM_EINVAL, "non-finite time step"));
  (void)hipSetDevice(b->cfg.device);
  auto run = [&]() -> Status {
    CFM_HIP(copy_to_device(b->stream, b->phase.p, phase, sizeof(double) * ncl));
    CFM_HIP(copy_to_device(b->stream, b->omega0.p, omega0, sizeof(double) * ncl));
    CFM_HIP(copy_to_device(b->stream, b->pdot.p, omega0, sizeof(double) * ncl));
    if (angle) {
      CFM_HIP(copy_to_device(b->stream, b->angle.p, angle, sizeof(double) * ncl));
    } else {
      hipLaunchKernelGGL(bfill_cilia_kernel, g256(ncl), dim3(256), 0, b->stream, b->angle.p, 0.0, (int)ncl);
    }
    hipLaunchKernelGGL(bfill_cilia_kernel, g256(ncl), dim3(256), 0, b->stream, b->adot.p, 0.0, (int)ncl);
    CFM_HIP(copy_to_device(b->stream, b->dt.p, dt, sizeof(double) * b->E));
    const CEK k = cek_of(b);
    hipLaunchKernelGGL(ce_init_guess_kernel, g256(b->L), dim3(256), 0, b->stream, k, (const double*)b->omega0.p,
                       b->sol.p);
    {
      const GeoLaunch gl = geometry_launch(b->geo);
      hipLaunchKernelGGL(cilia_geometry_kernel, gl.grid, gl.block, gl.lds, b->stream, b->geo);
    }
    CFM_CHECK_LAUNCH();
    CFM_HIP(hipStreamSynchronize(b->stream));
    return {};
  };
  Status s = run();
  if (!s.ok()) return cefail(b, s);
  b->have_state = true;
  return CFM_OK;
}

int cfm_cilia_batch_set_pair_dp(cfm_cilia_batch* b, double pair_dp) {
  if (!b) return CFM_EINVAL;
  if (!std::isfinite(pair_dp)) return cefail(b, make_status(CFM_EINVAL, "non-finite pair_dp"));
  (void)hipSetDevice(b->cfg.device);
  b->geo.pair_dp = pair_dp;
  if (!b->have_state) return CFM_OK;  // placed by the next cfm_cilia_batch_set_state
  auto run = [&]() -> Status {
    const GeoLaunch gl = geometry_launch(b->geo);
    hipLaunchKernelGGL(cilia_geometry_kernel, gl.grid, gl.block, gl.lds, b->stream, b->geo);
    CFM_CHECK_LAUNCH();
    CFM_HIP(hipStreamSynchronize(b->stream));
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : cefail(b, s);
}

int cfm_cilia_batch_step(cfm_cilia_batch* b, int32_t nt, double* q_phase, double* q_angle,
                         cfm_batch_sys_stats* per_member) {
  if (!b) return CFM_EINVAL;
  if (!b->have_state) return cefail(b, make_status(CFM_ESTATE, "cfm_cilia_batch_set_state has not been called"));
  (void)hipSetDevice(b->cfg.device);
  const size_t ncl = (size_t)b->E * b->nc;
  auto run = [&]() -> Status {
    const CEK k = cek_of(b);
    if (nt > 0)  // phase += dt phase_dot before the step (filament.cpp:1340-1356)
      hipLaunchKernelGGL(ce_advance_kernel, g256(ncl), dim3(256), 0, b->stream, (int)ncl, b->nc, b->dyn_angle,
                         (const double*)b->dt.p, (const double*)b->pdot.p, (const double*)b->adot.p, b->phase.p,
                         b->angle.p);
    {
      const GeoLaunch gl = geometry_launch(b->geo);
      hipLaunchKernelGGL(cilia_geometry_kernel, gl.grid, gl.block, gl.lds, b->stream, b->geo);
    }
    DriveOpts d{b->cfg.spring_model, b->cfg.noise, b->cfg.seed, nt};
    hipLaunchKernelGGL(ce_drive_kernel, g256(b->L), dim3(256), 0, b->stream, k, cilium_length(b->cfg.prm),
                       b->cfg.spring, (const double*)b->phase.p, (const double*)b->angle.p,
                       (const double*)b->omega0.p, (const double*)b->tab_phase.p, b->n_tab_phase,
                       (const double*)b->tab_angle.p, b->n_tab_angle, d, b->rhs.p);
    CFM_CHECK_LAUNCH();
    std::vector<double> si;
    std::vector<int> ii;
    CFM_TRY(ce_solve(b, &si, &ii));
    hipLaunchKernelGGL(ce_unpack_kernel, gc((int)ncl), dim3(64), 0, b->stream, k, (const double*)b->sol.p,
                       b->pdot.p, b->adot.p, b->q.p);
    CFM_CHECK_LAUNCH();
    CFM_HIP(hipMemcpyAsync(b->host, b->q.p, sizeof(double) * 2 * ncl, hipMemcpyDeviceToHost, b->stream));
    CFM_HIP(hipStreamSynchronize(b->stream));
    if (q_phase) std::memcpy(q_phase, b->host, sizeof(double) * ncl);
    if (q_angle) std::memcpy(q_angle, b->host + ncl, sizeof(double) * ncl);
    bool all = true;
    for (int e = 0; e < b->E; ++e) {
      const double nb = si[8 * e + CE_NB];
      const double rel = nb > 0.0 ? si[8 * e + CE_NR] / nb : 0.0;
      const bool conv = rel <= b->cfg.prm.solve.rtol;
      all = all && conv;
      if (per_member) {
        cfm_batch_sys_stats& P = per_member[e];
        P.iterations = ii[8 * e + CE_ITERS];
        P.converged = conv ? 1 : 0;
        P.rel_residual = rel;
        P.latency_ms = 0.0;
        P.reserved0 = ii[8 * e + CE_REJ];  // 1: the guess was rejected
        P.reserved1 = 0;
      }
    }
    if (!all) return make_status(CFM_ENOCONV, "ensemble GMRES: a member stopped above rtol %.3e",
                                 b->cfg.prm.solve.rtol);
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : cefail(b, s);
}

int cfm_cilia_batch_get_state(cfm_cilia_batch* b, double* phase, double* phase_dot, double* angle,
                              double* angle_dot) {
  if (!b) return CFM_EINVAL;
  (void)hipSetDevice(b->cfg.device);
  const size_t n = sizeof(double) * (size_t)b->E * b->nc;
  auto run = [&]() -> Status {
    CFM_HIP(hipStreamSynchronize(b->stream));
    if (phase) CFM_HIP(copy_to_host(b->stream, phase, b->phase.p, n));
    if (phase_dot) CFM_HIP(copy_to_host(b->stream, phase_dot, b->pdot.p, n));
    if (angle) CFM_HIP(copy_to_host(b->stream, angle, b->angle.p, n));
    if (angle_dot) CFM_HIP(copy_to_host(b->stream, angle_dot, b->adot.p, n));
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : cefail(b, s);
}

int cfm_cilia_batch_get_solution(cfm_cilia_batch* b, double* out) {
  if (!b || !out) return CFM_EINVAL;
  (void)hipSetDevice(b->cfg.device);
  auto run = [&]() -> Status {
    std::vector<double> t(b->L);
    CFM_HIP(hipStreamSynchronize(b->stream));
    CFM_HIP(copy_to_host(b->stream, t.data(), b->sol.p, sizeof(double) * b->L));
    const size_t ns = b->ns;
    for (int e = 0; e < b->E; ++e) {
      const double* src = t.data() + (size_t)e * b->Lm;
      double* dst = out + (size_t)e * b->Lm;
      for (size_t i = 0; i < ns; ++i)
        for (int d = 0; d < 3; ++d) dst[3 * i + d] = src[d * ns + i];
      for (int j = 0; j < b->ne; ++j) dst[3 * ns + j] = src[3 * ns + j];
    }
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : cefail(b, s);
}

int cfm_cilia_batch_get_positions(cfm_cilia_batch* b, double* x) {
  if (!b || !x) return CFM_EINVAL;
  (void)hipSetDevice(b->cfg.device);
  auto run = [&]() -> Status {
    std::vector<double> t(3 * b->N);
    CFM_HIP(hipStreamSynchronize(b->stream));
    CFM_HIP(copy_to_host(b->stream, t.data(), b->x.p, sizeof(double) * b->N));
    CFM_HIP(copy_to_host(b->stream, t.data() + b->N, b->y.p, sizeof(double) * b->N));
    CFM_HIP(copy_to_host(b->stream, t.data() + 2 * b->N, b->z.p, sizeof(double) * b->N));
    for (size_t i = 0; i < b->N; ++i)
      for (int d = 0; d < 3; ++d) x[3 * i + d] = t[d * b->N + i];
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : cefail(b, s);
}

}  // extern "C"
