// fil.hip -- elastic-filament mechanics of the implicit step on the device (FP64).
//
// Restates, for tethered filaments with a clamped base and a follower end force
// (the reference's INSTABILITY_CILIA model):
//   initial guess      filament.cpp:1336-1663 + segment.cpp:38-79 + robot_arm :629-651
//   internal loads     filament.cpp:1729-1853
//   BDF residual       mobility_solver.cpp:2760-2853 (filament rows)
//   approx. Jacobian   filament.cpp:9-60, :1855-2172 (built and inverted per filament)
//   J0^-1 product      swimmer.cpp:919-931
//   state update       filament.cpp:2174-2208, segment.cpp:81-90
//   barrier forces     cuda_functions.cu:1825-1930 (segments; wall repulsion)
//
// Threading: kinematics and loads use one thread per filament (the chain recursions
// are sequential along a filament); the residual uses one thread per segment; the
// 6 NSEG x 6 NSEG approximate Jacobian of a filament is assembled in LDS by one
// workgroup (a thread per segment block row) and inverted by Gauss-Jordan elimination
// held in registers (two lanes per row); its transpose is stored so that the inverse
// products read HBM coalesced.
#include "ctx.hpp"
#include <cfloat>
#include <mutex>
#include <vector>
#include "fil.hpp"
#include "fil_device.hpp"

namespace cfm {
using namespace fil;

__device__ __forceinline__ Q4 ldq(const double* q, size_t n, int s) {
  return {q[s], q[n + s], q[2 * n + s], q[3 * n + s]};
}
__device__ __forceinline__ void stq(double* q, size_t n, int s, Q4 v) {
  q[s] = v.s;
  q[n + s] = v.x;
  q[2 * n + s] = v.y;
  q[3 * n + s] = v.z;
}
__device__ __forceinline__ V3 ld3(const double* a, size_t n, int s) {
  return {a[s], a[n + s], a[2 * n + s]};
}
__device__ __forceinline__ void st3(double* a, size_t n, int s, V3 v) {
  a[s] = v.x;
  a[n + s] = v.y;
  a[2 * n + s] = v.z;
}

// accept_state_from_rigid_body (filament.cpp:653-678): the first segment of filament f
// at its body's base point, turned by the body's Lie-algebra element from its previous
// orientation.
__device__ void attach_to_body(const FilK& K, int f) {
  const size_t n = K.nseg_total;
  const int s0 = f * K.nseg, b = f / K.fpb;
  const Q4 qb = {K.bq[4 * b], K.bq[4 * b + 1], K.bq[4 * b + 2], K.bq[4 * b + 3]};
  const V3 r = mv(rot_mat(qb), v3(K.fref[3 * f], K.fref[3 * f + 1], K.fref[3 * f + 2]));
  K.x[s0] = K.bx[3 * b] + r.x;
  K.y[s0] = K.bx[3 * b + 1] + r.y;
  K.z[s0] = K.bx[3 * b + 2] + r.z;
  const V3 ub = {K.bu[3 * b], K.bu[3 * b + 1], K.bu[3 * b + 2]};
  st3(K.u, n, s0, ub);
  stq(K.q, n, s0, qmul(lie_exp(ub), ldq(K.qm1, n, s0)));
}

// robot arm from the base segment of filament f
__device__ void robot_arm(const FilK& K, int f) {
  const size_t n = K.nseg_total;
  const int s0 = f * K.nseg;
  V3 t1 = tangent(ldq(K.q, n, s0));
  V3 x = {K.x[s0], K.y[s0], K.z[s0]};
  for (int i = 1; i < K.nseg; ++i) {
    const V3 t2 = tangent(ldq(K.q, n, s0 + i));
    x = x + (0.5 * K.dl) * (t1 + t2);
    K.x[s0 + i] = x.x;
    K.y[s0 + i] = x.y;
    K.z[s0 + i] = x.z;
    t1 = t2;
  }
}

__global__ void fil_initial_guess_kernel(FilK K, int nt) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= K.nfil) return;
  const size_t n = K.nseg_total;
  const int s0 = f * K.nseg;
  for (int i = 0; i < K.nseg; ++i) {
    const int s = s0 + i;
    const V3 x = {K.x[s], K.y[s], K.z[s]};
    st3(K.xm2, n, s, ld3(K.xm1, n, s));
    st3(K.xm1, n, s, x);
    V3 u = ld3(K.u, n, s);
    if (nt <= K.num_euler) {
      st3(K.um1, n, s, u);
    } else {
      const V3 um1 = ld3(K.um1, n, s);
      st3(K.um1, n, s, u);
      u = 2.0 * u - um1;
      st3(K.u, n, s, u);
    }
    const Q4 q = ldq(K.q, n, s);
    stq(K.qm1, n, s, q);
    stq(K.q, n, s, qmul(lie_exp(u), q));
  }
  // Lagrange-multiplier extrapolation (row 0 of each filament is the tether)
  auto extrap = [&](double* a, double* a1, double* a2, size_t ld, int idx) {
    for (int c = 0; c < 3; ++c) {
      double& v = a[c * ld + idx];
      double& v1 = a1[c * ld + idx];
      double& v2 = a2[c * ld + idx];
      if (nt == 1) {
        v1 = v;
      } else if (nt == 2) {
        v2 = v1;
        v1 = v;
        v = 2.0 * v1 - v2;
      } else if (nt > 2) {
        const double t = 3.0 * (v - v1) + v2;
        v2 = v1;
        v1 = v;
        v = t;
      }
    }
  };
  for (int i = 0; i < K.nseg; ++i) extrap(K.lam, K.lam_m1, K.lam_m2, n, s0 + i);
  extrap(K.clamp, K.clamp_m1, K.clamp_m2, (size_t)K.nfil, f);
  if (K.bx) {
    attach_to_body(K, f);
  } else {  // base follows the (fixed) anchor: position reset, Lie algebra element zero
    K.x[s0] = K.base[f];
    K.y[s0] = K.base[K.nfil + f];
    K.z[s0] = K.base[2 * K.nfil + f];
    st3(K.u, n, s0, v3(0.0, 0.0, 0.0));
    stq(K.q, n, s0, ldq(K.qm1, n, s0));
  }
  robot_arm(K, f);
}

__device__ V3 body_frame_moment(const FilK& K, Q4 q1, Q4 q2) {
  Q4 m = qconj(midpoint(q1, q2));
  const double il = 1.0 / K.dl;
  const Q4 dq = {(q2.s - q1.s) * il, (q2.x - q1.x) * il, (q2.y - q1.y) * il, (q2.z - q1.z) * il};
  m = qmul(m, dq);
  return {K.kt * 2.0 * m.x, K.kb * 2.0 * m.y, K.kb * 2.0 * m.z};
}

// One thread per segment: the loads of segment i are the contribution of link (i-1, i)
// (its multiplier and the moment about its midpoint, the part a loop along the filament
// carries from the previous segment) followed by that of link (i, i+1), the tether and
// clamp on the base segment, the follower force at the tip and the step-0 perturbation
// of segment 4 -- added in that order, with the same expressions as the loop.
__global__ void fil_forces_kernel(FilK K, int nt, double* __restrict__ fout, int ld) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= (int)K.nseg_total) return;
  const size_t n = K.nseg_total;
  const int f = s / K.nseg, i = s - f * K.nseg, s0 = f * K.nseg;
  double F[6] = {0, 0, 0, 0, 0, 0};
  const Q4 qi = ldq(K.q, n, s);
  const V3 ti = tangent(qi);
  if (i > 0) {
    const Q4 qp = ldq(K.q, n, s - 1);
    const V3 lam = ld3(K.lam, n, s);
    const V3 c2 = cross(ti, lam);
    const V3 M = mv(rot_mat(midpoint(qp, qi)), body_frame_moment(K, qp, qi));
    F[0] += lam.x; F[1] += lam.y; F[2] += lam.z;
    F[3] += -0.5 * K.dl * c2.x - M.x;
    F[4] += -0.5 * K.dl * c2.y - M.y;
    F[5] += -0.5 * K.dl * c2.z - M.z;
  }
  if (i < K.nseg - 1) {
    const Q4 qn = ldq(K.q, n, s + 1);
    const V3 lam = ld3(K.lam, n, s + 1);
    const V3 c1 = cross(ti, lam);
    const V3 M = mv(rot_mat(midpoint(qi, qn)), body_frame_moment(K, qi, qn));
    F[0] -= lam.x; F[1] -= lam.y; F[2] -= lam.z;
    F[3] += -0.5 * K.dl * c1.x + M.x;
    F[4] += -0.5 * K.dl * c1.y + M.y;
    F[5] += -0.5 * K.dl * c1.z + M.z;
  }
  if (i == 0) {
    // tether force and clamping torque on the base segment
    const V3 tl = ld3(K.lam, n, s0);
    F[0] += tl.x; F[1] += tl.y; F[2] += tl.z;
    const V3 ct = dexpinv_transpose(ld3(K.u, n, s0), ld3(K.clamp, K.nfil, f));
    F[3] += ct.x; F[4] += ct.y; F[5] += ct.z;
  }
  if (i == K.nseg - 1) {
    // follower force at the tip
    F[0] -= K.end_force * ti.x;
    F[1] -= K.end_force * ti.y;
    F[2] -= K.end_force * ti.z;
  }
  if (nt == 0 && i == 4 && K.nseg > 4) {
    F[0] += K.pert[f];
    F[1] += K.pert[K.nfil + f];
    F[2] += K.pert[2 * K.nfil + f];
  }
  for (int c = 0; c < 6; ++c) {
    K.fint[c * n + s] = F[c];
    fout[(size_t)c * ld + s] = F[c];
  }
}

// error vector (reference layout per filament) and the per-segment badness
// max(|e_x|/(DL TOL), |e_u|/TOL) (NaN propagates).
__global__ void fil_errors_kernel(FilK K, int nt, const double* __restrict__ v, int ld,
                                  double* __restrict__ err, double* __restrict__ bad) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double b = 0.0;
  if (s < (int)K.nseg_total) {
    const size_t n = K.nseg_total;
    const int f = s / K.nseg, j = s - f * K.nseg;
    const V3 x = {K.x[s], K.y[s], K.z[s]};
    const V3 xm1 = ld3(K.xm1, n, s), xm2 = ld3(K.xm2, n, s);
    const V3 vel = {v[s], v[ld + s], v[2 * (size_t)ld + s]};
    const V3 om = {v[3 * (size_t)ld + s], v[4 * (size_t)ld + s], v[5 * (size_t)ld + s]};
    const V3 u = ld3(K.u, n, s), um1 = ld3(K.um1, n, s);
    const V3 W = dexpinv(u, om);
    V3 ex, eu;
    if (nt < K.num_euler) {
      ex = x - xm1 - K.dt * vel;
      eu = u - K.dt * W;
    } else {
      ex = x + (-4.0 / 3.0) * xm1 + (1.0 / 3.0) * xm2 + (-2.0 / 3.0 * K.dt) * vel;
      eu = u - (1.0 / 3.0) * um1 + (-2.0 / 3.0 * K.dt) * W;
    }
    double* e = err + (size_t)f * 6 * K.nseg;
    e[3 * j] = ex.x;
    e[3 * j + 1] = ex.y;
    e[3 * j + 2] = ex.z;
    e[3 * K.nseg + 3 * j] = eu.x;
    e[3 * K.nseg + 3 * j + 1] = eu.y;
    e[3 * K.nseg + 3 * j + 2] = eu.z;
    const double sx = 1.0 / (K.dl * K.tol), su = 1.0 / K.tol;
    b = fmax(fmax(fabs(ex.x), fmax(fabs(ex.y), fabs(ex.z))) * sx,
             fmax(fabs(eu.x), fmax(fabs(eu.y), fabs(eu.z))) * su);
    if (!(ex.x == ex.x && ex.y == ex.y && ex.z == ex.z && eu.x == eu.x && eu.y == eu.y &&
          eu.z == eu.z))
      b = __builtin_nan("");
  }
  __shared__ double red[256];
  red[threadIdx.x] = b;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      const double a = red[threadIdx.x], c = red[threadIdx.x + off];
      red[threadIdx.x] = (a != a || c != c) ? __builtin_nan("") : fmax(a, c);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) bad[blockIdx.x] = red[0];
}

__global__ void max_finish_kernel(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double red[256];
  double b = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) {
    const double a = part[i];
    b = (a != a || b != b) ? __builtin_nan("") : fmax(a, b);
  }
  red[threadIdx.x] = b;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      const double a = red[threadIdx.x], c = red[threadIdx.x + off];
      red[threadIdx.x] = (a != a || c != c) ? __builtin_nan("") : fmax(a, c);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// bad[e] = max over rows r of system e of |err[e Lm + r]| times 1/(DL TOL) (position
// rows: the first 3 nseg of each filament's 6 nseg) or 1/TOL (Lie-algebra rows).
__global__ void member_badness_kernel(FilK K, const double* __restrict__ err, int Lm,
                                      double* __restrict__ bad) {
  const double* x = err + (size_t)blockIdx.x * Lm;
  const int per = 6 * K.nseg, half = 3 * K.nseg;
  const double sx = 1.0 / (K.dl * K.tol), su = 1.0 / K.tol;
  double b = 0.0;
  for (int r = threadIdx.x; r < Lm; r += 256) {
    const double a = fabs(x[r]) * ((r % per) < half ? sx : su);
    b = (a != a || b != b) ? __builtin_nan("") : fmax(a, b);
  }
  __shared__ double red[256];
  red[threadIdx.x] = b;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      const double a = red[threadIdx.x], c = red[threadIdx.x + off];
      red[threadIdx.x] = (a != a || c != c) ? __builtin_nan("") : fmax(a, c);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) bad[blockIdx.x] = red[0];
}

// ---- approximate Jacobian -----------------------------------------------------
struct M43 {
  double a[4][3];
};
__device__ __forceinline__ M43 psi_mat(Q4 q) {
  M43 P;
  P.a[0][0] = -0.5 * q.x; P.a[0][1] = -0.5 * q.y; P.a[0][2] = -0.5 * q.z;
  const M3 r = rcross(v3(q.x, q.y, q.z));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) P.a[i + 1][j] = 0.5 * (r.a[i][j] + (i == j ? q.s : 0.0));
  return P;
}
// rows 1..3 of left_mult_mat(q) * A
__device__ __forceinline__ M3 lmul_rows(Q4 q, const M43& A) {
  // left_mult_mat rows 1..3: [v_i, -rcross(v) + s I]
  const M3 r = rcross(v3(q.x, q.y, q.z));
  const double v[3] = {q.x, q.y, q.z};
  M3 out;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = v[i] * A.a[0][j];
      for (int k = 0; k < 3; ++k) s += (-r.a[i][k] + (i == k ? q.s : 0.0)) * A.a[k + 1][j];
      out.a[i][j] = s;
    }
  return out;
}

__device__ M3 moment_lie_derivative(const FilK& K, Q4 q1, Q4 q2, bool upper) {
  M43 A = upper ? psi_mat(q2) : psi_mat(q1);
  const double sA = upper ? 1.0 / K.dl : -1.0 / K.dl;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) A.a[i][j] *= sA;
  const Q4 mq = qconj(midpoint(q1, q2));
  M3 d = lmul_rows(mq, A);
  const double il = 1.0 / K.dl;
  const Q4 dq = {(q2.s - q1.s) * il, (q2.x - q1.x) * il, (q2.y - q1.y) * il, (q2.z - q1.z) * il};
  const double sB = upper ? 0.5 * K.dl : -0.5 * K.dl;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) A.a[i][j] *= sB;
  d = madd(d, lmul_rows(dq, A));
  for (int j = 0; j < 3; ++j) {
    d.a[0][j] *= 2.0 * K.kt;
    d.a[1][j] *= 2.0 * K.kb;
    d.a[2][j] *= 2.0 * K.kb;
  }
  return d;
}

__device__ __forceinline__ void put_block(double* J, int N, int r0, int c0, const M3& B) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) J[(r0 + i) * N + c0 + j] = B.a[i][j];
}

__device__ __forceinline__ M3 frame_sum(V3 Kv, Q4 q) {
  // Kv0 rcross(t) + Kv1 rcross(n) + Kv2 rcross(b)
  M3 A = mscale(rcross(tangent(q)), Kv.x);
  A = madd(A, rcross(normal(q)), Kv.y);
  return madd(A, rcross(binormal(q)), Kv.z);
}

// Approximate Jacobian of one filament (filament.cpp:1855-2037 blocks, swimmer.cpp:
// 539-560 assembly), built in LDS and stored row-major in K.jinv; the in-register
// inversion (fil_jacobian_invert_kernel) then overwrites it with the inverse.
__global__ __launch_bounds__(256) void fil_jacobian_kernel(FilK K, int nt) {
  extern __shared__ double J[];
  const int f = blockIdx.x;
  const int n = K.nseg, N = 6 * n;
  const size_t ns = K.nseg_total;
  const int s0 = f * n;
  const int tid = threadIdx.x, nth = blockDim.x;
  const bool euler = nt < K.num_euler;
  const double pi = 3.14159265358979323846;
  const double a3 = K.a * K.a * K.a;
  for (int e = tid; e < N * N; e += nth) J[e] = 0.0;
  __syncthreads();
  const double fac = euler ? -K.dt / (6.0 * pi * K.mu * K.a) : -K.dt / (9.0 * pi * K.mu * K.a);
  for (int r = tid; r < 3 * n; r += nth) {
    J[r * N + r] = fac;
    if (r < 3 * n - 3) J[r * N + r + 3] = -fac;
  }
  __syncthreads();
  if (tid < n) {
    const int j = tid;
    if (j >= 1) {
      const M3 temp = mscale(rcross(tangent(ldq(K.q, ns, s0 + j))), K.dl);
      put_block(J, N, 3 * j, 3 * (n + j), mscale(temp, 0.5));
      for (int i = j + 1; i < n; ++i) put_block(J, N, 3 * i, 3 * (n + j), temp);
      if (j == n - 1) {
        const M3 tip = mscale(rcross(tangent(ldq(K.q, ns, s0 + n - 1))), fac * K.end_force);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) J[(3 * (n - 1) + a) * N + 3 * (2 * n - 1) + b] -= tip.a[a][b];
      }
      const double fac2 = euler ? K.dl * K.dt / (16.0 * pi * K.mu * a3) : K.dl * K.dt / (24.0 * pi * K.mu * a3);
      for (int kk = j - 1; kk <= j; ++kk) {
        const M3 um = rcross(ld3(K.u, ns, s0 + kk));
        const M3 tm = mscale(rcross(tangent(ldq(K.q, ns, s0 + kk))), -1.0);
        M3 B = madd(tm, mm(um, tm), 0.5);
        B = madd(B, mm(um, mm(um, tm)), 1.0 / 12.0);
        put_block(J, N, 3 * (n + kk), 3 * j, mscale(B, fac2));
      }
    }
    // Lie-algebra rows of segment k = j
    const int k = j;
    const double tsf = euler ? K.dt : 2.0 * K.dt / 3.0;
    const double Tfac = 1.0 / (8.0 * pi * K.mu * a3);
    const V3 uk = ld3(K.u, ns, s0 + k);
    const M3 Lie = mscale(rcross(uk), -1.0);
    const V3 om = {Tfac * K.fint[3 * ns + s0 + k], Tfac * K.fint[4 * ns + s0 + k],
                   Tfac * K.fint[5 * ns + s0 + k]};
    const M3 Om = rcross(om);
    const M3 LL = mm(Lie, Lie);
    auto diag_block = [&](const M3& dO) {
      M3 B = madd(dO, madd(Om, mm(Lie, dO)), -0.5);
      M3 C = madd(madd(rcross(mv(Om, uk)), mm(Lie, Om)), mm(LL, dO));
      B = madd(B, C, 1.0 / 12.0);
      return mscale(B, -tsf);
    };
    auto off_block = [&](const M3& dO) {
      M3 B = madd(dO, mm(Lie, dO), -0.5);
      B = madd(B, mm(LL, dO), 1.0 / 12.0);
      return mscale(B, -tsf);
    };
    const Q4 qk = ldq(K.q, ns, s0 + k);
    const V3 tk = tangent(qk);
    M3 D = zero3(), Lb = zero3(), Rb = zero3();
    M3 Cp;
    {
      V3 lamvec = v3(0, 0, 0);
      if (k >= 1) lamvec = lamvec + ld3(K.lam, ns, s0 + k);
      if (k <= n - 2) lamvec = lamvec + ld3(K.lam, ns, s0 + k + 1);
      Cp = mscale(mm(rcross(lamvec), rcross(tk)), -0.5 * K.dl);
    }
    M3 Eld = zero3();  // elastic part of the diagonal derivative
    if (k <= n - 2) {
      const Q4 qp = ldq(K.q, ns, s0 + k + 1);
      const V3 Kph = body_frame_moment(K, qk, qp);
      const M3 Rph = rot_mat(midpoint(qk, qp));
      Eld = madd(Eld, madd(mscale(frame_sum(Kph, qk), 0.5), mm(Rph, moment_lie_derivative(K, qk, qp, false))));
      const M3 Er = madd(mscale(frame_sum(Kph, qp), 0.5), mm(Rph, moment_lie_derivative(K, qk, qp, true)));
      Rb = off_block(mscale(Er, Tfac));
    }
    if (k >= 1) {
      const Q4 qm = ldq(K.q, ns, s0 + k - 1);
      const V3 Kmh = body_frame_moment(K, qm, qk);
      const M3 Rmh = rot_mat(midpoint(qm, qk));
      Eld = madd(Eld, madd(mscale(frame_sum(Kmh, qk), 0.5), mm(Rmh, moment_lie_derivative(K, qm, qk, true))), -1.0);
      const M3 El = madd(mscale(frame_sum(Kmh, qm), -0.5), mm(Rmh, moment_lie_derivative(K, qm, qk, false)), -1.0);
      Lb = off_block(mscale(El, Tfac));
    }
    D = diag_block(mscale(madd(Eld, Cp), Tfac));
    D = madd(D, eye3());
    const int r0 = 3 * (n + k);
    if (K.clampblk && k <= 1) {  // how rows 0 and 1 depend on u_0 (elastic_clamping_block1/2)
      const M3& E = k == 0 ? D : Lb;
      double* cb = K.clampblk + (size_t)f * 18 + 9 * k;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) cb[3 * a + b] = E.a[a][b];
    }
    if (k == 0) {
      // clamped base: the clamp multiplier replaces the base Lie-algebra unknowns
      const double fac3 = euler ? -K.dt / (8.0 * pi * K.mu * a3) : -K.dt / (12.0 * pi * K.mu * a3);
      const M3 um = rcross(uk);
      const M3 blk = madd(madd(eye3(), um, 0.5), mm(um, um), 1.0 / 12.0);
      const M3 blkT = madd(madd(eye3(), um, -0.5), mm(um, um), 1.0 / 12.0);
      D = mscale(mm(blk, blkT), fac3);
    }
    put_block(J, N, r0, 3 * (n + k), D);
    if (k >= 1) put_block(J, N, r0, 3 * (n + k - 1), k == 1 ? zero3() : Lb);
    if (k <= n - 2) put_block(J, N, r0, 3 * (n + k + 1), Rb);
  }
  __syncthreads();
  // row-major copy (element (r, c) at r N + c) for the in-register inversion
  double* out = K.jinv + (size_t)f * N * N;
  for (int e = tid; e < N * N; e += nth) out[e] = J[e];
}

// ---- in-register Gauss-Jordan inversion of the approximate Jacobians ------------------
// One workgroup of 256 threads per filament; row r of the N x N block (N = 6 NSEG <= 128)
// belongs to threads 2r (columns 0..63) and 2r + 1 (columns 64..127), each holding its
// 64 entries in registers.  Step k picks the pivot of column k among the rows not yet
// used (largest magnitude, lowest row on ties), the pivot row is scaled and broadcast
// through LDS, and every other row eliminates column k; rows are never moved.  After the
// N steps the row that pivoted column k holds row k of (P A)^-1 = A^-1 P^T, i.e.
// A^-1[k][p_j] in its column j (p_j the pivot row of column j), which is stored as the
// transposed inverse (element (r, c) of A^-1 at c N + r).  This is the in-place
// Gauss-Jordan elimination with partial pivoting; the reference inverts each block with
// LAPACK getrf/getri (matrix.cpp:538-571), the same inverse up to rounding.
constexpr int kGjHalf = 64;

// Register entry c of a thread's half-row, c uniform across the wave (a scalar branch).
#define CFM_GJ_CASES(X)                                                                          \
  X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)   \
  X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31)     \
  X(32) X(33) X(34) X(35) X(36) X(37) X(38) X(39) X(40) X(41) X(42) X(43) X(44) X(45) X(46)     \
  X(47) X(48) X(49) X(50) X(51) X(52) X(53) X(54) X(55) X(56) X(57) X(58) X(59) X(60) X(61)     \
  X(62) X(63)
__device__ __forceinline__ double gj_get(const double (&a)[kGjHalf], int c) {
  switch (c) {
#define CFM_GJ_GET(i) case i: return a[i];
    CFM_GJ_CASES(CFM_GJ_GET)
#undef CFM_GJ_GET
  }
  return 0.0;
}
__device__ __forceinline__ void gj_set(double (&a)[kGjHalf], int c, double v) {
  switch (c) {
#define CFM_GJ_SET(i) case i: a[i] = v; break;
    CFM_GJ_CASES(CFM_GJ_SET)
#undef CFM_GJ_SET
  }
}
#undef CFM_GJ_CASES

__global__ __launch_bounds__(256) void fil_jacobian_invert_kernel(FilK K, int* __restrict__ status) {
  __shared__ double prow[2 * kGjHalf];
  __shared__ double rval[4];
  __shared__ int rrow[4];
  __shared__ int perm[2 * kGjHalf];
  const int f = blockIdx.x;
  const int N = 6 * K.nseg;
  const int tid = threadIdx.x;
  const int r = tid >> 1, h = tid & 1;
  const bool row_ok = r < N;
  double* Jf = K.jinv + (size_t)f * N * N;
  double a[kGjHalf];
#pragma unroll
  for (int c = 0; c < kGjHalf; ++c) {
    const int col = h * kGjHalf + c;
    a[c] = (row_ok && col < N) ? Jf[(size_t)r * N + col] : 0.0;
  }
  bool used = !row_ok;
  int my_k = 0;
  bool bad = false;
  for (int k = 0; k < N; ++k) {
    const int hk = k >> 6, ck = k & (kGjHalf - 1);
    // column k of this row (held by the lane of half hk)
    const double mine = gj_get(a, ck);
    const double other = __shfl_xor(mine, 1, 64);
    const double bk = (h == hk) ? mine : other;
    // pivot: largest |B[r][k]| over unused rows, the lowest row on ties; a non-finite
    // entry ranks above every finite one (and flags the block), so the pivot is always
    // an unused row and the pivots form a permutation of the rows
    double v = used ? -1.0 : (isfinite(bk) ? fabs(bk) : DBL_MAX);
    int vr = r;
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(v, off, 64);
      const int orr = __shfl_xor(vr, off, 64);
      if (ov > v || (ov == v && orr < vr)) {
        v = ov;
        vr = orr;
      }
    }
    if ((tid & 63) == 0) {
      rval[tid >> 6] = v;
      rrow[tid >> 6] = vr;
    }
    __syncthreads();
    double pv = rval[0];
    int p = rrow[0];
    for (int w = 1; w < 4; ++w)
      if (rval[w] > pv || (rval[w] == pv && rrow[w] < p)) {
        pv = rval[w];
        p = rrow[w];
      }
    if (!(pv > 0.0) || pv == DBL_MAX) bad = true;
    if (r == p) {
      // pivot row: B[p][k] = 1, then the row scaled by 1 / piv
      const double inv = 1.0 / bk;
      if (h == hk) gj_set(a, ck, 1.0);
#pragma unroll
      for (int c = 0; c < kGjHalf; ++c) {
        a[c] *= inv;
        prow[h * kGjHalf + c] = a[c];
      }
      used = true;
      my_k = k;
      if (h == 0) perm[k] = r;
    }
    __syncthreads();
    if (r != p && row_ok) {
      // B[r][k] = 0, then B[r][:] -= B[r][k] (old) x pivot row
      if (h == hk) gj_set(a, ck, 0.0);
#pragma unroll
      for (int c = 0; c < kGjHalf; ++c) a[c] = fma(-bk, prow[h * kGjHalf + c], a[c]);
    }
  }
  if (bad && tid == 0) status[0] = 1;
  __syncthreads();  // perm complete; every row read its block before any store below
  if (row_ok) {
#pragma unroll
    for (int c = 0; c < kGjHalf; ++c) {
      const int col = h * kGjHalf + c;
      if (col < N && perm[col] < N) Jf[(size_t)perm[col] * N + my_k] = a[c];
    }
  }
}

// out_f = scale * J0^-1 in_f for every filament (vectors in the reference layout)
__global__ __launch_bounds__(128) void fil_jinv_mult_kernel(FilK K, const double* __restrict__ in,
                                                            double* __restrict__ out, double scale) {
  __shared__ double x[128];
  const int f = blockIdx.x;
  const int N = 6 * K.nseg;
  const double* v = in + (size_t)f * N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) x[i] = v[i];
  __syncthreads();
  const double* JT = K.jinv + (size_t)f * N * N;
  for (int r = threadIdx.x; r < N; r += blockDim.x) {
    double s = 0.0;
    for (int c = 0; c < N; ++c) s += JT[(size_t)c * N + r] * x[c];
    out[(size_t)f * N + r] = scale * s;
  }
}

// One workgroup per group of 256 / nseg filaments: first one thread per segment (its
// multiplier, the clamp on the base segment, the Lie-algebra element and quaternion of
// the others; with free bodies the base attached to its moved body, swimmer.cpp:1076-1100),
// then one thread per filament places the segments by the robot arm.
__global__ __launch_bounds__(256) void fil_update_kernel(FilK K, const double* __restrict__ upd) {
  const int n = K.nseg, per = 256 / n;
  const int f0 = blockIdx.x * per;
  const int fl = threadIdx.x / n, i = threadIdx.x - fl * n, f = f0 + fl;
  const size_t ns = K.nseg_total;
  if (fl < per && f < K.nfil) {
    const int s = f * n + i;
    const double* u = upd + (size_t)f * 6 * n;
    if (i == 0 && K.bx) attach_to_body(K, f);
    st3(K.lam, ns, s, ld3(K.lam, ns, s) + v3(u[3 * i], u[3 * i + 1], u[3 * i + 2]));
    if (i == 0) {
      st3(K.clamp, K.nfil, f, ld3(K.clamp, K.nfil, f) + v3(u[3 * n], u[3 * n + 1], u[3 * n + 2]));
    } else {
      const V3 un = ld3(K.u, ns, s) + v3(u[3 * (n + i)], u[3 * (n + i) + 1], u[3 * (n + i) + 2]);
      st3(K.u, ns, s, un);
      stq(K.q, ns, s, qmul(lie_exp(un), ldq(K.qm1, ns, s)));
    }
  }
  __syncthreads();  // the group's quaternions and bases are written
  if ((int)threadIdx.x < per && f0 + (int)threadIdx.x < K.nfil) robot_arm(K, f0 + threadIdx.x);
}

// ---- barrier forces over a hashed cell list -----------------------------------------
// Cells have the edge of the barrier cut-off (1.1 x 2a), so every partner of a
// segment lies in its own or one of the 26 adjacent cells.  Cells are hashed into a
// power-of-two table; a bucket can hold particles of several cells, so a visit keeps
// only the particles whose own cell is the cell being visited (no double counting).

__device__ __forceinline__ int cell_coord(double v, double inv_h) {
  return (int)floor(fmin(fmax(v * inv_h, -1.0e9), 1.0e9));
}

__device__ __forceinline__ int cell_bucket(int cx, int cy, int cz, int mask) {
  const unsigned h = ((unsigned)cx * 73856093u) ^ ((unsigned)cy * 19349663u) ^
                     ((unsigned)cz * 83492791u);
  return (int)(h & (unsigned)mask);
}

struct GridPos {
  const double *sx, *sy, *sz;  // segments
  const double *bx, *by, *bz;  // blobs
  int ns, n;
  __device__ __forceinline__ void get(int j, double& x, double& y, double& z) const {
    if (j < ns) {
      x = sx[j]; y = sy[j]; z = sz[j];
    } else {
      x = bx[j - ns]; y = by[j - ns]; z = bz[j - ns];
    }
  }
};

__global__ void grid_clear_kernel(int* __restrict__ a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

__global__ void grid_count_kernel(GridPos P, double inv_h, int mask, int* __restrict__ key,
                                  int* __restrict__ count) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= P.n) return;
  double x, y, z;
  P.get(j, x, y, z);
  const int b = cell_bucket(cell_coord(x, inv_h), cell_coord(y, inv_h), cell_coord(z, inv_h), mask);
  key[j] = b;
  atomicAdd(count + b, 1);
}

// Exclusive scan of the bucket counts in two passes over blocks of 1024 buckets
// (nb is a power of two >= 1024).  Pass 1: each block scans its buckets in place
// (int4 loads, wave prefix sums by lane shuffles) and writes its total.
__global__ __launch_bounds__(256) void grid_scan_blocks_kernel(int* __restrict__ start,
                                                               int* __restrict__ blocksum) {
  __shared__ int wsum[4];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  int4* p = reinterpret_cast<int4*>(start + (size_t)blockIdx.x * 1024) + t;
  int4 v = *p;
  const int tot = v.x + v.y + v.z + v.w;
  int inc = tot;  // inclusive prefix over the wave
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(inc, o, 64);
    if (lane >= o) inc += u;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  int base = 0;
  for (int w = 0; w < wave; ++w) base += wsum[w];
  const int ex = base + inc - tot;
  int4 o;
  o.x = ex;
  o.y = ex + v.x;
  o.z = o.y + v.y;
  o.w = o.z + v.z;
  *p = o;
  if (t == 0) blocksum[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// Pass 2: add the totals of the preceding blocks, copy to the scatter cursors and
// close the table with start[nb] = number of particles.
__global__ __launch_bounds__(256) void grid_scan_add_kernel(int* __restrict__ start,
                                                            int* __restrict__ cursor,
                                                            const int* __restrict__ blocksum,
                                                            int nb) {
  __shared__ int off;
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int b = 0; b < (int)blockIdx.x; ++b) acc += blocksum[b];
    off = acc;
    if (blockIdx.x == gridDim.x - 1) start[nb] = acc + blocksum[blockIdx.x];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 1024; e += 256) {
    const size_t i = (size_t)blockIdx.x * 1024 + e;
    const int v = start[i] + off;
    start[i] = v;
    cursor[i] = v;
  }
}

__global__ void grid_scatter_kernel(int n, const int* __restrict__ key, int* __restrict__ cursor,
                                    int* __restrict__ ids) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  ids[atomicAdd(cursor + key[j], 1)] = j;
}

// Ascending ids inside every bucket, so the visit order (and the rounding of the
// force sums) does not depend on the order the atomics handed out slots.
__global__ void grid_sort_kernel(int nb, const int* __restrict__ start, int* __restrict__ ids) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int s0 = start[b], s1 = start[b + 1];
  for (int i = s0 + 1; i < s1; ++i) {
    const int v = ids[i];
    int k = i - 1;
    while (k >= s0 && ids[k] > v) {
      ids[k + 1] = ids[k];
      --k;
    }
    ids[k + 1] = v;
  }
}

// Barrier (short-range repulsion) forces added to the segment loads in fout
// (cuda_functions.cu:1825-1930): wall repulsion below half a segment length, and a
// (1 - chi (r - 2a))^4 / r repulsion from every particle within 1.1 x 2a except the
// segment itself and its neighbours along the same filament.
// 32 lanes per segment, lane c < 27 walking neighbour cell c (dz, dy, dx = c / 9 - 1,
// c / 3 % 3 - 1, c % 3 - 1) and its bucket's particles in ascending id; the 27 cell sums
// are added by a butterfly in a fixed pattern, then the wall term.  (One thread walking
// the 27 cells in turn was bound by the chain of dependent loads of each cell.)
__global__ __launch_bounds__(256) void barrier_kernel(FilK K, GridPos P, double inv_h, int mask,
                                                      const int* __restrict__ start,
                                                      const int* __restrict__ ids, int wall,
                                                      double* __restrict__ fout, int ld) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = g >> 5, c = g & 31;
  if (i >= (int)K.nseg_total) return;  // all 32 lanes of a segment leave together
  const double xi = K.x[i], yi = K.y[i], zi = K.z[i];
  const int fi = i / K.nseg;
  const double fmag = K.rep_factor * K.end_force;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  const double a_sum = 2.0 * K.a;
  const double cut = 1.1 * a_sum, chi = 10.0 / a_sum;
  if (c < 27) {
    const int nx = cell_coord(xi, inv_h) + c % 3 - 1, ny = cell_coord(yi, inv_h) + (c / 3) % 3 - 1,
              nz = cell_coord(zi, inv_h) + c / 9 - 1;
    const int b = cell_bucket(nx, ny, nz, mask);
    const int t1 = start[b + 1];
    for (int t = start[b]; t < t1; ++t) {
      const int j = ids[t];
      double xj, yj, zj;
      P.get(j, xj, yj, zj);
      if (cell_coord(xj, inv_h) != nx || cell_coord(yj, inv_h) != ny || cell_coord(zj, inv_h) != nz)
        continue;
      if (j < P.ns && j / K.nseg == fi && abs(i - j) <= 1) continue;
      if (K.fpm && j < P.ns && (j / K.nseg) / K.fpm != fi / K.fpm) continue;  // other system
      const double ex = xi - xj, ey = yi - yj, ez = zi - zj;
      const double d2 = ex * ex + ey * ey + ez * ez;
      if (d2 < cut * cut) {
        const double dist = sqrt(d2);
        double fac = fmin(1.0, 1.0 - chi * (dist - a_sum));
        fac = fmag * fac * fac * fac * fac / dist;
        fx += fac * ex;
        fy += fac * ey;
        fz += fac * ez;
      }
    }
  }
  for (int off = 16; off > 0; off >>= 1) {
    fx += __shfl_xor(fx, off, 32);
    fy += __shfl_xor(fy, off, 32);
    fz += __shfl_xor(fz, off, 32);
  }
  if (c != 0) return;
  if (wall && zi < 0.5 * K.dl) {
    double w = fmin(1.0, 1.0 - (zi - K.a) / (0.5 * K.dl - K.a));
    fz += fmag * w * w * w * w;
  }
  fout[i] += fx;
  fout[(size_t)ld + i] += fy;
  fout[2 * (size_t)ld + i] += fz;
}

// Repulsion on blob ib from every segment and every blob of another body within
// 1.1 x 2a (cuda_functions.cu:1932-2012; the pairs of one body's own blobs cancel in
// its total force and torque, which is all the body balance uses, and are skipped).
__global__ __launch_bounds__(256) void blob_barrier_kernel(FilK K, GridPos P, double inv_h,
                                                           int mask, const int* __restrict__ start,
                                                           const int* __restrict__ ids,
                                                           int nb_per_body,
                                                           double* __restrict__ rep) {
  // 32 lanes per blob, one neighbour cell each, as barrier_kernel
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int ib = g >> 5, c = g & 31;
  const int nblob = P.n - P.ns;
  if (ib >= nblob) return;  // all 32 lanes of a blob leave together
  const double xi = P.bx[ib], yi = P.by[ib], zi = P.bz[ib];
  const int bi = ib / nb_per_body;
  const double fmag = K.rep_factor * K.end_force;
  const double a_sum = 2.0 * K.a;
  const double cut = 1.1 * a_sum, chi = 10.0 / a_sum;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  if (c < 27) {
    const int nx = cell_coord(xi, inv_h) + c % 3 - 1, ny = cell_coord(yi, inv_h) + (c / 3) % 3 - 1,
              nz = cell_coord(zi, inv_h) + c / 9 - 1;
    const int b = cell_bucket(nx, ny, nz, mask);
    const int t1 = start[b + 1];
    for (int t = start[b]; t < t1; ++t) {
      const int j = ids[t];
      if (j >= P.ns && (j - P.ns) / nb_per_body == bi) continue;
      double xj, yj, zj;
      P.get(j, xj, yj, zj);
      if (cell_coord(xj, inv_h) != nx || cell_coord(yj, inv_h) != ny || cell_coord(zj, inv_h) != nz)
        continue;
      const double ex = xi - xj, ey = yi - yj, ez = zi - zj;
      const double d2 = ex * ex + ey * ey + ez * ez;
      if (d2 < cut * cut) {
        const double dist = sqrt(d2);
        double fac = fmin(1.0, 1.0 - chi * (dist - a_sum));
        fac = fmag * fac * fac * fac * fac / dist;
        fx += fac * ex;
        fy += fac * ey;
        fz += fac * ez;
      }
    }
  }
  for (int off = 16; off > 0; off >>= 1) {
    fx += __shfl_xor(fx, off, 32);
    fy += __shfl_xor(fy, off, 32);
    fz += __shfl_xor(fz, off, 32);
  }
  if (c != 0) return;
  rep[ib] = fx;
  rep[(size_t)nblob + ib] = fy;
  rep[2 * (size_t)nblob + ib] = fz;
}

Status BarrierGrid::ensure(int n_particles) {
  int nb = 1024;
  while (nb < 2 * n_particles) nb <<= 1;
  if (nb != nbucket) {
    CFM_TRY(start.alloc((size_t)nb + 1));
    CFM_TRY(cursor.alloc((size_t)nb));
    CFM_TRY(blocksum.alloc((size_t)nb / 1024));
    nbucket = nb;
  }
  CFM_TRY(ids.alloc((size_t)std::max(1, n_particles)));
  CFM_TRY(key.alloc((size_t)std::max(1, n_particles)));
  return {};
}

// ---- launchers -------------------------------------------------------------------
static inline dim3 gsz(int n, int b) { return dim3((n + b - 1) / b); }

Status fil_initial_guess(cfm_ctx* c, const FilK& K, int nt) {
  hipLaunchKernelGGL(fil_initial_guess_kernel, gsz(K.nfil, 64), dim3(64), 0, c->stream, K, nt);
  CFM_CHECK_LAUNCH();
  return {};
}

Status fil_forces(cfm_ctx* c, const FilK& K, int nt) {
  hipLaunchKernelGGL(fil_forces_kernel, gsz((int)K.nseg_total, 256), dim3(256), 0, c->stream, K, nt,
                     c->p.f.p, c->p.n);
  CFM_CHECK_LAUNCH();
  return {};
}

Status fil_barrier(cfm_ctx* c, const FilK& K, BarrierGrid& g) {
  const Particles& p = c->p;
  const int n = p.n_seg + p.n_blob;
  CFM_TRY(g.ensure(n));
  const GridPos P{K.x, K.y, K.z, p.x.p + p.n_seg, p.y.p + p.n_seg, p.z.p + p.n_seg, p.n_seg, n};
  const double inv_h = 1.0 / (1.1 * 2.0 * K.a);
  const int nb = g.nbucket, mask = nb - 1;
  hipLaunchKernelGGL(grid_clear_kernel, gsz(nb + 1, 256), dim3(256), 0, c->stream, g.start.p, nb + 1);
  hipLaunchKernelGGL(grid_count_kernel, gsz(n, 256), dim3(256), 0, c->stream, P, inv_h, mask,
                     g.key.p, g.start.p);
  hipLaunchKernelGGL(grid_scan_blocks_kernel, dim3(nb / 1024), dim3(256), 0, c->stream, g.start.p,
                     g.blocksum.p);
  hipLaunchKernelGGL(grid_scan_add_kernel, dim3(nb / 1024), dim3(256), 0, c->stream, g.start.p,
                     g.cursor.p, g.blocksum.p, nb);
  hipLaunchKernelGGL(grid_scatter_kernel, gsz(n, 256), dim3(256), 0, c->stream, n, g.key.p,
                     g.cursor.p, g.ids.p);
  hipLaunchKernelGGL(grid_sort_kernel, gsz(nb, 256), dim3(256), 0, c->stream, nb, g.start.p, g.ids.p);
  hipLaunchKernelGGL(barrier_kernel, gsz((int)K.nseg_total * 32, 256), dim3(256), 0, c->stream, K, P,
                     inv_h, mask, g.start.p, g.ids.p, c->cfg.wall, p.f.p, p.n);
  CFM_CHECK_LAUNCH();
  return {};
}

Status fil_blob_barrier(cfm_ctx* c, const FilK& K, BarrierGrid& g, int nb_per_body, double* rep) {
  const Particles& p = c->p;
  const int n = p.n_seg + p.n_blob;
  if (p.n_blob == 0) return {};
  const GridPos P{K.x, K.y, K.z, p.x.p + p.n_seg, p.y.p + p.n_seg, p.z.p + p.n_seg, p.n_seg, n};
  const double inv_h = 1.0 / (1.1 * 2.0 * K.a);
  hipLaunchKernelGGL(blob_barrier_kernel, gsz(p.n_blob * 32, 256), dim3(256), 0, c->stream, K, P, inv_h,
                     g.nbucket - 1, g.start.p, g.ids.p, nb_per_body, rep);
  CFM_CHECK_LAUNCH();
  return {};
}

Status fil_member_badness(cfm_ctx* c, const FilK& K, const double* err, int Lm, int E, double* bad) {
  hipLaunchKernelGGL(member_badness_kernel, dim3(E), dim3(256), 0, c->stream, K, err, Lm, bad);
  CFM_CHECK_LAUNCH();
  return {};
}

Status fil_errors(cfm_ctx* c, const FilK& K, int nt, double* err, double* bad_part,
                  double* bad_out, int extra) {
  const int nb = ((int)K.nseg_total + 255) / 256;
  hipLaunchKernelGGL(fil_errors_kernel, dim3(nb), dim3(256), 0, c->stream, K, nt, c->p.v.p,
                     c->p.n, err, bad_part);
  hipLaunchKernelGGL(max_finish_kernel, dim3(1), dim3(256), 0, c->stream, bad_part, nb + extra,
                     bad_out);
  CFM_CHECK_LAUNCH();
  return {};
}

// Raises the Jacobian kernel's dynamic-LDS limit once per device (with that device
// current) to the largest block cfm_fil_init admits, after checking it against the
// device's per-workgroup LDS; contexts on several host threads share the setting.
static Status jacobian_lds_limit(int device) {
  static std::mutex mu;
  static std::vector<int> done;  // per device: 0 unset, 1 set
  const size_t need = (size_t)(6 * kMaxFilSeg) * (6 * kMaxFilSeg) * sizeof(double);
  std::lock_guard<std::mutex> lock(mu);
  if ((int)done.size() <= device) done.resize(device + 1, 0);
  if (done[device]) return {};
  int limit = 0;
  CFM_HIP(hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  if ((size_t)limit < need)
    return make_status(CFM_EINVAL, "the %d-segment Jacobian block needs %zu bytes of LDS; the device offers %d",
                       kMaxFilSeg, need, limit);
  CFM_HIP(hipFuncSetAttribute((const void*)fil_jacobian_kernel,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
  done[device] = 1;
  return {};
}

Status fil_jacobian(cfm_ctx* c, const FilK& K, int nt, int* status_dev, hipStream_t s) {
  const hipStream_t st = s ? s : c->stream;
  const int N = 6 * K.nseg;
  const size_t lds = (size_t)N * N * sizeof(double);
  if (K.nseg > kMaxFilSeg)
    return make_status(CFM_EINVAL, "filaments of %d segments exceed the limit of %d", K.nseg, kMaxFilSeg);
  CFM_TRY(jacobian_lds_limit(c->device));
  hipLaunchKernelGGL(fil_jacobian_kernel, dim3(K.nfil), dim3(256), lds, st, K, nt);
  CFM_CHECK_LAUNCH();
  // J0^-1 of every filament in place (the reference inverts each block with LAPACK
  // getrf/getri, matrix.cpp:538-571)
  hipLaunchKernelGGL(fil_jacobian_invert_kernel, dim3(K.nfil), dim3(256), 0, st, K, status_dev);
  CFM_CHECK_LAUNCH();
  return {};
}

Status fil_jinv_mult(cfm_ctx* c, const FilK& K, const double* in, double* out, double scale) {
  hipLaunchKernelGGL(fil_jinv_mult_kernel, dim3(K.nfil), dim3(128), 0, c->stream, K, in, out, scale);
  CFM_CHECK_LAUNCH();
  return {};
}

Status fil_update(cfm_ctx* c, const FilK& K, const double* upd) {
  const int per = 256 / K.nseg;  // filaments per workgroup (nseg <= kMaxFilSeg)
  hipLaunchKernelGGL(fil_update_kernel, dim3((K.nfil + per - 1) / per), dim3(256), 0, c->stream, K, upd);
  CFM_CHECK_LAUNCH();
  return {};
}

}  // namespace cfm
