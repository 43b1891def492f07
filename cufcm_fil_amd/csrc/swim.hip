// swim.hip -- free swimmers: rigid bodies of blobs carrying the elastic filaments, with
// the body position and orientation among the Broyden unknowns (BODY_VELOCITY_TYPE 0,
// blob forces and body velocities by GMRES).
//
//   body initial guess / update     rigid_body.cpp:410-458, 517-544
//   body load of the swimmer        swimmer.cpp:470-535, mobility_solver.cpp:215-236
//   body rows of the BDF residual   mobility_solver.cpp:3003-3040
//   body Schur complement of J0     swimmer.cpp:688-909 (no sphere-specific correction)
//   J0^-1 with the body             swimmer.cpp:1022-1068
// Unknown order of the Broyden vector: the filaments' 6 NSEG blocks, then 6 per body
// (position, Lie algebra).  Filament f belongs to body f / fpb.
#include <hip/hip_runtime.h>

#include "fil.hpp"
#include "fil_device.hpp"
#include "swim.hpp"
#include "body.hpp"

namespace cfm {
using namespace fil;

namespace {

__device__ __forceinline__ V3 ldv(const double* a) { return {a[0], a[1], a[2]}; }
__device__ __forceinline__ void stv(double* a, V3 v) {
  a[0] = v.x;
  a[1] = v.y;
  a[2] = v.z;
}
__device__ __forceinline__ Q4 ldq4(const double* a) { return {a[0], a[1], a[2], a[3]}; }
__device__ __forceinline__ Q4 segq(const FilK& K, int s) {
  const size_t n = K.nseg_total;
  return {K.q[s], K.q[n + s], K.q[2 * n + s], K.q[3 * n + s]};
}
__device__ __forceinline__ void stq4(double* a, Q4 q) {
  a[0] = q.s;
  a[1] = q.x;
  a[2] = q.y;
  a[3] = q.z;
}
__device__ __forceinline__ M3 ldm(const double* a) {  // row-major 3 x 3
  M3 m;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m.a[i][j] = a[3 * i + j];
  return m;
}
__device__ __forceinline__ M3 transpose(const M3& A) {
  M3 T;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T.a[i][j] = A.a[j][i];
  return T;
}
// 3 x 3 block (bi, bj) of a row-major 6 x 6
__device__ __forceinline__ M3 blk6(const double* A, int bi, int bj) {
  M3 m;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m.a[i][j] = A[(bi + i) * 6 + bj + j];
  return m;
}

// rigid_body::initial_guess: extrapolate position and Lie algebra, q = exp(u) q_-1
__global__ void swim_guess_kernel(SwimK S, int nt) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= S.nbody) return;
  const V3 x = ldv(S.x + 3 * b), xm1 = ldv(S.xm1 + 3 * b), xm2 = ldv(S.xm2 + 3 * b);
  const V3 u = ldv(S.u + 3 * b), um1 = ldv(S.um1 + 3 * b);
  V3 gx, gu;
  if (nt <= S.num_euler) {
    gx = 2.0 * x - xm1;
    gu = u;
  } else {
    gx = 3.0 * (x - xm1) + xm2;
    gu = 2.0 * u - um1;
  }
  stv(S.xm2 + 3 * b, xm1);
  stv(S.xm1 + 3 * b, x);
  stv(S.x + 3 * b, gx);
  stv(S.um1 + 3 * b, u);
  stv(S.u + 3 * b, gu);
  const Q4 q = ldq4(S.q + 4 * b);
  stq4(S.qm1 + 4 * b, q);
  const Q4 qn = qmul(lie_exp(gu), q);
  stq4(S.q + 4 * b, qn);
  const M3 R = rot_mat(qn);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) S.Qinit[9 * b + 3 * i + j] = R.a[i][j];
}

// rigid_body::update from the body block of a Broyden increment
__global__ void swim_update_kernel(SwimK S, const double* __restrict__ upd) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= S.nbody) return;
  const double* d = upd + 6 * b;
  stv(S.x + 3 * b, ldv(S.x + 3 * b) + ldv(d));
  const V3 u = ldv(S.u + 3 * b) + ldv(d + 3);
  stv(S.u + 3 * b, u);
  stq4(S.q + 4 * b, qmul(lie_exp(u), ldq4(S.qm1 + 4 * b)));
}

// Body load: minus the internal loads of its segments about the centre, the repulsion on
// its blobs about the centre, plus the external load.  One workgroup per body.
__global__ __launch_bounds__(256) void swim_force_kernel(SwimK S, FilK K, const double* __restrict__ rep,
                                                         const double* __restrict__ boff, int nblob) {
  __shared__ double red[4][6];
  const int b = blockIdx.x;
  const size_t ns = K.nseg_total;
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const V3 xb = ldv(S.x + 3 * b);
  const int s1 = (b + 1) * S.fpb * S.nseg;
  for (int i = b * S.fpb * S.nseg + threadIdx.x; i < s1; i += 256) {
    const V3 F = {K.fint[i], K.fint[ns + i], K.fint[2 * ns + i]};
    const V3 T = {K.fint[3 * ns + i], K.fint[4 * ns + i], K.fint[5 * ns + i]};
    const V3 d = v3(K.x[i], K.y[i], K.z[i]) - xb;
    const V3 tq = T + cross(d, F);
    s[0] -= F.x;
    s[1] -= F.y;
    s[2] -= F.z;
    s[3] -= tq.x;
    s[4] -= tq.y;
    s[5] -= tq.z;
  }
  if (rep) {
    const int j1 = (b + 1) * S.nb;
    for (int j = b * S.nb + threadIdx.x; j < j1; j += 256) {
      const V3 F = {rep[j], rep[nblob + j], rep[2 * (size_t)nblob + j]};
      const V3 r = {boff[j], boff[nblob + j], boff[2 * (size_t)nblob + j]};
      const V3 tq = cross(r, F);
      s[0] += F.x;
      s[1] += F.y;
      s[2] += F.z;
      s[3] += tq.x;
      s[4] += tq.y;
      s[5] += tq.z;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double v = s[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    S.force[6 * b + k] = ((red[0][k] + red[1][k]) + (red[2][k] + red[3][k])) + S.fext[6 * b + k];
  }
}

// Body rows of the BDF residual and their badness max(|e_x| / (DL TOL), |e_u| / TOL).
__global__ void swim_errors_kernel(SwimK S, int nt, double* __restrict__ err, double* __restrict__ bad) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= S.nbody) return;
  const V3 x = ldv(S.x + 3 * b), xm1 = ldv(S.xm1 + 3 * b), xm2 = ldv(S.xm2 + 3 * b);
  const V3 u = ldv(S.u + 3 * b), um1 = ldv(S.um1 + 3 * b);
  const V3 V = ldv(S.vel + 6 * b), W = dexpinv(u, ldv(S.vel + 6 * b + 3));
  V3 ex, eu;
  if (nt < S.num_euler) {
    ex = x - xm1 - S.dt * V;
    eu = u - S.dt * W;
  } else {
    ex = x + (-4.0 / 3.0) * xm1 + (1.0 / 3.0) * xm2 + (-2.0 / 3.0 * S.dt) * V;
    eu = u - (1.0 / 3.0) * um1 + (-2.0 / 3.0 * S.dt) * W;
  }
  stv(err + 6 * b, ex);
  stv(err + 6 * b + 3, eu);
  double v = fmax(fmax(fabs(ex.x), fmax(fabs(ex.y), fabs(ex.z))) / (S.dl * S.tol),
                  fmax(fabs(eu.x), fmax(fabs(eu.y), fabs(eu.z))) / S.tol);
  if (!(ex.x == ex.x && ex.y == ex.y && ex.z == ex.z && eu.x == eu.x && eu.y == eu.y && eu.z == eu.z))
    v = __builtin_nan("");
  bad[b] = v;
}

// Rotation part of a position row's dependence on the body: rcross((DL/2 + |d|) d / |d|)
__device__ __forceinline__ M3 arm_block(const SwimK& S, V3 d) {
  const double dn = sqrt(dot(d, d));
  return rcross(((0.5 * S.dl + dn) / dn) * d);
}

__device__ __forceinline__ V3 base_disp(const SwimK& S, int f) {
  const int b = f / S.fpb;
  return mv(ldm(S.Qinit + 9 * b), ldv(S.fref + 3 * f));
}

// Body-mobility block (bi, bj) rotated by Q_init of body b
__device__ __forceinline__ M3 rot_mob(const SwimK& S, int b, int bi, int bj) {
  const M3 Q = ldm(S.Qinit + 9 * b);
  return mm(mm(Q, blk6(S.mob, bi, bj)), transpose(Q));
}

// B J^-1 of filament f (swimmer.cpp:744-886, with the sphere correction when the body
// is a sphere) and its product with the filament's body columns A (:888-902).  One
// workgroup per filament.
// type0 (blob forces among the Broyden unknowns, swimmer.cpp:624-668): the body rows are
// the force and torque balance, whose block against the filament unknowns is D = -B with
// B the raw block below (no body mobility, time-step factor or sphere correction).
__global__ __launch_bounds__(256) void swim_bj_kernel(SwimK S, FilK K, int nt, int type0) {
  extern __shared__ double lds[];
  const int f = blockIdx.x;
  const int n = K.nseg, N = 6 * n;
  const int s0 = f * n, b = f / S.fpb;
  double* B = lds;          // 6 x N
  double* BJ = lds + 6 * N; // 6 x N
  double* C = lds + 12 * N; // 6 x N: sphere-segment coupling added to N B
  const int tid = threadIdx.x;
  const double dt_fac = nt < S.num_euler ? -S.dt : -2.0 * S.dt / 3.0;
  const V3 ub = ldv(S.u + 3 * b);
  const M3 blc = mscale(rcross(ub), -1.0);
  const M3 blc2 = mm(blc, blc);
  const M3 Dinv = madd(madd(eye3(), blc2, 1.0 / 12.0), blc, -0.5);
  const M3 DinvT = madd(madd(eye3(), blc2, 1.0 / 12.0), blc, 0.5);
  const V3 d = base_disp(S, f);
  for (int e = tid; e < 6 * N; e += 256) {
    B[e] = 0.0;
    C[e] = 0.0;
  }
  __syncthreads();
  if (tid == 1 && S.sphere_R > 0.0 && !type0) {
    // RPY of the sphere (radius R) and segment m (radius a): translation-translation,
    // translation-rotation and rotation-rotation blocks (swimmer.cpp:809-877)
    auto add = [&](int r0, int c0, const M3& M) {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[(r0 + i) * N + c0 + j] += M.a[i][j];
    };
    const double pi = 3.14159265358979323846, R2a2 = S.sphere_R * S.sphere_R + S.a_seg * S.a_seg;
    for (int m = 0; m < n; ++m) {
      const V3 rv = v3(K.x[s0 + m], K.y[s0 + m], K.z[s0 + m]) - ldv(S.x + 3 * b);
      const double r = sqrt(dot(rv, rv));
      const V3 e = (1.0 / r) * rv;
      M3 rr;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) rr.a[i][j] = (&e.x)[i] * (&e.x)[j];
      const M3 Mtt = mscale(madd(mscale(eye3(), 1.0 + R2a2 / (3.0 * r * r)), rr, 1.0 - R2a2 / (r * r)),
                            1.0 / (8.0 * pi * S.mu * r));
      const M3 Mtr = mscale(rcross(e), 1.0 / (8.0 * pi * S.mu * r * r));
      const M3 Mrr = mscale(madd(mscale(rr, 3.0), eye3(), -1.0), 1.0 / (16.0 * pi * S.mu * r * r * r));
      const M3 tc = rcross(tangent(segq(K, s0 + m)));
      const M3 ta = mscale(mm(Mtr, tc), 0.5 * S.dl), tb = mscale(mm(Mrr, tc), 0.5 * S.dl);
      if (m == 0) {
        add(0, 0, Mtt);  // tether constraint force
        add(3, 0, Mtr);
        add(0, 3 * n, mm(Mtr, DinvT));  // clamp constraint
        add(3, 3 * n, mm(Mrr, DinvT));
        add(0, 3, madd(ta, Mtt, -1.0));  // inextensibility force of segment 1
        add(3, 3, madd(tb, Mtr, -1.0));
      } else if (m == n - 1) {
        add(0, N - 3, mscale(mm(Mtt, tc), -S.end_force));  // tip Lie algebra element
        add(3, N - 3, mscale(mm(Mtr, tc), -S.end_force));
        add(0, 3 * n - 3, madd(ta, Mtt));
        add(3, 3 * n - 3, madd(tb, Mtr));
      } else {
        add(0, 3 * m, madd(ta, Mtt));
        add(3, 3 * m, madd(tb, Mtr));
        add(0, 3 * (m + 1), madd(ta, Mtt, -1.0));
        add(3, 3 * (m + 1), madd(tb, Mtr, -1.0));
      }
    }
  }
  if (tid == 0) {
    auto put = [&](int r0, int c0, const M3& M) {
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) B[(r0 + i) * N + c0 + j] = M.a[i][j];
    };
    put(0, 0, mscale(eye3(), -1.0));
    put(3, 0, rcross(d));
    put(3, 3 * n, mscale(DinvT, -1.0));
    const V3 tl = tangent(segq(K, s0 + n - 1));
    const M3 E = rcross(S.end_force * tl);
    put(0, N - 3, E);
    const V3 diff = v3(K.x[s0 + n - 1], K.y[s0 + n - 1], K.z[s0 + n - 1]) - ldv(S.x + 3 * b);
    put(3, N - 3, mm(rcross((0.5 * S.dl) * tl - diff), E));
    for (int m = 1; m < n - 1; ++m) put(3, 3 * (n + m), mm(E, rcross(S.dl * tangent(segq(K, s0 + m)))));
  }
  __syncthreads();
  if (type0) {
    for (int e = tid; e < 6 * N; e += 256) B[e] = -B[e];
  } else if (tid < N) {  // B <- dt_fac diag(I, Dinv) N B, column by column
    const int c = tid;
    double col[6], out[6];
    for (int r = 0; r < 6; ++r) col[r] = B[r * N + c];
    for (int bi = 0; bi < 6; bi += 3) {
      V3 acc = v3(0.0, 0.0, 0.0);
      for (int bj = 0; bj < 6; bj += 3) acc = acc + mv(rot_mob(S, b, bi, bj), v3(col[bj], col[bj + 1], col[bj + 2]));
      acc = acc + v3(C[bi * N + c], C[(bi + 1) * N + c], C[(bi + 2) * N + c]);
      if (bi == 3) acc = mv(Dinv, acc);
      out[bi] = dt_fac * acc.x;
      out[bi + 1] = dt_fac * acc.y;
      out[bi + 2] = dt_fac * acc.z;
    }
    for (int r = 0; r < 6; ++r) B[r * N + c] = out[r];
  }
  __syncthreads();
  // BJ = B J^-1 (J^-1 stored column-major: element (r, c) at c N + r)
  const double* Ji = K.jinv + (size_t)f * N * N;
  for (int e = tid; e < 6 * N; e += 256) {
    const int k = e / N, c = e - k * N;
    double s = 0.0;
    for (int r = 0; r < N; ++r) s += B[k * N + r] * Ji[(size_t)c * N + r];
    BJ[e] = s;
    S.BJ[(size_t)f * 6 * N + e] = s;
  }
  __syncthreads();
  // S_f = BJ A: position rows take I (translation) and the arm block (rotation); the
  // Lie rows of segments 0 and 1 take the clamping blocks (rotation)
  if (tid < 36) {
    const int k = tid / 6, j = tid - 6 * k;
    const double* row = BJ + k * N;
    double s = 0.0;
    if (j < 3) {
      for (int m = 0; m < n; ++m) s += row[3 * m + j];
    } else {
      const int jj = j - 3;
      const M3 Dm = arm_block(S, d);
      for (int m = 0; m < n; ++m)
        for (int a = 0; a < 3; ++a) s += row[3 * m + a] * Dm.a[a][jj];
      const double* cb = K.clampblk + (size_t)f * 18;
      for (int a = 0; a < 3; ++a) s += row[3 * n + a] * cb[3 * a + jj] + row[3 * n + 3 + a] * cb[9 + 3 * a + jj];
    }
    S.Sf[36 * (size_t)f + tid] = s;
  }
}

// In-place Gauss-Jordan with partial pivoting of a row-major 6 x 6: I6 <- A^-1; false
// when a pivot vanishes or is not finite.
__device__ bool invert6(double* A, double* I6) {
  for (int i = 0; i < 36; ++i) I6[i] = (i % 7 == 0) ? 1.0 : 0.0;
  for (int c = 0; c < 6; ++c) {
    int p = c;
    for (int r = c + 1; r < 6; ++r)
      if (fabs(A[r * 6 + c]) > fabs(A[p * 6 + c])) p = r;
    const double piv = A[p * 6 + c];
    if (!(fabs(piv) > 0.0) || piv != piv) return false;
    if (p != c)
      for (int j = 0; j < 6; ++j) {
        double t = A[p * 6 + j];
        A[p * 6 + j] = A[c * 6 + j];
        A[c * 6 + j] = t;
        t = I6[p * 6 + j];
        I6[p * 6 + j] = I6[c * 6 + j];
        I6[c * 6 + j] = t;
      }
    const double ip = 1.0 / piv;
    for (int j = 0; j < 6; ++j) {
      A[c * 6 + j] *= ip;
      I6[c * 6 + j] *= ip;
    }
    for (int r = 0; r < 6; ++r) {
      if (r == c) continue;
      const double g = A[r * 6 + c];
      for (int j = 0; j < 6; ++j) {
        A[r * 6 + j] -= g * A[c * 6 + j];
        I6[r * 6 + j] -= g * I6[c * 6 + j];
      }
    }
  }
  return true;
}

// Body block of J0 and its Schur complement (swimmer.cpp:690-742, 902-909), inverted.
__global__ void swim_schur_kernel(SwimK S, FilK K, int nt, int* __restrict__ status) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= S.nbody) return;
  const int n = K.nseg;
  const size_t ns = K.nseg_total;
  const double dt_fac = nt < S.num_euler ? -S.dt : -2.0 * S.dt / 3.0;
  const V3 u = ldv(S.u + 3 * b);
  const M3 blc = mscale(rcross(u), -1.0);
  const M3 cbo = mscale(blc, -1.0 / dt_fac);
  double A[36];
  for (int i = 0; i < 36; ++i) A[i] = (i % 7 == 0) ? 1.0 : 0.0;
  M3 T = mscale(madd(mscale(mm(blc, cbo), 1.0 / 12.0), cbo, -0.5), dt_fac);
  M3 old = zero3();
  for (int f = b * S.fpb; f < (b + 1) * S.fpb; ++f) {
    const V3 d = base_disp(S, f);
    const int s0 = f * n;
    const V3 lt = {K.lam[s0], K.lam[ns + s0], K.lam[2 * ns + s0]};
    const V3 cl = {K.clamp[f], K.clamp[K.nfil + f], K.clamp[2 * K.nfil + f]};
    old = madd(old, mm(rcross(lt), rcross(d)), -1.0);
    old = madd(old, rcross(cl), 0.5);
    old = madd(old, madd(rcross(cross(u, cl)), mm(blc, rcross(cl))), -1.0 / 12.0);
    const V3 tl = tangent(segq(K, s0 + n - 1));
    old = madd(old, mm(rcross(S.end_force * tl), arm_block(S, d)));
  }
  old = mm(rot_mob(S, b, 3, 3), old);
  T = madd(T, mscale(madd(madd(old, mm(blc, old), -0.5), mm(mm(blc, blc), old), 1.0 / 12.0), dt_fac));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[(3 + i) * 6 + 3 + j] += T.a[i][j];
  for (int i = 0; i < 36; ++i) S.S0[36 * b + i] = A[i];
  for (int f = b * S.fpb; f < (b + 1) * S.fpb; ++f)
    for (int i = 0; i < 36; ++i) A[i] -= S.Sf[36 * (size_t)f + i];
  double I6[36];
  if (!invert6(A, I6)) status[0] = 1;
  for (int i = 0; i < 36; ++i) S.Sinv[36 * b + i] = I6[i];
}

// Body block of J0^-1 in: t = S^-1 (in_b - sum_f BJ_f in_f); out_b = scale t.  One
// workgroup per body.
__global__ __launch_bounds__(256) void swim_b2_kernel(SwimK S, int N, const double* __restrict__ in,
                                                      double* __restrict__ out, double scale) {
  __shared__ double red[4][6];
  const int b = blockIdx.x;
  const size_t L0 = (size_t)S.nfil * N;
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const size_t e0 = (size_t)b * S.fpb * N, e1 = e0 + (size_t)S.fpb * N;
  for (size_t e = e0 + threadIdx.x; e < e1; e += 256) {
    const size_t f = e / N, c = e - f * N;
    const double* bj = S.BJ + f * 6 * N + c;
    const double v = in[e];
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] += bj[(size_t)k * N] * v;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double v = s[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double b2[6];
    for (int k = 0; k < 6; ++k)
      b2[k] = in[L0 + 6 * b + k] - ((red[0][k] + red[1][k]) + (red[2][k] + red[3][k]));
    for (int k = 0; k < 6; ++k) {
      double t = 0.0;
      for (int j = 0; j < 6; ++j) t += S.Sinv[36 * b + 6 * k + j] * b2[j];
      S.tb[6 * b + k] = t;
      out[L0 + 6 * b + k] = scale * t;
    }
  }
}

// tmp_f = in_f - A_f t_b for every filament row (then J_f^-1 tmp_f).  One thread per segment.
__global__ void swim_rhs_kernel(SwimK S, FilK K, const double* __restrict__ in, double* __restrict__ tmp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= (int)K.nseg_total) return;
  const int n = K.nseg, f = s / n, m = s - f * n, b = f / S.fpb;
  const size_t base = (size_t)f * 6 * n;
  const V3 tv = ldv(S.tb + 6 * b), tw = ldv(S.tb + 6 * b + 3);
  const V3 px = tv + mv(arm_block(S, base_disp(S, f)), tw);
  const size_t ip = base + 3 * m;
  tmp[ip] = in[ip] - px.x;
  tmp[ip + 1] = in[ip + 1] - px.y;
  tmp[ip + 2] = in[ip + 2] - px.z;
  const size_t il = base + 3 * n + 3 * m;
  V3 pl = v3(0.0, 0.0, 0.0);
  if (m <= 1) pl = mv(ldm(K.clampblk + (size_t)f * 18 + 9 * m), tw);
  tmp[il] = in[il] - pl.x;
  tmp[il + 1] = in[il + 1] - pl.y;
  tmp[il + 2] = in[il + 2] - pl.z;
}


// ---- blob forces among the Broyden unknowns (the reference's SOLVER_TYPE 0) ----------

// Offset of blob j from its body's centre with the orientation the step started from,
// Q_init ref (swimmer.cpp:582, 951).
__device__ __forceinline__ V3 bx_disp0(const BxK& X, const SwimK& S, int j) {
  const int b = j / X.nb, k = j - b * X.nb;
  return mv(ldm(S.Qinit + 9 * b), ldv(X.ref + 3 * k));
}

// rigid_body::initial_guess, blob-force branch (rigid_body.cpp:459-490).
__global__ void bx_guess_kernel(BxK X, int nt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * X.m || nt < 1) return;
  const double f = X.fb[e], f1 = X.fb_m1[e], f2 = X.fb_m2[e];
  if (nt == 1) {
    X.fb_m1[e] = f;
  } else if (nt == 2) {
    X.fb_m2[e] = f1;
    X.fb_m1[e] = f;
    X.fb[e] = 2.0 * f - f1;
  } else {
    X.fb_m2[e] = f1;
    X.fb_m1[e] = f;
    X.fb[e] = 3.0 * (f - f1) + f2;
  }
}

// blob forces += their block of a Broyden increment (3 per blob) (rigid_body.cpp:521-527)
__global__ void bx_update_kernel(BxK X, const double* __restrict__ upd) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * X.m) return;
  const int k = e / X.m, j = e - k * X.m;
  X.fb[e] += upd[3 * (size_t)j + k];
}

// Six per-thread sums reduced over a 256-thread workgroup in a fixed order.
__device__ __forceinline__ void bx_reduce6(double* s, double (*red)[6]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double v = s[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wave][k] = v;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 6; ++k) s[k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
}

// Body rows: -(body load) + sum of the blob forces and of their torques about the centre
// with the current orientation (mobility_solver.cpp:2770-2781, 2930-2941); badness
// max |e| / TOL.  One workgroup per body.
__global__ __launch_bounds__(256) void bx_body_err_kernel(BxK X, SwimK S, const double* __restrict__ boff,
                                                          double* __restrict__ err, double* __restrict__ bad) {
  __shared__ double red[4][6];
  const int b = blockIdx.x;
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int j = b * X.nb + threadIdx.x; j < (b + 1) * X.nb; j += 256) {
    const V3 F = {X.fb[j], X.fb[X.m + j], X.fb[2 * (size_t)X.m + j]};
    const V3 d = {boff[j], boff[X.m + j], boff[2 * (size_t)X.m + j]};
    const V3 t = cross(d, F);
    s[0] += F.x;
    s[1] += F.y;
    s[2] += F.z;
    s[3] += t.x;
    s[4] += t.y;
    s[5] += t.z;
  }
  bx_reduce6(s, red);
  if (threadIdx.x == 0) {
    double v = 0.0;
    bool nan = false;
    for (int k = 0; k < 6; ++k) {
      const double e = s[k] - S.force[6 * b + k];
      err[6 * b + k] = e;
      v = fmax(v, fabs(e) / X.tol);
      nan = nan || e != e;
    }
    bad[b] = nan ? __builtin_nan("") : v;
  }
}

// Blob rows: V + Omega x d - v_blob, V and Omega from the body's BDF2 (Euler) relation
// (mobility_solver.cpp:2873-2922); -v_blob for a body held fixed.  Badness per
// workgroup, max |e| / TOL.
__global__ __launch_bounds__(256) void bx_blob_err_kernel(BxK X, SwimK S, int nt, const double* __restrict__ v,
                                                          size_t ld, int off, const double* __restrict__ boff,
                                                          double* __restrict__ err, double* __restrict__ bad) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  double bj = 0.0;
  if (j < X.m) {
    V3 e = {-v[off + j], -v[ld + off + j], -v[2 * ld + off + j]};
    if (X.nbody > 0) {
      const int b = j / X.nb;
      const V3 x = ldv(S.x + 3 * b), xm1 = ldv(S.xm1 + 3 * b), xm2 = ldv(S.xm2 + 3 * b);
      const V3 u = ldv(S.u + 3 * b), um1 = ldv(S.um1 + 3 * b);
      V3 Vb, Om;
      if (nt < X.num_euler) {
        Vb = (1.0 / X.dt) * (x - xm1);
        Om = (1.0 / X.dt) * u;
      } else {
        const double c1 = -4.0 / 3.0, c2 = 1.0 / 3.0, c3 = -2.0 / 3.0;
        Vb = (-1.0 / (c3 * X.dt)) * (x + c1 * xm1 + c2 * xm2);
        Om = dexp(u, (1.0 / (c3 * X.dt)) * (c2 * um1 - u));
      }
      const V3 d = {boff[j], boff[X.m + j], boff[2 * (size_t)X.m + j]};
      e = e + Vb + cross(Om, d);
    }
    err[3 * (size_t)j] = e.x;
    err[3 * (size_t)j + 1] = e.y;
    err[3 * (size_t)j + 2] = e.z;
    bj = fmax(fabs(e.x), fmax(fabs(e.y), fabs(e.z))) / X.tol;
    if (!(e.x == e.x && e.y == e.y && e.z == e.z)) bj = __builtin_nan("");
  }
  __shared__ double red[256];
  red[threadIdx.x] = bj;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const double a = red[threadIdx.x], c = red[threadIdx.x + o];
      red[threadIdx.x] = (a != a || c != c) ? __builtin_nan("") : fmax(a, c);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) bad[blockIdx.x] = red[0];
}

// Body block of J0 with the blob forces among the unknowns and its Schur complement
// (swimmer.cpp:575-695), inverted.  One workgroup per body: the blob sums
// sum rcross(d), sum rcross(d)^2 and sum rcross(f) rcross(d) (d = Q_init ref, f the blob
// force) in a fixed order, then thread 0 adds the filaments' terms.
__global__ __launch_bounds__(256) void bx_schur_kernel(BxK X, SwimK S, FilK K, int nt, int* __restrict__ status) {
  __shared__ double red[4][27];
  const int b = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double acc[27];
  for (int i = 0; i < 27; ++i) acc[i] = 0.0;
  for (int j = b * X.nb + threadIdx.x; j < (b + 1) * X.nb; j += 256) {
    const M3 Rc = rcross(bx_disp0(X, S, j));
    const M3 Rc2 = mm(Rc, Rc);
    const M3 FR = mm(rcross(v3(X.fb[j], X.fb[X.m + j], X.fb[2 * (size_t)X.m + j])), Rc);
    for (int r = 0; r < 3; ++r)
      for (int q = 0; q < 3; ++q) {
        acc[3 * r + q] += Rc.a[r][q];
        acc[9 + 3 * r + q] += Rc2.a[r][q];
        acc[18 + 3 * r + q] += FR.a[r][q];
      }
  }
  for (int i = 0; i < 27; ++i) {
    double v = acc[i];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wave][i] = v;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int i = 0; i < 27; ++i) acc[i] = (red[0][i] + red[1][i]) + (red[2][i] + red[3][i]);
  const double bmf = (nt < X.num_euler ? -6.0 : -9.0) * 3.14159265358979323846 * X.mu * X.a / X.dt;
  double A[36];
  for (int i = 0; i < 36; ++i) A[i] = 0.0;
  for (int i = 0; i < 3; ++i) {
    A[i * 6 + i] = -X.nb * bmf;
    for (int j = 0; j < 3; ++j) {
      A[i * 6 + 3 + j] = -bmf * acc[3 * i + j];
      A[(3 + i) * 6 + j] = bmf * acc[3 * i + j];
      A[(3 + i) * 6 + 3 + j] = acc[18 + 3 * i + j] + bmf * acc[9 + 3 * i + j];
    }
  }
  const int n = K.nseg;
  const size_t ns = K.nseg_total;
  const V3 u = ldv(S.u + 3 * b);
  M3 T = zero3();
  for (int f = b * S.fpb; f < (b + 1) * S.fpb; ++f) {
    const V3 d = base_disp(S, f);
    const int s0 = f * n;
    const V3 lt = {K.lam[s0], K.lam[ns + s0], K.lam[2 * ns + s0]};
    const V3 cl = {K.clamp[f], K.clamp[K.nfil + f], K.clamp[2 * K.nfil + f]};
    T = madd(T, mm(rcross(lt), rcross(d)));
    T = madd(T, rcross(cl), 0.5);
    T = madd(T, madd(rcross(cross(u, cl)), mm(rcross(u), rcross(cl)), -1.0), 1.0 / 12.0);
    const V3 t0 = tangent(segq(K, s0)), tl = tangent(segq(K, s0 + n - 1));
    T = madd(T, mm(rcross((-S.end_force) * tl), rcross((0.5 * S.dl) * t0)));
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[(3 + i) * 6 + 3 + j] += T.a[i][j];
  for (int i = 0; i < 36; ++i) S.S0[36 * b + i] = A[i];
  for (int f = b * S.fpb; f < (b + 1) * S.fpb; ++f)
    for (int i = 0; i < 36; ++i) A[i] -= S.Sf[36 * (size_t)f + i];
  double I6[36];
  if (!invert6(A, I6)) status[0] = 1;
  for (int i = 0; i < 36; ++i) S.Sinv[36 * b + i] = I6[i];
}

// Body block of J0^-1 in with the blob rows eliminated (swimmer.cpp:946-971):
// t = S^-1 (in_b - sum_f D J_f^-1 in_f + 6 pi mu a (sum_j in_j, sum_j d_j x in_j));
// out_b = scale t.  One workgroup per body.
__global__ __launch_bounds__(256) void bx_b2_kernel(BxK X, SwimK S, int N, const double* __restrict__ in,
                                                    double* __restrict__ out, double scale) {
  __shared__ double red[4][6];
  const int b = blockIdx.x;
  const size_t L0 = (size_t)S.nfil * N, G0 = L0 + 6 * (size_t)S.nbody;
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const size_t e0 = (size_t)b * S.fpb * N, e1 = e0 + (size_t)S.fpb * N;
  for (size_t e = e0 + threadIdx.x; e < e1; e += 256) {
    const size_t f = e / N, c = e - f * N;
    const double* bj = S.BJ + f * 6 * N + c;
    const double v = in[e];
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] += bj[(size_t)k * N] * v;
  }
  bx_reduce6(s, red);
  double g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  __syncthreads();  // red is reused
  for (int j = b * X.nb + threadIdx.x; j < (b + 1) * X.nb; j += 256) {
    const V3 r = ldv(in + G0 + 3 * (size_t)j);
    const V3 t = cross(bx_disp0(X, S, j), r);
    g[0] += r.x;
    g[1] += r.y;
    g[2] += r.z;
    g[3] += t.x;
    g[4] += t.y;
    g[5] += t.z;
  }
  bx_reduce6(g, red);
  if (threadIdx.x == 0) {
    const double bmf0 = -6.0 * 3.14159265358979323846 * X.mu * X.a;
    double b3[6];
    for (int k = 0; k < 6; ++k) b3[k] = in[L0 + 6 * b + k] - s[k] - bmf0 * g[k];
    for (int k = 0; k < 6; ++k) {
      double t = 0.0;
      for (int j = 0; j < 6; ++j) t += S.Sinv[36 * b + 6 * k + j] * b3[j];
      S.tb[6 * b + k] = t;
      out[L0 + 6 * b + k] = scale * t;
    }
  }
}

// Blob block of J0^-1 in: -6 pi mu a (in_j - dt_fac (t_V + t_Omega x d_j)) for free
// bodies (swimmer.cpp:973-989), -6 pi mu a in_j with the body fixed (:940-944).
__global__ void bx_blob_out_kernel(BxK X, SwimK S, int nt, const double* __restrict__ in,
                                   double* __restrict__ out, double scale) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= X.m) return;
  V3 r = ldv(in + 3 * (size_t)j);
  if (X.nbody > 0) {
    const int b = j / X.nb;
    const double dtf = (nt < X.num_euler ? 1.0 : 1.5) / X.dt;
    r = r - dtf * (ldv(S.tb + 6 * b) + cross(ldv(S.tb + 6 * b + 3), bx_disp0(X, S, j)));
  }
  const double c = scale * (-6.0 * 3.14159265358979323846 * X.mu * X.a);
  stv(out + 3 * (size_t)j, c * r);
}

inline dim3 gs(int n, int b) { return dim3((unsigned)((n + b - 1) / b)); }

}  // namespace

Status swim_guess(cfm_ctx* c, const SwimK& S, int nt) {
  hipLaunchKernelGGL(swim_guess_kernel, gs(S.nbody, 64), dim3(64), 0, c->stream, S, nt);
  CFM_CHECK_LAUNCH();
  return {};
}

Status swim_update(cfm_ctx* c, const SwimK& S, const double* upd_body) {
  hipLaunchKernelGGL(swim_update_kernel, gs(S.nbody, 64), dim3(64), 0, c->stream, S, upd_body);
  CFM_CHECK_LAUNCH();
  return {};
}

Status swim_force(cfm_ctx* c, const SwimK& S, const FilK& K, const double* rep, const double* boff) {
  hipLaunchKernelGGL(swim_force_kernel, dim3(S.nbody), dim3(256), 0, c->stream, S, K, rep, boff,
                     c->p.n_blob);
  CFM_CHECK_LAUNCH();
  return {};
}

Status swim_errors(cfm_ctx* c, const SwimK& S, int nt, double* err_body, double* bad) {
  hipLaunchKernelGGL(swim_errors_kernel, gs(S.nbody, 64), dim3(64), 0, c->stream, S, nt, err_body, bad);
  CFM_CHECK_LAUNCH();
  return {};
}

Status swim_jacobian(cfm_ctx* c, const SwimK& S, const FilK& K, int nt, int* status_dev) {
  const int N = 6 * K.nseg;
  const size_t lds = sizeof(double) * 18 * N;
  hipLaunchKernelGGL(swim_bj_kernel, dim3(S.nfil), dim3(256), lds, c->stream, S, K, nt, 0);
  hipLaunchKernelGGL(swim_schur_kernel, gs(S.nbody, 64), dim3(64), 0, c->stream, S, K, nt, status_dev);
  CFM_CHECK_LAUNCH();
  return {};
}

Status swim_jinv_mult(cfm_ctx* c, const SwimK& S, const FilK& K, const double* in, double* out,
                      double scale, double* tmp) {
  const int N = 6 * K.nseg;
  hipLaunchKernelGGL(swim_b2_kernel, dim3(S.nbody), dim3(256), 0, c->stream, S, N, in, out, scale);
  hipLaunchKernelGGL(swim_rhs_kernel, gs((int)K.nseg_total, 256), dim3(256), 0, c->stream, S, K, in, tmp);
  CFM_CHECK_LAUNCH();
  return fil_jinv_mult(c, K, tmp, out, scale);
}


Status bx_guess(cfm_ctx* c, const BxK& X, int nt) {
  if (X.m == 0) return {};
  hipLaunchKernelGGL(bx_guess_kernel, gs(3 * X.m, 256), dim3(256), 0, c->stream, X, nt);
  CFM_CHECK_LAUNCH();
  return {};
}

Status bx_update(cfm_ctx* c, const BxK& X, const double* upd) {
  if (X.m == 0) return {};
  hipLaunchKernelGGL(bx_update_kernel, gs(3 * X.m, 256), dim3(256), 0, c->stream, X, upd);
  CFM_CHECK_LAUNCH();
  return {};
}

Status bx_errors(cfm_ctx* c, const BxK& X, const SwimK& S, int nt, double* err_body, double* err_blob,
                 double* bad) {
  const Particles& p = c->p;
  const double* boff = X.nbody > 0 ? body_offsets(c) : nullptr;
  if (X.nbody > 0) {
    if (!boff) return make_status(CFM_ESTATE, "no body placement for the blob rows");
    hipLaunchKernelGGL(bx_body_err_kernel, dim3(X.nbody), dim3(256), 0, c->stream, X, S, boff, err_body, bad);
  }
  hipLaunchKernelGGL(bx_blob_err_kernel, gs(X.m, 256), dim3(256), 0, c->stream, X, S, nt, (const double*)p.v.p,
                     (size_t)p.n, p.n_seg, boff, err_blob, bad + X.nbody);
  CFM_CHECK_LAUNCH();
  return {};
}

Status bx_jacobian(cfm_ctx* c, const BxK& X, const SwimK& S, const FilK& K, int nt, int* status_dev) {
  if (X.nbody == 0) return {};
  const int N = 6 * K.nseg;
  const size_t lds = sizeof(double) * 18 * N;
  hipLaunchKernelGGL(swim_bj_kernel, dim3(S.nfil), dim3(256), lds, c->stream, S, K, nt, 1);
  hipLaunchKernelGGL(bx_schur_kernel, dim3(S.nbody), dim3(256), 0, c->stream, X, S, K, nt, status_dev);
  CFM_CHECK_LAUNCH();
  return {};
}

Status bx_jinv_mult(cfm_ctx* c, const BxK& X, const SwimK& S, const FilK& K, int nt, const double* in,
                    double* out, double scale, double* tmp) {
  const size_t L0 = (size_t)K.nseg_total * 6;
  const size_t G0 = L0 + 6 * (size_t)X.nbody;
  if (X.nbody > 0) {
    const int N = 6 * K.nseg;
    hipLaunchKernelGGL(bx_b2_kernel, dim3(S.nbody), dim3(256), 0, c->stream, X, S, N, in, out, scale);
    hipLaunchKernelGGL(bx_blob_out_kernel, gs(X.m, 256), dim3(256), 0, c->stream, X, S, nt, in + G0, out + G0,
                       scale);
    hipLaunchKernelGGL(swim_rhs_kernel, gs((int)K.nseg_total, 256), dim3(256), 0, c->stream, S, K, in, tmp);
    CFM_CHECK_LAUNCH();
    return fil_jinv_mult(c, K, tmp, out, scale);
  }
  hipLaunchKernelGGL(bx_blob_out_kernel, gs(X.m, 256), dim3(256), 0, c->stream, X, S, nt, in + G0, out + G0,
                     scale);
  CFM_CHECK_LAUNCH();
  return fil_jinv_mult(c, K, in, out, scale);
}

}  // namespace cfm
