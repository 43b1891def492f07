// body.hip -- free rigid bodies made of the context's blobs.
//
// The reference's rigid-body blob system (USE_BROYDEN_FOR_EVERYTHING 0,
// PRESCRIBED_BODY_VELOCITIES 0): with n_body bodies of nb blobs each, the unknowns
// are the blob forces F (3 per blob) and the body velocities U_b = (V_b, Omega_b),
// and the system is the saddle-point problem
//     M_bb F - K U = rhs_blob,     -K^T F = rhs_body,
// (K U)_i = V_b + Omega_b x r_i with r_i = R(q_b) ref_i the rotated body-frame
// offset of blob i (mobility_solver.cpp:1822-1880 product, :455-632 right-hand side,
// rigid_body.cpp:497-515 placement).  Device layout of a system vector: the
// [3][n_blob] force block of the blob system, then 6 entries per body.
//
// Preconditioners (right, inside the library's GMRES):
//   CFM_PC_DIAG  the reference's block form (mobility_solver.cpp:1518-1590):
//                U = (K^T K)^-1 (-K^T f) - N u,  F = 6 pi mu a (f + K U), with the
//                body-frame (K^T K)^-1 and body mobility N rotated by R(q_b);
//   CFM_PC_SELF  the same with each blob's own wall-corrected self-block inverse;
//   CFM_PC_DENSE the exact Schur complement around the resident FP32 inverse P of
//                M_bb: U = S^-1 (-u - K^T P f), F = P f + (P K) U, S = K^T P K
//                (6 n_body square, inverted once per placement on the host).
// make_body_reference_matrices (mobility_solver.cpp:3049-3146) is
// cfm_body_reference_matrices: K^T K from the offsets, and N from six solves of one
// body alone with a unit force or torque.
//
// Prescribed rotation (BODY_VELOCITY_TYPE 2, cfm_body_set_motion): the reference drops
// the Omega columns from the blob rows and the torque rows from the product
// (mobility_solver.cpp:1947-1994, 2177-2181) and from the preconditioner (:1268-1291,
// 1506-1512).  Here the unknown vector keeps its 6 entries per body; the blob rows see
// only V_b, and the torque rows are identity rows on Omega_b, so the solve returns
// Omega_b = the torque entries of the right-hand side (zero, as the reference
// prescribes) and the blob forces and V_b of the reduced system [[M, -K_V], [-K_V^T, 0]].
// The preconditioners invert the same reduction: CFM_PC_DIAG / _SELF with
// (K_V^T K_V)^-1 = I / nb and the clamped translational mobility ((N^-1)_VV)^-1,
// CFM_PC_DENSE with the Schur complement's translation block (K_V^T P K_V)^-1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "body.hpp"
#include "ctx.hpp"
#include "densepc.hpp"
#include "krylov.hpp"
#include "precond.hpp"

namespace cfm {

Status upload_aos(cfm_ctx* c, const double* host, int cnt, int dof, double* dst, int ld, int off);

namespace {

constexpr int kMaxSchur = 768;  // 6 n_body limit of the dense Schur complement (128 bodies)

struct BodyState {
  int n_body = 0, nb = 0;   // bodies, blobs per body
  std::vector<double> ref;  // 3 nb body-frame offsets (AoS)
  std::vector<double> x, q; // centres 3 n_body, quaternions (w, x, y, z) 4 n_body
  bool placed = false;
  double ktk_inv[36];       // (K^T K)^-1 of one body, body frame, row-major
  double mob[36];           // body mobility, body frame (identity until made or set)
  DevBuf<double> r;         // rotated offsets [3][n_blob]
  DevBuf<double> mats;      // per body: rotated (K^T K)^-1 then rotated N (72, row-major)
  DevBuf<double> w, aux;    // scratch: 3 n_blob, 6 n_body
  DevBuf<double> kcol;      // one column of K (dense Schur build)
  DevBuf<double> PK;        // P K, 6 n_body columns of 3 n_blob
  DevBuf<double> Sinv;      // (K^T P K)^-1, row-major
  int schur_gen = -1;       // dense-inverse generation PK and Sinv belong to
  DevBuf<double> ref_dev;   // ref on the device (placement from device-held body states)
  DevBuf<double> frame;     // body-frame (K^T K)^-1 and N on the device (72)
  // Frozen dense preconditioner (free swimmers): built once and applied in the frame it
  // was built in, rotated by R(q) R(q_build)^T per body.
  bool frozen = false, built_frozen = false;
  DevBuf<double> q_cur, q_build;  // quaternions of the last placement / of the build (4 n_body)
  DevBuf<double> r_build;         // offsets at the build ([3][n_blob])
  DevBuf<double> rin, rout;       // rotated system vectors
  DevBuf<double> st;              // staging of an embedded blob block (3 n_blob + 6 n_body)
  std::vector<double> R0;         // K^T P K at the build, (6 n_body)^2 row-major
  // Prescribed rotation (the reference's BODY_VELOCITY_TYPE 2): Omega_b = 0 is not an
  // unknown; its columns leave the blob rows and each torque row reads Omega_b itself.
  bool rot_fixed = false;
};

BodyState* body_of(cfm_ctx* c) { return reinterpret_cast<BodyState*>(c->body); }

inline dim3 g1(int n) { return dim3((unsigned)((n + 255) / 256)); }

// Rotation matrix of a unit quaternion (w, x, y, z) (quaternion.cpp:305-347).
void rot_mat(const double* q, double R[9]) {
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

// Each 3 x 3 block of the 6 x 6 A conjugated: out_ij = R A_ij R^T (transpose: R^T A_ij R).
void rotate6(const double* A, const double R[9], bool transpose, double* out) {
  auto r = [&](int i, int j) { return transpose ? R[3 * j + i] : R[3 * i + j]; };
  for (int bi = 0; bi < 6; bi += 3)
    for (int bj = 0; bj < 6; bj += 3)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double s = 0.0;
          for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) s += r(i, k) * A[(bi + k) * 6 + bj + l] * r(j, l);
          out[(bi + i) * 6 + bj + j] = s;
        }
}

// In-place Gauss-Jordan inverse with partial pivoting of the n x n row-major A.
bool invert(std::vector<double>& A, int n) {
  std::vector<double> B((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) B[(size_t)i * n + i] = 1.0;
  for (int col = 0; col < n; ++col) {
    int piv = col;
    for (int i = col + 1; i < n; ++i)
      if (std::fabs(A[(size_t)i * n + col]) > std::fabs(A[(size_t)piv * n + col])) piv = i;
    const double d = A[(size_t)piv * n + col];
    if (!(std::fabs(d) > 0.0) || !std::isfinite(d)) return false;
    if (piv != col)
      for (int j = 0; j < n; ++j) {
        std::swap(A[(size_t)piv * n + j], A[(size_t)col * n + j]);
        std::swap(B[(size_t)piv * n + j], B[(size_t)col * n + j]);
      }
    const double inv = 1.0 / d;
    for (int j = 0; j < n; ++j) {
      A[(size_t)col * n + j] *= inv;
      B[(size_t)col * n + j] *= inv;
    }
    for (int i = 0; i < n; ++i) {
      if (i == col) continue;
      const double f = A[(size_t)i * n + col];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) {
        A[(size_t)i * n + j] -= f * A[(size_t)col * n + j];
        B[(size_t)i * n + j] -= f * B[(size_t)col * n + j];
      }
    }
  }
  A.swap(B);
  return true;
}

// K^T K of one body from its body-frame offsets (mobility_solver.cpp:3052-3070):
// rows/columns (V, Omega); with [r] the matrix of w -> w x r,
// K^T K = [[nb I, sum [r]], [-sum [r], -sum [r][r]]].
void ktk(const std::vector<double>& ref, double A[36]) {
  std::fill(A, A + 36, 0.0);
  const int nb = (int)(ref.size() / 3);
  A[0] = A[7] = A[14] = nb;
  for (int m = 0; m < nb; ++m) {
    const double* v = &ref[3 * m];
    const double C[9] = {0.0, v[2], -v[1], -v[2], 0.0, v[0], v[1], -v[0], 0.0};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double cc = 0.0;
        for (int k = 0; k < 3; ++k) cc += C[3 * i + k] * C[3 * k + j];
        A[i * 6 + 3 + j] += C[3 * i + j];
        A[(3 + i) * 6 + j] -= C[3 * i + j];
        A[(3 + i) * 6 + 3 + j] -= cc;
      }
  }
}

// ---------------------------------------------------------------- kernels

// y_F -= K U for every blob: V_b + Omega_b x r_i (mobility_solver.cpp:1838-1866, -K mult).
// rot: prescribed rotation, V_b only.
__global__ __launch_bounds__(256) void body_k_kernel(int m, int nb, int rot, const double* __restrict__ r,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double* U = x + 3 * (size_t)m + 6 * (size_t)(i / nb);
  if (rot) {
    y[i] -= U[0];
    y[m + i] -= U[1];
    y[2 * (size_t)m + i] -= U[2];
    return;
  }
  const double rx = r[i], ry = r[m + i], rz = r[2 * (size_t)m + i];
  y[i] -= U[0] + (U[4] * rz - U[5] * ry);
  y[m + i] -= U[1] + (U[5] * rx - U[3] * rz);
  y[2 * (size_t)m + i] -= U[2] + (U[3] * ry - U[4] * rx);
}

// One workgroup per body: t = K^T F over its blobs (total force, total torque about the
// centre), summed in a fixed order, then
//   mode 0: outU = -t                                   (system product, -K^T mult)
//   mode 1: outU = Kinv_b (-t) - N_b inU                (block preconditioner)
//   mode 2: outU = -inU - t                             (Schur right-hand side)
// With rot (prescribed rotation) the torque rows of modes 0 and 2 are inU_Omega and
// -inU_Omega (mode 1 takes the reduced matrices, whose Omega rows give inU_Omega).
__global__ __launch_bounds__(256) void body_kt_kernel(int mode, int rot, int m, int nb,
                                                      const double* __restrict__ r,
                                                      const double* __restrict__ F,
                                                      const double* __restrict__ inU,
                                                      const double* __restrict__ mats,
                                                      double* __restrict__ outU) {
  __shared__ double red[4][6];
  const int b = blockIdx.x;
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int i1 = (b + 1) * nb;
  for (int i = b * nb + threadIdx.x; i < i1; i += 256) {
    const double fx = F[i], fy = F[m + i], fz = F[2 * (size_t)m + i];
    const double rx = r[i], ry = r[m + i], rz = r[2 * (size_t)m + i];
    s[0] += fx;
    s[1] += fy;
    s[2] += fz;
    s[3] += ry * fz - rz * fy;
    s[4] += rz * fx - rx * fz;
    s[5] += rx * fy - ry * fx;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double v = s[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x >= 6) return;
  const int k = threadIdx.x;
  double t[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) t[j] = (red[0][j] + red[1][j]) + (red[2][j] + red[3][j]);
  if (mode == 0) {
    outU[6 * b + k] = (rot && k >= 3) ? inU[6 * b + k] : -t[k];
  } else if (mode == 1) {
    const double* Ki = mats + 72 * (size_t)b;
    const double* N = Ki + 36;
    double v = 0.0;
    for (int j = 0; j < 6; ++j) v -= Ki[6 * k + j] * t[j] + N[6 * k + j] * inU[6 * b + j];
    outU[6 * b + k] = v;
  } else {
    outU[6 * b + k] = (rot && k >= 3) ? -inU[6 * b + k] : -inU[6 * b + k] - t[k];
  }
}

// out_F = P_i (f_i + V_b + Omega_b x r_i), P_i the particle preconditioner of blob i
// (rot: P_i (f_i + V_b)).
__global__ __launch_bounds__(256) void body_pc_blob_kernel(int kind, int wall, int rot, int m, int nb,
                                                           const double* __restrict__ r,
                                                           const double* __restrict__ z,
                                                           double a, double mu,
                                                           const double* __restrict__ in,
                                                           double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double* U = out + 3 * (size_t)m + 6 * (size_t)(i / nb);
  const double rx = r[i], ry = r[m + i], rz = r[2 * (size_t)m + i];
  const double w0 = rot ? 0.0 : U[3], w1 = rot ? 0.0 : U[4], w2 = rot ? 0.0 : U[5];
  double v[3], o[3];
  v[0] = in[i] + U[0] + (w1 * rz - w2 * ry);
  v[1] = in[m + i] + U[1] + (w2 * rx - w0 * rz);
  v[2] = in[2 * (size_t)m + i] + U[2] + (w0 * ry - w1 * rx);
  particle_precond(kind, wall, 3, z[i], a, mu, v, o);
  out[i] = o[0];
  out[m + i] = o[1];
  out[2 * (size_t)m + i] = o[2];
}

// Column `col` of K as a blob-system vector: body col / 6, unknown col % 6.
__global__ __launch_bounds__(256) void body_kcol_kernel(int m, int nb, int col,
                                                        const double* __restrict__ r,
                                                        double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double v[3] = {0.0, 0.0, 0.0};
  if (i / nb == col / 6) {
    const int k = col % 6;
    if (k < 3) {
      v[k] = 1.0;
    } else {  // e_{k-3} x r_i
      const double rr[3] = {r[i], r[m + i], r[2 * (size_t)m + i]};
      const int a1 = (k - 3 + 1) % 3, a2 = (k - 3 + 2) % 3;
      v[a2] = rr[a1];
      v[a1] = -rr[a2];
    }
  }
  out[i] = v[0];
  out[m + i] = v[1];
  out[2 * (size_t)m + i] = v[2];
}

// out[row] = sum_j S[row n + j] v[j] (n = 6 n_body).
__global__ __launch_bounds__(256) void body_small_gemv_kernel(int n, const double* __restrict__ S,
                                                              const double* __restrict__ v,
                                                              double* __restrict__ out) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  double s = 0.0;
  for (int j = 0; j < n; ++j) s += S[(size_t)row * n + j] * v[j];
  out[row] = s;
}

// out_F[j] = w[j] + sum_c PK[c][j] U[c] (rot: over the V columns only).
__global__ __launch_bounds__(256) void body_dense_finish_kernel(int n3, int ncol, int rot,
                                                                const double* __restrict__ w,
                                                                const double* __restrict__ PK,
                                                                const double* __restrict__ U,
                                                                double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n3) return;
  double s = w[j];
  for (int cc = 0; cc < ncol; ++cc)
    if (!rot || cc % 6 < 3) s += PK[(size_t)cc * n3 + j] * U[cc];
  out[j] = s;
}

// dst[k ld + j] += src[k m + j] for the three components of every blob.
__global__ __launch_bounds__(256) void body_add_rows_kernel(int m, const double* __restrict__ src,
                                                            size_t ld, double* __restrict__ dst) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  for (int k = 0; k < 3; ++k) dst[(size_t)k * ld + j] += src[(size_t)k * m + j];
}

__device__ __forceinline__ void quat_rot(const double* q, double R[9]) {
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

// out = G in (transpose: G^T in) for a body-system vector, G = R(q_b) R(q0_b)^T on every
// blob 3-vector and on both 3-vectors (V, Omega) of every body.
__global__ __launch_bounds__(256) void body_rotate_kernel(int m, int nb, int nbody,
                                                          const double* __restrict__ qc,
                                                          const double* __restrict__ q0, int transpose,
                                                          const double* __restrict__ in,
                                                          double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m + 2 * nbody) return;
  int b;
  size_t i0, st;
  if (t < m) {
    b = t / nb;
    i0 = t;
    st = m;
  } else {
    const int k = t - m;
    b = k / 2;
    i0 = 3 * (size_t)m + 6 * b + 3 * (k % 2);
    st = 1;
  }
  double Rc[9], R0[9], G[9];
  quat_rot(qc + 4 * b, Rc);
  quat_rot(q0 + 4 * b, R0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      G[3 * i + j] = Rc[3 * i] * R0[3 * j] + Rc[3 * i + 1] * R0[3 * j + 1] + Rc[3 * i + 2] * R0[3 * j + 2];
  const double v0 = in[i0], v1 = in[i0 + st], v2 = in[i0 + 2 * st];
  for (int i = 0; i < 3; ++i) {
    const double o = transpose ? G[i] * v0 + G[3 + i] * v1 + G[6 + i] * v2
                               : G[3 * i] * v0 + G[3 * i + 1] * v1 + G[3 * i + 2] * v2;
    out[i0 + i * st] = o;
  }
}

// Blob positions and rotated offsets from body centres / orientations on the device.
__global__ __launch_bounds__(256) void body_place_kernel(int m, int nb, const double* __restrict__ ref,
                                                         const double* __restrict__ bx,
                                                         const double* __restrict__ bq,
                                                         double* __restrict__ r, double* __restrict__ px,
                                                         double* __restrict__ py, double* __restrict__ pz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int b = i / nb, k = i - b * nb;
  const double s = bq[4 * b], a = bq[4 * b + 1], bb = bq[4 * b + 2], cc = bq[4 * b + 3];
  const double R[9] = {1.0 - 2.0 * (bb * bb + cc * cc), 2.0 * (a * bb - s * cc), 2.0 * (a * cc + s * bb),
                       2.0 * (a * bb + s * cc), 1.0 - 2.0 * (a * a + cc * cc), 2.0 * (bb * cc - s * a),
                       2.0 * (a * cc - s * bb), 2.0 * (bb * cc + s * a), 1.0 - 2.0 * (a * a + bb * bb)};
  const double v0 = ref[3 * k], v1 = ref[3 * k + 1], v2 = ref[3 * k + 2];
  const double o0 = R[0] * v0 + R[1] * v1 + R[2] * v2;
  const double o1 = R[3] * v0 + R[4] * v1 + R[5] * v2;
  const double o2 = R[6] * v0 + R[7] * v1 + R[8] * v2;
  r[i] = o0;
  r[m + i] = o1;
  r[2 * (size_t)m + i] = o2;
  px[i] = bx[3 * b] + o0;
  py[i] = bx[3 * b + 1] + o1;
  pz[i] = bx[3 * b + 2] + o2;
}

// Per body: each 3 x 3 block of the body-frame matrices conjugated by R(q_b).
__global__ void body_mats_kernel(int nbody, const double* __restrict__ frame,
                                 const double* __restrict__ bq, double* __restrict__ mats) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbody * 72) return;
  const int b = t / 72, e = t - 72 * b, mat = e / 36, ij = e - 36 * mat;
  const int i = ij / 6, j = ij - 6 * i, bi = i / 3 * 3, bj = j / 3 * 3;
  const double s = bq[4 * b], a = bq[4 * b + 1], bb = bq[4 * b + 2], cc = bq[4 * b + 3];
  const double R[9] = {1.0 - 2.0 * (bb * bb + cc * cc), 2.0 * (a * bb - s * cc), 2.0 * (a * cc + s * bb),
                       2.0 * (a * bb + s * cc), 1.0 - 2.0 * (a * a + cc * cc), 2.0 * (bb * cc - s * a),
                       2.0 * (a * cc - s * bb), 2.0 * (bb * cc + s * a), 1.0 - 2.0 * (a * a + bb * bb)};
  const double* A = frame + 36 * mat;
  double v = 0.0;
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l < 3; ++l) v += R[3 * (i - bi) + k] * A[(bi + k) * 6 + bj + l] * R[3 * (j - bj) + l];
  mats[72 * (size_t)b + e] = v;
}

// Body-frame matrices of the block preconditioner: (K^T K)^-1 and N, or with the
// rotation prescribed, diag(I / nb, 0) and diag(((N^-1)_VV)^-1, -I) (mode 1 of
// body_kt_kernel then returns V from the translation rows and Omega = inU_Omega).
Status frame_mats(const BodyState* S, double fr[72]) {
  if (!S->rot_fixed) {
    std::copy(S->ktk_inv, S->ktk_inv + 36, fr);
    std::copy(S->mob, S->mob + 36, fr + 36);
    return {};
  }
  std::fill(fr, fr + 72, 0.0);
  for (int k = 0; k < 3; ++k) fr[7 * k] = 1.0 / S->nb;
  std::vector<double> Ni(S->mob, S->mob + 36);
  if (!invert(Ni, 6)) return make_status(CFM_ENONFINITE, "singular body mobility");
  std::vector<double> Nv(9);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Nv[3 * i + j] = Ni[6 * i + j];
  if (!invert(Nv, 3)) return make_status(CFM_ENONFINITE, "singular translational resistance");
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) fr[36 + 6 * i + j] = Nv[3 * i + j];
    fr[36 + 6 * (3 + i) + 3 + i] = -1.0;
  }
  return {};
}

Status upload_mats(cfm_ctx* c, BodyState* S) {
  double fr[72];
  CFM_TRY(frame_mats(S, fr));
  std::vector<double> h((size_t)72 * S->n_body);
  for (int b = 0; b < S->n_body; ++b) {
    double R[9];
    rot_mat(&S->q[4 * b], R);
    rotate6(fr, R, false, &h[72 * (size_t)b]);
    rotate6(fr + 36, R, false, &h[72 * (size_t)b + 36]);
  }
  CFM_TRY(S->mats.alloc(h.size()));
  CFM_HIP(hipMemcpyAsync(S->mats.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream));
  CFM_TRY(S->frame.alloc(72));
  CFM_HIP(hipMemcpyAsync(S->frame.p, fr, sizeof(fr), hipMemcpyHostToDevice, c->stream));
  CFM_HIP(hipStreamSynchronize(c->stream));
  return {};
}

// P K and (K^T P K)^-1 for the current placement and resident inverse.
Status build_schur(cfm_ctx* c, BodyState* S) {
  const int m = c->p.n_blob, n3 = 3 * m, nc = 6 * S->n_body;
  if (nc > kMaxSchur)
    return make_status(CFM_EINVAL, "CFM_PC_DENSE on the body system takes at most %d bodies", kMaxSchur / 6);
  if (!c->dinv.valid) CFM_TRY(dense_build(c, c->dinv));
  CFM_TRY(S->kcol.alloc(n3));
  CFM_TRY(S->PK.alloc((size_t)nc * n3));
  CFM_TRY(S->Sinv.alloc((size_t)nc * nc));
  DevBuf<double> St;
  struct Release {
    DevBuf<double>& b;
    ~Release() { b.release(); }
  } rel{St};
  CFM_TRY(St.alloc((size_t)nc * nc));
  for (int col = 0; col < nc; ++col) {
    hipLaunchKernelGGL(body_kcol_kernel, g1(m), dim3(256), 0, c->stream, m, S->nb, col, S->r.p, S->kcol.p);
    CFM_CHECK_LAUNCH();
    CFM_TRY(dense_apply(c, c->dinv, S->kcol.p, S->PK.p + (size_t)col * n3));
    // column col of -K^T P K
    hipLaunchKernelGGL(body_kt_kernel, dim3(S->n_body), dim3(256), 0, c->stream, 0, 0, m, S->nb, S->r.p,
                       S->PK.p + (size_t)col * n3, (const double*)nullptr, (const double*)nullptr,
                       St.p + (size_t)col * nc);
    CFM_CHECK_LAUNCH();
  }
  std::vector<double> A((size_t)nc * nc);
  CFM_HIP(hipMemcpyAsync(A.data(), St.p, sizeof(double) * A.size(), hipMemcpyDeviceToHost, c->stream));
  CFM_HIP(hipStreamSynchronize(c->stream));
  // St holds column col at [col * nc]: A(row, col) = -St[col * nc + row]; symmetrise
  std::vector<double> Sm((size_t)nc * nc);
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nc; ++j)
      Sm[(size_t)i * nc + j] = -0.5 * (A[(size_t)j * nc + i] + A[(size_t)i * nc + j]);
  S->R0 = Sm;
  if (S->rot_fixed) {
    // translation block (K_V^T P K_V)^-1; Omega rows -I against the -inU_Omega of mode 2
    const int nv = 3 * S->n_body;
    std::vector<double> T((size_t)nv * nv);
    for (int i = 0; i < nv; ++i)
      for (int j = 0; j < nv; ++j) T[(size_t)i * nv + j] = Sm[(size_t)(6 * (i / 3) + i % 3) * nc + 6 * (j / 3) + j % 3];
    if (!invert(T, nv)) return make_status(CFM_ENONFINITE, "singular body Schur complement K_V^T P K_V");
    std::fill(Sm.begin(), Sm.end(), 0.0);
    for (int i = 0; i < nv; ++i)
      for (int j = 0; j < nv; ++j) Sm[(size_t)(6 * (i / 3) + i % 3) * nc + 6 * (j / 3) + j % 3] = T[(size_t)i * nv + j];
    for (int b = 0; b < S->n_body; ++b)
      for (int k = 3; k < 6; ++k) Sm[(size_t)(6 * b + k) * nc + 6 * b + k] = -1.0;
  } else if (!invert(Sm, nc)) {
    return make_status(CFM_ENONFINITE, "singular body Schur complement K^T P K");
  }
  CFM_HIP(hipMemcpyAsync(S->Sinv.p, Sm.data(), sizeof(double) * Sm.size(), hipMemcpyHostToDevice, c->stream));
  CFM_HIP(hipStreamSynchronize(c->stream));
  S->schur_gen = c->dinv.generation;
  return {};
}

void body_release_state(BodyState* S) {
  for (auto* b : {&S->r, &S->mats, &S->w, &S->aux, &S->kcol, &S->PK, &S->Sinv, &S->ref_dev, &S->frame,
                 &S->q_cur, &S->q_build, &S->r_build, &S->rin, &S->rout, &S->st})
    b->release();
}

}  // namespace

Status body_system_desc(cfm_ctx* c, SystemDesc* s) {
  BodyState* S = body_of(c);
  if (!S) return make_status(CFM_ESTATE, "cfm_body_init has not been called");
  if (!S->placed) return make_status(CFM_ESTATE, "body state not set (cfm_body_set_state)");
  *s = system_desc(c, CFM_SYS_BLOB);
  s->body = true;
  s->extra = 6 * S->n_body;
  s->L = 3 * s->m + s->extra;
  return {};
}

Status body_couple_apply(cfm_ctx* c, const double* x, double* y) {
  BodyState* S = body_of(c);
  if (!S || !S->placed) return make_status(CFM_ESTATE, "no placed bodies");
  const int m = c->p.n_blob, rot = S->rot_fixed;
  hipLaunchKernelGGL(body_kt_kernel, dim3(S->n_body), dim3(256), 0, c->stream, 0, rot, m, S->nb, S->r.p, x,
                     x + 3 * (size_t)m, (const double*)nullptr, y + 3 * (size_t)m);
  hipLaunchKernelGGL(body_k_kernel, g1(m), dim3(256), 0, c->stream, m, S->nb, rot, S->r.p, x, y);
  CFM_CHECK_LAUNCH();
  return {};
}

// The resident inverse, P K and (K^T P K)^-1 for the current placement -- or, frozen,
// those of the build, applied in its frame.
static Status ensure_dense(cfm_ctx* c, BodyState* S) {
  const int m = c->p.n_blob;
  const bool reuse = S->frozen && S->built_frozen;
  if (!reuse && (!c->dinv.valid || S->schur_gen != c->dinv.generation)) {
    CFM_TRY(build_schur(c, S));
    if (S->frozen) {  // remember the frame the preconditioner belongs to
      CFM_TRY(S->q_build.alloc(4 * (size_t)S->n_body));
      CFM_TRY(S->r_build.alloc(3 * (size_t)m));
      CFM_HIP(hipMemcpyAsync(S->q_build.p, S->q_cur.p, sizeof(double) * 4 * S->n_body,
                             hipMemcpyDeviceToDevice, c->stream));
      CFM_HIP(hipMemcpyAsync(S->r_build.p, S->r.p, sizeof(double) * 3 * m, hipMemcpyDeviceToDevice,
                             c->stream));
      S->built_frozen = true;
    }
  }
  return {};
}

Status body_dense_ready(cfm_ctx* c) {
  BodyState* S = body_of(c);
  if (!S || !S->placed) return make_status(CFM_ESTATE, "no placed bodies");
  return ensure_dense(c, S);
}

// y_b = P x_b for a blob block embedded with row stride ld (component k of blob j at
// [k ld + j]); frozen: rotated into the frame of the build and back.
Status body_dense_blob_apply(cfm_ctx* c, const double* x, size_t ldx, double* y, size_t ldy) {
  BodyState* S = body_of(c);
  const int m = c->p.n_blob, nrot = m + 2 * S->n_body;
  const size_t L = 3 * (size_t)m + 6 * (size_t)S->n_body;
  CFM_TRY(S->st.alloc(L));
  CFM_TRY(S->rin.alloc(L));
  CFM_TRY(S->rout.alloc(L));
  CFM_TRY(dev_copy_rows(c, x, ldx, 0, S->st.p, m, 0, 3, m));
  CFM_TRY(dev_fill(c, S->st.p + 3 * (size_t)m, 0.0, 6 * S->n_body));
  const double* in = S->st.p;
  if (S->frozen) {
    hipLaunchKernelGGL(body_rotate_kernel, g1(nrot), dim3(256), 0, c->stream, m, S->nb, S->n_body,
                       S->q_cur.p, S->q_build.p, 1, (const double*)S->st.p, S->rin.p);
    in = S->rin.p;
  }
  CFM_TRY(dense_apply(c, c->dinv, in, S->rout.p));
  CFM_TRY(dev_fill(c, S->rout.p + 3 * (size_t)m, 0.0, 6 * S->n_body));
  const double* out = S->rout.p;
  if (S->frozen) {
    hipLaunchKernelGGL(body_rotate_kernel, g1(nrot), dim3(256), 0, c->stream, m, S->nb, S->n_body,
                       S->q_cur.p, S->q_build.p, 0, (const double*)S->rout.p, S->st.p);
    out = S->st.p;
  }
  CFM_CHECK_LAUNCH();
  return dev_copy_rows(c, out, m, 0, y, ldy, 0, 3, m);
}

// y_b += P K U (U: 6 per body, in the current frame; prescribed rotation: P K_V V).
Status body_dense_rigid_add(cfm_ctx* c, const double* U, double* y, size_t ldy) {
  BodyState* S = body_of(c);
  const int m = c->p.n_blob, nrot = m + 2 * S->n_body, nc = 6 * S->n_body;
  const size_t L = 3 * (size_t)m + nc;
  CFM_TRY(S->st.alloc(L));
  CFM_TRY(S->rin.alloc(L));
  CFM_TRY(S->rout.alloc(L));
  CFM_TRY(dev_fill(c, S->st.p, 0.0, 3 * m));
  CFM_HIP(hipMemcpyAsync(S->st.p + 3 * (size_t)m, U, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
  const double* in = S->st.p;
  if (S->frozen) {
    hipLaunchKernelGGL(body_rotate_kernel, g1(nrot), dim3(256), 0, c->stream, m, S->nb, S->n_body,
                       S->q_cur.p, S->q_build.p, 1, (const double*)S->st.p, S->rin.p);
    in = S->rin.p;
  }
  // blob part of `in` is zero: out = P K U'
  hipLaunchKernelGGL(body_dense_finish_kernel, g1(3 * m), dim3(256), 0, c->stream, 3 * m, nc,
                     (int)S->rot_fixed, in, S->PK.p, in + 3 * (size_t)m, S->rout.p);
  CFM_TRY(dev_fill(c, S->rout.p + 3 * (size_t)m, 0.0, nc));
  const double* out = S->rout.p;
  if (S->frozen) {
    hipLaunchKernelGGL(body_rotate_kernel, g1(nrot), dim3(256), 0, c->stream, m, S->nb, S->n_body,
                       S->q_cur.p, S->q_build.p, 0, (const double*)S->rout.p, S->st.p);
    out = S->st.p;
  }
  hipLaunchKernelGGL(body_add_rows_kernel, g1(m), dim3(256), 0, c->stream, m, out, ldy, y);
  CFM_CHECK_LAUNCH();
  return {};
}

// R = K^T P K of the current placement, (6 n_body)^2 row-major on the host; frozen: the
// build's, rotated by G_b = R(q_b) R(q_build,b)^T on each 3 x 3 block.
Status body_dense_resistance(cfm_ctx* c, std::vector<double>* R) {
  BodyState* S = body_of(c);
  const int nb = S->n_body, n = 6 * nb;
  *R = S->R0;
  if (!S->frozen) return {};
  std::vector<double> qc(4 * (size_t)nb), q0(4 * (size_t)nb);
  CFM_HIP(hipMemcpyAsync(qc.data(), S->q_cur.p, sizeof(double) * qc.size(), hipMemcpyDeviceToHost, c->stream));
  CFM_HIP(hipMemcpyAsync(q0.data(), S->q_build.p, sizeof(double) * q0.size(), hipMemcpyDeviceToHost, c->stream));
  CFM_HIP(hipStreamSynchronize(c->stream));
  std::vector<double> G(9 * (size_t)nb);
  for (int b = 0; b < nb; ++b) {
    double Rc[9], Rb[9];
    rot_mat(&qc[4 * b], Rc);
    rot_mat(&q0[4 * b], Rb);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        G[9 * b + 3 * i + j] = Rc[3 * i] * Rb[3 * j] + Rc[3 * i + 1] * Rb[3 * j + 1] + Rc[3 * i + 2] * Rb[3 * j + 2];
  }
  // block (I, J) of 3 x 3 entries: G_{b(I)} R0_IJ G_{b(J)}^T
  for (int I = 0; I < n; I += 3)
    for (int J = 0; J < n; J += 3) {
      const double* Gi = &G[9 * (I / 6)];
      const double* Gj = &G[9 * (J / 6)];
      double T[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double v = 0.0;
          for (int k = 0; k < 3; ++k) v += Gi[3 * i + k] * S->R0[(size_t)(I + k) * n + J + j];
          T[3 * i + j] = v;
        }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double v = 0.0;
          for (int k = 0; k < 3; ++k) v += T[3 * i + k] * Gj[3 * j + k];
          (*R)[(size_t)(I + i) * n + J + j] = v;
        }
    }
  return {};
}

Status body_precond(cfm_ctx* c, int kind, const double* x, double* y) {
  BodyState* S = body_of(c);
  if (!S || !S->placed) return make_status(CFM_ESTATE, "no placed bodies");
  const int m = c->p.n_blob;
  const size_t L = 3 * (size_t)m + 6 * (size_t)S->n_body;
  double* yU = y + 3 * (size_t)m;
  const double* xU = x + 3 * (size_t)m;
  if (kind == CFM_PC_NONE) {
    CFM_HIP(hipMemcpyAsync(y, x, sizeof(double) * L, hipMemcpyDeviceToDevice, c->stream));
    return {};
  }
  const int rot = S->rot_fixed;
  if (kind == CFM_PC_DIAG || kind == CFM_PC_SELF) {
    hipLaunchKernelGGL(body_kt_kernel, dim3(S->n_body), dim3(256), 0, c->stream, 1, rot, m, S->nb, S->r.p, x,
                       xU, S->mats.p, yU);
    hipLaunchKernelGGL(body_pc_blob_kernel, g1(m), dim3(256), 0, c->stream, kind, c->cfg.wall, rot, m, S->nb,
                       S->r.p, c->p.z.p + c->p.n_seg, c->cfg.a_blob, c->cfg.mu, x, y);
    CFM_CHECK_LAUNCH();
    return {};
  }
  if (kind == CFM_PC_DENSE) {
    CFM_TRY(ensure_dense(c, S));
    const double* xin = x;
    double* yout = y;
    const double* roff = S->r.p;
    const int nrot = m + 2 * S->n_body;
    if (S->frozen) {  // into the frame of the build
      CFM_TRY(S->rin.alloc(L));
      CFM_TRY(S->rout.alloc(L));
      hipLaunchKernelGGL(body_rotate_kernel, g1(nrot), dim3(256), 0, c->stream, m, S->nb, S->n_body,
                         S->q_cur.p, S->q_build.p, 1, x, S->rin.p);
      xin = S->rin.p;
      yout = S->rout.p;
      roff = S->r_build.p;
    }
    const int nc = 6 * S->n_body;
    CFM_TRY(S->w.alloc(3 * (size_t)m));
    CFM_TRY(S->aux.alloc(nc));
    CFM_TRY(dense_apply(c, c->dinv, xin, S->w.p));  // w = P f
    hipLaunchKernelGGL(body_kt_kernel, dim3(S->n_body), dim3(256), 0, c->stream, 2, rot, m, S->nb, roff,
                       S->w.p, xin + 3 * (size_t)m, (const double*)nullptr, S->aux.p);  // -u - K^T w
    hipLaunchKernelGGL(body_small_gemv_kernel, g1(nc), dim3(256), 0, c->stream, nc, S->Sinv.p, S->aux.p,
                       yout + 3 * (size_t)m);
    hipLaunchKernelGGL(body_dense_finish_kernel, g1(3 * m), dim3(256), 0, c->stream, 3 * m, nc, rot, S->w.p,
                       S->PK.p, yout + 3 * (size_t)m, yout);
    if (S->frozen)  // back to the current frame
      hipLaunchKernelGGL(body_rotate_kernel, g1(nrot), dim3(256), 0, c->stream, m, S->nb, S->n_body,
                         S->q_cur.p, S->q_build.p, 0, (const double*)S->rout.p, y);
    CFM_CHECK_LAUNCH();
    return {};
  }
  return make_status(CFM_EINVAL, "the body system takes CFM_PC_NONE, _DIAG, _SELF or _DENSE");
}

Status body_info(cfm_ctx* c, BodyInfo* out) {
  BodyState* S = body_of(c);
  if (!S) return make_status(CFM_ESTATE, "cfm_body_init has not been called");
  if (!S->placed) return make_status(CFM_ESTATE, "body state not set (cfm_body_set_state)");
  out->n_body = S->n_body;
  out->nb = S->nb;
  out->x = S->x;
  out->q = S->q;
  out->ref = S->ref;
  std::copy(S->mob, S->mob + 36, out->mob);
  return {};
}

const double* body_offsets(cfm_ctx* c) {
  BodyState* S = body_of(c);
  return S ? S->r.p : nullptr;
}

const double* body_ref_device(cfm_ctx* c) {
  BodyState* S = body_of(c);
  return S ? S->ref_dev.p : nullptr;
}

Status body_place_device(cfm_ctx* c, const double* bx, const double* bq) {
  BodyState* S = body_of(c);
  if (!S) return make_status(CFM_ESTATE, "cfm_body_init has not been called");
  const int m = c->p.n_blob, off = c->p.n_seg;
  if (!S->ref_dev.p) {
    CFM_TRY(S->ref_dev.alloc(S->ref.size()));
    CFM_HIP(hipMemcpyAsync(S->ref_dev.p, S->ref.data(), sizeof(double) * S->ref.size(),
                           hipMemcpyHostToDevice, c->stream));
  }
  CFM_TRY(S->r.alloc(3 * (size_t)m));
  hipLaunchKernelGGL(body_place_kernel, g1(m), dim3(256), 0, c->stream, m, S->nb, S->ref_dev.p, bx, bq,
                     S->r.p, c->p.x.p + off, c->p.y.p + off, c->p.z.p + off);
  hipLaunchKernelGGL(body_mats_kernel, g1(72 * S->n_body), dim3(256), 0, c->stream, S->n_body, S->frame.p,
                     bq, S->mats.p);
  CFM_TRY(S->q_cur.alloc(4 * (size_t)S->n_body));
  CFM_HIP(hipMemcpyAsync(S->q_cur.p, bq, sizeof(double) * 4 * S->n_body, hipMemcpyDeviceToDevice, c->stream));
  CFM_CHECK_LAUNCH();
  // cached blob-position products no longer apply (a frozen dense preconditioner is
  // applied rotated instead)
  c->blob_pos.clear();
  c->nf.valid = false;
  if (!(S->frozen && S->built_frozen)) {
    c->dinv.valid = false;
    S->schur_gen = -1;
  }
  S->placed = true;
  return {};
}

void body_freeze_dense(cfm_ctx* c, bool on) {
  BodyState* S = body_of(c);
  if (!S) return;
  S->frozen = on;
  S->built_frozen = false;
}

bool body_rotation_prescribed(cfm_ctx* c) {
  BodyState* S = body_of(c);
  return S && S->rot_fixed;
}

// Prescribed rotation: the torque rows of a body-system right-hand side hold the
// prescribed Omega_b, zero (the reference has no torque rows).
__global__ void body_zero_torque_kernel(int nbody, double* __restrict__ u) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 3 * nbody) u[6 * (t / 3) + 3 + t % 3] = 0.0;
}

Status body_prescribed_rhs(cfm_ctx* c, double* rhs) {
  BodyState* S = body_of(c);
  if (!S || !S->rot_fixed) return {};
  hipLaunchKernelGGL(body_zero_torque_kernel, g1(3 * S->n_body), dim3(256), 0, c->stream, S->n_body,
                     rhs + 3 * (size_t)c->p.n_blob);
  CFM_CHECK_LAUNCH();
  return {};
}

// Blob positions changed behind the bodies' back (cfm_set_positions): the body
// system needs cfm_body_set_state again.
void body_positions_changed(cfm_ctx* c) {
  BodyState* S = body_of(c);
  if (S) S->placed = false;
}

}  // namespace cfm

using namespace cfm;

void cfm_body_release(cfm_ctx* c) {
  BodyState* S = body_of(c);
  if (!S) return;
  body_release_state(S);
  delete S;
  c->body = nullptr;
}

static int bfail(cfm_ctx* c, const Status& s) {
  c->err = s.msg;
  return s.code;
}

extern "C" {

int cfm_body_init(cfm_ctx* c, int32_t n_body, const double* ref) {
  if (!c) return CFM_EINVAL;
  if (!ref) return bfail(c, make_status(CFM_EINVAL, "null argument"));
  const int m = c->p.n_blob;
  if (n_body <= 0 || m <= 0 || m % n_body != 0)
    return bfail(c, make_status(CFM_EINVAL, "n_body must divide the context's n_blob (%d) and be positive", m));
  const int nb = m / n_body;
  for (int i = 0; i < 3 * nb; ++i)
    if (!std::isfinite(ref[i])) return bfail(c, make_status(CFM_ENONFINITE, "non-finite blob offset"));
  (void)hipSetDevice(c->device);
  cfm_body_release(c);
  BodyState* S = new (std::nothrow) BodyState();
  if (!S) return CFM_ENOMEM;
  S->n_body = n_body;
  S->nb = nb;
  S->ref.assign(ref, ref + 3 * (size_t)nb);
  std::vector<double> A(36);
  ktk(S->ref, A.data());
  if (!invert(A, 6)) {
    delete S;
    return bfail(c, make_status(CFM_EINVAL, "K^T K is singular (blob offsets collinear?)"));
  }
  std::copy(A.begin(), A.end(), S->ktk_inv);
  std::fill(S->mob, S->mob + 36, 0.0);
  for (int k = 0; k < 6; ++k) S->mob[7 * k] = 1.0;
  c->body = S;
  return CFM_OK;
}

int cfm_body_set_state(cfm_ctx* c, const double* x, const double* q) {
  if (!c) return CFM_EINVAL;
  BodyState* S = body_of(c);
  if (!S) return bfail(c, make_status(CFM_ESTATE, "cfm_body_init has not been called"));
  if (!x || !q) return bfail(c, make_status(CFM_EINVAL, "null argument"));
  if (c->p.n_seg > 0 && !c->have_positions)
    return bfail(c, make_status(CFM_ESTATE, "set the segment positions first (cfm_set_positions)"));
  const int B = S->n_body, nb = S->nb, m = c->p.n_blob;
  std::vector<double> qn(4 * (size_t)B);
  for (int b = 0; b < B; ++b) {
    double n2 = 0.0;
    for (int k = 0; k < 4; ++k) {
      if (!std::isfinite(q[4 * b + k])) return bfail(c, make_status(CFM_ENONFINITE, "non-finite quaternion"));
      n2 += q[4 * b + k] * q[4 * b + k];
    }
    if (std::fabs(n2 - 1.0) > 1e-8)
      return bfail(c, make_status(CFM_EINVAL, "quaternion of body %d is not normalised (|q|^2 = %.12g)", b, n2));
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(x[3 * b + k])) return bfail(c, make_status(CFM_ENONFINITE, "non-finite body centre"));
    for (int k = 0; k < 4; ++k) qn[4 * b + k] = q[4 * b + k];
  }
  // positions x_b + R ref_i and offsets R ref_i (rigid_body.cpp:497-515)
  std::vector<double> pos(3 * (size_t)m), off(3 * (size_t)m);
  for (int b = 0; b < B; ++b) {
    double R[9];
    rot_mat(&qn[4 * b], R);
    for (int i = 0; i < nb; ++i) {
      const double* v = &S->ref[3 * (size_t)i];
      const size_t g = (size_t)b * nb + i;
      for (int k = 0; k < 3; ++k) {
        const double o = R[3 * k] * v[0] + R[3 * k + 1] * v[1] + R[3 * k + 2] * v[2];
        off[3 * g + k] = o;
        pos[3 * g + k] = x[3 * b + k] + o;
      }
      if (c->cfg.wall && !(pos[3 * g + 2] > 0.0))
        return bfail(c, make_status(CFM_EINVAL, "blob %d of body %d is not above the wall (z = %g)", i, b,
                                    pos[3 * g + 2]));
    }
  }
  (void)hipSetDevice(c->device);
  auto run = [&]() -> Status {
    CFM_TRY(S->r.alloc(3 * (size_t)m));
    CFM_TRY(S->w.alloc(3 * (size_t)m));
    CFM_TRY(upload_aos(c, off.data(), m, 3, S->r.p, m, 0));
    CFM_TRY(upload_aos(c, pos.data(), m, 3, S->w.p, m, 0));
    const int off0 = c->p.n_seg;
    CFM_HIP(hipMemcpyAsync(c->p.x.p + off0, S->w.p, sizeof(double) * m, hipMemcpyDeviceToDevice, c->stream));
    CFM_HIP(hipMemcpyAsync(c->p.y.p + off0, S->w.p + m, sizeof(double) * m, hipMemcpyDeviceToDevice, c->stream));
    CFM_HIP(hipMemcpyAsync(c->p.z.p + off0, S->w.p + 2 * (size_t)m, sizeof(double) * m, hipMemcpyDeviceToDevice,
                           c->stream));
    S->x.assign(x, x + 3 * (size_t)B);
    S->q = qn;
    CFM_TRY(S->q_cur.alloc(4 * (size_t)B));
    CFM_HIP(hipMemcpyAsync(S->q_cur.p, qn.data(), sizeof(double) * 4 * B, hipMemcpyHostToDevice, c->stream));
    CFM_TRY(upload_mats(c, S));
    return {};
  };
  Status st = run();
  if (!st.ok()) return bfail(c, st);
  if (c->blob_pos != pos) {  // near-field blocks and the resident inverse follow the blobs
    c->blob_pos = pos;
    c->nf.valid = false;
    c->dinv.valid = false;
  }
  if (c->p.n_seg == 0) c->have_positions = true;
  S->schur_gen = -1;  // K changed with the offsets
  S->built_frozen = false;
  S->placed = true;
  return CFM_OK;
}

int cfm_body_set_motion(cfm_ctx* c, int32_t type) {
  if (!c) return CFM_EINVAL;
  BodyState* S = body_of(c);
  if (!S) return bfail(c, make_status(CFM_ESTATE, "cfm_body_init has not been called"));
  if (type == CFM_BODY_PRESCRIBED_VELOCITY)
    return bfail(c, make_status(CFM_EINVAL, "bodies with prescribed (zero) velocities are the blob system "
                                            "CFM_SYS_BLOB, not a body motion"));
  if (type != CFM_BODY_FREE && type != CFM_BODY_PRESCRIBED_ROTATION)
    return bfail(c, make_status(CFM_EINVAL, "unknown body motion %d", type));
  const bool rot = type == CFM_BODY_PRESCRIBED_ROTATION;
  if (rot == S->rot_fixed) return CFM_OK;
  const bool was = S->rot_fixed;
  S->rot_fixed = rot;
  if (S->placed) {
    (void)hipSetDevice(c->device);
    Status st = upload_mats(c, S);
    if (!st.ok()) {
      S->rot_fixed = was;
      return bfail(c, st);
    }
  }
  S->schur_gen = -1;  // the dense Schur inverse depends on the motion
  S->built_frozen = false;
  return CFM_OK;
}

int cfm_body_get_motion(cfm_ctx* c, int32_t* type) {
  if (!c || !type) return CFM_EINVAL;
  BodyState* S = body_of(c);
  if (!S) return bfail(c, make_status(CFM_ESTATE, "cfm_body_init has not been called"));
  *type = S->rot_fixed ? CFM_BODY_PRESCRIBED_ROTATION : CFM_BODY_FREE;
  return CFM_OK;
}

int cfm_body_set_reference_matrices(cfm_ctx* c, const double* ktk_inv, const double* mobility) {
  if (!c) return CFM_EINVAL;
  BodyState* S = body_of(c);
  if (!S) return bfail(c, make_status(CFM_ESTATE, "cfm_body_init has not been called"));
  for (int k = 0; k < 36; ++k)
    if ((ktk_inv && !std::isfinite(ktk_inv[k])) || (mobility && !std::isfinite(mobility[k])))
      return bfail(c, make_status(CFM_ENONFINITE, "non-finite reference matrix"));
  if (ktk_inv) std::copy(ktk_inv, ktk_inv + 36, S->ktk_inv);
  if (mobility) std::copy(mobility, mobility + 36, S->mob);
  if (S->placed) {
    (void)hipSetDevice(c->device);
    Status st = upload_mats(c, S);
    if (!st.ok()) return bfail(c, st);
  }
  return CFM_OK;
}

int cfm_body_get_reference_matrices(cfm_ctx* c, double* ktk_inv, double* mobility) {
  if (!c) return CFM_EINVAL;
  BodyState* S = body_of(c);
  if (!S) return bfail(c, make_status(CFM_ESTATE, "cfm_body_init has not been called"));
  if (ktk_inv) std::copy(S->ktk_inv, S->ktk_inv + 36, ktk_inv);
  if (mobility) std::copy(S->mob, S->mob + 36, mobility);
  return CFM_OK;
}

int cfm_body_reference_matrices(cfm_ctx* c, int32_t body, const cfm_solve_opts* opts, double* ktk_inv,
                                double* mobility, cfm_solve_stats* stats) {
  if (!c) return CFM_EINVAL;
  BodyState* S = body_of(c);
  if (!S) return bfail(c, make_status(CFM_ESTATE, "cfm_body_init has not been called"));
  if (!S->placed) return bfail(c, make_status(CFM_ESTATE, "body state not set (cfm_body_set_state)"));
  if (!opts) return bfail(c, make_status(CFM_EINVAL, "null argument"));
  if (body < 0 || body >= S->n_body) return bfail(c, make_status(CFM_EINVAL, "no body %d", body));
  // One body alone (the reference moves the others 1e10 DL NSEG away): a context of
  // its nb blobs, the same fluid and pair precision.
  cfm_config cfg = c->cfg;
  cfg.n_seg = 0;
  cfg.n_blob = S->nb;
  cfm_ctx* sub = nullptr;
  int rc = cfm_create(&cfg, &sub);
  if (rc != CFM_OK) return bfail(c, make_status(rc, "sub-context: %s", cfm_last_error(nullptr)));
  struct Destroy {
    cfm_ctx* s;
    ~Destroy() { cfm_destroy(s); }
  } guard{sub};
  auto sub_fail = [&](int code) {
    c->err = std::string("single-body solve: ") + cfm_last_error(sub);
    return code;
  };
  if ((rc = cfm_body_init(sub, 1, S->ref.data())) != CFM_OK) return sub_fail(rc);
  if ((rc = cfm_body_set_state(sub, &S->x[3 * (size_t)body], &S->q[4 * (size_t)body])) != CFM_OK)
    return sub_fail(rc);
  // (K^T K)^-1 first, so that the preconditioner can use it (N = identity meanwhile)
  BodyState* T = body_of(sub);
  std::copy(S->ktk_inv, S->ktk_inv + 36, T->ktk_inv);
  cfm_solve_opts o = *opts;
  o.system = CFM_SYS_BODY;
  o.use_guess = 0;
  const int L = 3 * S->nb + 6;
  std::vector<double> rhs(L), sol(L);
  double N[36];
  cfm_solve_stats worst{};
  worst.converged = 1;
  for (int n = 0; n < 6; ++n) {
    std::fill(rhs.begin(), rhs.end(), 0.0);
    rhs[L - 1 - n] = -1.0;
    cfm_solve_stats st{};
    if ((rc = cfm_gmres(sub, &o, rhs.data(), sol.data(), &st)) != CFM_OK) return sub_fail(rc);
    for (int k = 0; k < 6; ++k) N[6 * k + (5 - n)] = sol[L - 6 + k];
    worst.iterations += st.iterations;
    worst.refinements += st.refinements;
    worst.rel_residual = std::max(worst.rel_residual, st.rel_residual);
  }
  // store in the body frame: rotated by R(q) whenever it is used
  double R[9];
  rot_mat(&S->q[4 * (size_t)body], R);
  rotate6(N, R, true, S->mob);
  (void)hipSetDevice(c->device);
  Status st = upload_mats(c, S);
  if (!st.ok()) return bfail(c, st);
  if (ktk_inv) std::copy(S->ktk_inv, S->ktk_inv + 36, ktk_inv);
  if (mobility) std::copy(S->mob, S->mob + 36, mobility);
  if (stats) *stats = worst;
  return CFM_OK;
}

}  // extern "C"
