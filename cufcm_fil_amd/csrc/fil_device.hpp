// fil_device.hpp -- device helpers for the elastic-filament mechanics (FP64).
//
// Quaternions are (s, v0, v1, v2) with the Hamilton product, as in
// src/general/quaternion.cpp:79-470; rcross(v) w = w x v as in
// src/general/matrix.cpp:665-705.
#pragma once
#include <hip/hip_runtime.h>

namespace cfm {
namespace fil {

struct Q4 {
  double s, x, y, z;
};
struct V3 {
  double x, y, z;
};
struct M3 {
  double a[3][3];
};

__device__ __forceinline__ V3 v3(double x, double y, double z) { return {x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

__device__ __forceinline__ Q4 qmul(Q4 p, Q4 q) {
  return {p.s * q.s - (p.x * q.x + p.y * q.y + p.z * q.z),
          p.s * q.x + q.s * p.x + p.y * q.z - p.z * q.y,
          p.s * q.y + q.s * p.y + p.z * q.x - p.x * q.z,
          p.s * q.z + q.s * p.z + p.x * q.y - p.y * q.x};
}
__device__ __forceinline__ Q4 qconj(Q4 q) { return {q.s, -q.x, -q.y, -q.z}; }
__device__ __forceinline__ Q4 qsqrt(Q4 q) {
  if (q.s < -0.999999999) return {0.0, 0.0, 0.0, 1.0};
  const double s = sqrt(0.5 * (1.0 + q.s));
  const double t = 2.0 * s;
  return {s, q.x / t, q.y / t, q.z / t};
}
__device__ __forceinline__ Q4 midpoint(Q4 q1, Q4 q2) { return qmul(qsqrt(qmul(q2, qconj(q1))), q1); }

__device__ __forceinline__ Q4 lie_exp(V3 u) {
  const double th = sqrt(dot(u, u));
  if (th < 1e-6) return {1.0, 0.0, 0.0, 0.0};
  const double sn = sin(0.5 * th) / th;
  return {cos(0.5 * th), sn * u.x, sn * u.y, sn * u.z};
}

// columns of the rotation matrix: tangent, normal, binormal
__device__ __forceinline__ V3 tangent(Q4 q) {
  return {1.0 - 2.0 * (q.y * q.y + q.z * q.z), 2.0 * (q.x * q.y + q.s * q.z),
          2.0 * (q.x * q.z - q.s * q.y)};
}
__device__ __forceinline__ V3 normal(Q4 q) {
  return {2.0 * (q.x * q.y - q.s * q.z), 1.0 - 2.0 * (q.x * q.x + q.z * q.z),
          2.0 * (q.y * q.z + q.s * q.x)};
}
__device__ __forceinline__ V3 binormal(Q4 q) {
  return {2.0 * (q.x * q.z + q.s * q.y), 2.0 * (q.y * q.z - q.s * q.x),
          1.0 - 2.0 * (q.x * q.x + q.y * q.y)};
}
__device__ __forceinline__ M3 rot_mat(Q4 q) {
  const V3 t = tangent(q), n = normal(q), b = binormal(q);
  M3 R;
  R.a[0][0] = t.x; R.a[1][0] = t.y; R.a[2][0] = t.z;
  R.a[0][1] = n.x; R.a[1][1] = n.y; R.a[2][1] = n.z;
  R.a[0][2] = b.x; R.a[1][2] = b.y; R.a[2][2] = b.z;
  return R;
}

__device__ __forceinline__ V3 mv(const M3& A, V3 v) {
  return {A.a[0][0] * v.x + A.a[0][1] * v.y + A.a[0][2] * v.z,
          A.a[1][0] * v.x + A.a[1][1] * v.y + A.a[1][2] * v.z,
          A.a[2][0] * v.x + A.a[2][1] * v.y + A.a[2][2] * v.z};
}
__device__ __forceinline__ M3 mm(const M3& A, const M3& B) {
  M3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      C.a[i][j] = A.a[i][0] * B.a[0][j] + A.a[i][1] * B.a[1][j] + A.a[i][2] * B.a[2][j];
  return C;
}
__device__ __forceinline__ M3 madd(const M3& A, const M3& B, double sb = 1.0) {
  M3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C.a[i][j] = A.a[i][j] + sb * B.a[i][j];
  return C;
}
__device__ __forceinline__ M3 mscale(const M3& A, double s) {
  M3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C.a[i][j] = s * A.a[i][j];
  return C;
}
__device__ __forceinline__ M3 eye3() {
  M3 I = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
  return I;
}
__device__ __forceinline__ M3 zero3() {
  M3 Z = {{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}};
  return Z;
}
// rcross(v) w = w x v
__device__ __forceinline__ M3 rcross(V3 v) {
  M3 M;
  M.a[0][0] = 0.0;  M.a[0][1] = v.z;  M.a[0][2] = -v.y;
  M.a[1][0] = -v.z; M.a[1][1] = 0.0;  M.a[1][2] = v.x;
  M.a[2][0] = v.y;  M.a[2][1] = -v.x; M.a[2][2] = 0.0;
  return M;
}

__device__ __forceinline__ double dexp_fac(V3 u) {
  const double th2 = dot(u, u);
  const double th = sqrt(th2);
  return (th < 1e-5) ? -1.0 / 12.0 : (0.5 * th * cos(0.5 * th) / sin(0.5 * th) - 1.0) / th2;
}
__device__ __forceinline__ V3 dexpinv(V3 u, V3 v) {
  const double f = dexp_fac(u);
  const V3 uv = cross(u, v);
  return v - 0.5 * uv - f * cross(u, uv);
}
__device__ __forceinline__ V3 dexpinv_transpose(V3 u, V3 v) {
  const double f = dexp_fac(u);
  const V3 uv = cross(u, v);
  return v + 0.5 * uv - f * cross(u, uv);
}
// dexp(u) v = v + alpha u x v + beta u x (u x v); below the small angle the
// coefficients invert dexpinv's series exactly (quaternion.cpp:552-586)
__device__ __forceinline__ V3 dexp(V3 u, V3 v) {
  const double th2 = dot(u, u);
  const double th = sqrt(th2);
  double alpha, beta;
  if (th < 1e-5) {
    const double th4 = th2 * th2;
    beta = (1.0 / 3.0 + th2 / 72.0) / (2.0 + th2 / 6.0 + th4 / 72.0);
    alpha = 1.0 / 6.0 + 2.0 * beta * (1.0 - th2 / 12.0);
  } else {
    alpha = sin(0.5 * th);
    alpha *= 2.0 * alpha / th2;
    beta = (th - sin(th)) / (th * th2);
  }
  const V3 uv = cross(u, v);
  return v + alpha * uv + beta * cross(u, uv);
}

}  // namespace fil
}  // namespace cfm
