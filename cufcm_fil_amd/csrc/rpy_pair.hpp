// rpy_pair.hpp -- device-side Rotne-Prager-Yamakawa pair blocks for gfx950.
//
// Mobility of sphere i (target) due to the load on sphere j (source):
//   * wall_pair  : no-slip plane at z = 0, a = 1, mu = 1 (Swan & Brady 2007 images),
//                  following src/cilia/cuda_functions.cu:65-201.  The far-field form is
//                  used at every separation, as in the reference (no overlap forms).
//   * wall_self  : single sphere above the wall, cuda_functions.cu:24-63.
//   * free_pair  : unbounded RPY with the Zuk et al. (2014) overlap forms,
//                  cuda_functions.cu:205-356.
// The arithmetic type T is float or double.  Inputs are the separation
// (dx, dy, dz) = x_i - x_j, the image separation Rz = z_i + z_j and the source
// height z_j; the caller forms dx, dy, dz in double before narrowing so that the
// FP32 kernels keep full relative accuracy for close pairs far from the origin.
#pragma once
#include <hip/hip_runtime.h>

namespace cfm {

template <typename T> __device__ __forceinline__ T rsq(T x);
template <> __device__ __forceinline__ float rsq<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <> __device__ __forceinline__ double rsq<double>(double x) { return rsqrt(x); }

template <typename T> __device__ __forceinline__ T rcp(T x);
template <> __device__ __forceinline__ float rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ __forceinline__ double rcp<double>(double x) { return 1.0 / x; }

// Adds M_ij F_j to V_i.  VD, FD in {3, 6}.
template <typename T, int VD, int FD>
__device__ __forceinline__ void wall_pair(T dx, T dy, T dz, T Rz, T zj, const T* __restrict__ F,
                                          T* __restrict__ V) {
  constexpr T c6 = T(0.053051647697298445);   // 1/(6 pi)
  constexpr T c8 = T(0.039788735772973836);   // 1/(8 pi)
  const T fx = F[0], fy = F[1], fz = F[2];
  const T tx = (FD > 3) ? F[3] : T(0), ty = (FD > 3) ? F[4] : T(0), tz = (FD > 3) ? F[5] : T(0);

  // ---- free-space part ----
  {
    const T is = rsq<T>(dx * dx + dy * dy + dz * dz);
    const T rx = dx * is, ry = dy * is, rz = dz * is;
    const T is2 = is * is;
    const T k1 = c8 * is;
    const T rf = rx * fx + ry * fy + rz * fz;
    const T iso = k1 * (T(1) + T(2.0 / 3.0) * is2);
    const T an = k1 * (T(1) - T(2) * is2) * rf;
    V[0] += iso * fx + an * rx;
    V[1] += iso * fy + an * ry;
    V[2] += iso * fz + an * rz;
    const T k2 = k1 * is;
    if (FD > 3) {
      V[0] += k2 * (ty * rz - tz * ry);
      V[1] += k2 * (tz * rx - tx * rz);
      V[2] += k2 * (tx * ry - ty * rx);
    }
    if (VD > 3) {
      V[3] += k2 * (fy * rz - fz * ry);
      V[4] += k2 * (fz * rx - fx * rz);
      V[5] += k2 * (fx * ry - fy * rx);
      if (FD > 3) {
        const T k3 = T(0.5) * k2 * is;
        const T rt3 = T(3) * k3 * (rx * tx + ry * ty + rz * tz);
        V[3] += rt3 * rx - k3 * tx;
        V[4] += rt3 * ry - k3 * ty;
        V[5] += rt3 * rz - k3 * tz;
      }
    }
  }

  // ---- image system ----
  const T p1 = rsq<T>(dx * dx + dy * dy + Rz * Rz);
  const T p2 = p1 * p1, p3 = p2 * p1, p4 = p2 * p2, p5 = p4 * p1;
  const T ex = dx * p1, ey = dy * p1, ez = Rz * p1;
  const T ez2 = ez * ez;
  const T h = zj * rcp<T>(Rz);
  const T hh = h * (T(1) - h);
  const T ef = ex * fx + ey * fy + ez * fz;
  const T q3 = T(1) - T(5) * ez2;  // recurring (1 - 5 ez^2)
  {
    const T A0 = -(T(0.75) * p1 * (T(1) + T(2) * hh * ez2) +
                   T(0.5) * (p3 * (T(1) - T(3) * ez2) - p5 * q3));
    const T A1 = -(T(0.75) * p1 * (T(1) - T(6) * hh * ez2) - T(1.5) * p3 * q3 +
                   T(2.5) * p5 * (T(1) - T(7) * ez2)) * ef;
    const T A2 = ez * fz *
                 (T(1.5) * h * p1 * (T(1) - T(6) * (T(1) - h) * ez2) - T(3) * p3 * q3 +
                  T(5) * p5 * (T(2) - T(7) * ez2));
    const T A3 = (T(1.5) * h * p1 - T(5) * p5) * ef;
    const T A4 = -((T(3) * h * h * p1 + T(3) * p3) * ez2 + (T(2) - T(15) * ez2) * p5) * fz;
    const T A12 = A1 + A2;
    V[0] += c6 * (A0 * fx + A12 * ex);
    V[1] += c6 * (A0 * fy + A12 * ey);
    V[2] += c6 * (A0 * fz + (A12 + A3) * ez + A4);
  }
  if (VD > 3) {
    const T B0 = T(0.75) * p2;
    {
      const T B1 = fz * (T(9) * h * ez2 * p2 + (T(1.5) - T(15) * ez2) * p4);
      const T B2 = -ez * ef * (T(4.5) * h * p2 - T(7.5) * p4);
      const T B3 = -T(1.5) * ez * (h * p2 - p4);
      const T B12 = B1 + B2;
      V[3] -= c6 * (B0 * (fy * ez - fz * ey) - B12 * ey + B3 * fy);
      V[4] -= c6 * (B0 * (fz * ex - fx * ez) + B12 * ex - B3 * fx);
      V[5] -= c6 * B0 * (fx * ey - fy * ex);
    }
    if (FD > 3) {
      const T et = ex * tx + ey * ty + ez * tz;
      const T C0 = p3 * (T(0.375) - T(2.25) * ez2);
      const T C1 = -T(1.125) * p3 * et;
      const T C2 = T(2.25) * ez * p3 * et;
      const T C3 = T(2.25) * p3 * (ex * ty - ey * tx);
      V[3] += c6 * (C0 * tx + C1 * ex - C3 * ey);
      V[4] += c6 * (C0 * ty + C1 * ey + C3 * ex);
      V[5] += c6 * (C0 * tz + C1 * ez + C2);
    }
  }
  if (FD > 3) {
    // translation <- torque: the rotation<-force expressions of the swapped pair
    // (h -> 1 - h, e_xy -> -e_xy), i.e. the transpose block.
    const T g = T(1) - h;
    const T mx = -ex, my = -ey;
    const T ecz = mx * ty - my * tx;
    const T D0 = T(0.75) * p2;
    const T D1 = (T(9) * g * ez2 * p2 + (T(1.5) - T(15) * ez2) * p4) * ecz;
    const T D2 = -ez * (T(4.5) * g * p2 - T(7.5) * p4) * ecz;
    const T D3 = -T(1.5) * ez * (g * p2 - p4);
    V[0] -= c6 * (D0 * (tz * my - ty * ez) + D2 * mx - D3 * ty);
    V[1] -= c6 * (D0 * (tx * ez - tz * mx) + D2 * my + D3 * tx);
    V[2] -= c6 * (D0 * (ty * mx - tx * my) + D2 * ez + D1);
  }
}

// Hot-loop form of wall_pair for loads already multiplied by 1/(6 pi) (the tiled
// products scale each source once when they stage it): the same blocks, regrouped
// so that every velocity component is accumulated by fused multiply-adds and the
// products shared between blocks (k2 r_hat, B0 e, (e x t)_z) are formed once.
template <typename T, int VD, int FD>
__device__ __forceinline__ void wall_pair_scaled(T dx, T dy, T dz, T Rz, T zj,
                                                 const T* __restrict__ F, T* __restrict__ V) {
  const T fx = F[0], fy = F[1], fz = F[2];
  const T tx = (FD > 3) ? F[3] : T(0), ty = (FD > 3) ? F[4] : T(0), tz = (FD > 3) ? F[5] : T(0);
  const T hxy = dx * dx + dy * dy;

  // ---- free-space part (1/(8 pi) = (6/8) / (6 pi)) ----
  {
    const T is = rsq<T>(hxy + dz * dz);
    const T rx = dx * is, ry = dy * is, rz = dz * is;
    const T is2 = is * is;
    const T k1 = T(0.75) * is;
    const T rf = rx * fx + ry * fy + rz * fz;
    const T iso = k1 * (T(1) + T(2.0 / 3.0) * is2);
    const T an = k1 * (T(1) - T(2) * is2) * rf;
    V[0] += iso * fx;
    V[0] += an * rx;
    V[1] += iso * fy;
    V[1] += an * ry;
    V[2] += iso * fz;
    V[2] += an * rz;
    if (FD > 3 || VD > 3) {
      const T k2 = k1 * is;
      const T qx = k2 * rx, qy = k2 * ry, qz = k2 * rz;
      if (FD > 3) {  // k2 (t x r_hat)
        V[0] += ty * qz;
        V[0] -= tz * qy;
        V[1] += tz * qx;
        V[1] -= tx * qz;
        V[2] += tx * qy;
        V[2] -= ty * qx;
      }
      if (VD > 3) {  // k2 (f x r_hat)
        V[3] += fy * qz;
        V[3] -= fz * qy;
        V[4] += fz * qx;
        V[4] -= fx * qz;
        V[5] += fx * qy;
        V[5] -= fy * qx;
        if (FD > 3) {
          const T k3 = T(0.5) * k2 * is;
          const T rt3 = T(3) * k3 * (rx * tx + ry * ty + rz * tz);
          V[3] += rt3 * rx;
          V[3] -= k3 * tx;
          V[4] += rt3 * ry;
          V[4] -= k3 * ty;
          V[5] += rt3 * rz;
          V[5] -= k3 * tz;
        }
      }
    }
  }

  // ---- image system ----
  const T p1 = rsq<T>(hxy + Rz * Rz);
  const T p2 = p1 * p1, p3 = p2 * p1, p4 = p2 * p2, p5 = p4 * p1;
  const T ex = dx * p1, ey = dy * p1, ez = Rz * p1;
  const T u = ez * ez;
  const T h = zj * rcp<T>(Rz);
  const T hh = h * (T(1) - h);
  const T ef = ex * fx + ey * fy + ez * fz;
  const T q3 = T(1) - T(5) * u;
  {
    const T A0 = -(T(0.75) * p1 * (T(1) + T(2) * hh * u) +
                   T(0.5) * (p3 * (T(1) - T(3) * u) - p5 * q3));
    const T A1 = -(T(0.75) * p1 * (T(1) - T(6) * hh * u) - T(1.5) * p3 * q3 +
                   T(2.5) * p5 * (T(1) - T(7) * u)) * ef;
    const T A2 = ez * fz *
                 (T(1.5) * h * p1 * (T(1) - T(6) * (T(1) - h) * u) - T(3) * p3 * q3 +
                  T(5) * p5 * (T(2) - T(7) * u));
    const T A3 = (T(1.5) * h * p1 - T(5) * p5) * ef;
    const T A4 = -((T(3) * h * h * p1 + T(3) * p3) * u + (T(2) - T(15) * u) * p5) * fz;
    const T A12 = A1 + A2;
    V[0] += A0 * fx;
    V[0] += A12 * ex;
    V[1] += A0 * fy;
    V[1] += A12 * ey;
    V[2] += A0 * fz;
    V[2] += (A12 + A3) * ez;
    V[2] += A4;
  }
  if (VD > 3 || FD > 3) {
    const T B0 = T(0.75) * p2;  // also D0 of the torque block
    const T bx = B0 * ex, by = B0 * ey, bz = B0 * ez;
    const T hp2 = h * p2;
    const T w4 = (T(1.5) - T(15) * u) * p4;
    const T cz = ex * ty - ey * tx;  // (e x t)_z
    if (VD > 3) {
      const T B1 = fz * (T(9) * u * hp2 + w4);
      const T B2 = -ez * ef * (T(4.5) * hp2 - T(7.5) * p4);
      const T B3 = -T(1.5) * ez * (hp2 - p4);
      const T B12 = B1 + B2;
      const T s3 = bz + B3;
      V[3] -= fy * s3;
      V[3] += fz * by;
      V[3] += B12 * ey;
      V[4] -= fz * bx;
      V[4] += fx * s3;
      V[4] -= B12 * ex;
      V[5] -= fx * by;
      V[5] += fy * bx;
      if (FD > 3) {
        const T et = ex * tx + ey * ty + ez * tz;
        const T p3et = p3 * et;
        const T C0 = p3 * (T(0.375) - T(2.25) * u);
        const T C1 = -T(1.125) * p3et;
        const T C2 = T(2.25) * ez * p3et;
        const T C3 = T(2.25) * p3 * cz;
        V[3] += C0 * tx;
        V[3] += C1 * ex;
        V[3] -= C3 * ey;
        V[4] += C0 * ty;
        V[4] += C1 * ey;
        V[4] += C3 * ex;
        V[5] += C0 * tz;
        V[5] += C1 * ez;
        V[5] += C2;
      }
    }
    if (FD > 3) {  // translation <- torque (transpose of rotation <- force, h -> 1 - h)
      const T gp2 = p2 - hp2;
      const T D1 = -(T(9) * u * gp2 + w4) * cz;
      const T D2 = ez * (T(4.5) * gp2 - T(7.5) * p4) * cz;
      const T D3 = -T(1.5) * ez * (gp2 - p4);
      const T s0 = bz + D3;
      V[0] += tz * by;
      V[0] += ty * s0;
      V[0] += D2 * ex;
      V[1] -= tx * s0;
      V[1] -= tz * bx;
      V[1] += D2 * ey;
      V[2] += ty * bx;
      V[2] -= tx * by;
      V[2] -= D2 * ez;
      V[2] -= D1;
    }
  }
}

// Self mobility of one sphere above the wall (a = 1, mu = 1); double only.
template <int VD, int FD>
__device__ __forceinline__ void wall_self(double z, const double* __restrict__ F,
                                          double* __restrict__ V) {
  const double c6 = 0.053051647697298445, c8 = 0.039788735772973836;
  const double h1 = 1.0 / z;
  const double h3 = h1 * h1 * h1, h4 = h3 * h1, h5 = h4 * h1;
  const double par = c6 * (1.0 - (9.0 * h1 - 2.0 * h3 + h5) * 0.0625);
  const double perp = c6 * (1.0 - (9.0 * h1 - 4.0 * h3 + h5) * 0.125);
  const double cpl = c6 * 0.09375 * h4;
  V[0] += par * F[0];
  V[1] += par * F[1];
  V[2] += perp * F[2];
  if (FD > 3) {
    V[0] += cpl * F[4];
    V[1] -= cpl * F[3];
  }
  if (VD > 3) {
    V[3] -= cpl * F[1];
    V[4] += cpl * F[0];
    if (FD > 3) {
      const double rpar = c8 * (1.0 - 0.3125 * h3);
      const double rperp = c8 * (1.0 - 0.125 * h3);
      V[3] += rpar * F[3];
      V[4] += rpar * F[4];
      V[5] += rperp * F[5];
    }
  }
}

// Unbounded RPY of unequal spheres including the overlap forms (any r >= 0).
template <typename T, int VD, int FD>
__device__ __forceinline__ void free_pair(T dx, T dy, T dz, T ai, T aj, T mu,
                                          const T* __restrict__ F, T* __restrict__ V) {
  constexpr T pi = T(3.14159265358979323846);
  const T r = sqrt(dx * dx + dy * dy + dz * dz);
  const T rm1 = T(1) / (r + T(1e-20));
  const T dmin = fabs(ai - aj);
  const T amax = ai > aj ? ai : aj;
  const T rx = dx * rm1, ry = dy * rm1, rz = dz * rm1;
  const T sum = ai + aj;
  T A, B;
  T rdot = rx * F[0] + ry * F[1] + rz * F[2];
  if (r > sum) {
    A = rm1 * rm1 * (ai * ai + aj * aj) / T(3);
    B = T(1) - T(3) * A;
    A += T(1);
    const T t = rm1 / (T(8) * pi * mu);
    A *= t;
    B *= t * rdot;
  } else if (r > dmin) {
    T t = T(32) * r * r * r;
    A = dmin * dmin + T(3) * r * r;
    A *= -A / t;
    A += T(0.5) * sum;
    B = dmin * dmin - r * r;
    B *= T(3) * B / t;
    t = T(1) / (T(6) * pi * mu * ai * aj);
    A *= t;
    B *= t * rdot;
  } else {
    A = T(1) / (T(6) * pi * mu * amax);
    B = T(0);
  }
  V[0] += A * F[0] + B * rx;
  V[1] += A * F[1] + B * ry;
  V[2] += A * F[2] + B * rz;
  if (FD > 3) {
    if (r > sum) {
      A = rm1 * rm1 / (T(8) * pi * mu);
    } else if (r > dmin) {
      A = aj - ai + r;
      A *= rm1 * rm1 * A * (ai * ai + T(2) * ai * (r + aj) - T(3) * (aj - r) * (aj - r)) /
           (T(128) * pi * mu * ai * aj * aj * aj);
    } else {
      A = (aj > ai) ? r / (T(8) * pi * mu * aj * aj * aj) : T(0);
    }
    V[0] += A * (F[4] * rz - F[5] * ry);
    V[1] += A * (F[5] * rx - F[3] * rz);
    V[2] += A * (F[3] * ry - F[4] * rx);
  }
  if (VD > 3) {
    if (r > sum) {
      A = rm1 * rm1 / (T(8) * pi * mu);
    } else if (r > dmin) {
      A = ai - aj + r;
      A *= rm1 * rm1 * A * (aj * aj + T(2) * aj * (r + ai) - T(3) * (ai - r) * (ai - r)) /
           (T(128) * pi * mu * aj * ai * ai * ai);
    } else {
      A = (ai > aj) ? r / (T(8) * pi * mu * ai * ai * ai) : T(0);
    }
    V[3] += A * (F[1] * rz - F[2] * ry);
    V[4] += A * (F[2] * rx - F[0] * rz);
    V[5] += A * (F[0] * ry - F[1] * rx);
    if (FD > 3) {
      rdot = rx * F[3] + ry * F[4] + rz * F[5];
      if (r > sum) {
        A = -rm1 * rm1 * rm1 / (T(16) * pi * mu);
        B = -T(3) * rdot * A;
      } else if (r > dmin) {
        const T r2 = r * r, r3 = r2 * r;
        const T ai2 = ai * ai, aj2 = aj * aj;
        A = T(5) * r3 * r3 - T(27) * r2 * r2 * (ai2 + aj2) + T(32) * r3 * (ai2 * ai + aj2 * aj) -
            T(9) * r2 * (ai2 - aj2) * (ai2 - aj2) -
            dmin * dmin * dmin * dmin * (ai2 + aj2 + T(4) * ai * aj);
        B = (ai - aj) * (ai - aj) - r2;
        B *= T(3) * B;
        B *= ai2 + aj2 + T(4) * ai * aj - r2;
        const T t = T(1) / (T(512) * pi * mu * r3 * (ai2 * ai) * (aj2 * aj));
        A *= t;
        B *= t * rdot;
      } else {
        A = T(1) / (T(8) * pi * mu * amax * amax * amax);
        B = T(0);
      }
      V[3] += A * F[3] + B * rx;
      V[4] += A * F[4] + B * ry;
      V[5] += A * F[5] + B * rz;
    }
  }
}

}  // namespace cfm
