// rpy_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of the reference's Rotne-Prager-Yamakawa pair interaction
// and of the all-pairs mobility products that call it.  It is the checker the
// HIP library is compared against; nothing in the product links or calls it.
//
// Followed reference code (racksa/CUFCM_FIL):
//   * src/cilia/cuda_functions.cu:12-203   wall branch of rpy_interaction<vd,fd>
//       - self term          :24-63
//       - unbounded pair     :65-116
//       - Swan-Brady images  :118-176
//       - torque->velocity   :178-200 (by symmetry, h -> 1-h, e_xy -> -e_xy)
//   * src/cilia/cuda_functions.cu:205-356  unbounded branch (Zuk et al. 2014
//       overlap forms for r < a_i + a_j)
//   * src/cilia/cuda_functions.cu:362-757  Mss/Mbb/Msb/Mbs_mult: V_i = sum_j M_ij F_j,
//       AoS layout x[3i+k], f/v segments [6i+k] (force 0-2, torque 3-5), blobs [3i+k]
//
// Deliberate difference (documented): the wall branch chooses the self term by
// comparing *global particle ids*, not the per-species loop indices the reference
// compares (cuda_functions.cu:24 with the index spaces of :605-609,719-723), so a
// segment and a blob that happen to share an index are never treated as the same
// particle.
//
// Parity status: pinned.  The wall pair block is checked in tests/ against an
// independent construction (Blake's image tensor with Faxen operators, evaluated
// by finite differences), the self term against the published Swan-Brady /
// Perkins-Jones closed forms, and the grand mobility matrix against symmetry and
// positive definiteness.
//
// Build: oracle/Makefile (g++ -O2 -fopenmp).  Exposed through oracle/oracle.py.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <omp.h>

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;

// ---------------------------------------------------------------------------
// Wall-bounded self mobility (a = 1, mu = 1) -- cuda_functions.cu:24-63.
void wall_self(int vd, int fd, double z, const double* F, double* V) {
  const double h1 = 1.0 / z;
  const double h3 = h1 * h1 * h1;
  const double h4 = h3 * h1;
  const double h5 = h4 * h1;
  const double c6 = 1.0 / (6.0 * kPi);
  const double c8 = 1.0 / (8.0 * kPi);
  const double par = c6 * (1.0 - (9.0 * h1 - 2.0 * h3 + h5) / 16.0);
  const double perp = c6 * (1.0 - (9.0 * h1 - 4.0 * h3 + h5) / 8.0);
  const double cpl = c6 * (3.0 / 32.0) * h4;
  V[0] += par * F[0];
  V[1] += par * F[1];
  V[2] += perp * F[2];
  if (fd > 3) {
    V[0] += cpl * F[4];
    V[1] -= cpl * F[3];
  }
  if (vd > 3) {
    V[3] -= cpl * F[1];
    V[4] += cpl * F[0];
    if (fd > 3) {
      const double rpar = c8 * (1.0 - (5.0 / 16.0) * h3);
      const double rperp = c8 * (1.0 - (1.0 / 8.0) * h3);
      V[3] += rpar * F[3];
      V[4] += rpar * F[4];
      V[5] += rperp * F[5];
    }
  }
}

// ---------------------------------------------------------------------------
// Wall-bounded pair mobility, i != j (a = 1, mu = 1) -- cuda_functions.cu:65-201.
void wall_pair(int vd, int fd, double xi, double yi, double zi, double xj, double yj,
               double zj, const double* F, double* V) {
  const double c6 = 1.0 / (6.0 * kPi);
  const double dx = xi - xj, dy = yi - yj, dz = zi - zj;
  const double fx = F[0], fy = F[1], fz = F[2];
  const double tx = (fd > 3) ? F[3] : 0.0;
  const double ty = (fd > 3) ? F[4] : 0.0;
  const double tz = (fd > 3) ? F[5] : 0.0;

  // --- free-space RPY part (far-field form used for every separation) ---
  {
    const double s = std::sqrt(dx * dx + dy * dy + dz * dz);
    const double is = 1.0 / s;
    const double rx = dx * is, ry = dy * is, rz = dz * is;
    const double rf = rx * fx + ry * fy + rz * fz;
    const double k1 = is / (8.0 * kPi);                 // 1/(8 pi s)
    const double iso = 1.0 + (2.0 / 3.0) * is * is;
    const double aniso = 1.0 - 2.0 * is * is;
    V[0] += k1 * (iso * fx + aniso * rf * rx);
    V[1] += k1 * (iso * fy + aniso * rf * ry);
    V[2] += k1 * (iso * fz + aniso * rf * rz);
    const double k2 = k1 * is;                          // 1/(8 pi s^2)
    if (fd > 3) {  // v += k2 (tau x rhat)
      V[0] += k2 * (ty * rz - tz * ry);
      V[1] += k2 * (tz * rx - tx * rz);
      V[2] += k2 * (tx * ry - ty * rx);
    }
    if (vd > 3) {  // w += k2 (f x rhat)
      V[3] += k2 * (fy * rz - fz * ry);
      V[4] += k2 * (fz * rx - fx * rz);
      V[5] += k2 * (fx * ry - fy * rx);
      if (fd > 3) {  // w += (3 (rhat.tau) rhat - tau)/(16 pi s^3)
        const double k3 = 0.5 * k2 * is;
        const double rt = rx * tx + ry * ty + rz * tz;
        V[3] += k3 * (3.0 * rt * rx - tx);
        V[4] += k3 * (3.0 * rt * ry - ty);
        V[5] += k3 * (3.0 * rt * rz - tz);
      }
    }
  }

  // --- image system (Swan & Brady 2007) ---
  const double Rz = zi + zj;
  const double rho = std::sqrt(dx * dx + dy * dy + Rz * Rz);
  const double p1 = 1.0 / rho;
  const double p2 = p1 * p1, p3 = p2 * p1, p4 = p3 * p1, p5 = p4 * p1;
  double ex = dx * p1, ey = dy * p1;
  const double ez = Rz * p1;
  const double ez2 = ez * ez;
  double h = zj / Rz;

  {  // translation <- force
    const double ef = ex * fx + ey * fy + ez * fz;
    const double A0 = -(0.75 * p1 * (1.0 + 2.0 * h * (1.0 - h) * ez2) +
                        0.5 * (p3 * (1.0 - 3.0 * ez2) - p5 * (1.0 - 5.0 * ez2)));
    const double A1 = -(0.75 * p1 * (1.0 - 6.0 * h * (1.0 - h) * ez2) -
                        1.5 * p3 * (1.0 - 5.0 * ez2) + 2.5 * p5 * (1.0 - 7.0 * ez2)) * ef;
    const double A2 = ez * fz *
                      (1.5 * h * p1 * (1.0 - 6.0 * (1.0 - h) * ez2) - 3.0 * p3 * (1.0 - 5.0 * ez2) +
                       5.0 * p5 * (2.0 - 7.0 * ez2));
    const double A3 = (1.5 * h * p1 - 5.0 * p5) * ef;
    const double A4 = -((3.0 * h * h * p1 + 3.0 * p3) * ez2 + (2.0 - 15.0 * ez2) * p5) * fz;
    V[0] += c6 * (A0 * fx + (A1 + A2) * ex);
    V[1] += c6 * (A0 * fy + (A1 + A2) * ey);
    V[2] += c6 * (A0 * fz + (A1 + A2 + A3) * ez + A4);
  }
  if (vd > 3) {
    {  // rotation <- force (sign convention of the reference: subtracted)
      const double ef = ex * fx + ey * fy + ez * fz;
      const double B0 = 0.75 * p2;
      const double B1 = fz * (9.0 * h * ez2 * p2 + (1.5 - 15.0 * ez2) * p4);
      const double B2 = -ez * ef * (4.5 * h * p2 - 7.5 * p4);
      const double B3 = -1.5 * ez * (h * p2 - p4);
      V[3] -= c6 * (B0 * (fy * ez - fz * ey) - (B1 + B2) * ey + B3 * fy);
      V[4] -= c6 * (B0 * (fz * ex - fx * ez) + (B1 + B2) * ex - B3 * fx);
      V[5] -= c6 * B0 * (fx * ey - fy * ex);
    }
    if (fd > 3) {  // rotation <- torque
      const double et = ex * tx + ey * ty + ez * tz;
      const double C0 = p3 * (0.375 - 2.25 * ez2);
      const double C1 = -1.125 * p3 * et;
      const double C2 = 2.25 * ez * p3 * et;
      const double C3 = 2.25 * p3 * (ex * ty - ey * tx);
      V[3] += c6 * (C0 * tx + C1 * ex - C3 * ey);
      V[4] += c6 * (C0 * ty + C1 * ey + C3 * ex);
      V[5] += c6 * (C0 * tz + C1 * ez + C2);
    }
  }
  if (fd > 3) {
    // translation <- torque: transpose of the rotation<-force block of the
    // swapped pair, i.e. the same expressions with h -> 1-h and e_xy -> -e_xy.
    h = 1.0 - h;
    ex = -ex;
    ey = -ey;
    const double ecz = ex * ty - ey * tx;
    const double D0 = 0.75 * p2;
    const double D1 = (9.0 * h * ez2 * p2 + (1.5 - 15.0 * ez2) * p4) * ecz;
    const double D2 = -ez * (4.5 * h * p2 - 7.5 * p4) * ecz;
    const double D3 = -1.5 * ez * (h * p2 - p4);
    V[0] -= c6 * (D0 * (tz * ey - ty * ez) + D2 * ex - D3 * ty);
    V[1] -= c6 * (D0 * (tx * ez - tz * ex) + D2 * ey + D3 * tx);
    V[2] -= c6 * (D0 * (ty * ex - tx * ey) + D2 * ez + D1);
  }
}

// ---------------------------------------------------------------------------
// Unbounded RPY with Zuk et al. (2014) overlap forms -- cuda_functions.cu:205-356.
void free_pair(int vd, int fd, double xi, double yi, double zi, double ai, double xj, double yj,
               double zj, double aj, double mu, const double* F, double* V) {
  double rx = xi - xj, ry = yi - yj, rz = zi - zj;
  const double r = std::sqrt(rx * rx + ry * ry + rz * rz);
  const double rm1 = 1.0 / (r + 1e-20);
  const double dmin = std::fabs(ai - aj);
  const double amax = (ai > aj) ? ai : aj;
  rx *= rm1;
  ry *= rm1;
  rz *= rm1;
  double A, B;
  double rdot = rx * F[0] + ry * F[1] + rz * F[2];
  if (r > ai + aj) {
    A = rm1 * rm1 * (ai * ai + aj * aj) / 3.0;
    B = 1.0 - 3.0 * A;
    A += 1.0;
    const double t = rm1 / (8.0 * kPi * mu);
    A *= t;
    B *= t * rdot;
  } else if (r > dmin) {
    double t = 32.0 * r * r * r;
    A = dmin * dmin + 3.0 * r * r;
    A *= -A / t;
    A += 0.5 * (ai + aj);
    B = dmin * dmin - r * r;
    B *= 3.0 * B / t;
    t = 1.0 / (6.0 * kPi * mu * ai * aj);
    A *= t;
    B *= t * rdot;
  } else {
    A = 1.0 / (6.0 * kPi * mu * amax);
    B = 0.0;
  }
  V[0] += A * F[0] + B * rx;
  V[1] += A * F[1] + B * ry;
  V[2] += A * F[2] + B * rz;
  if (fd > 3) {
    if (r > ai + aj) {
      A = rm1 * rm1 / (8.0 * kPi * mu);
    } else if (r > dmin) {
      A = aj - ai + r;
      A *= rm1 * rm1 * A * (ai * ai + 2.0 * ai * (r + aj) - 3.0 * (aj - r) * (aj - r)) /
           (128.0 * kPi * mu * ai * aj * aj * aj);
    } else {
      A = (aj > ai) * r / (8.0 * kPi * mu * aj * aj * aj);
    }
    V[0] += A * (F[4] * rz - F[5] * ry);
    V[1] += A * (F[5] * rx - F[3] * rz);
    V[2] += A * (F[3] * ry - F[4] * rx);
  }
  if (vd > 3) {
    if (r > ai + aj) {
      A = rm1 * rm1 / (8.0 * kPi * mu);
    } else if (r > dmin) {
      A = ai - aj + r;
      A *= rm1 * rm1 * A * (aj * aj + 2.0 * aj * (r + ai) - 3.0 * (ai - r) * (ai - r)) /
           (128.0 * kPi * mu * aj * ai * ai * ai);
    } else {
      A = (ai > aj) * r / (8.0 * kPi * mu * ai * ai * ai);
    }
    V[3] += A * (F[1] * rz - F[2] * ry);
    V[4] += A * (F[2] * rx - F[0] * rz);
    V[5] += A * (F[0] * ry - F[1] * rx);
    if (fd > 3) {
      rdot = rx * F[3] + ry * F[4] + rz * F[5];
      if (r > ai + aj) {
        A = -rm1 * rm1 * rm1 / (16.0 * kPi * mu);
        B = -3.0 * rdot * A;
      } else if (r > dmin) {
        A = 5.0 * (r * r * r) * (r * r * r);
        B = (ai - aj) * (ai - aj) - r * r;
        A -= 27.0 * r * r * r * r * (ai * ai + aj * aj);
        B *= 3.0 * B;
        A += 32.0 * (r * r * r) * ((ai * ai * ai) + (aj * aj * aj));
        B *= ai * ai + aj * aj + 4.0 * ai * aj - r * r;
        A -= 9.0 * r * r * (ai * ai - aj * aj) * (ai * ai - aj * aj);
        A -= dmin * dmin * dmin * dmin * (ai * ai + aj * aj + 4.0 * ai * aj);
        const double t = 1.0 / (512.0 * kPi * mu * (r * r * r) * (ai * ai * ai) * (aj * aj * aj));
        A *= t;
        B *= t * rdot;
      } else {
        A = 1.0 / (8.0 * kPi * mu * amax * amax * amax);
        B = 0.0;
      }
      V[3] += A * F[3] + B * rx;
      V[4] += A * F[4] + B * ry;
      V[5] += A * F[5] + B * rz;
    }
  }
}

}  // namespace

extern "C" {

int oracle_abi_version() { return 1; }

void oracle_wall_self(int vd, int fd, double z, const double* F, double* V) {
  wall_self(vd, fd, z, F, V);
}

void oracle_wall_pair(int vd, int fd, const double* xi, const double* xj, const double* F,
                      double* V) {
  wall_pair(vd, fd, xi[0], xi[1], xi[2], xj[0], xj[1], xj[2], F, V);
}

void oracle_free_pair(int vd, int fd, const double* xi, double ai, const double* xj, double aj,
                      double mu, const double* F, double* V) {
  free_pair(vd, fd, xi[0], xi[1], xi[2], ai, xj[0], xj[1], xj[2], aj, mu, F, V);
}

// V_tgt (=|+=) M(tgt <- src) F_src by direct double loop (Mss/Mbb/Msb/Mbs_mult).
//   wall      : 1 = wall branch (a = 1 assumed, result scaled by 1/mu), 0 = unbounded
//   *_id      : global particle ids; equal ids mean the same particle (self term)
//   *_dof     : 3 or 6 (AoS stride of the force / velocity arrays)
//   accumulate: 0 overwrite, 1 add
//   nthreads  : OpenMP threads (<= 0: runtime default)
void oracle_apply(int wall, double mu, int n_tgt, int vd, const double* x_tgt,
                  const int64_t* id_tgt, double a_tgt, int n_src, int fd, const double* x_src,
                  const int64_t* id_src, double a_src, const double* f_src, double* v_tgt,
                  int accumulate, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n_tgt; ++i) {
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const double* xi = &x_tgt[3 * i];
    for (int j = 0; j < n_src; ++j) {
      const double* xj = &x_src[3 * j];
      double Fj[6] = {0, 0, 0, 0, 0, 0};
      for (int k = 0; k < fd; ++k) Fj[k] = f_src[fd * j + k];
      if (wall) {
        if (id_tgt[i] == id_src[j])
          wall_self(vd, fd, xi[2], Fj, acc);
        else
          wall_pair(vd, fd, xi[0], xi[1], xi[2], xj[0], xj[1], xj[2], Fj, acc);
      } else {
        free_pair(vd, fd, xi[0], xi[1], xi[2], a_tgt, xj[0], xj[1], xj[2], a_src, mu, Fj, acc);
      }
    }
    const double s = wall ? 1.0 / mu : 1.0;
    for (int k = 0; k < vd; ++k) {
      if (accumulate)
        v_tgt[vd * i + k] += s * acc[k];
      else
        v_tgt[vd * i + k] = s * acc[k];
    }
  }
}

double oracle_wtime() { return omp_get_wtime(); }
int oracle_max_threads() { return omp_get_max_threads(); }

}  // extern "C"
