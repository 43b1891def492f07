// precond.hpp -- per-particle right preconditioners shared by the single-system and
// batched solvers.
#pragma once
#include "../../include/cufcm_fil.h"

namespace cfm {

// o = P^-1 v for one particle with `dof` unknowns (3 or 6) at height z.
//   CFM_PC_NONE: identity.
//   CFM_PC_DIAG (or no wall): 6 pi mu a on forces, 8 pi mu a^3 on torques
//     (mobility_solver.cpp:1108-1114,1550-1556).
//   CFM_PC_SELF: inverse of the particle's own wall-corrected self-mobility block
//     (a = 1), which couples (F_x, T_y) and (F_y, T_x).
__host__ __device__ inline void particle_precond(int kind, int wall, int dof, double z, double a,
                                                 double mu, const double* v, double* o) {
  const double pi = 3.14159265358979323846;
  if (kind == CFM_PC_NONE) {
    for (int k = 0; k < dof; ++k) o[k] = v[k];
  } else if (kind == CFM_PC_DIAG || !wall) {
    const double ct = 6.0 * pi * mu * a, cr = 8.0 * pi * mu * a * a * a;
    for (int k = 0; k < dof; ++k) o[k] = (k < 3 ? ct : cr) * v[k];
  } else {
    const double c6 = 1.0 / (6.0 * pi), c8 = 1.0 / (8.0 * pi);
    const double h1 = 1.0 / z;
    const double h3 = h1 * h1 * h1, h4 = h3 * h1, h5 = h4 * h1;
    const double par = c6 * (1.0 - (9.0 * h1 - 2.0 * h3 + h5) * 0.0625) / mu;
    const double perp = c6 * (1.0 - (9.0 * h1 - 4.0 * h3 + h5) * 0.125) / mu;
    o[2] = v[2] / perp;
    if (dof == 3) {
      o[0] = v[0] / par;
      o[1] = v[1] / par;
    } else {
      const double cpl = c6 * 0.09375 * h4 / mu;
      const double rpar = c8 * (1.0 - 0.3125 * h3) / mu;
      const double rperp = c8 * (1.0 - 0.125 * h3) / mu;
      const double det = par * rpar - cpl * cpl;
      // [vx, wy] = [[par, cpl], [cpl, rpar]] [Fx, Ty]
      o[0] = (rpar * v[0] - cpl * v[4]) / det;
      o[4] = (par * v[4] - cpl * v[0]) / det;
      // [vy, wx] = [[par, -cpl], [-cpl, rpar]] [Fy, Tx]
      o[1] = (rpar * v[1] + cpl * v[3]) / det;
      o[3] = (par * v[3] + cpl * v[1]) / det;
      o[5] = v[5] / rperp;
    }
  }
}

}  // namespace cfm
