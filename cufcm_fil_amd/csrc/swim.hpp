// swim.hpp -- free swimmers: rigid bodies carrying the filaments, the body state among
// the Broyden unknowns (swim.hip).
#pragma once
#include "ctx.hpp"
#include "fil.hpp"

namespace cfm {

struct SwimK {
  int nbody = 0, fpb = 0, nfil = 0, nseg = 0;  // bodies, filaments per body, filaments, NSEG
  int nb = 0;                                  // blobs per body
  int num_euler = 1;
  double dl = 2.2, dt = 0.1, tol = 1e-4, end_force = 0.0;
  // > 0: the body is a sphere of this radius, and the approximate Jacobian carries the
  // RPY coupling of sphere and segments (swimmer.cpp:800-877); a_seg, mu for it
  double sphere_R = 0.0, a_seg = 1.0, mu = 1.0;
  // per body, AoS: position 3 (x, x_-1, x_-2), Lie algebra 3 (u, u_-1), quaternion 4
  // (q, q_-1), R(q) at the step's initial guess 9 (row-major)
  double *x = nullptr, *xm1 = nullptr, *xm2 = nullptr, *u = nullptr, *um1 = nullptr;
  double *q = nullptr, *qm1 = nullptr, *Qinit = nullptr;
  const double* fref = nullptr;  // 3 per filament: base offset in the body frame
  double* vel = nullptr;         // 6 per body: V, Omega of the last blob solve
  double* force = nullptr;       // 6 per body: load the blobs must balance
  const double* fext = nullptr;  // 6 per body: external force and torque
  const double* mob = nullptr;   // 36: body-frame body mobility (row-major)
  double* BJ = nullptr;          // per filament 6 x 6 NSEG: B J^-1
  double* Sf = nullptr;          // 36 per filament: B J^-1 A
  double* S0 = nullptr;          // 36 per body: body block of J0
  double* Sinv = nullptr;        // 36 per body: inverse Schur complement
  double* tb = nullptr;          // 6 per body: body part of the last J0^-1 product
};

// Blob forces among the Broyden unknowns (the reference's SOLVER_TYPE 0,
// config.hpp:360-364): every blob's force ([3][m]: component k of blob j at k m + j) and
// its two previous steps; free bodies (nbody > 0, nb blobs each, their unknowns and rows
// as for SwimK) or a body held fixed (nbody = 0).  Broyden vector: the filaments, then
// 6 per free body, then 3 per blob.
struct BxK {
  int m = 0, nbody = 0, nb = 0, num_euler = 1;
  double dt = 0.1, tol = 1e-4, mu = 1.0, a = 1.0;
  double *fb = nullptr, *fb_m1 = nullptr, *fb_m2 = nullptr;
  const double* ref = nullptr;  // free bodies: body-frame offsets of one body's blobs (AoS)
};

Status swim_guess(cfm_ctx* c, const SwimK& S, int nt);
Status swim_update(cfm_ctx* c, const SwimK& S, const double* upd_body);
Status swim_force(cfm_ctx* c, const SwimK& S, const FilK& K, const double* rep, const double* boff);
Status swim_errors(cfm_ctx* c, const SwimK& S, int nt, double* err_body, double* bad);
Status swim_jacobian(cfm_ctx* c, const SwimK& S, const FilK& K, int nt, int* status_dev);
Status swim_jinv_mult(cfm_ctx* c, const SwimK& S, const FilK& K, const double* in, double* out,
                      double scale, double* tmp);

Status bx_guess(cfm_ctx* c, const BxK& X, int nt);
Status bx_update(cfm_ctx* c, const BxK& X, const double* upd);  // upd: 3 per blob
// Body rows (6 per free body) into err_body, blob rows (3 per blob) into err_blob; the
// badness of each body and of each 256 blobs into bad[0 .. nbody + ceil(m / 256)).
Status bx_errors(cfm_ctx* c, const BxK& X, const SwimK& S, int nt, double* err_body, double* err_blob,
                 double* bad);
Status bx_jacobian(cfm_ctx* c, const BxK& X, const SwimK& S, const FilK& K, int nt, int* status_dev);
Status bx_jinv_mult(cfm_ctx* c, const BxK& X, const SwimK& S, const FilK& K, int nt, const double* in,
                    double* out, double scale, double* tmp);

}  // namespace cfm
