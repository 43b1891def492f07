// body.hpp -- what the implicit step needs from the free rigid bodies (body.hip).
#pragma once
#include <vector>

#include "ctx.hpp"

namespace cfm {

struct BodyInfo {
  int n_body = 0, nb = 0;    // bodies, blobs per body
  std::vector<double> x, q;  // centres (3 n_body), quaternions (4 n_body) of the last placement
  std::vector<double> ref;   // body-frame offsets of one body's blobs (3 nb, AoS)
  double mob[36];            // body-frame body mobility (identity until made or set)
};

Status body_info(cfm_ctx* c, BodyInfo* out);
// Places every blob from body centres bx (3 n_body) and quaternions bq (4 n_body) held in
// device memory, with the rotated reference matrices; invalidates caches of blob products.
Status body_place_device(cfm_ctx* c, const double* bx, const double* bq);
// Keep CFM_PC_DENSE of the body system across body motions: built once, then applied in
// the frame it was built in (exact without a wall, where rigid motions leave the blob
// mobility invariant up to rotation; a preconditioner above one).
void body_freeze_dense(cfm_ctx* c, bool on);
// Rotated blob offsets R(q_b) ref of the last placement ([3][n_blob], device), or null.
const double* body_offsets(cfm_ctx* c);
// Body-frame offsets of one body's blobs on the device (3 nb, AoS; set by the first
// placement from device-held body states), or null.
const double* body_ref_device(cfm_ctx* c);
// Prescribed rotation (cfm_body_set_motion): Omega_b is held at zero.
bool body_rotation_prescribed(cfm_ctx* c);
// Prescribed rotation: zero the torque rows (the prescribed Omega_b) of a body-system
// right-hand side in device memory; nothing for free bodies.
Status body_prescribed_rhs(cfm_ctx* c, double* rhs);

// Dense blob block for systems that embed the bodies (cilia on free bodies): the
// resident inverse P of M_bb, its rigid columns P K and R = K^T P K, built for the
// current placement or, frozen, once and applied in the frame of the build.
Status body_dense_ready(cfm_ctx* c);
// y_b = P x_b; blob blocks with component k of blob j at [k ld + j].
Status body_dense_blob_apply(cfm_ctx* c, const double* x, size_t ldx, double* y, size_t ldy);
// y_b += P K U, U: (V, Omega) of every body.
Status body_dense_rigid_add(cfm_ctx* c, const double* U, double* y, size_t ldy);
// R = K^T P K, (6 n_body)^2 row-major, on the host.
Status body_dense_resistance(cfm_ctx* c, std::vector<double>* R);

}  // namespace cfm
