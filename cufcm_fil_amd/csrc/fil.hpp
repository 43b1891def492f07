// fil.hpp -- device state of the filaments and the kernels of the implicit step.
#pragma once
#include "ctx.hpp"

namespace cfm {

// Longest filament the implicit step takes: its (6 NSEG)^2 FP64 approximate Jacobian
// is assembled in one workgroup's LDS (127,008 bytes at 21 segments).
constexpr int kMaxFilSeg = 21;

// Kernel-side view of the filament state.  Per-segment arrays are structure of
// arrays over the nseg_total segments (component-major, e.g. q + k*nseg_total);
// per-filament arrays over nfil.  lam row 0 of each filament holds the tether
// multiplier, row k >= 1 the inextensibility multiplier of link (k-1, k).
struct FilK {
  int nseg = 0, nfil = 0;
  size_t nseg_total = 0;
  int num_euler = 1;
  double dl = 2.2, dt = 0.1, kb = 1800.0, kt = 1800.0, mu = 1.0, a = 1.0;
  double end_force = 0.0, rep_factor = 2.0, tol = 1e-4;
  double* x = nullptr;   // segment positions (the context's particle arrays)
  double* y = nullptr;
  double* z = nullptr;
  double *q = nullptr, *qm1 = nullptr;           // 4 x nseg_total
  double *u = nullptr, *um1 = nullptr;           // 3 x nseg_total
  double *xm1 = nullptr, *xm2 = nullptr;         // 3 x nseg_total
  double *lam = nullptr, *lam_m1 = nullptr, *lam_m2 = nullptr;  // 3 x nseg_total
  double *clamp = nullptr, *clamp_m1 = nullptr, *clamp_m2 = nullptr;  // 3 x nfil
  double* base = nullptr;  // 3 x nfil
  double* pert = nullptr;  // 3 x nfil
  double* fint = nullptr;  // 6 x nseg_total internal loads (no barrier)
  double* jinv = nullptr;  // nfil x (6 nseg)^2, transposed inverse Jacobians
  // Free bodies (null when the bases are fixed anchors): filament f belongs to body
  // f / fpb, its base sits at bx + R(bq) fref[f] and its first segment carries the
  // body's Lie-algebra element bu (AoS per body: 3, 4, 3 values).
  const double *bx = nullptr, *bq = nullptr, *bu = nullptr, *fref = nullptr;
  int fpb = 0;
  double* clampblk = nullptr;  // 18 per filament: Lie rows 0 and 1 of J against u_0
  // Equal independent systems of fpm consecutive filaments each (0: one system): the
  // barrier skips pairs of segments of different systems.
  int fpm = 0;
};

// Hashed uniform cell list over segments and blobs for the barrier: cells of edge
// equal to the barrier cut-off, hashed into nbucket (power of two) buckets; ids holds
// the particle ids grouped by bucket, ascending inside each bucket, and start the
// bucket offsets (nbucket + 1 entries).
struct BarrierGrid {
  int nbucket = 0;
  DevBuf<int> start, cursor, ids, key, blocksum;
  Status ensure(int n_particles);
  void release() {
    blocksum.release();
    start.release();
    cursor.release();
    ids.release();
    key.release();
    nbucket = 0;
  }
};

Status fil_initial_guess(cfm_ctx* c, const FilK& K, int nt);
Status fil_forces(cfm_ctx* c, const FilK& K, int nt);
Status fil_barrier(cfm_ctx* c, const FilK& K, BarrierGrid& grid);
// Residual of the filament rows; bad_part[nb .. nb + extra) already hold the badness
// of further rows (free bodies), nb = ceil(nseg_total / 256).
Status fil_errors(cfm_ctx* c, const FilK& K, int nt, double* err, double* bad_part,
                  double* bad_out, int extra = 0);
// Repulsion on every blob from the segments and from the blobs of other bodies, on the
// cell list fil_barrier built (rep: 3 x n_blob).
Status fil_blob_barrier(cfm_ctx* c, const FilK& K, BarrierGrid& grid, int nb_per_body, double* rep);
// Builds and inverts every filament's approximate Jacobian on stream s (the context's
// stream when null).
Status fil_jacobian(cfm_ctx* c, const FilK& K, int nt, int* status_dev, hipStream_t s = nullptr);
Status fil_jinv_mult(cfm_ctx* c, const FilK& K, const double* in, double* out, double scale);
Status fil_update(cfm_ctx* c, const FilK& K, const double* upd);
// Badness of each of E equal systems from the residual err (Lm rows each, laid out as
// fil_errors writes them): the max of |e_x| / (DL TOL) and |e_u| / TOL, NaN if a row is
// not finite; bad[e].
Status fil_member_badness(cfm_ctx* c, const FilK& K, const double* err, int Lm, int E, double* bad);

}  // namespace cfm
