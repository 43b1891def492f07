/* cufcm_fil.h -- C ABI of the MI355X wall-RPY mobility library.
 *
 * One context = one set of particles (filament segments, then rigid-body blobs)
 * on one HIP device with one HIP stream.  Every call is synchronous on return,
 * returns 0 on success and a nonzero CFM_E* code on failure (the message is
 * available from cfm_last_error).  Host arrays are caller-owned, double, in the
 * reference's interleaved (AoS) layout; the library copies them and keeps no
 * caller pointer.  A context must not be used from two threads at once;
 * separate contexts are independent.
 *
 * Reference interface each entry point replaces (racksa/CUFCM_FIL):
 *   mobility_solver (src/cilia/mobility/mobility_solver.hpp:26-46) pure virtuals,
 *   implemented by rpy_mobility_solver (src/cilia/mobility/rpy_mobility_solver.cu).
 */
#ifndef CUFCM_FIL_H
#define CUFCM_FIL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFM_ABI_VERSION 1

/* status codes */
#define CFM_OK 0
#define CFM_EINVAL 1    /* bad argument or configuration */
#define CFM_EHIP 2      /* HIP runtime error */
#define CFM_ENOMEM 3    /* allocation failed */
#define CFM_ESTATE 4    /* call out of order (e.g. apply before positions) */
#define CFM_ENOCONV 5   /* iterative solver did not reach its tolerance */
#define CFM_ENONFINITE 6 /* NaN/Inf in inputs or iterates */

/* precision of the O(N^2) pair sums */
#define CFM_FP64 0      /* FP64 pair arithmetic */
#define CFM_FP32C 1     /* FP32 pair arithmetic, compensated (two-sum) accumulation */

/* which block of the grand mobility to apply (cfm_apply) */
#define CFM_SS 0   /* v_seg  = Mss f_seg   (evaluate_segment_segment_mobility, rpy_mobility_solver.cu:258-265) */
#define CFM_SB 1   /* v_seg += Msb f_blob  (evaluate_segment_blob_mobility,    rpy_mobility_solver.cu:267-283) */
#define CFM_BB 2   /* v_blob = Mbb f_blob  (evaluate_blob_blob_mobility,       rpy_mobility_solver.cu:285-301) */
#define CFM_BS 3   /* v_blob = Mbs f_seg   (evaluate_blob_segment_mobility,    rpy_mobility_solver.cu:303-319) */
#define CFM_FULL 4 /* both species from both species (evaluate_full_mobility, rpy_mobility_solver.cu:330-338) */

typedef struct cfm_ctx cfm_ctx;

typedef struct cfm_config {
  int32_t n_seg;      /* filament segments (NSWIM*NFIL*NSEG in the reference) */
  int32_t n_blob;     /* rigid-body blobs (NSWIM*NBLOB) */
  int32_t seg_dof;    /* 6: force+torque per segment; 3: force only (PRESCRIBED_CILIA) */
  int32_t wall;       /* 1: no-slip wall at z = 0 (INFINITE_PLANE_WALL), 0: unbounded */
  double a_seg;       /* segment radius (RSEG); must be 1 with wall = 1 */
  double a_blob;      /* blob radius (RBLOB); must be 1 with wall = 1 */
  double mu;          /* viscosity (MU); velocities scale as 1/mu */
  int32_t precision;  /* CFM_FP64 or CFM_FP32C */
  int32_t device;     /* HIP device ordinal */
  int32_t tile_split; /* source-range split across workgroups: > 0 fixed; 0 = size heuristic
                         (deterministic); -1 = autotuned: the first product of each shape
                         times the candidate splits on the device and keeps the fastest */
  int32_t reserved;
} cfm_config;

/* Library ABI version (CFM_ABI_VERSION). */
int cfm_abi_version(void);

/* Create / destroy a context.  Replaces allocate_host_memory/allocate_device_memory
 * and free_host_memory/free_device_memory (rpy_mobility_solver.cu:14-117). */
int cfm_create(const cfm_config* cfg, cfm_ctx** out);
int cfm_destroy(cfm_ctx* ctx);

/* Message of the last failure on this context ("" if none).  ctx may be NULL
 * for failures of cfm_create. */
const char* cfm_last_error(const cfm_ctx* ctx);

/* Upload positions: x_seg[3*n_seg], x_blob[3*n_blob] (either may be NULL when its
 * count is 0).  Replaces copy_segment_positions_to_device / copy_blob_positions_to_device
 * (rpy_mobility_solver.cu:119-148). */
int cfm_set_positions(cfm_ctx* ctx, const double* x_seg, const double* x_blob);

/* Switch the pair-sum precision of an existing context (CFM_FP64 / CFM_FP32C). */
int cfm_set_precision(cfm_ctx* ctx, int32_t precision);

/* Apply a mobility block: f_seg[seg_dof*n_seg], f_blob[3*n_blob] in,
 * v_seg[seg_dof*n_seg], v_blob[3*n_blob] out, with the write/accumulate rule of
 * the reference call named next to each CFM_* block id.  Replaces
 * copy_*_forces_to_device + evaluate_*_mobility + copy_*_velocities_to_host +
 * wait_for_device (rpy_mobility_solver.cu:150-338). */
int cfm_apply(cfm_ctx* ctx, int32_t which, const double* f_seg, const double* f_blob,
              double* v_seg, double* v_blob);

/* Time `reps` launches of the pair kernel(s) for `which` on the context's stream
 * with HIP events, using the forces most recently passed to cfm_apply.  Writes the
 * mean milliseconds per launch.  (Measurement helper; no reference counterpart.) */
int cfm_time_apply(cfm_ctx* ctx, int32_t which, int32_t reps, double* ms_per_apply);

/* Allocation-free variant of cfm_apply.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  f_seg, f_blob, v_seg, v_blob are
 * DEVICE arrays in cfm_apply's layout and write/accumulate rules; positions come from
 * cfm_set_positions as for cfm_apply.  The pair kernel's split partial sums go to one
 * static device array of CFM_NOALLOC_SCRATCH_BYTES per GPU (held until the process
 * ends) instead of context memory; the split is the context's fixed tile_split when
 * > 0 -- rejected with CFM_EINVAL when split x 6 x (n_seg + n_blob) doubles exceed the
 * array -- and otherwise the size heuristic, lowered to fit.  Single-species wall
 * products run the one-directional kernels (those of cfm_set_pair_symmetry(ctx, 0)),
 * whose partial sums fit the array; they agree with cfm_apply's symmetric default to
 * rounding.  The call allocates
 * nothing and returns after the context's stream has finished.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other contexts, other host
 * threads) race and must be serialised by the caller. */
#define CFM_NOALLOC_SCRATCH_BYTES (32u << 20)
int cfm_apply_noalloc(cfm_ctx* ctx, int32_t which, const double* f_seg, const double* f_blob,
                      double* v_seg, double* v_blob);


/* ---- linear solves (GMRES) ------------------------------------------------ */

/* linear systems */
#define CFM_SYS_SEG 0   /* Mss f_seg = v_seg  (segment resistance problem) */
#define CFM_SYS_BLOB 1  /* Mbb f_blob = v_blob (rigid-body blob forces, body held fixed) */
#define CFM_SYS_BODY 2  /* free rigid bodies: blob forces and body velocities (cfm_body_*) */

/* right preconditioners P^-1 */
#define CFM_PC_NONE 0
#define CFM_PC_DIAG 1   /* scalar 6 pi mu a (and 8 pi mu a^3 on torques): mobility_solver.cpp:1108-1114,1550-1556 */
#define CFM_PC_SELF 2   /* inverse of each particle's own (wall-corrected) self-mobility block */
#define CFM_PC_NEAR 3   /* blob system only: local inverse over each blob and its 16 nearest
                           blobs (near-field block preconditioner), plus the learned correction
                           when one is loaded (cfm_gnn_*) */
#define CFM_PC_DENSE 4  /* blob system only: resident FP32 inverse of the whole blob-blob
                           mobility (FP64 Cholesky, built on first use and kept until the
                           blob positions change; 4 (3 n_blob)^2 bytes of device memory) */

/* arithmetic of the solve */
#define CFM_SOLVE_FP64 0   /* FP64 pair sums throughout */
#define CFM_SOLVE_MIXED 1  /* FP32C pair sums inside FP64 iterative refinement */

typedef struct cfm_solve_opts {
  int32_t system;     /* CFM_SYS_SEG or CFM_SYS_BLOB */
  int32_t precond;    /* CFM_PC_* */
  int32_t mode;       /* CFM_SOLVE_FP64 or CFM_SOLVE_MIXED */
  int32_t restart;    /* Krylov vectors per cycle (0: 50) */
  int32_t max_iter;   /* cap on Arnoldi steps over all cycles (0: 1000) */
  int32_t use_guess;  /* 1: x holds the initial guess, rejected if its residual exceeds
                         ||rhs|| (mobility_solver.cpp:2296-2417); 0: start from zero */
  double rtol;        /* target ||rhs - A x|| / ||rhs||, checked with an FP64 product */
} cfm_solve_opts;

typedef struct cfm_solve_stats {
  int32_t iterations;   /* Arnoldi steps (matrix products inside Krylov cycles) */
  int32_t refinements;  /* FP64 residual evaluations after the first */
  int32_t converged;    /* 1 if rel_residual <= rtol */
  int32_t reserved;
  double rel_residual;  /* final ||rhs - A x|| / ||rhs|| from an FP64 product */
  double initial_rel_residual; /* same for the initial guess (1 for a zero guess) */
} cfm_solve_stats;

/* Solve A x = rhs by right-preconditioned restarted GMRES on the device (FP64
 * Arnoldi with classical Gram-Schmidt applied twice, Givens rotations, back
 * substitution).  rhs and x are host AoS arrays of dof*n values (dof = seg_dof for
 * CFM_SYS_SEG, 3 for CFM_SYS_BLOB).  Returns CFM_ENOCONV (with x holding the best
 * iterate) when rtol is not reached within max_iter.  Replaces
 * mobility_solver::solve_linear_system (mobility_solver.cpp:2187-2669). */
int cfm_gmres(cfm_ctx* ctx, const cfm_solve_opts* opts, const double* rhs, double* x,
              cfm_solve_stats* stats);

/* out = P^-1 in for one system (host AoS arrays).  Replaces
 * mobility_solver::apply_preconditioner_new (mobility_solver.cpp:1102-1594). */
int cfm_apply_preconditioner(cfm_ctx* ctx, int32_t system, int32_t precond, const double* in,
                             double* out);
/* y = A x for a whole system (mobility_solver.cpp:1596-1881 system_matrix_mult): the pair
 * products of the system in `precision` (CFM_FP64 / CFM_FP32C) plus, for CFM_SYS_BODY,
 * the coupling rows and columns -K U and -K^T F in FP64.  Vectors as for cfm_gmres. */
int cfm_system_apply(cfm_ctx* ctx, int32_t system, int32_t precision, const double* x, double* y);

/* ---- implicit filament step (Broyden) ------------------------------------- */
/* Tethered elastic filaments with a clamped base and a follower force at the tip
 * (the reference's INSTABILITY_CILIA model), optionally next to a rigid body that
 * is held fixed (prescribed zero velocity, the reference's default
 * BODY_VELOCITY_TYPE 1).  Segments are numbered filament by filament, base first;
 * n_seg = nfil * nseg. */

#define CFM_GUESS_PREVIOUS 0   /* blob-force GMRES guess = previous solution (mobility_solver.cpp:2296-2388) */
#define CFM_GUESS_LEARNED 1    /* history projection; before any history exists, the network's guess
                                  (W + Z c) b with the learned coarse space (CFM_COARSE_LEARNED) */
#define CFM_GUESS_PROJECTED 2  /* minimal-residual combination of the last 16 (rhs, solution) pairs
                                  (the blob matrix is constant while the body is held fixed) */

typedef struct cfm_fil_params {
  int32_t nseg;            /* segments per filament (NSEG), <= 21 */
  int32_t max_broyden;     /* MAX_BROYDEN_ITER (config.hpp:357) */
  int32_t num_euler_steps; /* NUM_EULER_STEPS (config.hpp:382) */
  int32_t barrier;         /* 1: add barrier forces before each product (apply_interparticle_forces) */
  double dl;               /* DL: segment spacing */
  double dt;               /* DT: time step */
  double kb, kt;           /* bending and twist moduli (config.hpp:320-321) */
  double end_force;        /* END_FORCE_MAGNITUDE (follower force at the tip) */
  double repulsive_factor; /* REPULSIVE_FORCE_FACTOR (config.hpp:462) */
  double tol;              /* Broyden TOL (config.hpp:358) */
  double gamma;            /* JACOBIAN_CONFIDENCE_FACTOR (config.hpp:355) */
  int32_t pair_precision;  /* CFM_FP64 / CFM_FP32C for the velocity products */
  int32_t guess;           /* CFM_GUESS_* for the blob-force GMRES */
  cfm_solve_opts blob_solve; /* inner GMRES on the body blobs (system field ignored) */
  int32_t overlap;         /* 1: M_ss f_seg runs on a second stream while the blob solve runs
                              (bit-identical results; the blob solve's latency-bound kernels
                              fill the pair kernel's gaps: 4-5% faster steps at 26,000
                              segments + 6,000 blobs) */
  int32_t broyden_only;    /* 1: the blob forces -- and, after cfm_fil_attach_bodies, the bodies'
                              positions and orientations -- are Broyden unknowns, every product
                              is taken from the current loads and no blob system is solved (the
                              reference's SOLVER_TYPE 0, config.hpp:360-364, 443-447); the
                              Broyden vector is the filaments, 6 per free body, 3 per blob.
                              blob_solve, guess and overlap are then unused.  0: GMRES for the
                              blob system inside every Broyden iteration (SOLVER_TYPE 1). */
} cfm_fil_params;

typedef struct cfm_step_stats {
  int32_t broyden_iterations;
  int32_t gmres_iterations;  /* Arnoldi steps summed over the step's blob solves */
  int32_t gmres_solves;
  int32_t converged;         /* 1 if the Broyden residual met TOL */
  double final_error;        /* max of |e_x|/(DL TOL), |e_u|/TOL at exit (<= 1 when converged) */
  double max_gmres_rel_residual;  /* worst FP64 relative residual of the step's blob solves */
  double seconds;            /* host wall time of the step (stream synchronised) */
} cfm_step_stats;

/* Set up the filament state on a context whose segment positions were given with
 * cfm_set_positions: q_seg[4*n_seg] unit quaternions (s, v0, v1, v2) per segment,
 * pert[3*nfil] (or NULL) the force added to segment 4 of each filament at nt = 0.
 * Replaces filament::initial_setup (filament.cpp:62-627, elastic branch). */
int cfm_fil_init(cfm_ctx* ctx, const cfm_fil_params* params, const double* q_seg,
                 const double* pert);

/* One implicit BDF time step nt (backward Euler for nt < num_euler_steps) solved by
 * Broyden's method: the loop of cilia_sim_main.cu:405-552 with compute_velocities
 * (mobility_solver.cpp:2673-2758), compute_errors (:2760-3045), find_update /
 * end_of_iter (broyden_solver.cpp:71-174) and the filament updates, all on the
 * device.  Returns CFM_ENONFINITE if the iteration diverges. */
int cfm_fil_step(cfm_ctx* ctx, int32_t nt, cfm_step_stats* stats);

/* Ensembles: the context's filaments form `members` equal independent systems (member e
 * owns filaments [e F, (e+1) F), F = nfil / members) -- the runs of a parameter sweep,
 * bisection or periodic-orbit search (pyfile/driver/func.py:544, one executable per run in
 * the reference).  Wall-bounded filaments only (no blobs or bodies).  From the next
 * cfm_fil_step on, one product launch gives every member its own M_ss f (no coupling
 * between members; the barrier skips pairs of different members) and each member runs its
 * own Broyden iteration (its own low-rank pairs, its own convergence test; a converged
 * member is held while the others iterate): every member's state is that of the same
 * system stepped alone.  The step's stats report the largest iteration count, converged
 * when every member converged and the largest final error; cfm_fil_member_stats gives
 * each member's.  members = 1 returns to one system. */
int cfm_fil_set_members(cfm_ctx* ctx, int32_t members);
int cfm_fil_member_stats(cfm_ctx* ctx, cfm_step_stats* per_member /* members */);

/* Read / overwrite the filament state (any pointer may be NULL).  Arrays are AoS:
 * x[3*n_seg], q/qm1[4*n_seg], u/um1/xm1/xm2/lam[3*n_seg] (lam row 0 of a filament
 * is its tether multiplier), clamp[3*nfil]. */
int cfm_fil_get_state(cfm_ctx* ctx, double* x, double* q, double* u, double* lam, double* clamp);
int cfm_fil_set_state(cfm_ctx* ctx, const double* q, const double* qm1, const double* u,
                      const double* um1, const double* lam, const double* clamp,
                      const double* xm1, const double* xm2);

/* Pieces of the step, exposed for parity checks:
 *   internal loads f_seg[6*n_seg] of the current state (internal_forces_and_torques),
 *   barrier forces added to zero loads f_seg[3*n_seg] (apply_interparticle_forces),
 *   inverse approximate Jacobians jinv[nfil*(6 nseg)^2], row-major per filament
 *     (filament::invert_approx_jacobian; uses the current internal loads),
 *   BDF residual err[6*n_seg] for given velocities v_seg[6*n_seg] (compute_errors)
 *     and its scaled max norm,
 *   state update by upd[6*n_seg] (swimmer::update), initial guess of step nt. */
int cfm_fil_internal_forces(cfm_ctx* ctx, int32_t nt, double* f_seg);
int cfm_barrier_forces(cfm_ctx* ctx, double* f_seg);
int cfm_fil_jacobian_inverse(cfm_ctx* ctx, int32_t nt, double* jinv);
int cfm_fil_errors(cfm_ctx* ctx, int32_t nt, const double* v_seg, double* err, double* badness);
/*   (cfm_fil_errors refuses a broyden_only context: its residual needs the blob velocities) */
int cfm_fil_update(cfm_ctx* ctx, const double* upd);
int cfm_fil_initial_guess(cfm_ctx* ctx, int32_t nt);

/* Free swimmers (BODY_VELOCITY_TYPE 0 with the blob system solved by GMRES): the
 * bodies of cfm_body_init / cfm_body_set_state carry the filaments, n_fil / n_body
 * consecutive filaments per body, each attached at its current first segment (kept
 * rigidly in the body frame; the first segment turns with the body,
 * filament.cpp:653-678).  From the next cfm_fil_step on, each body's position and
 * orientation are Broyden unknowns (rigid_body.cpp:410-544; body rows of
 * mobility_solver.cpp:3003-3040), the blob solve is CFM_SYS_BODY with the body load
 * (swimmer.cpp:470-535: minus the segments' internal loads about the centre, plus the
 * repulsion on the blobs, plus f_ext) as its last rows, and J0^-1 includes the body
 * Schur complement (swimmer.cpp:688-909, 1022-1068) built with the body mobility of
 * cfm_body_reference_matrices.  f_ext: 6 per body (force, torque; may be null).  The
 * Broyden vectors of cfm_fil_errors / cfm_fil_update then hold 6 n_seg + 6 n_body
 * values (bodies last: position increment, Lie-algebra increment).  Needs equal blob
 * and segment radii and a blob preconditioner NONE, DIAG or SELF; the blob-force guess
 * becomes the previous solution.  With cfm_fil_params.broyden_only the blob forces are
 * Broyden unknowns too (SOLVER_TYPE 0): the body rows become the force and torque
 * balance -f + sum of the blob forces and their torques about the centre, 3 blob rows
 * V + Omega x (R r) - v_blob follow (mobility_solver.cpp:2768-2999), J0^-1 takes the body
 * Schur complement of swimmer.cpp:575-695 and the blob block -6 pi mu a
 * (swimmer.cpp:929-1018), and the vectors hold 6 n_seg + 6 n_body + 3 n_blob values. */
int cfm_fil_attach_bodies(cfm_ctx* ctx, const double* f_ext);
/* Body centres (3 n_body), quaternions (4), Lie-algebra elements of the step (3) and the
 * last solve's (V, Omega) (6); each pointer may be null. */
int cfm_fil_get_body_state(cfm_ctx* ctx, double* x, double* q, double* u, double* vel);
/* Blob forces (3 n_blob, AoS): the Broyden state with broyden_only, else the last blob
 * solve's (the reference's f_blobs_host / rigid_body::blob_forces). */
int cfm_fil_get_blob_forces(cfm_ctx* ctx, double* f_blob);
/* The bodies are spheres of this radius (0: off, the default): the approximate Jacobian's
 * body rows then also carry the RPY coupling of sphere and segments -- the unequal-sphere
 * translation-translation, translation-rotation and rotation-rotation blocks of every
 * segment (swimmer.cpp:800-877, surface-of-revolution bodies whose generatrix is a
 * sphere).  Changes Broyden's convergence only, not the converged state. */
int cfm_fil_set_body_sphere(cfm_ctx* ctx, double radius);

/* ---- measurement --------------------------------------------------------- */
typedef struct cfm_profile_stats {
  int64_t launches;        /* pair-kernel launches recorded since the last reset */
  double kernel_ms;        /* summed HIP-event durations of those launches */
  double alg_flops;        /* algorithmic flops of those launches (FLOPS_PAIR_* x pairs) */
  double alg_bytes;        /* algorithmic HBM bytes (positions + loads in, velocities out) */
  int64_t launches_fp32;   /* of which FP32C, one-directional tiled kernel */
  double kernel_ms_fp32;
  double alg_flops_fp32;
  int64_t launches_sym;    /* of which FP32C products that evaluate each unordered pair once
                              (kernel and finalisation; flops counted for both directions,
                              as the one-directional kernel would) */
  double kernel_ms_sym;
  double alg_flops_sym;
} cfm_profile_stats;

/* Algorithmic flops per pair of the wall branch (add/sub/mul = 1, FMA = 2,
 * rsqrt = 1, div = 1), counted on the canonical formula of cuda_functions.cu:65-201. */
#define CFM_FLOPS_PAIR_66 321
#define CFM_FLOPS_PAIR_63 197
#define CFM_FLOPS_PAIR_36 203
#define CFM_FLOPS_PAIR_33 136

/* Record every pair-kernel launch between HIP events on the context's stream
 * (enable = 1), stop (0).  Enabling resets the counters. (No reference counterpart.) */
int cfm_profile(cfm_ctx* ctx, int32_t enable);
int cfm_profile_read(cfm_ctx* ctx, cfm_profile_stats* stats);
/* Wall products of one species from itself over all of its targets (segments from
 * segments, blobs from blobs; FP32C and FP64) evaluate each unordered pair once,
 * applying its block both ways (on = 1, the default); 0 restores the one-directional
 * kernel (for comparison).  (No reference counterpart.) */
int cfm_set_pair_symmetry(cfm_ctx* ctx, int32_t on);

/* ---- coarse correction of the near-field preconditioner (no reference counterpart) ----
 * CFM_PC_NEAR applies the local inverse W of each blob's 17-blob neighbourhood; its
 * iterations are dominated by collective modes no local block holds (rigid and
 * linear motions of whole bodies, and the near-null "squeeze" of a closed surface).
 * A coarse correction over k vectors Z of the blob system adds
 *     y = W r + Z c,  c = argmin_c ||(r - M W r) - M Z c||^2 + delta ||c||^2
 * (the residual of W r minimised over span(M Z)); M Z, M^2 Z and the Gram inverse are
 * built once per blob geometry with FP64 products, so an application costs 2k dot
 * products and k axpys.  CFM_COARSE_LINEAR: the 12 linear fields of the blob positions
 * (3 directions x {1, y_x, y_y, y_z}, y = (x - centroid) / radius).  CFM_COARSE_LEARNED:
 * those plus the learned network's coarse vectors (cfm_gnn_*; rebuilt after training). */
#define CFM_COARSE_NONE 0
#define CFM_COARSE_LINEAR 1
#define CFM_COARSE_LEARNED 2
int cfm_nearfield_coarse(cfm_ctx* ctx, int32_t mode);

/* ---- learned near-field correction (no reference counterpart) -------------- */
/* A message-passing network on the blobs' 16-nearest-neighbour graph predicts a
 * 3x3 correction block per edge; with it enabled, CFM_PC_NEAR applies the analytic
 * near-field local inverse plus the learned correction.  Dense layers run on the
 * matrix cores (BF16 operands, FP32 accumulation); training minimises the
 * preconditioned residual mean_t ||b_t - A P^-1 b_t||^2 / ||b_t||^2 over recorded
 * (rhs, solution) pairs of the plain solver.  Plain GMRES stays the judge: every
 * solve still checks its FP64 residual against rtol. */
int cfm_gnn_init(cfm_ctx* ctx, uint32_t seed, int32_t capacity);   /* capacity: training pairs kept */
int cfm_gnn_num_params(cfm_ctx* ctx);
int cfm_gnn_record(cfm_ctx* ctx, int32_t on);          /* record (rhs, solution) of every blob solve of
                                                          cfm_fil_step */
int cfm_gnn_samples(cfm_ctx* ctx);                     /* recorded pairs available */
int cfm_gnn_add_samples(cfm_ctx* ctx, int32_t n, const double* rhs, const double* sol); /* AoS [n][3 n_blob] */
int cfm_gnn_train(cfm_ctx* ctx, int32_t epochs, int32_t batch, double lr, double* loss_before,
                  double* loss_after);
int cfm_gnn_enable(cfm_ctx* ctx, int32_t on);          /* add the correction inside CFM_PC_NEAR */
int cfm_gnn_get_params(cfm_ctx* ctx, float* params);   /* cfm_gnn_num_params floats */
int cfm_gnn_set_params(cfm_ctx* ctx, const float* params);
int cfm_gnn_correction(cfm_ctx* ctx, double* corr, int32_t* nbr); /* corr [n_blob][17][9], nbr [n_blob][16] */
int cfm_gnn_gradient(cfm_ctx* ctx, int32_t s0, int32_t s1, float* grad, double* loss);
int cfm_gemm_bf16(cfm_ctx* ctx, int32_t M, int32_t N, int32_t K, const float* A, int32_t transA,
                  const float* B, int32_t transB, float* C);  /* C = op(A) op(B), row-major */
/* The network's coarse head: per blob, [h_i, positional monomials] (80 -> 48) gives 16
 * learned coarse vectors of the blob system that extend the linear fields of
 * CFM_COARSE_LINEAR (CFM_COARSE_LEARNED).  Training minimises, over the recorded
 * right-hand sides b_t with s_t = b_t - M W b_t (W the near-field inverse),
 *   L = mean_t (||s_t - M Z c_t||^2 + delta ||c_t||^2) / ||s_t||^2,  Z = [linear | learned],
 * c_t its minimiser, delta = 1e-10 tr((M Zlin)^T M Zlin) / 12 (geometry only); the
 * gradient dL/dZ = -2/T M sum_t (s_t - M Z c_t) c_t^T / ||s_t||^2 is back-propagated
 * through the network (BF16 matrix-core GEMMs) and applied with Adam.  batch <= 128. */
int cfm_gnn_train_coarse(cfm_ctx* ctx, int32_t epochs, int32_t batch, double lr, double* loss_before,
                         double* loss_after);
int cfm_gnn_coarse_gradient(cfm_ctx* ctx, int32_t s0, int32_t s1, float* grad, double* loss,
                            double* delta);
int cfm_gnn_coarse_vectors(cfm_ctx* ctx, double* zl); /* 16 x [3][n_blob] */

/* ---- prescribed-shape cilia above the wall ------------------------------------------
 * The reference's CILIA_TYPE 3 path in the mode that writes its generalised-force
 * reference files (WRITE_GENERALISED_FORCES): every cilium follows the fitted
 * Fulford-Blake beat (filament.cpp:684-1160) at a prescribed phase speed; each step
 * places the segments (filament.cpp:1336-1490), solves the wall-RPY resistance problem
 * M F = phase_dot vel_dir_phase over all segments (mobility_solver.cpp:301-420 rhs,
 * GMRES), and returns each cilium's phase and angle generalised forces
 * sum_n vel_dir . F_n (cilia_sim_main.cu:405-438).  The context must have the wall,
 * seg_dof = 3, no blobs and n_seg = n_cilia * nseg (cilium c owns segments
 * [c nseg, (c+1) nseg)). */
#define CFM_BEAT_FULFORD_BLAKE 1   /* Ito, Omori & Ishikawa (2019) fit, filament.cpp:684-758 */
#define CFM_BEAT_FULFORD_BLAKE_ORIGINAL 2 /* the fit's unaltered coefficients, zero cosine mode
                                             weighted 0.5 (SHAPE_SEQUENCE 4, filament.cpp:813-868,977-1008) */
#define CFM_BEAT_VOLVOX 3          /* volvox fit, four polynomial degrees (SHAPE_SEQUENCE 3,
                                      filament.cpp:748-812) */
#define CFM_BEAT_CORAL_LARVAE 4    /* coral-larvae fit (SHAPE_SEQUENCE 2, filament.cpp:869-893) */
#define CFM_BEAT_BUILD_A_BEAT 5    /* tangent-angle model with amplitude 1.99, recovery window
                                      1/12, effective stroke 0.75, zero-velocity avoidance 0.05
                                      (SHAPE_SEQUENCE 0, filament.cpp:1238-1334,1401-1546,
                                      config.hpp:133-139): segments at uniform arc fraction,
                                      trapezoid sums of the unit tangent; s_tol and the s table
                                      are unused */
#define CFM_BEAT_BICILIA 6         /* paired cilia (SHAPE_SEQUENCE 5, deprecated in the reference):
                                      each unit of nseg segments is two cilia of nseg / 2 beating
                                      the volvox fit (filament.cpp:895-960), the second at phase
                                      psi + 2 pi pair_dp with its base 2.2 a_seg along the frame's
                                      third axis (lab -x on the wall); the unit's phase and angle
                                      generalised forces sum over both (filament.cpp:371-375,
                                      1096-1185, 1341-1490) */
#define CFM_BEAT_BICILIA_LONGT 7   /* the same with the second phase pair_dp psi
                                      (SHAPE_SEQUENCE 6) */

typedef struct {
  int32_t n_cilia;
  int32_t nseg;          /* segments per cilium, 2..64 */
  int32_t beat;          /* CFM_BEAT_* */
  int32_t reserved;
  double dl;             /* segment spacing (filament length dl (nseg - 1)) */
  double dt;             /* time step: phase += phase_dot dt at every step after the first */
  double s_tol;          /* bisection tolerance on the arc-length fraction of the segment
                            parameters (0: 0.1 / nseg, filament.cpp:1152) */
  cfm_solve_opts solve;  /* segment solve: precond NONE/DIAG/SELF, mode, rtol, ... */
} cfm_cilia_params;

/* filament::initial_setup, prescribed branch (filament.cpp:225-470); base 3 n_cilia
 * (AoS, above the wall), phase and phase_dot n_cilia.  Places the segments. */
int cfm_cilia_init(cfm_ctx* ctx, const cfm_cilia_params* p, const double* base,
                   const double* phase, const double* phase_dot);
/* One step nt (the reference's time-loop body): q_phase, q_angle (n_cilia each, may be
 * null) receive the generalised forces; st (may be null) the solve statistics. */
int cfm_cilia_step(cfm_ctx* ctx, int32_t nt, double* q_phase, double* q_angle,
                   cfm_solve_stats* st);
/* Dynamic evolution (DYNAMIC_PHASE_EVOLUTION, and DYNAMIC_SHAPE_ROTATION when an angle
 * table is given): from the next step on, the phase speeds (and shape-angle speeds)
 * are unknowns of the coupled system of mobility_solver.cpp:1891-2115, driven by the
 * recorded generalised forces of one beat (tables of n values at phase 2 pi k / n,
 * e.g. the reference's input/forcing files), interpolated periodically
 * (mobility_solver.cpp:242-281) and scaled by phase_dot / (2 pi) of cfm_cilia_init
 * (each cilium's natural frequency); the shape angle feels the linear torsional spring
 * spring * phase_dot * L^3 * angle (mobility_solver.cpp:306-337, FENE_MODEL 0).
 * Preconditioner: the block form of mobility_solver.cpp:1104-1244 with the
 * particle preconditioner of opts.precond (CFM_PC_DIAG reproduces it exactly). */
int cfm_cilia_set_driving(cfm_ctx* ctx, const double* q_phase_table, int32_t n_phase,
                          const double* q_angle_table /* may be null */, int32_t n_angle,
                          double spring);
/* Options of the dynamic driving (mobility_solver.cpp:306-356): spring_model 0 is the
 * linear torsional spring, 1 the FENE form (FENE_MODEL 1): beyond |angle| = 0.78 the
 * spring's angle factor becomes 0.5 angle (1 + angle^2 / 0.78^2 exp(|angle| - 0.78)).
 * noise (FORCE_NOISE_MAG) adds noise * N(0, 1) to each cilium's phase driving force at
 * every step; the deviate is a counter-based function of (seed, step, cilium) --
 * SplitMix64 mixing into two 53-bit uniforms, then Box-Muller -- so runs repeat exactly
 * (the reference draws from an unseeded std::random_device). */
int cfm_cilia_set_driving_options(cfm_ctx* ctx, int32_t spring_model, double noise, uint64_t seed);
/* Segment parameters from a recorded table instead of the bisection, as the reference
 * does whenever it does not write its reference files (filament.cpp:1190-1235; the table
 * is the s_values file of input/forcing, read at cilia_sim_main.cu:339-378): rows at phases
 * 2 pi k / rows, cols recorded segments (>= 2), linear in phase and in segment position.
 * A null table returns to the bisection.  Re-places the segments. */
int cfm_cilia_set_s_table(cfm_ctx* ctx, const double* table, int32_t rows, int32_t cols);
/* Replaces the phases and / or shape angles (n_cilia each, either may be null) and
 * re-places the segments: a run restarted from a recorded state (the psi.dat initial
 * state the reference's drivers write, read at swimmer.cpp:43-60). */
int cfm_cilia_set_phases(cfm_ctx* ctx, const double* phase, const double* angle);
/* PAIR_DP of the paired beats (globals.ini "pair_dp", read at cilia_sim_main.cu:69;
 * default 0): the second cilium's phase offset (CFM_BEAT_BICILIA, in cycles) or phase
 * factor (CFM_BEAT_BICILIA_LONGT).  Re-places the segments.  (The reference's other pairing,
 * PAIR 1, only multiplies the intrinsic frequency of the second half of the filaments by
 * PAIR_DP (filament.cpp:240-244): pass such phase speeds to cfm_cilia_init.) */
int cfm_cilia_set_pair_dp(cfm_ctx* ctx, double pair_dp);
/* Phases, phase speeds, shape angles and their speeds (n_cilia each, may be null). */
int cfm_cilia_get_speeds(cfm_ctx* ctx, double* phase, double* phase_dot, double* angle,
                         double* angle_dot);
/* Current geometry and forces (each pointer may be null): positions and velocity
 * directions 3 n_seg (AoS), segment parameters n_cilia nseg, phases n_cilia, segment
 * forces of the last step 3 n_seg. */
int cfm_cilia_get_state(cfm_ctx* ctx, double* x, double* vel_dir_phase, double* vel_dir_angle,
                        double* s_values, double* phase, double* f_seg);

/* Prescribed cilia carried by free rigid bodies: the reference's PRESCRIBED_CILIA
 * swimmers with PRESCRIBED_BODY_VELOCITIES 0.  The context's segments are the cilia
 * (seg_dof = 3, n_seg = n_cilia nseg) and its blobs the bodies, set up first with
 * cfm_body_init and cfm_body_set_state; every body carries n_cilia / n_body cilia (body
 * b owns cilia [b per, (b+1) per)).  fil_ref (3 per cilium of one body, AoS) are the
 * cilia's base offsets and fil_q (4 per cilium, unit) their frames in the body frame (the
 * reference's filament_references and body_q_ref, written to _fil_references.dat and
 * _fil_q.dat; swimmer.cpp:448-462, filament.cpp:370-432): cilium c of body b has its
 * base at x_b + R(q_b) fil_ref_l and its frame R(q_b) R(fil_q_l), whose first two axes
 * carry the beat's filament axis and beat direction.
 * Each cfm_cilia_step solves the system of mobility_solver.cpp:1891-2115: unknowns the
 * forces of all segments and blobs, (V, Omega) of every body, and with
 * cfm_cilia_set_driving the phase (and shape-angle) speeds; rows
 *     M F - K U - sum_c (phi_dot_c k_phi,c + theta_dot_c k_theta,c) = rhs,
 *     -K^T F = 0 (force- and torque-free bodies),  -k_phi,c . F = -q_phase_c, ...
 * (rhs: phase_dot k_phi with prescribed phase speeds, zero with driving;
 * mobility_solver.cpp:301-458).  Preconditioner CFM_PC_DIAG is the reference's block
 * inverse with 6 pi mu a and make_precon_mat's 6 x 6 Schur complement
 * (mobility_solver.cpp:1104-1515, swimmer.cpp:1175-1347); CFM_PC_SELF the same with each
 * particle's wall self block; CFM_PC_NEAR replaces 6 pi mu a on each cilium's segments by
 * the inverse of that cilium's own mobility block; CFM_PC_DENSE also uses the resident
 * FP32 inverse of the blob block (built once, applied in the bodies' frame) with a full
 * 6 n_body Schur complement.  Steps after the first first advance phases, shape angles
 * and bodies explicitly with the previous step's speeds: x_b += dt V_b,
 * q_b = exp(dt Omega_b) q_b (cilia_sim_main.cu:724-740). */
int cfm_cilia_init_bodies(cfm_ctx* ctx, const cfm_cilia_params* p, const double* fil_ref,
                          const double* fil_q, const double* phase, const double* phase_dot);
/* Body centres 3 n_body, quaternions 4 n_body, velocities (V, Omega) 6 n_body of the last
 * step and blob forces 3 n_blob (AoS); each pointer may be null. */
int cfm_cilia_get_bodies(cfm_ctx* ctx, double* x, double* q, double* vel, double* f_blob);

/* ---- free rigid bodies made of the blobs ---------------------------------------------
 * The reference's rigid-body blob system (USE_BROYDEN_FOR_EVERYTHING 0,
 * PRESCRIBED_BODY_VELOCITIES 0): n_body bodies of nb = n_blob / n_body blobs each (body b
 * owns blobs [b nb, (b+1) nb)), all with the same body-frame offsets `ref` (3 nb, AoS;
 * rigid_body blob_references), centre x_b and unit quaternion q_b = (w, x, y, z):
 * blob i of body b sits at x_b + R(q_b) ref_i (rigid_body.cpp:497-515).
 * System CFM_SYS_BODY (cfm_gmres, cfm_apply_preconditioner, cfm_system_apply): unknowns
 * F (3 n_blob blob forces, AoS) then U (6 per body: V, Omega), in the reference's order;
 *     M_bb F - K U = rhs_blob,   -K^T F = rhs_body,   (K U)_i = V_b + Omega_b x r_i
 * (mobility_solver.cpp:1822-1880).  rhs_blob is minus the velocity the segments induce
 * at the blobs and rhs_body minus the total force and torque the body must balance
 * (mobility_solver.cpp:455-632).  Preconditioners: CFM_PC_DIAG is the reference's block
 * form with 6 pi mu a on the blobs and the rotated reference matrices
 * (mobility_solver.cpp:1518-1590); CFM_PC_SELF the same with each blob's wall-corrected
 * self block; CFM_PC_DENSE the exact Schur complement around the resident FP32 inverse
 * of M_bb (at most 128 bodies).  cfm_set_positions un-places the bodies. */
int cfm_body_init(cfm_ctx* ctx, int32_t n_body, const double* ref);
/* Places every body (x 3 n_body, q 4 n_body); with segments in the context, their
 * positions must have been set first. */
int cfm_body_set_state(cfm_ctx* ctx, const double* x, const double* q);
/* make_body_reference_matrices (mobility_solver.cpp:3049-3146): (K^T K)^-1 of one body
 * and the 6 x 6 mobility of body `body` alone (six solves of CFM_SYS_BODY with a unit
 * force or torque, options `opts`; the reference's sphere check compares it with
 * 1/(6 pi mu R), 1/(8 pi mu R^3)).  Both are returned (36 each, row-major, may be null)
 * in the body frame and kept for CFM_PC_DIAG / _SELF, rotated by R(q_b) for each body;
 * st (may be null) sums the six solves' steps and keeps their worst residual. */
int cfm_body_reference_matrices(cfm_ctx* ctx, int32_t body, const cfm_solve_opts* opts,
                                double* ktk_inv, double* mobility, cfm_solve_stats* st);
/* The reference matrices without the solves (the reference copies them between swimmers,
 * cilia_sim_main.cu:293-297); either pointer may be null to keep the current one. */
int cfm_body_set_reference_matrices(cfm_ctx* ctx, const double* ktk_inv, const double* mobility);
int cfm_body_get_reference_matrices(cfm_ctx* ctx, double* ktk_inv, double* mobility);
/* How the bodies move: the reference's BODY_VELOCITY_TYPE (config.hpp:94-101).
 * CFM_BODY_FREE (0, the default) solves for V and Omega.  CFM_BODY_PRESCRIBED_VELOCITY
 * (1, all velocities zero, rigid_body.cpp:599-614) is the blob system CFM_SYS_BLOB and is
 * refused here.  CFM_BODY_PRESCRIBED_ROTATION (2) holds Omega_b at zero: the reference
 * drops the Omega columns and the torque rows of CFM_SYS_BODY (mobility_solver.cpp:
 * 1947-1994, 2177-2181; preconditioner :1268-1291, 1506-1512); here the unknown vector
 * keeps 6 entries per body, the blob rows see V_b only and each torque row is the
 * identity on Omega_b, so a solve returns the reduced system's F and V and Omega = the
 * right-hand side's torque entries (pass zeros; the library's own right-hand sides --
 * free swimmers, squirmers, cilia on bodies -- zero them).  CFM_PC_DIAG / _SELF / _DENSE
 * invert the reduced system as they do the free one, with (K_V^T K_V)^-1 = I / nb, the
 * clamped mobility ((N^-1)_VV)^-1 of the reference matrices and (K_V^T P K_V)^-1. */
#define CFM_BODY_FREE 0
#define CFM_BODY_PRESCRIBED_VELOCITY 1
#define CFM_BODY_PRESCRIBED_ROTATION 2
int cfm_body_set_motion(cfm_ctx* ctx, int32_t type);
int cfm_body_get_motion(cfm_ctx* ctx, int32_t* type);

/* ---- squirmer bodies ------------------------------------------------------------------
 * The reference's CILIA_TYPE 4 (NO_CILIA_SQUIRMER, config.hpp:87,407): bodies of blobs
 * without cilia, driven by a surface slip velocity (mobility_solver.cpp:470-505,
 * util.cu:69-115).  For a blob at body-frame offset r (polar angle theta from the body's
 * z axis, azimuth phi): u_r = sum_n A_n P_n(cos theta), u_theta = sum_n B_n V_n(cos theta)
 * with V_n(x) = 2 sin(theta) P_n'(x) / (n (n + 1)), in Cartesian components turned by
 * R(q_b).  (The reference evaluates this slip but leaves its addition to the right-hand
 * side commented out; here it is applied.)  A lone spherical squirmer swims along its
 * axis at 2 B_1 / 3.  The context holds blobs only, with the bodies of cfm_body_init /
 * cfm_body_set_state.  Each cfm_squirmer_step solves the force- and torque-free body
 * system M F - K U = u_s, -K^T F = 0 (CFM_SYS_BODY with opts `solve`; the previous
 * solution is the guess) and, from the second call on, first advances every body with
 * the previous velocities, x += dt V, q = exp(dt Omega) q (cilia_sim_main.cu:724-740);
 * vel receives (V, Omega) per body (6 n_body, may be null). */
typedef struct {
  int32_t n_modes;      /* 1..8 */
  int32_t reserved;
  double A[8];          /* radial modes A_n, n = 1..n_modes */
  double B[8];          /* tangential modes B_n */
  double dt;
  cfm_solve_opts solve;
} cfm_squirmer_params;
int cfm_squirmer_init(cfm_ctx* ctx, const cfm_squirmer_params* p);
int cfm_squirmer_step(cfm_ctx* ctx, int32_t nt, double* vel, cfm_solve_stats* st);

/* ---- batched ensembles of independent wall systems ----------------------------------
 * No counterpart inside the reference's mobility_solver: the reference runs each
 * member of an ensemble (parameter sweep, bisection, periodic-orbit search) as its
 * own process (pyfile/driver/func.py:544 launches one executable per parameter set).
 * A batch holds n_sys independent systems of segments (sizes n_seg[s], any mix);
 * particles interact only within their own system.  Host arrays concatenate the
 * systems in order, AoS as everywhere else: positions 3 per particle, loads and
 * velocities seg_dof per particle. */
typedef struct cfm_batch cfm_batch;

typedef struct {
  int32_t n_sys;
  int32_t seg_dof;    /* 6 (force + torque) or 3 */
  int32_t wall;       /* no-slip wall z = 0 (requires a_seg == 1) */
  int32_t device;
  double a_seg, mu;
  int32_t reserved[4];
} cfm_batch_config;

/* Work decomposition of the persistent pair kernel: items are 64-target tiles times
 * slices of at most max_sources sources of the same system (0: whole system, a
 * multiple of 256 otherwise), pulled largest first by `grid` resident workgroups. */
typedef struct {
  int32_t max_sources;
  int32_t grid;
  int32_t reserved[2];
  double ms;          /* cfm_batch_tune: measured time of one batched product */
} cfm_batch_tuning;

typedef struct {
  int32_t iterations;   /* Arnoldi steps taken by this system */
  int32_t converged;
  int32_t reserved0, reserved1;
  double rel_residual;  /* FP64 ||b - M x|| / ||b|| at return */
  double latency_ms;    /* from the start of the call to the end of the GMRES step at
                           which the system's residual estimate met its target */
} cfm_batch_sys_stats;

int cfm_batch_create(const cfm_batch_config* cfg, const int32_t* n_seg, cfm_batch** out);
int cfm_batch_destroy(cfm_batch* b);
const char* cfm_batch_last_error(const cfm_batch* b);
int cfm_batch_set_positions(cfm_batch* b, const double* x /* 3 sum(n_seg) */);
/* v = M f per system, pair sums in `precision` (CFM_FP64 / CFM_FP32C) */
int cfm_batch_apply(cfm_batch* b, int32_t precision, const double* f, double* v);
int cfm_batch_set_tuning(cfm_batch* b, const cfm_batch_tuning* t);
int cfm_batch_get_tuning(cfm_batch* b, cfm_batch_tuning* t);
/* Autotuner: times the batched product for every candidate decomposition on the
 * batch's own positions and keeps the fastest; tried[] (optional, max_tried
 * entries) receives every candidate with its time. */
int cfm_batch_tune(cfm_batch* b, int32_t precision, cfm_batch_tuning* chosen,
                   cfm_batch_tuning* tried, int32_t max_tried, int32_t* n_tried);
/* Every system's segment resistance problem M_s x_s = rhs_s, solved together by
 * restarted right-preconditioned GMRES in lockstep (mobility_solver.cpp:2187-2669
 * per system).  opts->system is ignored; precond NONE, DIAG or SELF; mode FP64 or
 * MIXED (FP32C pair sums inside FP64 refinement).  One Arnoldi step of the batch is
 * one hipGraph replay.  per_sys (optional) has n_sys entries. */
int cfm_batch_gmres(cfm_batch* b, const double* rhs, double* x /* in: guess if use_guess */,
                    const cfm_solve_opts* opts, cfm_batch_sys_stats* per_sys, double* total_ms);

/* ---- ensembles of dynamic-phase cilia lattices ---------------------------------------
 * The reference's outer loops run one process per flow map (newton_solver.py:49-74,
 * bisection.py:150-162, pyfile/driver/func.py:544).  An ensemble holds n_members copies
 * of one lattice of prescribed cilia above the wall (bases `base`, 3 n_cilia AoS; beat,
 * nseg, dl, s_tol and the segment solve of cfg.prm) driven by the same recorded
 * generalised-force tables (cfm_cilia_set_driving's meaning; q_angle_table may be null:
 * no shape rotation), each member with its own phases, natural frequencies, shape angles
 * and time step.  A step advances every member as cfm_cilia_step does for one context
 * -- phase += dt phase_dot (from nt = 1), geometry, the coupled solve of forces and
 * speeds (restarted GMRES, true residual at each cycle's end, the previous solution as
 * guess unless it is worse than zero), generalised forces -- with one product launch for
 * all members (each unordered pair of a member's segments once, FP64) and one GMRES in
 * lockstep whose Arnoldi step is a captured hipGraph.  cfg.prm.solve: precond
 * CFM_PC_DIAG or CFM_PC_SELF, mode CFM_SOLVE_FP64; cfg.prm.dt and n_cilia's
 * phase_dot are not used (per-member values come with cfm_cilia_batch_set_state). */
typedef struct cfm_cilia_batch cfm_cilia_batch;
typedef struct {
  int32_t n_members;
  int32_t device;
  int32_t spring_model;   /* 0 linear torsional spring, 1 FENE (cfm_cilia_set_driving_options) */
  int32_t reserved;
  double mu;
  double spring;          /* the shape angle's torsional spring constant */
  double noise;           /* FORCE_NOISE_MAG, deviates keyed on (seed, step, cilium of the member) */
  uint64_t seed;
  cfm_cilia_params prm;
} cfm_cilia_batch_config;

int cfm_cilia_batch_create(const cfm_cilia_batch_config* cfg, const double* base,
                           const double* q_phase_table, int32_t n_phase,
                           const double* q_angle_table, int32_t n_angle, cfm_cilia_batch** out);
int cfm_cilia_batch_destroy(cfm_cilia_batch* b);
const char* cfm_cilia_batch_last_error(const cfm_cilia_batch* b);
/* Member states (n_members x n_cilia each, member-major): phases, natural frequencies
 * (the phase speeds of the first guess and the drive's scale), shape angles (null: 0);
 * dt: one time step per member.  Resets every member's guess (forces 0, speeds =
 * natural frequencies) and places the cilia. */
int cfm_cilia_batch_set_state(cfm_cilia_batch* b, const double* phase, const double* omega0,
                              const double* angle, const double* dt);
/* PAIR_DP of the paired beats for every member (cfm_cilia_set_pair_dp; default 0);
 * re-places the cilia when the members have a state. */
int cfm_cilia_batch_set_pair_dp(cfm_cilia_batch* b, double pair_dp);
/* One step nt of every member; q_phase, q_angle (n_members x n_cilia, may be null)
 * receive the generalised forces, per_member (n_members, may be null) the solve
 * statistics (reserved0 = 1 when the member's guess was rejected). */
int cfm_cilia_batch_step(cfm_cilia_batch* b, int32_t nt, double* q_phase, double* q_angle,
                         cfm_batch_sys_stats* per_member);
int cfm_cilia_batch_get_state(cfm_cilia_batch* b, double* phase, double* phase_dot, double* angle,
                              double* angle_dot);
/* Segment positions of every member (3 per segment, AoS, member-major). */
int cfm_cilia_batch_get_positions(cfm_cilia_batch* b, double* x);
/* The last step's solution of every member, member-major: segment forces (3 per segment,
 * AoS), then n_cilia phase speeds and, with shape rotation, n_cilia angle speeds. */
int cfm_cilia_batch_get_solution(cfm_cilia_batch* b, double* z);

#ifdef __cplusplus
}
#endif

#endif /* CUFCM_FIL_H */
