// ctx.hpp -- library context, error plumbing and shared device-buffer layout.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <array>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/cufcm_fil.h"

namespace cfm {

// Status carrier used inside the library; converted to int at the C boundary.
struct Status {
  int code = CFM_OK;
  std::string msg;
  bool ok() const { return code == CFM_OK; }
};

// Copies between host and device ordered after the work already queued on a stream.
// The library's streams are non-blocking, so a plain hipMemcpy (null stream) neither
// waits for their kernels nor is waited for by them.
inline hipError_t copy_to_host(hipStream_t s, void* dst, const void* src, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
  return e != hipSuccess ? e : hipStreamSynchronize(s);
}
inline hipError_t copy_to_device(hipStream_t s, void* dst, const void* src, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
  return e != hipSuccess ? e : hipStreamSynchronize(s);
}

inline Status make_status(int code, const char* fmt, ...) {
  Status s;
  s.code = code;
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  s.msg = buf;
  return s;
}

#define CFM_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ::cfm::make_status(CFM_EHIP, "%s failed: %s (%s:%d)", #call,               \
                                hipGetErrorString(e_), __FILE__, __LINE__);             \
  } while (0)

#define CFM_TRY(expr)              \
  do {                             \
    ::cfm::Status s_ = (expr);     \
    if (!s_.ok()) return s_;       \
  } while (0)

#define CFM_CHECK_LAUNCH() CFM_HIP(hipGetLastError())

// Device buffer owned by a context.
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  Status alloc(size_t count) {
    if (count <= n && p) return {};
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (count == 0) return {};
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess)
      return make_status(CFM_ENOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T),
                         hipGetErrorString(e));
    n = count;
    return {};
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Particle storage on the device, structure of arrays over the combined index
// space [segments 0..n_seg) ++ [blobs n_seg..n_seg+n_blob).
struct Particles {
  int n_seg = 0, n_blob = 0, n = 0;
  int seg_dof = 6;
  DevBuf<double> x, y, z;     // positions
  DevBuf<double> f;           // 6 x n (component-major: f + k*n); blobs use k < 3
  DevBuf<double> v;           // 6 x n
  DevBuf<double> partial;     // split x 6 x n partial sums
};

// Near-field block preconditioner data (nearfield.hip): NF_K nearest blobs of each
// blob and the 3x3 blocks of its local inverse, [m][NF_K+1][3][3].
#define NF_K 16
struct NearField {
  bool valid = false;
  int m = 0;
  DevBuf<int> nbr;
  DevBuf<double> W;
  DevBuf<int> status;
  void release() {
    nbr.release();
    W.release();
    status.release();
    valid = false;
  }
};

// Coarse correction of the near-field preconditioner (coarse.hip): k vectors Z of the
// blob system ([k][3][m]) -- the 12 linear fields of the blob positions, optionally
// followed by the learned network's vectors -- with M Z, M^2 Z and the inverse of the
// regularised Gram matrix (M Z)^T (M Z) + delta I, built once per blob geometry (and
// per network state).  Applied after the near-field inverse W:
//   y = W r + Z c,  c = argmin ||(r - M W r) - M Z c||  (+ delta ||c||^2),
// with (M Z)^T M W r = (M^2 Z)^T W r, so no mobility product is needed per application.
struct CoarseSpace {
  int mode = 0;       // CFM_COARSE_*
  bool valid = false;
  int k = 0, L = 0;
  DevBuf<double> Z, MZ, M2Z, Ginv, coef, part;
  std::vector<double> G;   // host copy of the Gram matrix (k x k)
  void release() {
    Z.release(); MZ.release(); M2Z.release(); Ginv.release(); coef.release(); part.release();
    valid = false;
    k = 0;
  }
};

// Resident inverse of the blob-blob mobility (densepc.hip): an n x n FP32 matrix,
// n = 3 n_blob, rows and columns in the blob-system vector order [component][blob],
// leading dimension ld (a multiple of 64 floats, padding zero).  Built once per set
// of blob positions from an FP64 Cholesky factorisation.
// P = M_bb^-1 kept as the upper triangle of square tiles (tile (I, J), I <= J, stored
// contiguously, row-major inside), FP32.
struct DenseInverse {
  bool valid = false;
  int n = 0;
  int nt = 0;               // tiles per side
  DevBuf<float> P;          // packed tiles
  DevBuf<int> tiles;        // (I, J) of every stored tile
  DevBuf<double> part;      // per (tile row, tile column) partial products
  double build_seconds = 0.0;
  int generation = 0;  // bumped by every build: caches derived from P compare it
  void release() {
    P.release();
    tiles.release();
    part.release();
    valid = false;
    n = 0;
    nt = 0;
  }
};

// Kernel-side launch description of one mobility product.
struct ApplyPlan {
  int tgt_begin = 0, tgt_end = 0;    // target range (combined ids)
  bool src_seg = false, src_blob = false;
  bool acc_seg = false, acc_blob = false;  // accumulate into v instead of overwrite
  // Optional external load / velocity blocks: row k of particle j (combined id)
  // lives at ext[k * ld + j - off].  Null: the context's own f / v arrays.
  const double* f_ext = nullptr;
  int f_ld = 0, f_off = 0;
  double* v_ext = nullptr;
  int v_ld = 0, v_off = 0;
  // Launch on the context's side stream (with its own split partials), to overlap
  // with work queued on the main stream; the caller orders the two with events.
  bool side = false;
  // Split partials in device memory the caller provides (split x 6 x n doubles): the
  // launch then allocates nothing.  Null: the context's own buffer, grown on demand.
  double* partial_ext = nullptr;
};

}  // namespace cfm

namespace cfm { struct Krylov; }
struct cfm_ctx;
namespace cfm {
// The context's rocBLAS handle (rocSOLVER calls), bound to the context's stream.
Status ctx_rocblas(cfm_ctx* c, void** handle);
}

struct cfm_ctx {
  cfm_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;           // second stream (default priority), created on demand
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  cfm::DevBuf<double> partial_side;     // split partials of side-stream launches
  cfm::Particles p;
  bool have_positions = false;
  int split = 1;                        // source split across workgroups
  cfm::DevBuf<double> stage;            // AoS staging for host transfers
  double* host_pinned = nullptr;        // pinned AoS staging on the host
  size_t host_pinned_n = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  cfm::Krylov* krylov = nullptr;        // GMRES workspace (created on first solve)
  cfm::DevBuf<double> vec_a, vec_b;     // system-vector staging
  void* fil = nullptr;                  // filament state of the implicit step (step.hip)
  cfm::NearField nf;                    // near-field preconditioner of the blob system
  const double* nf_corr = nullptr;      // learned correction blocks (same layout as nf.W) or null
  cfm::DenseInverse dinv;               // resident inverse of the blob system (CFM_PC_DENSE)
  cfm::CoarseSpace coarse;              // coarse correction of CFM_PC_NEAR (cfm_nearfield_coarse)
  std::vector<double> blob_pos;         // blob positions (AoS) of the last cfm_set_positions
  void* rocblas = nullptr;              // rocblas_handle on `stream` (created on first use)
  // autotuned source splits (tile_split = -1), keyed on the product's shape
  std::map<std::array<long long, 6>, int> split_cache;
  cfm::DevBuf<double> tune_v;           // scratch output of the tuning launches
  // arrival counter of the single-launch reductions on `stream` (zero between launches:
  // the last workgroup of each reduction resets it)
  cfm::DevBuf<unsigned> red_count;
  // single-species wall products evaluating each unordered pair once (apply_sym.hip)
  bool sym_pairs = true;
  cfm::DevBuf<double> sym_own, sym_own_side;  // per-workgroup target partials
  cfm::DevBuf<float> sym_src, sym_src_side;   // per-block-pair source partials (FP32C)
  cfm::DevBuf<double> sym_src64, sym_src64_side;  // the same, FP64
  // pair-kernel profiling (HIP events around each launch)
  bool prof_on = false;
  std::vector<hipEvent_t> prof_events;  // pool, two per launch
  size_t prof_used = 0;
  std::vector<double> prof_flops, prof_bytes;
  std::vector<int> prof_fp32;
  void* gnn = nullptr;                  // learned initial guess (gnn.hip)
  void* cilia = nullptr;                // prescribed-shape cilia (cilia.hip)
  void* body = nullptr;                 // free rigid bodies made of the blobs (body.hip)
  void* squirmer = nullptr;             // squirmer slip driving of the bodies (squirmer.hip)
  double* scratch = nullptr;            // this device's copy of the allocation-free entry
                                        // point's static scratch array (noalloc.hip)
  std::string err;
};
