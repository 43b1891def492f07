// noalloc.hip -- allocation-free variant of cfm_apply (cfm_apply_noalloc).
//
// For callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).  Loads and velocities are
// device arrays in the reference's interleaved layout; they are transposed into the
// context's own particle rows (allocated by cfm_create), the pair kernel writes its
// split partial sums into `scratch` below instead of the context's on-demand buffer,
// and the finalisation writes the context's velocity rows, transposed back into the
// caller's array.  Nothing here calls hipMalloc, hipFree or an allocating library
// routine.
//
// `scratch` is one file-scope device array per GPU, held until the process ends.  In
// host code its name is only a stand-in: its device address is looked up once per
// device (noalloc_scratch, called by cfm_create with that device current) and only
// that address is ever passed to kernels.  Every call on a GPU shares the array, which
// keeps the last call's contents; the pair kernel overwrites every entry the
// finalisation reads, so nothing needs clearing.  Calls that overlap on one GPU (from
// different contexts or threads) race on it.
#include <algorithm>
#include <mutex>
#include <vector>

#include "ctx.hpp"

namespace cfm {

Status launch_apply(cfm_ctx* c, const ApplyPlan& plan);
ApplyPlan plan_for(const cfm_ctx* c, int which);
int heuristic_split(const cfm_ctx* c, const ApplyPlan& pl);

constexpr size_t kScratchDoubles = CFM_NOALLOC_SCRATCH_BYTES / sizeof(double);

static __device__ __align__(256) double scratch[kScratchDoubles];

// The device address of `scratch` on `device` (which must be current), looked up on
// the first call for that device and kept.
Status noalloc_scratch(int device, double** ptr) {
  static std::mutex mu;
  static std::vector<double*> addr;
  std::lock_guard<std::mutex> lock(mu);
  if (device < 0) return make_status(CFM_EINVAL, "device %d out of range", device);
  if ((int)addr.size() <= device) addr.resize(device + 1, nullptr);
  if (!addr[device]) {
    double* p = nullptr;
    CFM_HIP(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)));
    if (!p) return make_status(CFM_EHIP, "no device address for the scratch array on device %d", device);
    addr[device] = p;
  }
  *ptr = addr[device];
  return {};
}

namespace {

// interleaved [dof*cnt] -> rows k of dst (dst + k*ld + off), k < dof
__global__ void interleaved_to_rows_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                           int cnt, int dof, int ld, int off) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  for (int k = 0; k < dof; ++k) dst[(size_t)k * ld + off + i] = src[(size_t)dof * i + k];
}

__global__ void rows_to_interleaved_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                           int cnt, int dof, int ld, int off) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  for (int k = 0; k < dof; ++k) dst[(size_t)dof * i + k] = src[(size_t)k * ld + off + i];
}

Status to_rows(cfm_ctx* c, const double* src, int cnt, int dof, double* dst, int ld, int off) {
  if (cnt == 0) return {};
  hipLaunchKernelGGL(interleaved_to_rows_kernel, dim3((cnt + 255) / 256), dim3(256), 0, c->stream,
                     src, dst, cnt, dof, ld, off);
  CFM_CHECK_LAUNCH();
  return {};
}

Status from_rows(cfm_ctx* c, const double* src, int cnt, int dof, double* dst, int ld, int off) {
  if (cnt == 0) return {};
  hipLaunchKernelGGL(rows_to_interleaved_kernel, dim3((cnt + 255) / 256), dim3(256), 0, c->stream,
                     src, dst, cnt, dof, ld, off);
  CFM_CHECK_LAUNCH();
  return {};
}

}  // namespace
}  // namespace cfm

using namespace cfm;

extern "C" int cfm_apply_noalloc(cfm_ctx* c, int32_t which, const double* f_seg, const double* f_blob,
                                 double* v_seg, double* v_blob) {
  if (!c) return CFM_EINVAL;
  auto fail = [c](const Status& s) {
    c->err = s.msg;
    return s.code;
  };
  if (which < CFM_SS || which > CFM_FULL)
    return fail(make_status(CFM_EINVAL, "unknown block id %d", which));
  if (!c->have_positions) return fail(make_status(CFM_ESTATE, "positions not set"));
  if (!c->scratch) return fail(make_status(CFM_ESTATE, "the scratch array of this device is not set up"));
  Particles& p = c->p;
  ApplyPlan pl = plan_for(c, which);
  const int sd = p.seg_dof;
  const bool tgt_seg = pl.tgt_begin < p.n_seg && p.n_seg > 0;
  const bool tgt_blob = pl.tgt_end > p.n_seg && p.n_blob > 0;
  if ((pl.src_seg && p.n_seg && !f_seg) || (pl.src_blob && p.n_blob && !f_blob) ||
      (tgt_seg && !v_seg) || (tgt_blob && !v_blob))
    return fail(make_status(CFM_EINVAL, "null load or velocity array for block %d", which));
  // split partials: split x 6 x n doubles must fit the scratch array
  const size_t per_slice = (size_t)6 * p.n;
  const size_t max_split = kScratchDoubles / per_slice;
  int split = 1;
  if (c->cfg.tile_split > 0) {
    split = c->cfg.tile_split;
    if (split > 1 && (size_t)split > max_split)
      return fail(make_status(CFM_EINVAL,
                              "split %d needs %zu bytes of partial sums for %d particles; the "
                              "allocation-free scratch holds %zu bytes",
                              split, (size_t)split * per_slice * sizeof(double), p.n,
                              (size_t)CFM_NOALLOC_SCRATCH_BYTES));
  } else {
    split = (int)std::min<size_t>((size_t)heuristic_split(c, pl), std::max<size_t>(1, max_split));
  }
  (void)hipSetDevice(c->device);
  auto run = [&]() -> Status {
    if (pl.src_seg) CFM_TRY(to_rows(c, f_seg, p.n_seg, sd, p.f.p, p.n, 0));
    if (pl.src_blob) CFM_TRY(to_rows(c, f_blob, p.n_blob, 3, p.f.p, p.n, p.n_seg));
    if (pl.acc_seg && tgt_seg) CFM_TRY(to_rows(c, v_seg, p.n_seg, sd, p.v.p, p.n, 0));
    if (pl.acc_blob && tgt_blob) CFM_TRY(to_rows(c, v_blob, p.n_blob, 3, p.v.p, p.n, p.n_seg));
    pl.partial_ext = c->scratch;
    c->split = split;
    const bool prof = c->prof_on;  // no event pool growth here
    c->prof_on = false;
    Status st = launch_apply(c, pl);
    c->prof_on = prof;
    CFM_TRY(st);
    if (tgt_seg) CFM_TRY(from_rows(c, p.v.p, p.n_seg, sd, v_seg, p.n, 0));
    if (tgt_blob) CFM_TRY(from_rows(c, p.v.p, p.n_blob, 3, v_blob, p.n, p.n_seg));
    CFM_HIP(hipStreamSynchronize(c->stream));
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : fail(s);
}
