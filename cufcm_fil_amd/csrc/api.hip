// api.hip -- C ABI entry points (declared in include/cufcm_fil.h).
//
// Host side of the mobility back-end: owns the HIP stream, the device particle
// buffers and the pinned staging buffers, converts between the reference's
// interleaved host layout and the device structure-of-arrays layout, and turns
// every failure into a status code plus a message (the reference's
// rpy_mobility_solver checks no return codes and exits on fatal paths).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include <rocblas/rocblas.h>

#include "ctx.hpp"
#include "krylov.hpp"

namespace cfm {

Status launch_apply(cfm_ctx* c, const ApplyPlan& plan);
int plan_tiles(const cfm_ctx* c, const ApplyPlan& plan);
Status noalloc_scratch(int device, double** ptr);

// AoS [dof*cnt] (stage) -> SoA rows k of dst (dst + k*ld + off), k < dof.
__global__ void aos_to_soa_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                  int cnt, int dof, int ld, int off) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  for (int k = 0; k < dof; ++k) dst[(size_t)k * ld + off + i] = src[(size_t)dof * i + k];
}

__global__ void soa_to_aos_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                  int cnt, int dof, int ld, int off) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  for (int k = 0; k < dof; ++k) dst[(size_t)dof * i + k] = src[(size_t)k * ld + off + i];
}

static Status ensure_pinned(cfm_ctx* c, size_t n) {
  if (n <= c->host_pinned_n && c->host_pinned) return {};
  if (c->host_pinned) (void)hipHostFree(c->host_pinned);
  c->host_pinned = nullptr;
  c->host_pinned_n = 0;
  CFM_HIP(hipHostMalloc((void**)&c->host_pinned, n * sizeof(double), hipHostMallocDefault));
  c->host_pinned_n = n;
  return {};
}

// host AoS -> device SoA rows [0, dof) of `dst` (leading dimension ld) at offset off
Status upload_aos(cfm_ctx* c, const double* host, int cnt, int dof, double* dst, int ld,
                  int off) {
  if (cnt == 0) return {};
  const size_t m = (size_t)cnt * dof;
  CFM_TRY(ensure_pinned(c, m));
  CFM_TRY(c->stage.alloc(std::max<size_t>(m, (size_t)6 * c->p.n)));
  std::memcpy(c->host_pinned, host, m * sizeof(double));
  CFM_HIP(hipMemcpyAsync(c->stage.p, c->host_pinned, m * sizeof(double), hipMemcpyHostToDevice,
                         c->stream));
  hipLaunchKernelGGL(aos_to_soa_kernel, dim3((cnt + 255) / 256), dim3(256), 0, c->stream,
                     c->stage.p, dst, cnt, dof, ld, off);
  CFM_CHECK_LAUNCH();
  CFM_HIP(hipStreamSynchronize(c->stream));  // pinned buffer is reused by the next transfer
  return {};
}

Status download_aos(cfm_ctx* c, double* host, int cnt, int dof, const double* src, int ld,
                    int off) {
  if (cnt == 0) return {};
  const size_t m = (size_t)cnt * dof;
  CFM_TRY(ensure_pinned(c, m));
  CFM_TRY(c->stage.alloc(std::max<size_t>(m, (size_t)6 * c->p.n)));
  hipLaunchKernelGGL(soa_to_aos_kernel, dim3((cnt + 255) / 256), dim3(256), 0, c->stream, src,
                     c->stage.p, cnt, dof, ld, off);
  CFM_CHECK_LAUNCH();
  CFM_HIP(hipMemcpyAsync(c->host_pinned, c->stage.p, m * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  CFM_HIP(hipStreamSynchronize(c->stream));
  std::memcpy(host, c->host_pinned, m * sizeof(double));
  return {};
}

static bool all_finite(const double* a, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(a[i])) return false;
  return true;
}

// Host <-> device system vectors: the particle block is AoS on the host and [dof][m]
// on the device; the extra entries (6 per free body) follow in the same order on both.
static Status resolve_system(cfm_ctx* c, int system, SystemDesc* s) {
  if (system == CFM_SYS_BODY) return body_system_desc(c, s);
  if (system != CFM_SYS_SEG && system != CFM_SYS_BLOB)
    return make_status(CFM_EINVAL, "unknown system %d", system);
  *s = system_desc(c, system);
  return {};
}

static Status upload_system(cfm_ctx* c, const SystemDesc& s, const double* host, double* dev) {
  CFM_TRY(upload_aos(c, host, s.m, s.dof, dev, s.m, 0));
  if (s.extra > 0) {
    const size_t n0 = (size_t)s.m * s.dof;
    CFM_HIP(hipMemcpyAsync(dev + n0, host + n0, sizeof(double) * s.extra, hipMemcpyHostToDevice, c->stream));
    CFM_HIP(hipStreamSynchronize(c->stream));
  }
  return {};
}

static Status download_system(cfm_ctx* c, const SystemDesc& s, const double* dev, double* host) {
  CFM_TRY(download_aos(c, host, s.m, s.dof, dev, s.m, 0));
  if (s.extra > 0) {
    const size_t n0 = (size_t)s.m * s.dof;
    CFM_HIP(hipMemcpyAsync(host + n0, dev + n0, sizeof(double) * s.extra, hipMemcpyDeviceToHost, c->stream));
    CFM_HIP(hipStreamSynchronize(c->stream));
  }
  return {};
}

ApplyPlan plan_for(const cfm_ctx* c, int which) {
  ApplyPlan pl;
  const int ns = c->p.n_seg, n = c->p.n;
  switch (which) {
    case CFM_SS: pl.tgt_begin = 0; pl.tgt_end = ns; pl.src_seg = true; break;
    case CFM_SB: pl.tgt_begin = 0; pl.tgt_end = ns; pl.src_blob = true; pl.acc_seg = true; break;
    case CFM_BB: pl.tgt_begin = ns; pl.tgt_end = n; pl.src_blob = true; break;
    case CFM_BS: pl.tgt_begin = ns; pl.tgt_end = n; pl.src_seg = true; break;
    default: pl.tgt_begin = 0; pl.tgt_end = n; pl.src_seg = true; pl.src_blob = true; break;
  }
  return pl;
}

// Source split across workgroups so that a launch has enough workgroups to fill
// 256 CUs (about 4 per CU), bounded by the number of 256-source tiles.
Status tune_split(cfm_ctx* c, const ApplyPlan& pl, int heuristic, int* best);

int heuristic_split(const cfm_ctx* c, const ApplyPlan& pl) {
  const int tiles = std::max(1, plan_tiles(c, pl));
  int nsrc = 0;
  if (pl.src_seg) nsrc += c->p.n_seg;
  if (pl.src_blob) nsrc += c->p.n_blob;
  const int stiles = std::max(1, (nsrc + 255) / 256);
  // about 8k workgroups: several per CU slot, so the partial last round is short
  int s = (8192 + tiles - 1) / tiles;
  s = std::min(s, std::max(1, stiles / 2));
  return std::max(1, std::min(s, 64));
}

int choose_split(cfm_ctx* c, const ApplyPlan& pl) {
  if (c->cfg.tile_split > 0) return c->cfg.tile_split;
  const int heur = heuristic_split(c, pl);
  if (c->cfg.tile_split == 0) return heur;
  const std::array<long long, 6> key = {pl.tgt_begin, pl.tgt_end, pl.src_seg, pl.src_blob,
                                        c->cfg.precision, c->p.n};
  auto it = c->split_cache.find(key);
  if (it != c->split_cache.end()) return it->second;
  int best = heur;
  if (!tune_split(c, pl, heur, &best).ok()) best = heur;  // timing failed: keep the heuristic
  c->split_cache[key] = best;
  return best;
}

}  // namespace cfm

using namespace cfm;

static thread_local std::string g_create_error;

void cfm_fil_release(cfm_ctx* c);
void cfm_gnn_release(cfm_ctx* c);
void cfm_cilia_release(cfm_ctx* c);
void cfm_body_release(cfm_ctx* c);
namespace cfm {
void cfm_squirmer_release(cfm_ctx* c);
void body_positions_changed(cfm_ctx* c);
}

static int fail(cfm_ctx* c, const Status& s) {
  if (c) c->err = s.msg;
  else g_create_error = s.msg;
  return s.code;
}

namespace cfm {
Status ctx_rocblas(cfm_ctx* c, void** handle) {
  if (!c->rocblas) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success)
      return make_status(CFM_EHIP, "rocblas_create_handle failed");
    if (rocblas_set_stream(h, c->stream) != rocblas_status_success) {
      (void)rocblas_destroy_handle(h);
      return make_status(CFM_EHIP, "rocblas_set_stream failed");
    }
    c->rocblas = h;
  }
  *handle = c->rocblas;
  return {};
}
}  // namespace cfm

extern "C" {

int cfm_abi_version(void) { return CFM_ABI_VERSION; }

const char* cfm_last_error(const cfm_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int cfm_create(const cfm_config* cfg, cfm_ctx** out) {
  if (!cfg || !out) return fail(nullptr, make_status(CFM_EINVAL, "null argument"));
  *out = nullptr;
  if (cfg->n_seg < 0 || cfg->n_blob < 0 || cfg->n_seg + cfg->n_blob <= 0)
    return fail(nullptr, make_status(CFM_EINVAL, "need n_seg >= 0, n_blob >= 0, n > 0"));
  if (cfg->seg_dof != 3 && cfg->seg_dof != 6)
    return fail(nullptr, make_status(CFM_EINVAL, "seg_dof must be 3 or 6 (got %d)", cfg->seg_dof));
  if (!(cfg->mu > 0.0) || !(cfg->a_seg > 0.0) || !(cfg->a_blob > 0.0))
    return fail(nullptr, make_status(CFM_EINVAL, "mu and radii must be positive"));
  if (cfg->wall && (cfg->a_seg != 1.0 || cfg->a_blob != 1.0))
    return fail(nullptr, make_status(CFM_EINVAL,
                                     "the wall-corrected branch assumes unit radii (as the "
                                     "reference's cuda_functions.cu:22)"));
  if (cfg->precision != CFM_FP64 && cfg->precision != CFM_FP32C)
    return fail(nullptr, make_status(CFM_EINVAL, "unknown precision %d", cfg->precision));
  if (cfg->tile_split < -1 || cfg->tile_split > 64)
    return fail(nullptr, make_status(CFM_EINVAL, "tile_split must be -1 (autotune), 0 or 1..64"));
  if ((int64_t)cfg->n_seg + cfg->n_blob > (int64_t)1 << 28)
    return fail(nullptr, make_status(CFM_EINVAL, "too many particles"));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(nullptr, make_status(CFM_EHIP, "no HIP device available"));
  if (cfg->device < 0 || cfg->device >= ndev)
    return fail(nullptr, make_status(CFM_EINVAL, "device %d out of range (%d)", cfg->device, ndev));
  cfm_ctx* c = new (std::nothrow) cfm_ctx();
  if (!c) return fail(nullptr, make_status(CFM_ENOMEM, "context allocation failed"));
  c->cfg = *cfg;
  c->device = cfg->device;
  auto init = [&]() -> Status {
    CFM_HIP(hipSetDevice(c->device));
    CFM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CFM_HIP(hipEventCreate(&c->ev0));
    CFM_HIP(hipEventCreate(&c->ev1));
    Particles& p = c->p;
    p.n_seg = cfg->n_seg;
    p.n_blob = cfg->n_blob;
    p.n = p.n_seg + p.n_blob;
    p.seg_dof = cfg->seg_dof;
    CFM_TRY(p.x.alloc(p.n));
    CFM_TRY(p.y.alloc(p.n));
    CFM_TRY(p.z.alloc(p.n));
    CFM_TRY(p.f.alloc((size_t)6 * p.n));
    CFM_TRY(p.v.alloc((size_t)6 * p.n));
    CFM_HIP(hipMemsetAsync(p.f.p, 0, sizeof(double) * 6 * p.n, c->stream));
    CFM_HIP(hipMemsetAsync(p.v.p, 0, sizeof(double) * 6 * p.n, c->stream));
    CFM_HIP(hipStreamSynchronize(c->stream));
    // this device's address of the allocation-free entry point's scratch array
    CFM_TRY(noalloc_scratch(c->device, &c->scratch));
    return {};
  };
  Status s = init();
  if (!s.ok()) {
    cfm_destroy(c);
    return fail(nullptr, s);
  }
  *out = c;
  return CFM_OK;
}

int cfm_destroy(cfm_ctx* c) {
  if (!c) return CFM_OK;
  (void)hipSetDevice(c->device);
  cfm_fil_release(c);
  cfm_gnn_release(c);
  cfm_cilia_release(c);
  cfm::cfm_squirmer_release(c);
  cfm_body_release(c);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->rocblas) (void)rocblas_destroy_handle((rocblas_handle)c->rocblas);
  c->tune_v.release();
  c->red_count.release();
  c->p.x.release();
  c->p.y.release();
  c->p.z.release();
  c->p.f.release();
  c->p.v.release();
  c->p.partial.release();
  c->stage.release();
  c->nf.release();
  c->coarse.release();
  c->dinv.release();
  c->vec_a.release();
  c->vec_b.release();
  if (c->krylov) {
    c->krylov->release();
    delete c->krylov;
  }
  if (c->host_pinned) (void)hipHostFree(c->host_pinned);
  for (hipEvent_t e : c->prof_events) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->side) {
    (void)hipStreamSynchronize(c->side);
    (void)hipEventDestroy(c->ev_fork);
    (void)hipEventDestroy(c->ev_join);
    (void)hipStreamDestroy(c->side);
  }
  c->partial_side.release();
  c->sym_own.release();
  c->sym_own_side.release();
  c->sym_src.release();
  c->sym_src_side.release();
  c->sym_src64.release();
  c->sym_src64_side.release();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CFM_OK;
}

int cfm_set_positions(cfm_ctx* c, const double* x_seg, const double* x_blob) {
  if (!c) return CFM_EINVAL;
  Particles& p = c->p;
  if ((p.n_seg > 0 && !x_seg) || (p.n_blob > 0 && !x_blob))
    return fail(c, make_status(CFM_EINVAL, "null position array"));
  if ((p.n_seg && !all_finite(x_seg, 3 * (size_t)p.n_seg)) ||
      (p.n_blob && !all_finite(x_blob, 3 * (size_t)p.n_blob)))
    return fail(c, make_status(CFM_ENONFINITE, "non-finite position"));
  if (c->cfg.wall) {
    for (int i = 0; i < p.n_seg; ++i)
      if (!(x_seg[3 * i + 2] > 0.0))
        return fail(c, make_status(CFM_EINVAL, "segment %d is not above the wall (z = %g)", i,
                                   x_seg[3 * i + 2]));
    for (int i = 0; i < p.n_blob; ++i)
      if (!(x_blob[3 * i + 2] > 0.0))
        return fail(c, make_status(CFM_EINVAL, "blob %d is not above the wall (z = %g)", i,
                                   x_blob[3 * i + 2]));
  }
  (void)hipSetDevice(c->device);
  // The three coordinate arrays are allocated separately; stage them as rows of
  // a 3 x n block, then scatter.
  auto up = [&](const double* h, int cnt, int off) -> Status {
    if (cnt == 0) return {};
    CFM_TRY(c->stage.alloc((size_t)6 * p.n + (size_t)3 * cnt));
    double* rows = c->stage.p + (size_t)6 * p.n;  // 3 x cnt scratch after the AoS stage
    CFM_TRY(upload_aos(c, h, cnt, 3, rows, cnt, 0));
    CFM_HIP(hipMemcpyAsync(p.x.p + off, rows, sizeof(double) * cnt, hipMemcpyDeviceToDevice, c->stream));
    CFM_HIP(hipMemcpyAsync(p.y.p + off, rows + cnt, sizeof(double) * cnt, hipMemcpyDeviceToDevice, c->stream));
    CFM_HIP(hipMemcpyAsync(p.z.p + off, rows + 2 * (size_t)cnt, sizeof(double) * cnt, hipMemcpyDeviceToDevice, c->stream));
    CFM_HIP(hipStreamSynchronize(c->stream));
    return {};
  };
  Status s = up(x_seg, p.n_seg, 0);
  if (s.ok()) s = up(x_blob, p.n_blob, p.n_seg);
  if (!s.ok()) return fail(c, s);
  c->have_positions = true;
  body_positions_changed(c);
  // the near-field blocks and the resident inverse depend on the blob positions only
  const size_t nb3 = (size_t)3 * p.n_blob;
  if (c->blob_pos.size() != nb3 || (nb3 && std::memcmp(c->blob_pos.data(), x_blob, sizeof(double) * nb3))) {
    c->blob_pos.assign(x_blob, x_blob + nb3);
    c->nf.valid = false;
    c->dinv.valid = false;
    c->coarse.valid = false;
  }
  return CFM_OK;
}

int cfm_set_precision(cfm_ctx* c, int32_t precision) {
  if (!c) return CFM_EINVAL;
  if (precision != CFM_FP64 && precision != CFM_FP32C)
    return fail(c, make_status(CFM_EINVAL, "unknown precision %d", precision));
  c->cfg.precision = precision;
  return CFM_OK;
}

int cfm_apply(cfm_ctx* c, int32_t which, const double* f_seg, const double* f_blob,
              double* v_seg, double* v_blob) {
  if (!c) return CFM_EINVAL;
  if (which < CFM_SS || which > CFM_FULL)
    return fail(c, make_status(CFM_EINVAL, "unknown block id %d", which));
  if (!c->have_positions) return fail(c, make_status(CFM_ESTATE, "positions not set"));
  Particles& p = c->p;
  const ApplyPlan pl = plan_for(c, which);
  const int sd = p.seg_dof;
  const bool tgt_seg = pl.tgt_begin < p.n_seg && p.n_seg > 0;
  const bool tgt_blob = pl.tgt_end > p.n_seg && p.n_blob > 0;
  if ((pl.src_seg && p.n_seg && !f_seg) || (pl.src_blob && p.n_blob && !f_blob) ||
      (tgt_seg && !v_seg) || (tgt_blob && !v_blob))
    return fail(c, make_status(CFM_EINVAL, "null load or velocity array for block %d", which));
  if ((pl.src_seg && p.n_seg && !all_finite(f_seg, (size_t)sd * p.n_seg)) ||
      (pl.src_blob && p.n_blob && !all_finite(f_blob, 3 * (size_t)p.n_blob)))
    return fail(c, make_status(CFM_ENONFINITE, "non-finite load"));
  (void)hipSetDevice(c->device);
  auto run = [&]() -> Status {
    if (pl.src_seg) CFM_TRY(upload_aos(c, f_seg, p.n_seg, sd, p.f.p, p.n, 0));
    if (pl.src_blob) CFM_TRY(upload_aos(c, f_blob, p.n_blob, 3, p.f.p, p.n, p.n_seg));
    if (pl.acc_seg && tgt_seg) CFM_TRY(upload_aos(c, v_seg, p.n_seg, sd, p.v.p, p.n, 0));
    if (pl.acc_blob && tgt_blob) CFM_TRY(upload_aos(c, v_blob, p.n_blob, 3, p.v.p, p.n, p.n_seg));
    c->split = choose_split(c, pl);
    CFM_TRY(launch_apply(c, pl));
    if (tgt_seg) CFM_TRY(download_aos(c, v_seg, p.n_seg, sd, p.v.p, p.n, 0));
    if (tgt_blob) CFM_TRY(download_aos(c, v_blob, p.n_blob, 3, p.v.p, p.n, p.n_seg));
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : fail(c, s);
}

int cfm_time_apply(cfm_ctx* c, int32_t which, int32_t reps, double* ms_per_apply) {
  if (!c || !ms_per_apply || reps <= 0) return CFM_EINVAL;
  if (!c->have_positions) return fail(c, make_status(CFM_ESTATE, "positions not set"));
  (void)hipSetDevice(c->device);
  const ApplyPlan pl = plan_for(c, which);
  auto run = [&]() -> Status {
    c->split = choose_split(c, pl);
    CFM_TRY(launch_apply(c, pl));  // warm-up (also allocates split partials)
    CFM_HIP(hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < reps; ++r) CFM_TRY(launch_apply(c, pl));
    CFM_HIP(hipEventRecord(c->ev1, c->stream));
    CFM_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    CFM_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    *ms_per_apply = double(ms) / reps;
    return {};
  };
  Status s = run();
  return s.ok() ? CFM_OK : fail(c, s);
}

}  // extern "C"

extern "C" {

int cfm_gmres(cfm_ctx* c, const cfm_solve_opts* o, const double* rhs, double* x,
              cfm_solve_stats* stats) {
  if (!c || !o || !rhs || !x) return c ? fail(c, make_status(CFM_EINVAL, "null argument")) : CFM_EINVAL;
  if (o->precond < CFM_PC_NONE || o->precond > CFM_PC_DENSE)
    return fail(c, make_status(CFM_EINVAL, "unknown preconditioner %d", o->precond));
  if (o->mode != CFM_SOLVE_FP64 && o->mode != CFM_SOLVE_MIXED)
    return fail(c, make_status(CFM_EINVAL, "unknown solve mode %d", o->mode));
  if (!(o->rtol > 0.0) || o->restart < 0 || o->max_iter < 0)
    return fail(c, make_status(CFM_EINVAL, "need rtol > 0, restart >= 0, max_iter >= 0"));
  if (!c->have_positions) return fail(c, make_status(CFM_ESTATE, "positions not set"));
  SystemDesc s;
  Status sd = resolve_system(c, o->system, &s);
  if (!sd.ok()) return fail(c, sd);
  if (s.m == 0) return fail(c, make_status(CFM_EINVAL, "empty system"));
  if (s.body && o->precond == CFM_PC_NEAR)
    return fail(c, make_status(CFM_EINVAL, "the body system takes CFM_PC_NONE, _DIAG, _SELF or _DENSE"));
  if (!all_finite(rhs, s.L) || (o->use_guess && !all_finite(x, s.L)))
    return fail(c, make_status(CFM_ENONFINITE, "non-finite rhs or guess"));
  (void)hipSetDevice(c->device);
  if (!c->krylov) c->krylov = new Krylov();
  auto run = [&]() -> Status {
    CFM_TRY(c->vec_a.alloc(s.L));
    CFM_TRY(c->vec_b.alloc(s.L));
    CFM_TRY(upload_system(c, s, rhs, c->vec_a.p));
    CFM_TRY(upload_system(c, s, x, c->vec_b.p));
    Status st = solve(c, *c->krylov, s, *o, c->vec_a.p, c->vec_b.p, stats, nullptr);
    if (st.code != CFM_OK && st.code != CFM_ENOCONV) return st;
    CFM_TRY(download_system(c, s, c->vec_b.p, x));
    return st;
  };
  Status st = run();
  return st.ok() ? CFM_OK : fail(c, st);
}

int cfm_apply_preconditioner(cfm_ctx* c, int32_t system, int32_t precond, const double* in,
                             double* out) {
  if (!c || !in || !out) return c ? fail(c, make_status(CFM_EINVAL, "null argument")) : CFM_EINVAL;
  if (precond < CFM_PC_NONE || precond > CFM_PC_DENSE)
    return fail(c, make_status(CFM_EINVAL, "unknown preconditioner %d", precond));
  if (!c->have_positions) return fail(c, make_status(CFM_ESTATE, "positions not set"));
  SystemDesc s;
  Status sd = resolve_system(c, system, &s);
  if (!sd.ok()) return fail(c, sd);
  (void)hipSetDevice(c->device);
  auto run = [&]() -> Status {
    CFM_TRY(c->vec_a.alloc(s.L));
    CFM_TRY(c->vec_b.alloc(s.L));
    CFM_TRY(upload_system(c, s, in, c->vec_a.p));
    CFM_TRY(system_precond(c, s, precond, c->vec_a.p, c->vec_b.p));
    CFM_TRY(download_system(c, s, c->vec_b.p, out));
    return {};
  };
  Status st = run();
  return st.ok() ? CFM_OK : fail(c, st);
}

int cfm_system_apply(cfm_ctx* c, int32_t system, int32_t precision, const double* x, double* y) {
  if (!c || !x || !y) return c ? fail(c, make_status(CFM_EINVAL, "null argument")) : CFM_EINVAL;
  if (precision != CFM_FP64 && precision != CFM_FP32C)
    return fail(c, make_status(CFM_EINVAL, "unknown precision %d", precision));
  if (!c->have_positions) return fail(c, make_status(CFM_ESTATE, "positions not set"));
  SystemDesc s;
  Status sd = resolve_system(c, system, &s);
  if (!sd.ok()) return fail(c, sd);
  if (!all_finite(x, s.L)) return fail(c, make_status(CFM_ENONFINITE, "non-finite input"));
  (void)hipSetDevice(c->device);
  auto run = [&]() -> Status {
    CFM_TRY(c->vec_a.alloc(s.L));
    CFM_TRY(c->vec_b.alloc(s.L));
    CFM_TRY(upload_system(c, s, x, c->vec_a.p));
    CFM_TRY(system_apply(c, s, precision, c->vec_a.p, c->vec_b.p));
    CFM_TRY(download_system(c, s, c->vec_b.p, y));
    return {};
  };
  Status st = run();
  return st.ok() ? CFM_OK : fail(c, st);
}

}  // extern "C"

extern "C" {

int cfm_profile(cfm_ctx* c, int32_t enable) {
  if (!c) return CFM_EINVAL;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->prof_on = enable != 0;
  if (enable) {
    c->prof_used = 0;
    c->prof_flops.clear();
    c->prof_bytes.clear();
    c->prof_fp32.clear();
  }
  return CFM_OK;
}

int cfm_set_pair_symmetry(cfm_ctx* c, int32_t on) {
  if (!c) return CFM_EINVAL;
  c->sym_pairs = on != 0;
  c->split_cache.clear();  // the remaining launches change shape
  return CFM_OK;
}

int cfm_profile_read(cfm_ctx* c, cfm_profile_stats* st) {
  if (!c || !st) return CFM_EINVAL;
  (void)hipSetDevice(c->device);
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, make_status(CFM_EHIP, "%s", hipGetErrorString(e)));
  cfm_profile_stats S{};
  for (size_t k = 0; k < c->prof_flops.size(); ++k) {
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, c->prof_events[2 * k], c->prof_events[2 * k + 1]);
    if (e != hipSuccess) return fail(c, make_status(CFM_EHIP, "%s", hipGetErrorString(e)));
    S.launches += 1;
    S.kernel_ms += ms;
    S.alg_flops += c->prof_flops[k];
    S.alg_bytes += c->prof_bytes[k];
    if (c->prof_fp32[k] == 1) {
      S.launches_fp32 += 1;
      S.kernel_ms_fp32 += ms;
      S.alg_flops_fp32 += c->prof_flops[k];
    } else if (c->prof_fp32[k] == 2) {
      S.launches_sym += 1;
      S.kernel_ms_sym += ms;
      S.alg_flops_sym += c->prof_flops[k];
    }
  }
  *st = S;
  return CFM_OK;
}

}  // extern "C"
