// apply.hip -- tiled all-pairs wall-RPY mobility products on gfx950.
//
// Replaces Mss/Mbb/Msb/Mbs_mult (src/cilia/cuda_functions.cu:362-757).
//
// Layout: positions x,y,z and loads f[k] are structure-of-arrays in HBM over the
// combined particle index (segments, then blobs).  A workgroup of 256 threads
// (4 wave64s) owns 64 targets, one per lane.  Sources are streamed through LDS in
// tiles of 256 (one coalesced load per thread); within a tile wave w evaluates
// sources [64w, 64w+64) against its 64 targets with the source record broadcast
// from LDS.  At the end the four waves' partial sums are reduced through LDS in a
// fixed order (deterministic).  With split > 1 the source tiles are additionally
// dealt round-robin to `split` workgroups per target tile and a second kernel adds
// the partials in slice order.
//
// FP64 path: pair arithmetic and accumulation in double.
// FP32C path: separations are formed in double and narrowed, the pair block is
// evaluated in float, each 64-source tile is summed in float and folded into a
// per-target two-sum (Kahan-Babuska) accumulator; the final value is s + c in
// double.
//
// The self term of the wall branch is keyed on the global particle id and is
// evaluated in double during finalisation.
#include <algorithm>
#include <vector>

#include "ctx.hpp"
#include "apply_tile.hpp"
#include "rpy_pair.hpp"

namespace cfm {

struct KArgs {
  const double* __restrict__ x;
  const double* __restrict__ y;
  const double* __restrict__ z;
  const double* __restrict__ f;   // loads: row k of particle j at f[k*fld + j - foff]
  double* __restrict__ v;         // velocities: v[k*vld + j - voff]
  double* __restrict__ partial;   // split x 6 x n
  int n, n_seg, n_blob, seg_dof;
  int fld, foff, vld, voff;
  int tgt_begin, tgt_end;
  int seg_tiles;                  // target tiles that hold segments
  int src_seg, src_blob;
  int acc_seg, acc_blob;
  int split;
  int wall;
  double mu, a_seg, a_blob;
};

// Loops the source tiles of one species [s_begin, s_end) (ids) for this workgroup.
template <typename T, int TD, int FD, bool WALL>
__device__ __forceinline__ void sweep_species(const KArgs& a, SrcRec<T>* tile, int s_begin,
                                              int s_end, int slice, int t_lo, int t_hi,
                                              int iid, double xi, double yi, double zi, T ai,
                                              T aj, double* accd, float* s32, float* c32) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const T zif = T(zi);
  const int ntiles = (s_end - s_begin + kTS - 1) / kTS;
  for (int tix = slice; tix < ntiles; tix += a.split) {
    const int t0 = s_begin + tix * kTS;
    const int j = t0 + threadIdx.x;
    __syncthreads();
    if (j < s_end) {
      SrcRec<T> r;
      r.x = a.x[j];
      r.y = a.y[j];
      r.z = a.z[j];
      if constexpr (sizeof(T) == 4) r.zf = float(r.z);
#pragma unroll
      for (int k = 0; k < FD; ++k)
        r.f[k] = T((WALL ? kStageScale : 1.0) * a.f[(size_t)k * a.fld + j - a.foff]);
#pragma unroll
      for (int k = FD; k < 6; ++k) r.f[k] = T(0);
      tile[threadIdx.x] = r;
    }
    __syncthreads();
    const int w0 = t0 + wave * 64;
    const int cnt = min(64, s_end - w0);
    if (cnt <= 0) continue;
    const SrcRec<T>* rec = tile + wave * 64;
    // the only tiles that may contain the target itself
    const bool check = WALL && (w0 < t_hi) && (w0 + cnt > t_lo);
    if constexpr (sizeof(T) == 4) {
      T part[TD];
#pragma unroll
      for (int k = 0; k < TD; ++k) part[k] = T(0);
      if (check)
        tile_sweep<T, TD, FD, WALL, true>(rec, cnt, w0, iid, xi, yi, zi, zif, ai, aj, T(a.mu), part);
      else
        tile_sweep<T, TD, FD, WALL, false>(rec, cnt, w0, iid, xi, yi, zi, zif, ai, aj, T(a.mu), part);
#pragma unroll
      for (int k = 0; k < TD; ++k) two_sum_add(s32[k], c32[k], part[k]);
    } else {
      T part[TD];
#pragma unroll
      for (int k = 0; k < TD; ++k) part[k] = T(0);
      if (check)
        tile_sweep<T, TD, FD, WALL, true>(rec, cnt, w0, iid, xi, yi, zi, zif, ai, aj, T(a.mu), part);
      else
        tile_sweep<T, TD, FD, WALL, false>(rec, cnt, w0, iid, xi, yi, zi, zif, ai, aj, T(a.mu), part);
#pragma unroll
      for (int k = 0; k < TD; ++k) accd[k] += double(part[k]);
    }
    (void)lane;
  }
}

template <int TD>
__device__ __forceinline__ void add_self(const KArgs& a, int i, double zi, double* out) {
  // self mobility (wall branch), only when the target's own species is a source
  const bool is_seg = i < a.n_seg;
  const bool own = is_seg ? a.src_seg : a.src_blob;
  if (!a.wall || !own) return;
  double F[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < TD; ++k) F[k] = a.f[(size_t)k * a.fld + i - a.foff];
  wall_self<TD, TD>(zi, F, out);
}

template <typename T, int TD, int SD, bool WALL>
__device__ void apply_body(const KArgs& a, SrcRec<T>* tile, double (*red)[kTT], int t_lo,
                           int t_hi, int slice) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int i = t_lo + lane;
  const bool valid = i < t_hi;
  const int ic = valid ? i : t_lo;
  const double xi = a.x[ic], yi = a.y[ic], zi = a.z[ic];
  const bool tgt_seg = t_lo < a.n_seg;
  const T ai = T(tgt_seg ? a.a_seg : a.a_blob);
  double accd[TD];
  float s32[TD], c32[TD];
#pragma unroll
  for (int k = 0; k < TD; ++k) {
    accd[k] = 0.0;
    s32[k] = 0.0f;
    c32[k] = 0.0f;
  }
  const int iid = valid ? i : -1;
  if (a.src_seg && a.n_seg > 0)
    sweep_species<T, TD, SD, WALL>(a, tile, 0, a.n_seg, slice, t_lo, t_hi, iid, xi, yi, zi, ai,
                                   T(a.a_seg), accd, s32, c32);
  if (a.src_blob && a.n_blob > 0)
    sweep_species<T, TD, 3, WALL>(a, tile, a.n_seg, a.n, slice, t_lo, t_hi, iid, xi, yi, zi, ai,
                                  T(a.a_blob), accd, s32, c32);
  // cross-wave reduction in a fixed order
  __syncthreads();
#pragma unroll
  for (int k = 0; k < TD; ++k)
    red[wave * 6 + k][lane] = (sizeof(T) == 4) ? double(s32[k]) + double(c32[k]) : accd[k];
  __syncthreads();
  if (wave != 0 || !valid) return;
  double out[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < TD; ++k) {
    double s = red[k][lane];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) s += red[w * 6 + k][lane];
    out[k] = s;
  }
  const size_t n = a.n;
  if (a.split > 1) {
#pragma unroll
    for (int k = 0; k < TD; ++k) a.partial[((size_t)slice * 6 + k) * n + i] = out[k];
    return;
  }
  if (slice == 0) add_self<TD>(a, i, zi, out);
  const double scale = a.wall ? 1.0 / a.mu : 1.0;
  const bool acc = tgt_seg ? a.acc_seg : a.acc_blob;
#pragma unroll
  for (int k = 0; k < TD; ++k) {
    double* dst = a.v + (size_t)k * a.vld + i - a.voff;
    *dst = acc ? *dst + scale * out[k] : scale * out[k];
  }
}

template <typename T, int SD, bool WALL>
__global__ __launch_bounds__(kWG) void rpy_apply_kernel(KArgs a) {
  __shared__ SrcRec<T> tile[kTS];
  __shared__ double red[kWaves * 6][kTT];
  const int t = blockIdx.x;
  const int slice = blockIdx.y;
  int t_lo, t_hi;
  // target tiles: segment tiles first (never straddling the species boundary)
  const int seg_lo = max(a.tgt_begin, 0), seg_hi = min(a.tgt_end, a.n_seg);
  if (t < a.seg_tiles) {
    t_lo = seg_lo + t * kTT;
    t_hi = min(t_lo + kTT, seg_hi);
    apply_body<T, SD, SD, WALL>(a, tile, red, t_lo, t_hi, slice);
  } else {
    const int blob_lo = max(a.tgt_begin, a.n_seg), blob_hi = min(a.tgt_end, a.n);
    t_lo = blob_lo + (t - a.seg_tiles) * kTT;
    t_hi = min(t_lo + kTT, blob_hi);
    apply_body<T, 3, SD, WALL>(a, tile, red, t_lo, t_hi, slice);
  }
}

// Sum the split partials, add the self term, scale and store.  A workgroup takes 64
// targets; wave w sums the slices s = w mod 4 in order, the four sums are added in
// wave order (a fixed order for every launch).
__global__ __launch_bounds__(256) void rpy_finalize_kernel(KArgs a) {
  __shared__ double red[kWaves][6][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = a.tgt_begin + blockIdx.x * 64 + lane;
  const bool valid = i < a.tgt_end;
  const bool is_seg = i < a.n_seg;
  const int td = is_seg ? a.seg_dof : 3;
  const size_t n = a.n;
  double part[6] = {0, 0, 0, 0, 0, 0};
  if (valid)
    for (int s = w; s < a.split; s += kWaves)
      for (int k = 0; k < td; ++k) part[k] += a.partial[((size_t)s * 6 + k) * n + i];
  for (int k = 0; k < 6; ++k) red[w][k][lane] = part[k];
  __syncthreads();
  if (w != 0 || !valid) return;
  double out[6];
  for (int k = 0; k < 6; ++k) out[k] = (red[0][k][lane] + red[1][k][lane]) + (red[2][k][lane] + red[3][k][lane]);
  const double zi = a.z[i];
  if (td == 6) add_self<6>(a, i, zi, out); else add_self<3>(a, i, zi, out);
  const double scale = a.wall ? 1.0 / a.mu : 1.0;
  const bool acc = is_seg ? a.acc_seg : a.acc_blob;
  for (int k = 0; k < td; ++k) {
    double* dst = a.v + (size_t)k * a.vld + i - a.voff;
    *dst = acc ? *dst + scale * out[k] : scale * out[k];
  }
}

template <typename T, int SD, bool WALL>
static Status launch_t(const KArgs& a, int tiles, int split, hipStream_t st) {
  dim3 grid(tiles, split);
  hipLaunchKernelGGL((rpy_apply_kernel<T, SD, WALL>), grid, dim3(kWG), 0, st, a);
  CFM_CHECK_LAUNCH();
  return {};
}

// Algorithmic flops and bytes of one product (pairs counted over the full
// target x source rectangle, self pairs included).
static void plan_cost(const cfm_ctx* c, const ApplyPlan& plan, bool fp32, double* flops,
                      double* bytes) {
  const Particles& p = c->p;
  const double nts = std::max(0, std::min(plan.tgt_end, p.n_seg) - std::max(plan.tgt_begin, 0));
  const double ntb = std::max(0, std::min(plan.tgt_end, p.n) - std::max(plan.tgt_begin, p.n_seg));
  const double nss = plan.src_seg ? p.n_seg : 0, nsb = plan.src_blob ? p.n_blob : 0;
  const bool s6 = p.seg_dof == 6;
  const double f_ss = s6 ? CFM_FLOPS_PAIR_66 : CFM_FLOPS_PAIR_33;
  const double f_sb = s6 ? CFM_FLOPS_PAIR_63 : CFM_FLOPS_PAIR_33;
  const double f_bs = s6 ? CFM_FLOPS_PAIR_36 : CFM_FLOPS_PAIR_33;
  *flops = nts * nss * f_ss + nts * nsb * f_sb + ntb * nss * f_bs + ntb * nsb * CFM_FLOPS_PAIR_33;
  const double w = fp32 ? 4.0 : 8.0;
  // sources: positions (3 FP64) + loads; targets: positions (3 FP64) + velocities written
  *bytes = (nss + nsb + nts + ntb) * 3.0 * 8.0 + (nss * p.seg_dof + nsb * 3.0) * 8.0 +
           (nts * p.seg_dof + ntb * 3.0) * 8.0;
  (void)w;
}

static Status prof_begin(cfm_ctx* c, hipEvent_t* e0, hipEvent_t* e1, hipStream_t st) {
  if (c->prof_used + 2 > c->prof_events.size()) {
    for (int k = 0; k < 64; ++k) {
      hipEvent_t e;
      CFM_HIP(hipEventCreate(&e));
      c->prof_events.push_back(e);
    }
  }
  *e0 = c->prof_events[c->prof_used];
  *e1 = c->prof_events[c->prof_used + 1];
  c->prof_used += 2;
  CFM_HIP(hipEventRecord(*e0, st));
  return {};
}

Status side_fork(cfm_ctx* c) {
  if (!c->side) {
    CFM_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    CFM_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    CFM_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  }
  CFM_HIP(hipEventRecord(c->ev_fork, c->stream));
  CFM_HIP(hipStreamWaitEvent(c->side, c->ev_fork, 0));
  return {};
}

Status side_join(cfm_ctx* c) {
  if (!c->side) return {};
  CFM_HIP(hipEventRecord(c->ev_join, c->side));
  CFM_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  return {};
}

int sym_species(const cfm_ctx* c, const ApplyPlan& plan);
Status launch_sym(cfm_ctx* c, const ApplyPlan& plan, int species, hipStream_t stream);

Status launch_apply(cfm_ctx* c, const ApplyPlan& plan_in) {
  Particles& p = c->p;
  // Wall products of one species from itself: the symmetric kernels take that species'
  // targets, this launch the rest (the other species' targets, if any)
  const int species = sym_species(c, plan_in);
  const bool sym = species != 0;
  ApplyPlan plan = plan_in;
  if (species == 1) plan.tgt_begin = std::max(plan_in.tgt_begin, p.n_seg);
  if (species == 2) plan.tgt_end = std::min(plan_in.tgt_end, p.n_seg);
  KArgs a{};
  a.x = p.x.p;
  a.y = p.y.p;
  a.z = p.z.p;
  a.f = p.f.p;
  a.v = p.v.p;
  a.fld = p.n;
  a.foff = 0;
  a.vld = p.n;
  a.voff = 0;
  if (plan.f_ext) {
    a.f = plan.f_ext;
    a.fld = plan.f_ld;
    a.foff = plan.f_off;
  }
  if (plan.v_ext) {
    a.v = plan.v_ext;
    a.vld = plan.v_ld;
    a.voff = plan.v_off;
  }
  a.partial = p.partial.p;
  a.n = p.n;
  a.n_seg = p.n_seg;
  a.n_blob = p.n_blob;
  a.seg_dof = p.seg_dof;
  a.tgt_begin = plan.tgt_begin;
  a.tgt_end = plan.tgt_end;
  a.src_seg = plan.src_seg;
  a.src_blob = plan.src_blob;
  a.acc_seg = plan.acc_seg;
  a.acc_blob = plan.acc_blob;
  a.wall = c->cfg.wall;
  a.mu = c->cfg.mu;
  a.a_seg = c->cfg.a_seg;
  a.a_blob = c->cfg.a_blob;
  const int seg_lo = std::max(plan.tgt_begin, 0), seg_hi = std::min(plan.tgt_end, p.n_seg);
  const int blob_lo = std::max(plan.tgt_begin, p.n_seg), blob_hi = std::min(plan.tgt_end, p.n);
  a.seg_tiles = seg_hi > seg_lo ? (seg_hi - seg_lo + kTT - 1) / kTT : 0;
  const int blob_tiles = blob_hi > blob_lo ? (blob_hi - blob_lo + kTT - 1) / kTT : 0;
  const int tiles = a.seg_tiles + blob_tiles;
  if (tiles == 0 && !sym) return {};
  int split = std::max(1, c->split);
  a.split = split;
  if (plan.side && !c->side) return make_status(CFM_ESTATE, "side stream launch before side_fork");
  const hipStream_t stream = plan.side ? c->side : c->stream;
  if (sym) {
    hipEvent_t se0 = nullptr, se1 = nullptr;
    if (c->prof_on) CFM_TRY(prof_begin(c, &se0, &se1, stream));
    CFM_TRY(launch_sym(c, plan_in, species, stream));
    if (c->prof_on) {
      CFM_HIP(hipEventRecord(se1, stream));
      ApplyPlan ss = plan_in;
      ss.tgt_begin = species == 1 ? 0 : p.n_seg;
      ss.tgt_end = species == 1 ? p.n_seg : p.n;
      double fl = 0.0, by = 0.0;
      const bool f32 = c->cfg.precision == CFM_FP32C;
      plan_cost(c, ss, f32, &fl, &by);
      c->prof_flops.push_back(fl);
      c->prof_bytes.push_back(by);
      c->prof_fp32.push_back(f32 ? 2 : 3);  // symmetric products: FP32C 2, FP64 3
    }
    if (tiles == 0) return {};
  }
  if (split > 1) {
    if (plan.partial_ext) {
      a.partial = plan.partial_ext;
    } else {
      DevBuf<double>& part = plan.side ? c->partial_side : p.partial;
      CFM_TRY(part.alloc((size_t)split * 6 * p.n));
      a.partial = part.p;
    }
  }
  const bool fp32 = c->cfg.precision == CFM_FP32C;
  const bool wall = c->cfg.wall != 0;
  const int sd = p.seg_dof;
  hipEvent_t pe0 = nullptr, pe1 = nullptr;
  if (c->prof_on) CFM_TRY(prof_begin(c, &pe0, &pe1, stream));
  Status s;
  if (fp32) {
    if (sd == 6) s = wall ? launch_t<float, 6, true>(a, tiles, split, stream)
                          : launch_t<float, 6, false>(a, tiles, split, stream);
    else s = wall ? launch_t<float, 3, true>(a, tiles, split, stream)
                  : launch_t<float, 3, false>(a, tiles, split, stream);
  } else {
    if (sd == 6) s = wall ? launch_t<double, 6, true>(a, tiles, split, stream)
                          : launch_t<double, 6, false>(a, tiles, split, stream);
    else s = wall ? launch_t<double, 3, true>(a, tiles, split, stream)
                  : launch_t<double, 3, false>(a, tiles, split, stream);
  }
  CFM_TRY(s);
  if (c->prof_on) {
    CFM_HIP(hipEventRecord(pe1, stream));
    double fl = 0.0, by = 0.0;
    plan_cost(c, plan, fp32, &fl, &by);
    c->prof_flops.push_back(fl);
    c->prof_bytes.push_back(by);
    c->prof_fp32.push_back(fp32 ? 1 : 0);
  }
  if (split > 1) {
    const int cnt = plan.tgt_end - plan.tgt_begin;
    hipLaunchKernelGGL(rpy_finalize_kernel, dim3((cnt + 63) / 64), dim3(256), 0, stream, a);
    CFM_CHECK_LAUNCH();
  }
  return {};
}

// Number of workgroups launched for a plan (used to choose the split).
int plan_tiles(const cfm_ctx* c, const ApplyPlan& plan) {
  const Particles& p = c->p;
  const int seg_lo = std::max(plan.tgt_begin, 0), seg_hi = std::min(plan.tgt_end, p.n_seg);
  const int blob_lo = std::max(plan.tgt_begin, p.n_seg), blob_hi = std::min(plan.tgt_end, p.n);
  int t = 0;
  if (seg_hi > seg_lo) t += (seg_hi - seg_lo + kTT - 1) / kTT;
  if (blob_hi > blob_lo) t += (blob_hi - blob_lo + kTT - 1) / kTT;
  return t;
}

// Autotuner of the source split: every candidate writes a scratch output (the
// product is never accumulated twice) and the fastest of two timed launches wins.
Status tune_split(cfm_ctx* c, const ApplyPlan& pl, int heuristic, int* best) {
  const Particles& p = c->p;
  int nsrc = (pl.src_seg ? p.n_seg : 0) + (pl.src_blob ? p.n_blob : 0);
  const int stiles = std::max(1, (nsrc + kTS - 1) / kTS);
  std::vector<int> cand = {heuristic};
  for (int s : {1, 2, 3, 4, 6, 8, 12, 16, 24, 32})
    if (s <= stiles && s != heuristic) cand.push_back(s);
  CFM_TRY(c->tune_v.alloc((size_t)6 * p.n));
  ApplyPlan t = pl;
  t.side = false;  // timed on the main stream
  t.acc_seg = t.acc_blob = false;
  t.v_ext = c->tune_v.p;
  t.v_ld = p.n;
  t.v_off = 0;
  const bool prof = c->prof_on;
  c->prof_on = false;
  float best_ms = 1e30f;
  Status st;
  for (int s : cand) {
    c->split = s;
    st = launch_apply(c, t);  // warm-up (sizes the split partials)
    if (!st.ok()) break;
    if (hipEventRecord(c->ev0, c->stream) != hipSuccess) { st = make_status(CFM_EHIP, "event"); break; }
    st = launch_apply(c, t);
    if (st.ok()) st = launch_apply(c, t);
    if (!st.ok()) break;
    if (hipEventRecord(c->ev1, c->stream) != hipSuccess || hipEventSynchronize(c->ev1) != hipSuccess) {
      st = make_status(CFM_EHIP, "event");
      break;
    }
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    if (ms < best_ms) {
      best_ms = ms;
      *best = s;
    }
  }
  c->prof_on = prof;
  c->split = *best;
  return st;
}

}  // namespace cfm
