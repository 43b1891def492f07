// apply_sym.hip -- segment-segment wall-RPY product that evaluates each unordered pair
// once (FP32C, 6 <- 6).
//
// The wall-corrected grand mobility is symmetric: the 6 x 6 block of target i and
// source j is the transpose of that of target j and source i.  The one-directional
// kernel (apply.hip) evaluates the pair formula for both orders; here the block M_ij
// is formed explicitly once and applied both ways, v_i += M_ij F_j and
// v_j += M_ij^T F_i, which costs about 0.6 of the two directed evaluations.
//
// Work decomposition: the segments are cut into B blocks of 64.  Block pair (I, J) with
// J = I + d (mod B) is taken by the workgroups of row I for ring offsets d = 1 .. B/2
// (d = B/2, when B is even, only for I < B/2), so every unordered block pair is taken
// exactly once; d = 0 (the block with itself) is evaluated one way round, skipping the
// self pair.  A workgroup (I, c) owns the 64 targets of block I, one per lane, and the
// ring offsets [c dpw, (c+1) dpw); its four waves take one offset each per round.  In
// step s of a block pair, lane l pairs its target with source (l + s) mod 64 of block J
// (records in LDS, structure-of-arrays so that the rotated reads are conflict-free) and
// carries the running sum of that source's velocity, handed to the lane below after
// every step: after 64 steps lane l holds the complete contribution of block I to
// source l of block J.
//
// Precision as the one-directional FP32C kernel: separations in double narrowed to
// float, the block in float; each 64-pair sum in float, folded per target into a
// two-sum accumulator, every partial stored and summed in double in a fixed order
// (deterministic).
//
// Partials (device memory of the context): own[I][c][6][64] (double) holds workgroup
// (I, c)'s target sums, src[J][d-1][6][64] (float) block I = J - d's contribution to
// the sources of block J.  rpy_sym_finalize_kernel adds them, the self term, scales by
// 1/mu and stores or accumulates the segment velocities.
#include <algorithm>

#include "apply_tile.hpp"
#include "ctx.hpp"
#include "rpy_pair.hpp"

namespace cfm {

namespace {

constexpr int kBlk = 64;  // segments per block

struct SymArgs {
  const double* __restrict__ x;
  const double* __restrict__ y;
  const double* __restrict__ z;
  const double* __restrict__ f;  // loads: row k of segment j at f[k fld + j - foff]
  double* __restrict__ v;        // velocities: v[k vld + j - voff]
  double* __restrict__ own;      // B x S x 6 x 64
  float* __restrict__ src;       // B x Dn x 6 x 64
  int nseg, nblk, S, Dn, dpw;
  int fld, foff, vld, voff;
  int acc;
  double mu;
};

// Does workgroup row I take ring offset d?
__device__ __forceinline__ bool takes(int I, int d, int B) {
  if (d == 0) return true;
  if (2 * d < B) return true;
  return 2 * d == B && I < d;
}

// The 6 x 6 wall block (target i, source j) for loads staged with 1/(6 pi): the same
// terms as wall_pair_scaled written as matrix entries.  With r the unit separation,
// e the image direction (dx, dy, Rz) p1, n = z x e = (-e_y, e_x, 0):
//   TT = (iso + A0) I + a r r^T + a1 e e^T + a2 e_z e z^T + a3 e_z z e^T + a4 z z^T
//   TR = [m - q]x + (d2 (e_x, e_y, -e_z) - d1 z) n^T
//   RT = [b - q]x - n (b1 z + b2 e)^T
//   RR = (C0 - k3) I + 3 k3 r r^T - 1.125 p3 e e^T + 2.25 p3 e_z z e^T + 2.25 p3 n n^T
// with q = k2 r, b = B0 e + B3 z, m = B0 (e_x, e_y, -e_z) - D3 z.
struct Blk66 {
  float tt[9], tr[9], rt[9], rr[9];
};

__device__ __forceinline__ void wall_block66(float dx, float dy, float dz, float Rz, float zj, Blk66& M) {
  const float hxy = dx * dx + dy * dy;
  // free space
  const float is = rsq<float>(hxy + dz * dz);
  const float rx = dx * is, ry = dy * is, rz = dz * is;
  const float is2 = is * is;
  const float k1 = 0.75f * is;
  const float iso = k1 * (1.0f + (2.0f / 3.0f) * is2);
  const float a_ = k1 * (1.0f - 2.0f * is2);
  const float k2 = k1 * is;
  const float qx = k2 * rx, qy = k2 * ry, qz = k2 * rz;
  const float k3 = 0.5f * k2 * is;
  // images
  const float p1 = rsq<float>(hxy + Rz * Rz);
  const float p2 = p1 * p1, p3 = p2 * p1, p4 = p2 * p2, p5 = p4 * p1;
  const float ex = dx * p1, ey = dy * p1, ez = Rz * p1;
  const float u = ez * ez;
  const float h = zj * rcp<float>(Rz);
  const float hh = h * (1.0f - h);
  const float q3 = 1.0f - 5.0f * u;
  const float A0 = -(0.75f * p1 * (1.0f + 2.0f * hh * u) + 0.5f * (p3 * (1.0f - 3.0f * u) - p5 * q3));
  const float a1 = -(0.75f * p1 * (1.0f - 6.0f * hh * u) - 1.5f * p3 * q3 + 2.5f * p5 * (1.0f - 7.0f * u));
  const float a2 = 1.5f * h * p1 * (1.0f - 6.0f * (1.0f - h) * u) - 3.0f * p3 * q3 + 5.0f * p5 * (2.0f - 7.0f * u);
  const float a3 = 1.5f * h * p1 - 5.0f * p5;
  const float a4 = -((3.0f * h * h * p1 + 3.0f * p3) * u + (2.0f - 15.0f * u) * p5);
  const float B0 = 0.75f * p2;
  const float hp2 = h * p2;
  const float w4 = (1.5f - 15.0f * u) * p4;
  const float b1 = 9.0f * u * hp2 + w4;
  const float b2 = -ez * (4.5f * hp2 - 7.5f * p4);
  const float B3 = -1.5f * ez * (hp2 - p4);
  const float C0 = p3 * (0.375f - 2.25f * u);
  const float gp2 = p2 - hp2;
  const float d1 = -(9.0f * u * gp2 + w4);
  const float d2 = ez * (4.5f * gp2 - 7.5f * p4);
  const float D3 = -1.5f * ez * (gp2 - p4);
  // TT
  const float c0 = iso + A0;
  const float arx = a_ * rx, ary = a_ * ry, arz = a_ * rz;
  const float aex = a1 * ex, aey = a1 * ey;
  const float gc = (a1 + a2) * ez, gr = (a1 + a3) * ez;
  M.tt[0] = c0 + arx * rx + aex * ex;
  M.tt[1] = arx * ry + aex * ey;
  M.tt[2] = arx * rz + ex * gc;
  M.tt[3] = M.tt[1];
  M.tt[4] = c0 + ary * ry + aey * ey;
  M.tt[5] = ary * rz + ey * gc;
  M.tt[6] = arz * rx + ex * gr;
  M.tt[7] = arz * ry + ey * gr;
  M.tt[8] = c0 + arz * rz + ez * (gc + a3 * ez) + a4;
  // TR (velocity <- torque) and RT (rotation <- force)
  const float wx = B0 * ex - qx, wy = B0 * ey - qy;
  const float wz = -B0 * ez - D3 - qz;
  const float tuz = -d2 * ez - d1;
  const float d2x = d2 * ex, d2y = d2 * ey;
  M.tr[0] = -d2x * ey;
  M.tr[1] = d2x * ex - wz;
  M.tr[2] = wy;
  M.tr[3] = wz - d2y * ey;
  M.tr[4] = d2y * ex;
  M.tr[5] = -wx;
  M.tr[6] = -wy - ey * tuz;
  M.tr[7] = wx + ex * tuz;
  M.tr[8] = 0.0f;
  const float vz = B0 * ez + B3 - qz;
  const float ruz = b2 * ez + b1;
  const float b2x = b2 * ex, b2y = b2 * ey;
  M.rt[0] = ey * b2x;
  M.rt[1] = ey * b2y - vz;
  M.rt[2] = wy + ey * ruz;
  M.rt[3] = vz - ex * b2x;
  M.rt[4] = -ex * b2y;
  M.rt[5] = -wx - ex * ruz;
  M.rt[6] = -wy;
  M.rt[7] = wx;
  M.rt[8] = 0.0f;
  // RR
  const float cr = C0 - k3;
  const float k33 = 3.0f * k3;
  const float r3x = k33 * rx, r3y = k33 * ry, r3z = k33 * rz;
  const float pe = -1.125f * p3, pn = 2.25f * p3;
  const float eex = pe * ex, eey = pe * ey;
  const float nx = -ey, ny = ex;
  const float nnx = pn * nx, nny = pn * ny;
  const float ezr = pe * ez + pn * ez;
  M.rr[0] = cr + r3x * rx + eex * ex + nnx * nx;
  M.rr[1] = r3x * ry + eex * ey + nnx * ny;
  M.rr[2] = r3x * rz + eex * ez;
  M.rr[3] = M.rr[1];
  M.rr[4] = cr + r3y * ry + eey * ey + nny * ny;
  M.rr[5] = r3y * rz + eey * ez;
  M.rr[6] = r3z * rx + ex * ezr;
  M.rr[7] = r3z * ry + ey * ezr;
  M.rr[8] = cr + r3z * rz + ez * ezr;
}

// V += M F (target side).
__device__ __forceinline__ void blk_apply(const Blk66& M, const float* F, float* V) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    V[a] += M.tt[3 * a] * F[0] + M.tt[3 * a + 1] * F[1] + M.tt[3 * a + 2] * F[2] +
            M.tr[3 * a] * F[3] + M.tr[3 * a + 1] * F[4] + M.tr[3 * a + 2] * F[5];
    V[3 + a] += M.rt[3 * a] * F[0] + M.rt[3 * a + 1] * F[1] + M.rt[3 * a + 2] * F[2] +
                M.rr[3 * a] * F[3] + M.rr[3 * a + 1] * F[4] + M.rr[3 * a + 2] * F[5];
  }
}

// W += M^T F (source side).
__device__ __forceinline__ void blk_apply_t(const Blk66& M, const float* F, float* W) {
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    W[b] += M.tt[b] * F[0] + M.tt[3 + b] * F[1] + M.tt[6 + b] * F[2] +
            M.rt[b] * F[3] + M.rt[3 + b] * F[4] + M.rt[6 + b] * F[5];
    W[3 + b] += M.tr[b] * F[0] + M.tr[3 + b] * F[1] + M.tr[6 + b] * F[2] +
                M.rr[b] * F[3] + M.rr[3 + b] * F[4] + M.rr[6 + b] * F[5];
  }
}

// Source records of one block, structure of arrays (rotated reads hit consecutive words).
struct SymTile {
  double x[kBlk], y[kBlk], z[kBlk];
  float zf[kBlk];
  float f[6][kBlk];
};

// Stages segment j (or, past the end, a distant unloaded placeholder) into slot l.
__device__ __forceinline__ void stage(const SymArgs& a, SymTile& t, int j, int l) {
  if (j < a.nseg) {
    const double zz = a.z[j];
    t.x[l] = a.x[j];
    t.y[l] = a.y[j];
    t.z[l] = zz;
    t.zf[l] = float(zz);
#pragma unroll
    for (int k = 0; k < 6; ++k) t.f[k][l] = float(kStageScale * a.f[(size_t)k * a.fld + j - a.foff]);
  } else {
    // far from every segment and from the other placeholders: finite blocks, zero loads
    t.x[l] = 1.0e6 + 64.0 * l;
    t.y[l] = 1.0e6;
    t.z[l] = 1.0e6;
    t.zf[l] = 1.0e6f;
#pragma unroll
    for (int k = 0; k < 6; ++k) t.f[k][l] = 0.0f;
  }
}

__global__ __launch_bounds__(256) void rpy_sym_kernel(SymArgs a) {
  __shared__ SymTile tile[kWaves];
  __shared__ double red[kWaves * 6][kBlk];
  const int I = blockIdx.x, c = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int B = a.nblk;
  // this lane's target (placeholder past the end, as staged)
  const int i = I * kBlk + lane;
  double xi, yi, zi;
  float fi[6];
  if (i < a.nseg) {
    xi = a.x[i];
    yi = a.y[i];
    zi = a.z[i];
#pragma unroll
    for (int k = 0; k < 6; ++k) fi[k] = float(kStageScale * a.f[(size_t)k * a.fld + i - a.foff]);
  } else {
    xi = 1.0e6 + 64.0 * lane + 32.0;  // apart from the source placeholders as well
    yi = 1.0e6 + 32.0;
    zi = 1.0e6;
#pragma unroll
    for (int k = 0; k < 6; ++k) fi[k] = 0.0f;
  }
  const float zif = float(zi);
  float s32[6], c32[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    s32[k] = 0.0f;
    c32[k] = 0.0f;
  }
  const int d_begin = c * a.dpw, d_end = min(d_begin + a.dpw, a.Dn + 1);
  SymTile& T = tile[wave];
  for (int d0 = d_begin; d0 < d_end; d0 += kWaves) {
    const int d = d0 + wave;
    const bool mine = d < d_end && takes(I, d, B);
    const int J = (I + d) % B;
    __syncthreads();
    if (mine) stage(a, T, J * kBlk + lane, lane);
    __syncthreads();
    if (!mine) continue;
    float part[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (d == 0) {
      // the block with itself, one way round, without the self pair
      for (int s = 1; s < kBlk; ++s) {
        const int js = (lane + s) & (kBlk - 1);
        const float dx = float(xi - T.x[js]), dy = float(yi - T.y[js]), dz = float(zi - T.z[js]);
        const float zj = T.zf[js];
        Blk66 M;
        wall_block66(dx, dy, dz, zif + zj, zj, M);
        float F[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) F[k] = T.f[k][js];
        blk_apply(M, F, part);
      }
    } else {
      float W[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      for (int s = 0; s < kBlk; ++s) {
        const int js = (lane + s) & (kBlk - 1);
        const float dx = float(xi - T.x[js]), dy = float(yi - T.y[js]), dz = float(zi - T.z[js]);
        const float zj = T.zf[js];
        Blk66 M;
        wall_block66(dx, dy, dz, zif + zj, zj, M);
        float F[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) F[k] = T.f[k][js];
        blk_apply(M, F, part);
        blk_apply_t(M, fi, W);
        // hand the running source sum to the lane that meets this source next
#pragma unroll
        for (int k = 0; k < 6; ++k) W[k] = __shfl(W[k], (lane + 1) & (kBlk - 1), 64);
      }
      // lane l now holds source l of block J
      float* dst = a.src + ((size_t)J * a.Dn + (d - 1)) * 6 * kBlk;
#pragma unroll
      for (int k = 0; k < 6; ++k) dst[k * kBlk + lane] = W[k];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) two_sum_add(s32[k], c32[k], part[k]);
  }
  // the four waves' target sums, added in wave order
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 6; ++k) red[wave * 6 + k][lane] = double(s32[k]) + double(c32[k]);
  __syncthreads();
  if (wave != 0) return;
  double* o = a.own + ((size_t)I * a.S + c) * 6 * kBlk;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double s = red[k][lane];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) s += red[w * 6 + k][lane];
    o[k * kBlk + lane] = s;
  }
}

// v_i = (sum of the own and source partials of segment i + self term) / mu, stored or
// accumulated.  One workgroup per block; wave w sums partial slots w, w+4, ...
__global__ __launch_bounds__(256) void rpy_sym_finalize_kernel(SymArgs a) {
  __shared__ double red[kWaves][6][kBlk];
  const int I = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int B = a.nblk;
  const int i = I * kBlk + lane;
  double part[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int nslot = a.S + a.Dn;
  for (int t = w; t < nslot; t += kWaves) {
    if (t < a.S) {
      const double* o = a.own + ((size_t)I * a.S + t) * 6 * kBlk;
#pragma unroll
      for (int k = 0; k < 6; ++k) part[k] += o[k * kBlk + lane];
    } else {
      const int d = t - a.S + 1;
      if (!takes((I - d + B) % B, d, B)) continue;
      const float* s = a.src + ((size_t)I * a.Dn + (d - 1)) * 6 * kBlk;
#pragma unroll
      for (int k = 0; k < 6; ++k) part[k] += double(s[k * kBlk + lane]);
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) red[w][k][lane] = part[k];
  __syncthreads();
  if (w != 0 || i >= a.nseg) return;
  double out[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) out[k] = (red[0][k][lane] + red[1][k][lane]) + (red[2][k][lane] + red[3][k][lane]);
  double F[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) F[k] = a.f[(size_t)k * a.fld + i - a.foff];
  wall_self<6, 6>(a.z[i], F, out);
  const double scale = 1.0 / a.mu;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    double* dst = a.v + (size_t)k * a.vld + i - a.voff;
    *dst = a.acc ? *dst + scale * out[k] : scale * out[k];
  }
}

}  // namespace

bool sym_applicable(const cfm_ctx* c, const ApplyPlan& plan) {
  const Particles& p = c->p;
  return c->sym_pairs && c->cfg.precision == CFM_FP32C && c->cfg.wall && p.seg_dof == 6 &&
         plan.src_seg && !plan.src_blob && !plan.partial_ext && plan.tgt_begin <= 0 &&
         plan.tgt_end >= p.n_seg && p.n_seg >= 2 * kBlk;
}

// Workgroups per block row: enough for a few workgroups per CU, offsets per workgroup a
// multiple of the four waves.
static void sym_shape(int nseg, int* B, int* Dn, int* S, int* dpw) {
  *B = (nseg + kBlk - 1) / kBlk;
  *Dn = *B / 2;
  const int want = std::max(1, (8 * 256 + *B - 1) / *B);
  int per = (*Dn + 1 + want - 1) / want;
  per = std::max(kWaves, (per + kWaves - 1) / kWaves * kWaves);
  *dpw = per;
  *S = (*Dn + 1 + per - 1) / per;
}

Status launch_sym(cfm_ctx* c, const ApplyPlan& plan, hipStream_t stream) {
  const Particles& p = c->p;
  SymArgs a{};
  a.x = p.x.p;
  a.y = p.y.p;
  a.z = p.z.p;
  a.f = plan.f_ext ? plan.f_ext : p.f.p;
  a.fld = plan.f_ext ? plan.f_ld : p.n;
  a.foff = plan.f_ext ? plan.f_off : 0;
  a.v = plan.v_ext ? plan.v_ext : p.v.p;
  a.vld = plan.v_ext ? plan.v_ld : p.n;
  a.voff = plan.v_ext ? plan.v_off : 0;
  a.acc = plan.acc_seg ? 1 : 0;
  a.mu = c->cfg.mu;
  a.nseg = p.n_seg;
  sym_shape(p.n_seg, &a.nblk, &a.Dn, &a.S, &a.dpw);
  DevBuf<double>& own = plan.side ? c->sym_own_side : c->sym_own;
  DevBuf<float>& src = plan.side ? c->sym_src_side : c->sym_src;
  CFM_TRY(own.alloc((size_t)a.nblk * a.S * 6 * kBlk));
  CFM_TRY(src.alloc((size_t)a.nblk * std::max(a.Dn, 1) * 6 * kBlk));
  a.own = own.p;
  a.src = src.p;
  hipLaunchKernelGGL(rpy_sym_kernel, dim3(a.nblk, a.S), dim3(256), 0, stream, a);
  CFM_CHECK_LAUNCH();
  hipLaunchKernelGGL(rpy_sym_finalize_kernel, dim3(a.nblk), dim3(256), 0, stream, a);
  CFM_CHECK_LAUNCH();
  return {};
}

}  // namespace cfm
