// apply_sym.hip -- wall-RPY products of one species from itself that evaluate each
// unordered pair once: segments from segments (6 or 3 DOF) and blobs from blobs, in
// FP32C or FP64.
//
// The wall-corrected grand mobility is symmetric: the block of target i and source j is
// the transpose of that of target j and source i.  The one-directional kernel
// (apply.hip) evaluates the pair formula for both orders; here the block M_ij is formed
// explicitly once and applied both ways, v_i += M_ij F_j and v_j += M_ij^T F_i.  For 6
// DOF that costs about 0.7 of the two directed evaluations; for the 3 x 3 blocks of the
// blobs, whose application is cheap beside the formula, about 0.4.
//
// Work decomposition: the particles of the species are cut into B blocks of 64.  Block
// pair (I, J) with J = I + d (mod B) is taken by the workgroups of row I for ring offsets
// d = 1 .. B/2 (d = B/2, when B is even, only for I < B/2), so every unordered block
// pair is taken exactly once; d = 0 (the block with itself) is evaluated one way round,
// skipping the self pair.  A workgroup (I, c) owns the 64 targets of block I, one per
// lane, and the ring offsets [c dpw, (c+1) dpw); its four waves take one offset each per
// round.  In step s of a block pair, lane l pairs its target with source (l + s) mod 64
// of block J (records in LDS, structure-of-arrays so that the rotated reads are
// conflict-free) and carries the running sum of that source's velocity, handed to the
// lane below after every step: after 64 steps lane l holds the complete contribution of
// block I to source l of block J.
//
// Precision as the one-directional kernels: FP32C forms separations in double, narrows
// them, evaluates the block in float, sums each 64-pair run in float and folds it per
// target into a two-sum accumulator; FP64 works in double throughout.  Every partial is
// stored and summed in a fixed order (deterministic).
//
// Partials (device memory of the context): own[I][c][D][64] (double) holds workgroup
// (I, c)'s target sums, src[J][d-1][D][64] (the arithmetic type) block I = J - d's
// contribution to the sources of block J.  rpy_sym_finalize_kernel adds them, the self
// term, scales by 1/mu and stores or accumulates the velocities.
#include <algorithm>

#include "apply_tile.hpp"
#include "ctx.hpp"
#include "rpy_pair.hpp"

namespace cfm {

namespace {

constexpr int kBlk = 64;  // segments per block

template <typename T>
struct SymArgs {
  const double* __restrict__ x;  // positions of the species (already offset to its first id)
  const double* __restrict__ y;
  const double* __restrict__ z;
  const double* __restrict__ f;  // loads: row k of species particle j at f[k fld + j]
  double* __restrict__ v;        // velocities: v[k vld + j]
  double* __restrict__ own;      // B x S x D x 64
  T* __restrict__ src;           // B x Dn x D x 64
  int cnt, nblk, S, Dn, dpw;
  size_t fld, vld;
  int acc;
  double mu;
  // ensembles of equal members (blockIdx.z = member): strides of the positions, of the
  // load / velocity vectors and of the partial sets, and a per-member status filter
  size_t mpos, mvec, mown, msrc;
  const int* status;  // member e runs when status[e * sstride] == want (null: all)
  int sstride, want;
};

// The arguments of member blockIdx.z, or false when that member is not evaluated.
template <typename T>
__device__ __forceinline__ bool member_args(SymArgs<T>& a) {
  const int e = blockIdx.z;
  if (a.status && a.status[(size_t)e * a.sstride] != a.want) return false;
  a.x += e * a.mpos;
  a.y += e * a.mpos;
  a.z += e * a.mpos;
  a.f += e * a.mvec;
  a.v += e * a.mvec;
  a.own += e * a.mown;
  a.src += e * a.msrc;
  return true;
}

// Does workgroup row I take ring offset d?
__device__ __forceinline__ bool takes(int I, int d, int B) {
  if (d == 0) return true;
  if (2 * d < B) return true;
  return 2 * d == B && I < d;
}

// The 6 x 6 wall block (target i, source j) for loads staged with 1/(6 pi), in the
// arithmetic type T: the same terms as wall_pair_scaled written as matrix entries.  With r the unit separation,
// e the image direction (dx, dy, Rz) p1, n = z x e = (-e_y, e_x, 0):
//   TT = (iso + A0) I + a r r^T + a1 e e^T + a2 e_z e z^T + a3 e_z z e^T + a4 z z^T
//   TR = [m - q]x + (d2 (e_x, e_y, -e_z) - d1 z) n^T
//   RT = [b - q]x - n (b1 z + b2 e)^T
//   RR = (C0 - k3) I + 3 k3 r r^T - 1.125 p3 e e^T + 2.25 p3 e_z z e^T + 2.25 p3 n n^T
// with q = k2 r, b = B0 e + B3 z, m = B0 (e_x, e_y, -e_z) - D3 z.
template <typename T>
struct Blk66 {
  T tt[9], tr[9], rt[9], rr[9];
};

// D = 3: the translational block TT only.
template <typename T, int D>
__device__ __forceinline__ void wall_block(T dx, T dy, T dz, T Rz, T zj, Blk66<T>& M) {
  const T hxy = dx * dx + dy * dy;
  // free space
  const T is = rsq<T>(hxy + dz * dz);
  const T rx = dx * is, ry = dy * is, rz = dz * is;
  const T is2 = is * is;
  const T k1 = T(0.75) * is;
  const T iso = k1 * (T(1) + T(2.0 / 3.0) * is2);
  const T a_ = k1 * (T(1) - T(2) * is2);
  // images
  const T p1 = rsq<T>(hxy + Rz * Rz);
  const T p2 = p1 * p1, p3 = p2 * p1, p4 = p2 * p2, p5 = p4 * p1;
  const T ex = dx * p1, ey = dy * p1, ez = Rz * p1;
  const T u = ez * ez;
  const T h = zj * rcp<T>(Rz);
  const T hh = h * (T(1) - h);
  const T q3 = T(1) - T(5) * u;
  const T A0 = -(T(0.75) * p1 * (T(1) + T(2) * hh * u) + T(0.5) * (p3 * (T(1) - T(3) * u) - p5 * q3));
  const T a1 = -(T(0.75) * p1 * (T(1) - T(6) * hh * u) - T(1.5) * p3 * q3 + T(2.5) * p5 * (T(1) - T(7) * u));
  const T a2 = T(1.5) * h * p1 * (T(1) - T(6) * (T(1) - h) * u) - T(3) * p3 * q3 + T(5) * p5 * (T(2) - T(7) * u);
  const T a3 = T(1.5) * h * p1 - T(5) * p5;
  const T a4 = -((T(3) * h * h * p1 + T(3) * p3) * u + (T(2) - T(15) * u) * p5);
  // TT
  const T c0 = iso + A0;
  const T arx = a_ * rx, ary = a_ * ry, arz = a_ * rz;
  const T aex = a1 * ex, aey = a1 * ey;
  const T gc = (a1 + a2) * ez, gr = (a1 + a3) * ez;
  M.tt[0] = c0 + arx * rx + aex * ex;
  M.tt[1] = arx * ry + aex * ey;
  M.tt[2] = arx * rz + ex * gc;
  M.tt[3] = M.tt[1];
  M.tt[4] = c0 + ary * ry + aey * ey;
  M.tt[5] = ary * rz + ey * gc;
  M.tt[6] = arz * rx + ex * gr;
  M.tt[7] = arz * ry + ey * gr;
  M.tt[8] = c0 + arz * rz + ez * (gc + a3 * ez) + a4;
  if constexpr (D == 6) {
    const T k2 = k1 * is;
    const T qx = k2 * rx, qy = k2 * ry, qz = k2 * rz;
    const T k3 = T(0.5) * k2 * is;
    const T B0 = T(0.75) * p2;
    const T hp2 = h * p2;
    const T w4 = (T(1.5) - T(15) * u) * p4;
    const T b1 = T(9) * u * hp2 + w4;
    const T b2 = -ez * (T(4.5) * hp2 - T(7.5) * p4);
    const T B3 = T(-1.5) * ez * (hp2 - p4);
    const T C0 = p3 * (T(0.375) - T(2.25) * u);
    const T gp2 = p2 - hp2;
    const T d1 = -(T(9) * u * gp2 + w4);
    const T d2 = ez * (T(4.5) * gp2 - T(7.5) * p4);
    const T D3 = T(-1.5) * ez * (gp2 - p4);
    // TR (velocity <- torque) and RT (rotation <- force)
    const T wx = B0 * ex - qx, wy = B0 * ey - qy;
    const T wz = -B0 * ez - D3 - qz;
    const T tuz = -d2 * ez - d1;
    const T d2x = d2 * ex, d2y = d2 * ey;
    M.tr[0] = -d2x * ey;
    M.tr[1] = d2x * ex - wz;
    M.tr[2] = wy;
    M.tr[3] = wz - d2y * ey;
    M.tr[4] = d2y * ex;
    M.tr[5] = -wx;
    M.tr[6] = -wy - ey * tuz;
    M.tr[7] = wx + ex * tuz;
    M.tr[8] = T(0);
    const T vz = B0 * ez + B3 - qz;
    const T ruz = b2 * ez + b1;
    const T b2x = b2 * ex, b2y = b2 * ey;
    M.rt[0] = ey * b2x;
    M.rt[1] = ey * b2y - vz;
    M.rt[2] = wy + ey * ruz;
    M.rt[3] = vz - ex * b2x;
    M.rt[4] = -ex * b2y;
    M.rt[5] = -wx - ex * ruz;
    M.rt[6] = -wy;
    M.rt[7] = wx;
    M.rt[8] = T(0);
    // RR
    const T cr = C0 - k3;
    const T k33 = T(3) * k3;
    const T r3x = k33 * rx, r3y = k33 * ry, r3z = k33 * rz;
    const T pe = T(-1.125) * p3, pn = T(2.25) * p3;
    const T eex = pe * ex, eey = pe * ey;
    const T nx = -ey, ny = ex;
    const T nnx = pn * nx, nny = pn * ny;
    const T ezr = pe * ez + pn * ez;
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
}

// V += M F (target side), every term a fused multiply-add into the accumulator.
template <typename T, int D>
__device__ __forceinline__ void blk_apply(const Blk66<T>& M, const T* F, T* V) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      V[a] = fma(M.tt[3 * a + b], F[b], V[a]);
      if constexpr (D == 6) {
        V[a] = fma(M.tr[3 * a + b], F[3 + b], V[a]);
        V[3 + a] = fma(M.rt[3 * a + b], F[b], V[3 + a]);
        V[3 + a] = fma(M.rr[3 * a + b], F[3 + b], V[3 + a]);
      }
    }
  }
}

// W += M^T F (source side).
template <typename T, int D>
__device__ __forceinline__ void blk_apply_t(const Blk66<T>& M, const T* F, T* W) {
#pragma unroll
  for (int b = 0; b < 3; ++b) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      W[b] = fma(M.tt[3 * a + b], F[a], W[b]);
      if constexpr (D == 6) {
        W[b] = fma(M.rt[3 * a + b], F[3 + a], W[b]);
        W[3 + b] = fma(M.tr[3 * a + b], F[a], W[3 + b]);
        W[3 + b] = fma(M.rr[3 * a + b], F[3 + a], W[3 + b]);
      }
    }
  }
}

// Source records of one block, structure of arrays (rotated reads hit consecutive words).
template <typename T>
struct SymTile {
  double x[kBlk], y[kBlk], z[kBlk];
  T zf[kBlk];
  T f[6][kBlk];
};

// Stages particle j of the species (or, past the end, a distant unloaded placeholder)
// into slot l.
template <typename T, int D>
__device__ __forceinline__ void stage(const SymArgs<T>& a, SymTile<T>& t, int j, int l) {
  if (j < a.cnt) {
    const double zz = a.z[j];
    t.x[l] = a.x[j];
    t.y[l] = a.y[j];
    t.z[l] = zz;
    t.zf[l] = T(zz);
#pragma unroll
    for (int k = 0; k < D; ++k) t.f[k][l] = T(kStageScale * a.f[(size_t)k * a.fld + j]);
  } else {
    // far from every particle and from the other placeholders: finite blocks, zero loads
    t.x[l] = 1.0e6 + 64.0 * l;
    t.y[l] = 1.0e6;
    t.z[l] = 1.0e6;
    t.zf[l] = T(1.0e6);
#pragma unroll
    for (int k = 0; k < D; ++k) t.f[k][l] = T(0);
  }
}

// Waves that share one block pair: in FP64 two waves take 32 steps each of a block
// pair's 64 (twice the waves of half the length: the FP64 blob products have few block
// pairs, and short waves fill the CUs evenly); in FP32 one wave takes all 64 steps.
template <typename T>
constexpr int sym_halves() { return sizeof(T) == 8 ? 2 : 1; }

template <typename T, int D>
__global__ __launch_bounds__(256) void rpy_sym_kernel(SymArgs<T> a) {
  if (!member_args(a)) return;
  constexpr int H = sym_halves<T>();
  constexpr int kOff = kWaves / H;    // ring offsets per round
  constexpr int kSteps = kBlk / H;    // steps of a block pair per wave
  __shared__ SymTile<T> tile[kOff];
  __shared__ double red[kWaves * D][kBlk];
  constexpr bool kF32 = sizeof(T) == 4;
  const int I = blockIdx.x, c = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int half = wave % H, slot = wave / H;
  const int B = a.nblk;
  // this lane's target (placeholder past the end, as staged)
  const int i = I * kBlk + lane;
  double xi, yi, zi;
  T fi[D];
  if (i < a.cnt) {
    xi = a.x[i];
    yi = a.y[i];
    zi = a.z[i];
#pragma unroll
    for (int k = 0; k < D; ++k) fi[k] = T(kStageScale * a.f[(size_t)k * a.fld + i]);
  } else {
    xi = 1.0e6 + 64.0 * lane + 32.0;  // apart from the source placeholders as well
    yi = 1.0e6 + 32.0;
    zi = 1.0e6;
#pragma unroll
    for (int k = 0; k < D; ++k) fi[k] = T(0);
  }
  const T zit = T(zi);
  // FP32C: two-sum (s, c) per component; FP64: s alone
  T s_acc[D], c_acc[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    s_acc[k] = T(0);
    c_acc[k] = T(0);
  }
  const int d_begin = c * a.dpw, d_end = min(d_begin + a.dpw, a.Dn + 1);
  SymTile<T>& Tl = tile[slot];
  // H = 2: the second half's source sums, [slot][k][lane] (red is free until the end)
  T* wsum = reinterpret_cast<T*>(&red[0][0]);
  for (int d0 = d_begin; d0 < d_end; d0 += kOff) {
    const int d = d0 + slot;
    const bool mine = d < d_end && takes(I, d, B);
    const int J = (I + d) % B;
    __syncthreads();
    if (mine && half == 0) stage<T, D>(a, Tl, J * kBlk + lane, lane);
    __syncthreads();
    T W[D];
#pragma unroll
    for (int k = 0; k < D; ++k) W[k] = T(0);
    if (mine) {
      T part[D];
#pragma unroll
      for (int k = 0; k < D; ++k) part[k] = T(0);
      const int s0 = half * kSteps, s1 = s0 + kSteps;
      if (d == 0) {
        // the block with itself, one way round, without the self pair
        for (int s = max(s0, 1); s < s1; ++s) {
          const int js = (lane + s) & (kBlk - 1);
          const T dx = T(xi - Tl.x[js]), dy = T(yi - Tl.y[js]), dz = T(zi - Tl.z[js]);
          const T zj = Tl.zf[js];
          Blk66<T> M;
          wall_block<T, D>(dx, dy, dz, zit + zj, zj, M);
          T F[D];
#pragma unroll
          for (int k = 0; k < D; ++k) F[k] = Tl.f[k][js];
          blk_apply<T, D>(M, F, part);
        }
      } else {
        for (int s = s0; s < s1; ++s) {
          const int js = (lane + s) & (kBlk - 1);
          const T dx = T(xi - Tl.x[js]), dy = T(yi - Tl.y[js]), dz = T(zi - Tl.z[js]);
          const T zj = Tl.zf[js];
          Blk66<T> M;
          wall_block<T, D>(dx, dy, dz, zit + zj, zj, M);
          T F[D];
#pragma unroll
          for (int k = 0; k < D; ++k) F[k] = Tl.f[k][js];
          blk_apply<T, D>(M, F, part);
          blk_apply_t<T, D>(M, fi, W);
          // hand the running source sum to the lane that meets this source next
#pragma unroll
          for (int k = 0; k < D; ++k) W[k] = __shfl(W[k], (lane + 1) & (kBlk - 1), 64);
        }
        if constexpr (H > 1) {
          // lane l holds source (l + s1) mod 64: rotate so that lane l holds source l
          const int r = s1 & (kBlk - 1);
#pragma unroll
          for (int k = 0; k < D; ++k) W[k] = __shfl(W[k], (lane - r) & (kBlk - 1), 64);
        }
      }
#pragma unroll
      for (int k = 0; k < D; ++k) {
        if constexpr (kF32)
          two_sum_add(s_acc[k], c_acc[k], part[k]);
        else
          s_acc[k] += part[k];
      }
    }
    // lane l of the first half now holds source l of block J (plus, for H = 2, the second
    // half's sum for it, added after the first's)
    if constexpr (H > 1) {
      if (mine && d != 0 && half == 1) {
#pragma unroll
        for (int k = 0; k < D; ++k) wsum[(slot * D + k) * kBlk + lane] = W[k];
      }
      __syncthreads();
      if (mine && d != 0 && half == 0) {
        T* dst = a.src + ((size_t)J * a.Dn + (d - 1)) * D * kBlk;
#pragma unroll
        for (int k = 0; k < D; ++k) dst[k * kBlk + lane] = W[k] + wsum[(slot * D + k) * kBlk + lane];
      }
    } else {
      if (mine && d != 0) {
        T* dst = a.src + ((size_t)J * a.Dn + (d - 1)) * D * kBlk;
#pragma unroll
        for (int k = 0; k < D; ++k) dst[k * kBlk + lane] = W[k];
      }
    }
  }
  // the four waves' target sums, added in wave order
  __syncthreads();
#pragma unroll
  for (int k = 0; k < D; ++k) red[wave * D + k][lane] = double(s_acc[k]) + double(c_acc[k]);
  __syncthreads();
  if (wave != 0) return;
  double* o = a.own + ((size_t)I * a.S + c) * D * kBlk;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double sum = red[k][lane];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) sum += red[w * D + k][lane];
    o[k * kBlk + lane] = sum;
  }
}

// v_i = (sum of the own and source partials of particle i + self term) / mu, stored or
// accumulated.  One workgroup per block; wave w sums partial slots w, w+4, ...
template <typename T, int D>
__global__ __launch_bounds__(256) void rpy_sym_finalize_kernel(SymArgs<T> a) {
  if (!member_args(a)) return;
  __shared__ double red[kWaves][D][kBlk];
  const int I = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int B = a.nblk;
  const int i = I * kBlk + lane;
  double part[D];
#pragma unroll
  for (int k = 0; k < D; ++k) part[k] = 0.0;
  const int nslot = a.S + a.Dn;
  for (int t = w; t < nslot; t += kWaves) {
    if (t < a.S) {
      const double* o = a.own + ((size_t)I * a.S + t) * D * kBlk;
#pragma unroll
      for (int k = 0; k < D; ++k) part[k] += o[k * kBlk + lane];
    } else {
      const int d = t - a.S + 1;
      if (!takes((I - d + B) % B, d, B)) continue;
      const T* s = a.src + ((size_t)I * a.Dn + (d - 1)) * D * kBlk;
#pragma unroll
      for (int k = 0; k < D; ++k) part[k] += double(s[k * kBlk + lane]);
    }
  }
#pragma unroll
  for (int k = 0; k < D; ++k) red[w][k][lane] = part[k];
  __syncthreads();
  if (w != 0 || i >= a.cnt) return;
  double out[D];
#pragma unroll
  for (int k = 0; k < D; ++k) out[k] = (red[0][k][lane] + red[1][k][lane]) + (red[2][k][lane] + red[3][k][lane]);
  double F[D];
#pragma unroll
  for (int k = 0; k < D; ++k) F[k] = a.f[(size_t)k * a.fld + i];
  wall_self<D, D>(a.z[i], F, out);
  const double scale = 1.0 / a.mu;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double* dst = a.v + (size_t)k * a.vld + i;
    *dst = a.acc ? *dst + scale * out[k] : scale * out[k];
  }
}

}  // namespace

// The species a plan's product may take from the symmetric kernels: 1 segments,
// 2 blobs, 0 none.  Wall products (the Swan-Brady blocks are the symmetric ones here)
// whose sources are one species and whose targets cover all of it.
int sym_species(const cfm_ctx* c, const ApplyPlan& plan) {
  const Particles& p = c->p;
  if (!c->sym_pairs || !c->cfg.wall || plan.partial_ext || plan.src_seg == plan.src_blob) return 0;
  if (plan.src_seg)
    return (plan.tgt_begin <= 0 && plan.tgt_end >= p.n_seg && p.n_seg >= 2 * kBlk) ? 1 : 0;
  return (plan.tgt_begin <= p.n_seg && plan.tgt_end >= p.n && p.n_blob >= 2 * kBlk) ? 2 : 0;
}

// Workgroups per block row: enough for a few workgroups per CU (twice as many in FP64,
// whose waves take half a block pair each), offsets per workgroup a multiple of the
// offsets of one round and at most eight.  Short workgroups (about
// 25 us at 26,000 segments) leave no long tail at the end of a launch and, when a
// product runs on a second stream, free a slot for the other stream's kernels every few
// microseconds: 36 offsets per workgroup ran the 26,000-segment product 4% slower and
// held the blob solve beside it waiting for slots.
template <typename T>
static void sym_shape(int cnt, int* B, int* Dn, int* S, int* dpw) {
  constexpr int kOff = kWaves / sym_halves<T>();  // ring offsets per round
  *B = (cnt + kBlk - 1) / kBlk;
  *Dn = *B / 2;
  const int want = std::max(1, (8 * 256 * (kWaves / kOff) + *B - 1) / *B);
  int per = (*Dn + 1 + want - 1) / want;
  per = std::max(kOff, (per + kOff - 1) / kOff * kOff);
  per = std::min(per, 2 * kWaves);
  *dpw = per;
  *S = (*Dn + 1 + per - 1) / per;
}

template <typename T, int D>
static Status launch_sym_t(cfm_ctx* c, const ApplyPlan& plan, int base, int cnt, hipStream_t stream,
                           DevBuf<T>& src) {
  const Particles& p = c->p;
  SymArgs<T> a{};
  a.x = p.x.p + base;
  a.y = p.y.p + base;
  a.z = p.z.p + base;
  if (plan.f_ext) {
    a.f = plan.f_ext + (base - plan.f_off);
    a.fld = plan.f_ld;
  } else {
    a.f = p.f.p + base;
    a.fld = p.n;
  }
  if (plan.v_ext) {
    a.v = plan.v_ext + (base - plan.v_off);
    a.vld = plan.v_ld;
  } else {
    a.v = p.v.p + base;
    a.vld = p.n;
  }
  a.acc = (base == 0 ? plan.acc_seg : plan.acc_blob) ? 1 : 0;
  a.mu = c->cfg.mu;
  a.cnt = cnt;
  sym_shape<T>(cnt, &a.nblk, &a.Dn, &a.S, &a.dpw);
  DevBuf<double>& own = plan.side ? c->sym_own_side : c->sym_own;
  CFM_TRY(own.alloc((size_t)a.nblk * a.S * D * kBlk));
  CFM_TRY(src.alloc((size_t)a.nblk * std::max(a.Dn, 1) * D * kBlk));
  a.own = own.p;
  a.src = src.p;
  hipLaunchKernelGGL((rpy_sym_kernel<T, D>), dim3(a.nblk, a.S), dim3(256), 0, stream, a);
  CFM_CHECK_LAUNCH();
  hipLaunchKernelGGL((rpy_sym_finalize_kernel<T, D>), dim3(a.nblk), dim3(256), 0, stream, a);
  CFM_CHECK_LAUNCH();
  return {};
}

// Ensembles of equal members (wall): member e's particles sit at [e ns, (e+1) ns) of the
// position arrays and its loads / velocities at f + e mvec, v + e mvec with component
// strides fld, vld; members whose status differs from `want` are skipped (status may be
// null).  Velocities are overwritten (scaled by 1/mu, self term included).  Each member
// is decomposed as a lone system of ns particles would be, so its sums are the same.
template <typename T, int D>
static Status sym_members_t(hipStream_t stream, const double* x, const double* y, const double* z,
                            int ns, int members, const double* f, size_t fld, size_t mvec,
                            double* v, size_t vld, double mu, const int* status, int sstride,
                            int want, DevBuf<double>& own, DevBuf<T>& src) {
  if (ns <= 0 || members <= 0) return {};
  SymArgs<T> a{};
  a.x = x;
  a.y = y;
  a.z = z;
  a.f = f;
  a.v = v;
  a.fld = fld;
  a.vld = vld;
  a.acc = 0;
  a.mu = mu;
  a.cnt = ns;
  sym_shape<T>(ns, &a.nblk, &a.Dn, &a.S, &a.dpw);
  a.mpos = ns;
  a.mvec = mvec;
  a.mown = (size_t)a.nblk * a.S * D * kBlk;
  a.msrc = (size_t)a.nblk * std::max(a.Dn, 1) * D * kBlk;
  a.status = status;
  a.sstride = sstride;
  a.want = want;
  CFM_TRY(own.alloc(a.mown * members));
  CFM_TRY(src.alloc(a.msrc * members));
  a.own = own.p;
  a.src = src.p;
  hipLaunchKernelGGL((rpy_sym_kernel<T, D>), dim3(a.nblk, a.S, members), dim3(256), 0, stream, a);
  CFM_CHECK_LAUNCH();
  hipLaunchKernelGGL((rpy_sym_finalize_kernel<T, D>), dim3(a.nblk, 1, members), dim3(256), 0, stream, a);
  CFM_CHECK_LAUNCH();
  return {};
}

// 3 DOF, FP64, member vectors contiguous (component stride ns).
Status sym_members_apply(hipStream_t stream, const double* x, const double* y, const double* z, int ns,
                         int members, const double* f, size_t mvec, double* v, double mu,
                         const int* status, int sstride, int want, DevBuf<double>& own,
                         DevBuf<double>& src) {
  return sym_members_t<double, 3>(stream, x, y, z, ns, members, f, ns, mvec, v, ns, mu, status,
                                  sstride, want, own, src);
}

// The context's segments as `members` equal systems of ns segments each (6 or 3 DOF, in
// the context's precision): v_seg = M_ss f_seg per member, no coupling between members;
// members with active[e] == 0 keep their velocities.
Status sym_members_segments(cfm_ctx* c, int members, const int* active) {
  const Particles& p = c->p;
  if (members <= 0 || p.n_seg % members != 0)
    return make_status(CFM_EINVAL, "members must divide the segments");
  const int ns = p.n_seg / members;
  const bool fp32 = c->cfg.precision == CFM_FP32C;
  const size_t fld = p.n, vld = p.n;
  if (fp32)
    return p.seg_dof == 6
               ? sym_members_t<float, 6>(c->stream, p.x.p, p.y.p, p.z.p, ns, members, p.f.p, fld, ns,
                                         p.v.p, vld, c->cfg.mu, active, 1, 1, c->sym_own, c->sym_src)
               : sym_members_t<float, 3>(c->stream, p.x.p, p.y.p, p.z.p, ns, members, p.f.p, fld, ns,
                                         p.v.p, vld, c->cfg.mu, active, 1, 1, c->sym_own, c->sym_src);
  return p.seg_dof == 6
             ? sym_members_t<double, 6>(c->stream, p.x.p, p.y.p, p.z.p, ns, members, p.f.p, fld, ns,
                                        p.v.p, vld, c->cfg.mu, active, 1, 1, c->sym_own, c->sym_src64)
             : sym_members_t<double, 3>(c->stream, p.x.p, p.y.p, p.z.p, ns, members, p.f.p, fld, ns,
                                        p.v.p, vld, c->cfg.mu, active, 1, 1, c->sym_own, c->sym_src64);
}

Status launch_sym(cfm_ctx* c, const ApplyPlan& plan, int species, hipStream_t stream) {
  const Particles& p = c->p;
  const bool fp32 = c->cfg.precision == CFM_FP32C;
  const int base = species == 1 ? 0 : p.n_seg, cnt = species == 1 ? p.n_seg : p.n_blob;
  const int dof = species == 1 ? p.seg_dof : 3;
  if (fp32) {
    DevBuf<float>& src = plan.side ? c->sym_src_side : c->sym_src;
    return dof == 6 ? launch_sym_t<float, 6>(c, plan, base, cnt, stream, src)
                    : launch_sym_t<float, 3>(c, plan, base, cnt, stream, src);
  }
  DevBuf<double>& src = plan.side ? c->sym_src64_side : c->sym_src64;
  return dof == 6 ? launch_sym_t<double, 6>(c, plan, base, cnt, stream, src)
                  : launch_sym_t<double, 3>(c, plan, base, cnt, stream, src);
}

}  // namespace cfm
