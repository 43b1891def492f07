// apply_tile.hpp -- the LDS source record and the inner pair loop shared by the
// single-system and batched mobility products.
#pragma once
#include "rpy_pair.hpp"

namespace cfm {

constexpr int kWG = 256;    // threads per workgroup
constexpr int kTT = 64;     // targets per workgroup
constexpr int kTS = 256;    // sources per LDS tile
constexpr int kWaves = kWG / 64;

// Wall-branch sources are staged as (1/(6 pi)) F (the form wall_pair_scaled takes).
constexpr double kStageScale = 0.053051647697298445;

template <typename T> struct alignas(16) SrcRec;
template <> struct alignas(16) SrcRec<float> {
  double x, y, z;
  float zf;
  float f[6];
  float pad_;
};
template <> struct alignas(16) SrcRec<double> {
  double x, y, z, pad_;
  double f[6];
};


// Two-sum accumulation of a float partial into (s, c).
__device__ __forceinline__ void two_sum_add(float& s, float& c, float x) {
  const float t = s + x;
  const float bp = t - s;
  const float err = (s - (t - bp)) + (x - bp);
  s = t;
  c += err;
}

template <typename T, int TD, int FD, bool WALL, bool CHECK>
__device__ __forceinline__ void tile_sweep(const SrcRec<T>* __restrict__ rec, int cnt, int j0,
                                           int iid, double xi, double yi, double zi, T zif,
                                           T ai, T aj, T mu, T* __restrict__ acc) {
#pragma unroll 2
  for (int jj = 0; jj < cnt; ++jj) {
    const SrcRec<T>& s = rec[jj];
    T dx = T(xi - s.x), dy = T(yi - s.y), dz = T(zi - s.z);
    T F[FD];
#pragma unroll
    for (int k = 0; k < FD; ++k) F[k] = s.f[k];
    if constexpr (WALL) {
      T zj;
      if constexpr (sizeof(T) == 4) zj = s.zf; else zj = s.z;
      if constexpr (CHECK) {
        if (j0 + jj == iid) {  // same particle: handled by the self term
          dx = T(1);
          dy = T(0);
          dz = T(0);
#pragma unroll
          for (int k = 0; k < FD; ++k) F[k] = T(0);
        }
      }
      wall_pair_scaled<T, TD, FD>(dx, dy, dz, zif + zj, zj, F, acc);
    } else {
      free_pair<T, TD, FD>(dx, dy, dz, ai, aj, mu, F, acc);
    }
  }
}

}  // namespace cfm
