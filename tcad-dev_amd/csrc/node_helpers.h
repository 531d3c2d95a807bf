// Device helpers shared by the Krylov and AC translation units: where a
// workgroup's device is, fixed-order block and device reductions, and the
// node a thread owns with its stencil neighbours.
#pragma once

#include "tcad_internal.h"

namespace tcad {

// ---------------------------------------------------------------------------
// shared device helpers
// ---------------------------------------------------------------------------
// Values a kernel needs after its first branches, pinned to the top of the
// kernel: an empty asm that consumes them (uniform values, in scalar
// registers) makes their loads join one early round instead of each
// waiting behind the branch that precedes its first use.  Loads after it
// are issued as vector loads (the asm counts as a possible store).
#define KARGS_EARLY(...) asm volatile("" ::KARGS_S(__VA_ARGS__));
#define KARGS_S1(a) "s"(a)
#define KARGS_S(...) KARGS_SN(__VA_ARGS__, KARGS_S5, KARGS_S4, KARGS_S3, KARGS_S2, KARGS_S1)(__VA_ARGS__)
#define KARGS_SN(_1, _2, _3, _4, _5, NAME, ...) NAME
#define KARGS_S2(a, b) KARGS_S1(a), KARGS_S1(b)
#define KARGS_S3(a, b, c) KARGS_S1(a), KARGS_S2(b, c)
#define KARGS_S4(a, b, c, d) KARGS_S1(a), KARGS_S3(b, c, d)
#define KARGS_S5(a, b, c, d, e) KARGS_S1(a), KARGS_S4(b, c, d, e)

struct Where {
  const DevInfo *info;              // batch table
  const int *blk2dev;
  int single;
};

__device__ __forceinline__ int dev_index(const Where &W) {
  return W.single ? 0 : W.blk2dev[blockIdx.x];
}

// By value from the table, copied where the kernel starts: the workgroup's
// device index is uniform and nothing is stored yet, so the fields arrive
// in one round of batched scalar loads.  (A reference that selected a
// kernel-argument copy for single-device batches made every field access
// a conditional load of its own, a chain of dependent scalar loads.)
__device__ __forceinline__ DevInfo dev_info(const Where &W, int d) { return W.info[d]; }

// v of lane i + O within its row of 16 lanes (O = 1, 2, 4, 8), by DPP row
// shifts of the two halves (lanes whose source leaves the row read 0).
template <int O>
__device__ __forceinline__ double row_down(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x100 + O, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x100 + O, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// Wavefront sum into lane 0 by the fixed tree v += v[lane + o] for
// o = 32, 16, 8, 4, 2, 1.  The last four steps only need lanes inside
// lane 0's row of 16, so they are DPP moves instead of LDS permutes.
__device__ __forceinline__ double wave_sum0(double v) {
  v += __shfl_down(v, 32, 64);
  v += __shfl_down(v, 16, 64);
  v += row_down<8>(v);
  v += row_down<4>(v);
  v += row_down<2>(v);
  v += row_down<1>(v);
  return v;
}

// Block-wide FP64 sum, fixed tree; result valid in every thread.
__device__ __forceinline__ double block_sum(double v, double *sh) {
  v = wave_sum0(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; ++w) s += sh[w];
  return s;
}

// Block-wide FP64 sums of NQ values with one pair of barriers, fixed tree;
// sh holds at least 16 * NQ doubles.  Results valid in every thread.
template <int NQ>
__device__ __forceinline__ void block_sums(double v[NQ], double *sh, double out[NQ]) {
#pragma unroll
  for (int q = 0; q < NQ; ++q) v[q] = wave_sum0(v[q]);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) sh[16 * q + wid] = v[q];
  }
  __syncthreads();
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += sh[16 * q + w];
    out[q] = s;
  }
}

// Sums of a device's per-workgroup partials of quantities q0 .. q0+NQ-1,
// fixed order: four independent accumulators per quantity, so the loads
// of one round are in flight together (a device spans up to ~1000
// workgroups, and one dependent load per partial was a latency chain in
// every consumer workgroup), then one block tree.  sh: 16 * NQ doubles.
template <int NQ>
__device__ __forceinline__ void device_sums(const double *part, int nblk_total, int q0,
                                            const DevInfo &I, double *sh, double out[NQ]) {
  double a[NQ][4];
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int k = 0; k < 4; ++k) a[q][k] = 0.0;
  const int n = I.nblk, st = blockDim.x;
  const double *p = part + I.blk0;
  int t = threadIdx.x;
  for (; t + 3 * st < n; t += 4 * st) {
    double v[NQ][4];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int k = 0; k < 4; ++k) v[q][k] = p[(long long)(q0 + q) * nblk_total + t + k * st];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int k = 0; k < 4; ++k) a[q][k] += v[q][k];
  }
  // the rest, at most three partials per thread (t + 3 st >= n here): all
  // loads issued first, then added in the same order as one at a time
  double r[3][NQ];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      r[m][q] = t + m * st < n ? p[(long long)(q0 + q) * nblk_total + t + m * st] : 0.0;
#pragma unroll
  for (int m = 0; m < 3; ++m)
    if (t + m * st < n) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) a[q][0] += r[m][q];
    }
  double s[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) s[q] = (a[q][0] + a[q][1]) + (a[q][2] + a[q][3]);
  block_sums<NQ>(s, sh, out);
}

// Sum of a device's per-workgroup partials of quantity q (sh: 16 doubles).
__device__ __forceinline__ double device_sum(const double *part, int nblk_total, int q,
                                             const DevInfo &I, double *sh) {
  double o[1];
  device_sums<1>(part, nblk_total, q, I, sh, o);
  return o[0];
}

struct NodeHood {
  bool valid;
  long long g;
  int i, j, kk;                     // device-local coordinates
  bool nb[6];
  long long nbg[6];
};

__device__ __forceinline__ NodeHood hood(const DevInfo &I, int blk) {
  NodeHood H;
  const NodeAt P = node_at(I, blk);
  H.valid = P.valid;
  H.g = I.off + P.k;
  const int i = P.i, j = P.j, kk = P.kk;
  H.i = i; H.j = j; H.kk = kk;
  const long long sy = I.nx, sz = (long long)I.nx * I.ny;
  H.nb[0] = i > 0;          H.nbg[0] = H.g - 1;
  H.nb[1] = i < I.nx - 1;   H.nbg[1] = H.g + 1;
  H.nb[2] = j > 0;          H.nbg[2] = H.g - sy;
  H.nb[3] = j < I.ny - 1;   H.nbg[3] = H.g + sy;
  H.nb[4] = I.dim == 3 && kk > 0;                      H.nbg[4] = H.g - sz;
  H.nb[5] = I.dim == 3 && (int)kk < I.nz - 1;          H.nbg[5] = H.g + sz;
  return H;
}


// Red-black tile ordering: a tile is black when the sum of its tile
// coordinates is odd.  Tiles that share a face differ in colour, so every
// coupling that leaves a tile joins a red and a black one.
__device__ __forceinline__ bool tile_black_at(int dim, int i, int j, int kk) {
  if (dim == 3) return ((i / T3X + j / T3Y + kk / T3Z) & 1) != 0;
  return ((i / TILE_X + j / TILE_Y) & 1) != 0;
}

// Whether the neighbour of node (i, j, kk) in direction dir (x-, x+, y-,
// y+, z-, z+) lies in another tile (at the device's edge: where it would).
__device__ __forceinline__ bool leaves_tile_at(int dim, int i, int j, int kk, int dir) {
  const int tx = dim == 3 ? T3X : TILE_X, ty = dim == 3 ? T3Y : TILE_Y;
  switch (dir) {
    case 0: return i % tx == 0;
    case 1: return i % tx == tx - 1;
    case 2: return j % ty == 0;
    case 3: return j % ty == ty - 1;
    case 4: return dim == 3 && kk % T3Z == 0;
    default: return dim == 3 && kk % T3Z == T3Z - 1;
  }
}

// Tile slot of this thread inside its node patch (tile index within the
// patch) and its node's place in the 16-node tile (x fastest).
__device__ __forceinline__ void tile_slot(int dim, int &tile, int &l) {
  if (dim == 3) {
    const int il = threadIdx.x % PX3, jl = (threadIdx.x / PX3) % PY3,
              kl = threadIdx.x / (PX3 * PY3);
    tile = il / T3X;                        // PY3 = T3Y, PZ3 = T3Z: one tile row
    l = (kl * T3Y + jl) * T3X + il % T3X;
  } else {
    const int il = threadIdx.x % PX, jl = threadIdx.x / PX;
    tile = (jl / TILE_Y) * (PX / TILE_X) + il / TILE_X;
    l = (jl % TILE_Y) * TILE_X + il % TILE_X;
  }
}

inline Where make_where(Batch &b) {
  Where W{};
  W.info = b.d_info;
  W.blk2dev = b.d_blk2dev;
  W.single = b.ndev == 1 ? 1 : 0;
  return W;
}

}  // namespace tcad
