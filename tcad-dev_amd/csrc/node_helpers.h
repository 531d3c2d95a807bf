// Device helpers shared by the Krylov and AC translation units: where a
// workgroup's device is, fixed-order block and device reductions, and the
// node a thread owns with its stencil neighbours.
#pragma once

#include "tcad_internal.h"

namespace tcad {

// ---------------------------------------------------------------------------
// shared device helpers
// ---------------------------------------------------------------------------
struct Where {
  const DevInfo *info;              // batch table
  const int *blk2dev;
  DevInfo one;                      // the device when the batch holds one
  int single;
};

__device__ __forceinline__ int dev_index(const Where &W) {
  return W.single ? 0 : W.blk2dev[blockIdx.x];
}

__device__ __forceinline__ const DevInfo &dev_info(const Where &W, int d) {
  return W.single ? W.one : W.info[d];
}

// Block-wide FP64 sum, fixed tree; result valid in every thread.
__device__ __forceinline__ double block_sum(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; ++w) s += sh[w];
  return s;
}

// Sum of a device's per-workgroup partials of quantity q, fixed order.
__device__ __forceinline__ double device_sum(const double *part, int nblk_total, int q,
                                             const DevInfo &I, double *sh) {
  double s = 0.0;
  for (int t = threadIdx.x; t < I.nblk; t += blockDim.x)
    s += part[(long long)q * nblk_total + I.blk0 + t];
  return block_sum(s, sh);
}

struct NodeHood {
  bool valid;
  long long g;
  bool nb[6];
  long long nbg[6];
};

__device__ __forceinline__ NodeHood hood(const DevInfo &I, int blk) {
  NodeHood H;
  const NodeAt P = node_at(I, blk);
  H.valid = P.valid;
  H.g = I.off + P.k;
  const int i = P.i, j = P.j, kk = P.kk;
  const long long sy = I.nx, sz = (long long)I.nx * I.ny;
  H.nb[0] = i > 0;          H.nbg[0] = H.g - 1;
  H.nb[1] = i < I.nx - 1;   H.nbg[1] = H.g + 1;
  H.nb[2] = j > 0;          H.nbg[2] = H.g - sy;
  H.nb[3] = j < I.ny - 1;   H.nbg[3] = H.g + sy;
  H.nb[4] = I.dim == 3 && kk > 0;                      H.nbg[4] = H.g - sz;
  H.nb[5] = I.dim == 3 && (int)kk < I.nz - 1;          H.nbg[5] = H.g + sz;
  return H;
}


inline Where make_where(Batch &b) {
  Where W{};
  W.info = b.d_info;
  W.blk2dev = b.d_blk2dev;
  W.single = b.ndev == 1 ? 1 : 0;
  W.one = b.info[0];
  return W;
}

}  // namespace tcad
