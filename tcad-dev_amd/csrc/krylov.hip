// Krylov solvers on the stencil-block Newton matrix.
//
// * BiCGStab (right node-block-Jacobi preconditioning), templated on the
//   storage precision T (double or float); dot products accumulate in
//   FP64.  One iteration is five kernels:
//     K1  reduce (rhat,r),(r,r) -> rho, stop test; p = r + beta(p - w v);
//         phat = Dinv p
//     K2  v = A phat, partial (rhat, v)
//     K3  reduce -> alpha; s = r - alpha v; shat = Dinv s
//     K4  t = A shat, partials (t,s), (t,t)
//     K5  reduce -> omega; x += alpha phat + omega shat; r = s - omega t;
//         partials (rhat,r), (r,r)
//   Every workgroup of a device re-reduces that device's per-workgroup
//   partials in a fixed order at the top of the next kernel, so the
//   scalars are identical in all workgroups and bit-reproducible run to
//   run, with no atomics and no separate reduction launches.
// * Mixed precision: BiCGStab on the FP32-stored matrix and block
//   inverses (FP64 vector arithmetic) inside FP64 iterative refinement
//   (residual r = b - A64 y in FP64, inner solve on r/||r||).  FP32
//   vector arithmetic was measured unstable on these Jacobians (BiCGStab
//   residual growth by 1e20+ on the second refinement sweep of a MOSFET
//   system), so only the storage of the matrix, the dominant stream, is
//   reduced.
// * PCG (Jacobi) for the equilibrium Poisson system, FP64.
//
// Replaces linsolve.solve_linear's node-block-Jacobi Krylov path
// (pytcad/pytcad/linsolve.py:363-427, 725-985).
//
// Roofline: HBM/L2 bound.  Algorithmic bytes per node and BiCGStab
// iteration, 2D (29 entries), matrix width w (8 or 4): 2 SpMV x
// (29 w + 24 + 24 + 24) + 2 precond x (9 w + 48) + ~12 vector passes x 24.
#include "tcad_internal.h"

namespace tcad {

enum { VX = 0, VR, VRH, VP, VV, VS, VT, VPH, VSH };

template <typename T>
struct KArgs {
  const DevInfo *info;
  const int *blk2dev;
  long long N;
  int nblk;
  int ne;
  const T *A;                       // matrix storage precision
  const T *Dinv;                    // preconditioner storage precision
  double *v[9];                     // Krylov vectors, always FP64
  const double *b;
  double *part;
  Scal *scal;
  double tol2;
  int maxit;
  int parity;
};

// Block-wide FP64 sum of one value per thread, fixed tree; result valid in
// every thread.
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

__device__ __forceinline__ double block_max(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; ++w) s = fmax(s, sh[w]);
  return s;
}

// Sum of a device's per-block partials of quantity q, fixed order.
__device__ __forceinline__ double device_sum(const double *part, int nblk_total, int q,
                                             const DevInfo &I, double *sh) {
  double s = 0.0;
  for (int t = threadIdx.x; t < I.nblk; t += blockDim.x) s += part[(long long)q * nblk_total + I.blk0 + t];
  return block_sum(s, sh);
}

__device__ __forceinline__ double device_max(const double *part, int nblk_total, int q,
                                             const DevInfo &I, double *sh) {
  double s = 0.0;
  for (int t = threadIdx.x; t < I.nblk; t += blockDim.x) s = fmax(s, part[(long long)q * nblk_total + I.blk0 + t]);
  return block_max(s, sh);
}

struct NodeHood {
  bool valid;
  long long g;
  bool nb[6];
  long long nbg[6];
};

__device__ __forceinline__ NodeHood hood(const DevInfo &I, int blk) {
  NodeHood H;
  const long long k = (long long)(blk - I.blk0) * BS + threadIdx.x;
  H.valid = k < I.n;
  H.g = I.off + k;
  const int i = (int)(k % I.nx);
  const int j = (int)((k / I.nx) % I.ny);
  const int kk = (int)(k / ((long long)I.nx * I.ny));
  const long long sy = I.nx, sz = (long long)I.nx * I.ny;
  H.nb[0] = i > 0;          H.nbg[0] = H.g - 1;
  H.nb[1] = i < I.nx - 1;   H.nbg[1] = H.g + 1;
  H.nb[2] = j > 0;          H.nbg[2] = H.g - sy;
  H.nb[3] = j < I.ny - 1;   H.nbg[3] = H.g + sy;
  H.nb[4] = I.dim == 3 && kk > 0;           H.nbg[4] = H.g - sz;
  H.nb[5] = I.dim == 3 && kk < I.nz - 1;    H.nbg[5] = H.g + sz;
  return H;
}

template <typename T>
__device__ __forceinline__ void spmv_node(const T *A, long long N, const NodeHood &H,
                                          const double *x, double y[3]) {
  const long long g = H.g;
  const double x0 = x[g], x1 = x[N + g], x2 = x[2 * N + g];
#pragma unroll
  for (int r = 0; r < 3; ++r)
    y[r] = (double)A[(3 * r) * N + g] * x0 + (double)A[(3 * r + 1) * N + g] * x1 +
           (double)A[(3 * r + 2) * N + g] * x2;
#pragma unroll
  for (int dir = 0; dir < 6; ++dir) {
    if (!H.nb[dir]) continue;
    const long long gn = H.nbg[dir];
    const T *o = A + (long long)(9 + 5 * dir) * N + g;
    const double xp = x[gn], xn = x[N + gn], xq = x[2 * N + gn];
    y[0] += (double)o[0] * xp;
    y[1] += (double)o[N] * xp + (double)o[2 * N] * xn;
    y[2] += (double)o[3 * N] * xp + (double)o[4 * N] * xq;
  }
}

template <typename T>
__device__ __forceinline__ void prec_node(const T *Di, long long N, long long g, const double *x,
                                          double *z) {
  const double x0 = x[g], x1 = x[N + g], x2 = x[2 * N + g];
#pragma unroll
  for (int r = 0; r < 3; ++r)
    z[r * N + g] = (double)Di[(3 * r) * N + g] * x0 + (double)Di[(3 * r + 1) * N + g] * x1 +
                   (double)Di[(3 * r + 2) * N + g] * x2;
}

#define KPROLOGUE                                   \
  __shared__ double sh[BS / 64];                    \
  const int d = K.blk2dev[blockIdx.x];              \
  const DevInfo &I = K.info[d];                     \
  Scal &S = K.scal[d];                              \
  if (S.done) return;                               \
  const bool lead = (blockIdx.x == I.blk0) && threadIdx.x == 0;

template <typename T>
__global__ void __launch_bounds__(BS) k_bicg_init(KArgs<T> K) {
  KPROLOGUE
  (void)lead;
  const NodeHood H = hood(I, blockIdx.x);
  const long long N = K.N;
  double rr = 0.0;
  if (H.valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long ix = c * N + H.g;
      const double bv = K.b[ix];
      K.v[VX][ix] = 0.0;
      K.v[VR][ix] = bv;
      K.v[VRH][ix] = bv;
      K.v[VP][ix] = 0.0;
      K.v[VV][ix] = 0.0;
      rr += bv * bv;
    }
  }
  const double s = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    K.part[0 * (long long)K.nblk + blockIdx.x] = s;
    K.part[1 * (long long)K.nblk + blockIdx.x] = s;
  }
}

template <typename T>
__global__ void __launch_bounds__(BS) k_bicg_1(KArgs<T> K) {
  KPROLOGUE
  const double rho = device_sum(K.part, K.nblk, 0, I, sh);
  const double rr = device_sum(K.part, K.nblk, 1, I, sh);
  const double bb = (S.it == 0) ? rr : S.bb;
  if (rr <= K.tol2 * bb || S.it >= K.maxit || !isfinite(rr) || !isfinite(rho) || rho == 0.0) {
    if (lead) {
      S.rr = rr;
      if (S.it == 0) S.bb = bb;
      if (!(rr <= K.tol2 * bb)) S.status = 1;
      S.done = 1;
    }
    return;
  }
  const double beta = (rho / S.rho[K.parity ^ 1]) * (S.alpha / S.omega);
  const NodeHood H = hood(I, blockIdx.x);
  if (H.valid) {
    const long long N = K.N;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long ix = c * N + H.g;
      K.v[VP][ix] = K.v[VR][ix] + beta * (K.v[VP][ix] - S.omega * K.v[VV][ix]);
    }
    prec_node<T>(K.Dinv, N, H.g, K.v[VP], K.v[VPH]);
  }
  if (lead) {
    S.rho[K.parity] = rho;
    S.rr = rr;
    if (S.it == 0) S.bb = bb;
  }
}

template <typename T>
__global__ void __launch_bounds__(BS) k_bicg_spmv(KArgs<T> K, int src, int dst, int dot_a,
                                                  int dot_b, int q0) {
  KPROLOGUE
  (void)lead;
  const NodeHood H = hood(I, blockIdx.x);
  const long long N = K.N;
  double s0 = 0.0, s1 = 0.0;
  if (H.valid) {
    double y[3];
    spmv_node<T>(K.A, N, H, K.v[src], y);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long ix = c * N + H.g;
      const double yt = y[c];
      K.v[dst][ix] = yt;
      s0 += K.v[dot_a][ix] * yt;
      if (dot_b >= 0) s1 += yt * yt;
    }
  }
  const double a = block_sum(s0, sh);
  if (threadIdx.x == 0) K.part[(long long)q0 * K.nblk + blockIdx.x] = a;
  if (dot_b >= 0) {
    const double b2 = block_sum(s1, sh);
    if (threadIdx.x == 0) K.part[(long long)(q0 + 1) * K.nblk + blockIdx.x] = b2;
  }
}

template <typename T>
__global__ void __launch_bounds__(BS) k_bicg_3(KArgs<T> K) {
  KPROLOGUE
  const double rv = device_sum(K.part, K.nblk, 2, I, sh);
  const double rho = S.rho[K.parity];
  const double alpha = rho / rv;
  if (!isfinite(alpha) || rv == 0.0) {
    if (lead) { S.status = 1; S.done = 1; }
    return;
  }
  const NodeHood H = hood(I, blockIdx.x);
  if (H.valid) {
    const long long N = K.N;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long ix = c * N + H.g;
      K.v[VS][ix] = K.v[VR][ix] - alpha * K.v[VV][ix];
    }
    prec_node<T>(K.Dinv, N, H.g, K.v[VS], K.v[VSH]);
  }
  if (lead) S.alpha = alpha;
}

template <typename T>
__global__ void __launch_bounds__(BS) k_bicg_5(KArgs<T> K) {
  KPROLOGUE
  const double ts = device_sum(K.part, K.nblk, 2, I, sh);
  const double tt = device_sum(K.part, K.nblk, 3, I, sh);
  const double omega = (tt > 0.0) ? ts / tt : 0.0;
  const double alpha = S.alpha;
  const NodeHood H = hood(I, blockIdx.x);
  const long long N = K.N;
  double srh = 0.0, srr = 0.0;
  if (H.valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long ix = c * N + H.g;
      K.v[VX][ix] = K.v[VX][ix] + alpha * K.v[VPH][ix] + omega * K.v[VSH][ix];
      const double r = K.v[VS][ix] - omega * K.v[VT][ix];
      K.v[VR][ix] = r;
      srh += K.v[VRH][ix] * r;
      srr += r * r;
    }
  }
  const double a = block_sum(srh, sh);
  const double b2 = block_sum(srr, sh);
  if (threadIdx.x == 0) {
    K.part[0 * (long long)K.nblk + blockIdx.x] = a;
    K.part[1 * (long long)K.nblk + blockIdx.x] = b2;
  }
  if (lead) {
    S.omega = omega;
    S.it = S.it + 1;
    if (omega == 0.0) {
      // stagnation: the next beta would divide by zero; the stop test of
      // the next K1 decides between converged and failed
      S.omega = 1e-300;
    }
  }
}

// Reset the per-device Krylov state; devices not active are marked done.
__global__ void k_krylov_reset(Scal *scal, int ndev) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndev) return;
  Scal &S = scal[d];
  S.rho[0] = S.rho[1] = 1.0;
  S.alpha = 1.0;
  S.omega = 1.0;
  S.it = 0;
  if (S.active) {
    S.status = 0;
    S.rr = 0.0;
    S.bb = 0.0;
    S.done = 0;
  } else {
    S.done = 1;
  }
}

// Host: launch BiCGStab on the batch (vector b = K.b, solution in v[VX]).
template <typename T>
static int run_bicgstab(Batch &b, KArgs<T> K, int maxit, int chunk) {
  hipLaunchKernelGGL(k_krylov_reset, dim3((b.ndev + 63) / 64), dim3(64), 0, b.stream, b.d_scal,
                     b.ndev);
  hipLaunchKernelGGL(k_bicg_init<T>, dim3(b.nblk), dim3(BS), 0, b.stream, K);
  b.launches += 2;
  int it = 0;
  while (true) {
    for (int c = 0; c < chunk; ++c, ++it) {
      K.parity = it & 1;
      hipLaunchKernelGGL(k_bicg_1<T>, dim3(b.nblk), dim3(BS), 0, b.stream, K);
      hipLaunchKernelGGL(k_bicg_spmv<T>, dim3(b.nblk), dim3(BS), 0, b.stream, K, (int)VPH,
                         (int)VV, (int)VRH, -1, 2);
      hipLaunchKernelGGL(k_bicg_3<T>, dim3(b.nblk), dim3(BS), 0, b.stream, K);
      hipLaunchKernelGGL(k_bicg_spmv<T>, dim3(b.nblk), dim3(BS), 0, b.stream, K, (int)VSH,
                         (int)VT, (int)VS, 1, 2);
      hipLaunchKernelGGL(k_bicg_5<T>, dim3(b.nblk), dim3(BS), 0, b.stream, K);
      b.launches += 5;
    }
    TCAD_HIP_CHECK(hipGetLastError());
    if (it > maxit) {
      // one more stop test so every device records its final state
      K.parity = it & 1;
      hipLaunchKernelGGL(k_bicg_1<T>, dim3(b.nblk), dim3(BS), 0, b.stream, K);
      b.launches++;
    }
    int rc = sync_scalars(b);
    if (rc) return rc;
    int open = 0;
    for (int d = 0; d < b.ndev; ++d) open += b.h_scal[d].done ? 0 : 1;
    if (open == 0 || it > maxit) break;
    if (chunk < 64) chunk *= 2;
  }
  return TCAD_OK;
}

int bicgstab(Batch &b, bool fp32, double rtol, int maxit, bool zero_guess, int chunk) {
  (void)zero_guess;
  if (fp32) {
    // FP32-stored matrix and block inverses, FP64 vectors: the inner solve
    // of the iterative refinement (rhs in ir_rhs, solution in v64[VX])
    KArgs<float> K{};
    K.info = b.d_info; K.blk2dev = b.d_blk2dev; K.N = b.ntot; K.nblk = b.nblk; K.ne = b.ne;
    K.A = b.A32; K.Dinv = b.Dinv32;
    for (int i = 0; i < 9; ++i) K.v[i] = b.v64[i];
    K.b = b.ir_rhs;
    K.part = b.part; K.scal = b.d_scal; K.tol2 = rtol * rtol; K.maxit = maxit;
    return run_bicgstab<float>(b, K, maxit, chunk);
  }
  KArgs<double> K{};
  K.info = b.d_info; K.blk2dev = b.d_blk2dev; K.N = b.ntot; K.nblk = b.nblk; K.ne = b.ne;
  K.A = b.A64; K.Dinv = b.Dinv64;
  for (int i = 0; i < 9; ++i) K.v[i] = b.v64[i];
  K.v[VX] = b.y64;
  K.b = b.bs;
  K.part = b.part; K.scal = b.d_scal; K.tol2 = rtol * rtol; K.maxit = maxit;
  return run_bicgstab<double>(b, K, maxit, chunk);
}

// One stencil SpMV launch (v = A phat with the (rhat, v) partial), as
// issued inside BiCGStab; used by the bench's event timing.  The caller
// has cleared the per-device `done` flags.
int launch_spmv_only(Batch &b, bool fp32) {
  if (fp32) {
    KArgs<float> K{};
    K.info = b.d_info; K.blk2dev = b.d_blk2dev; K.N = b.ntot; K.nblk = b.nblk; K.ne = b.ne;
    K.A = b.A32; K.Dinv = b.Dinv32;
    for (int i = 0; i < 9; ++i) K.v[i] = b.v64[i];
    K.part = b.part; K.scal = b.d_scal;
    hipLaunchKernelGGL(k_bicg_spmv<float>, dim3(b.nblk), dim3(BS), 0, b.stream, K, (int)VPH,
                       (int)VV, (int)VRH, -1, 2);
  } else {
    KArgs<double> K{};
    K.info = b.d_info; K.blk2dev = b.d_blk2dev; K.N = b.ntot; K.nblk = b.nblk; K.ne = b.ne;
    K.A = b.A64; K.Dinv = b.Dinv64;
    for (int i = 0; i < 9; ++i) K.v[i] = b.v64[i];
    K.part = b.part; K.scal = b.d_scal;
    hipLaunchKernelGGL(k_bicg_spmv<double>, dim3(b.nblk), dim3(BS), 0, b.stream, K, (int)VPH,
                       (int)VV, (int)VRH, -1, 2);
  }
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

__global__ void k_clear_done(Scal *scal, int ndev) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d < ndev) scal[d].done = 0;
}

int clear_done(Batch &b) {
  hipLaunchKernelGGL(k_clear_done, dim3((b.ndev + 63) / 64), dim3(64), 0, b.stream, b.d_scal,
                     b.ndev);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

// ---------------------------------------------------------------------------
// FP64 iterative refinement around the FP32 BiCGStab.
// ---------------------------------------------------------------------------
struct IRArgs {
  const DevInfo *info;
  const int *blk2dev;
  long long N;
  int nblk;
  const double *A64;
  const double *bs;
  double *y64;
  double *r64;
  double *rhs;                      // normalised residual, inner rhs
  const double *xin;                // inner solution
  double *part;
  Scal *scal;
  double tol2;
};

// r64 = bs - A64 y64, partials (r,r) and (b,b)
__global__ void __launch_bounds__(BS) k_ir_residual(IRArgs R, int first) {
  __shared__ double sh[BS / 64];
  const int d = R.blk2dev[blockIdx.x];
  const DevInfo &I = R.info[d];
  Scal &S = R.scal[d];
  if (!S.active) return;
  const NodeHood H = hood(I, blockIdx.x);
  const long long N = R.N;
  double rr = 0.0, bb = 0.0;
  if (H.valid) {
    double y[3];
    if (first) {
      y[0] = y[1] = y[2] = 0.0;
    } else {
      spmv_node<double>(R.A64, N, H, R.y64, y);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long ix = c * N + H.g;
      const double bv = R.bs[ix];
      if (first) R.y64[ix] = 0.0;
      const double r = bv - y[c];
      R.r64[ix] = r;
      rr += r * r;
      bb += bv * bv;
    }
  }
  const double a = block_sum(rr, sh);
  const double b2 = block_sum(bb, sh);
  if (threadIdx.x == 0) {
    R.part[0 * (long long)R.nblk + blockIdx.x] = a;
    R.part[1 * (long long)R.nblk + blockIdx.x] = b2;
  }
}

// stop test; r32 = r64 / ||r64||
__global__ void __launch_bounds__(BS) k_ir_check(IRArgs R, int maxsweep) {
  __shared__ double sh[BS / 64];
  const int d = R.blk2dev[blockIdx.x];
  const DevInfo &I = R.info[d];
  Scal &S = R.scal[d];
  if (!S.active) return;
  const bool lead = (blockIdx.x == I.blk0) && threadIdx.x == 0;
  const double rr = device_sum(R.part, R.nblk, 0, I, sh);
  const double bb = device_sum(R.part, R.nblk, 1, I, sh);
  if (rr <= R.tol2 * bb || S.ir_it >= maxsweep || !isfinite(rr)) {
    if (lead) {
      S.active = 0;
      S.rr = rr;
      S.bb = bb;
      S.status = (rr <= R.tol2 * bb) ? 0 : 2;
    }
    return;
  }
  const double nrm = sqrt(rr);
  const NodeHood H = hood(I, blockIdx.x);
  if (H.valid) {
    const long long N = R.N;
#pragma unroll
    for (int c = 0; c < 3; ++c) R.rhs[c * N + H.g] = R.r64[c * N + H.g] / nrm;
  }
  if (lead) {
    S.nrm = nrm;
    S.rr = rr;
    S.bb = bb;
  }
}

__global__ void __launch_bounds__(BS) k_ir_update(IRArgs R) {
  const int d = R.blk2dev[blockIdx.x];
  const DevInfo &I = R.info[d];
  Scal &S = R.scal[d];
  if (!S.active) return;
  const bool lead = (blockIdx.x == I.blk0) && threadIdx.x == 0;
  const NodeHood H = hood(I, blockIdx.x);
  if (H.valid) {
    const long long N = R.N;
    const double nrm = S.nrm;
#pragma unroll
    for (int c = 0; c < 3; ++c) R.y64[c * N + H.g] += nrm * R.xin[c * N + H.g];
  }
  if (lead) {
    S.kry_total += S.it;
  }
}

__global__ void k_ir_count(Scal *scal, int ndev) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndev) return;
  if (scal[d].active) scal[d].ir_it += 1;
}

// Save / restore the per-device "active" flags around the inner solves.
__global__ void k_active_copy(Scal *scal, int ndev, int *save, int dir) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndev) return;
  if (dir == 0) save[d] = scal[d].active;
  else scal[d].active = save[d];
}

int ir_solve(Batch &b, double rtol, int maxsweep, int inner_maxit, double inner_rtol) {
  IRArgs R{};
  R.info = b.d_info; R.blk2dev = b.d_blk2dev; R.N = b.ntot; R.nblk = b.nblk;
  R.A64 = b.A64; R.bs = b.bs; R.y64 = b.y64; R.r64 = b.ir_r;
  R.rhs = b.ir_rhs; R.xin = b.v64[VX];
  R.part = b.part; R.scal = b.d_scal; R.tol2 = rtol * rtol;
  // Newton-level active flags are saved in the last slot of d_flag
  int *save = b.d_flag + 1;
  hipLaunchKernelGGL(k_active_copy, dim3((b.ndev + 63) / 64), dim3(64), 0, b.stream, b.d_scal,
                     b.ndev, save, 0);
  b.launches++;
  for (int sweep = 0; sweep <= maxsweep; ++sweep) {
    hipLaunchKernelGGL(k_ir_residual, dim3(b.nblk), dim3(BS), 0, b.stream, R, sweep == 0 ? 1 : 0);
    hipLaunchKernelGGL(k_ir_check, dim3(b.nblk), dim3(BS), 0, b.stream, R, maxsweep);
    b.launches += 2;
    TCAD_HIP_CHECK(hipGetLastError());
    int rc = sync_scalars(b);
    if (rc) return rc;
    int open = 0;
    for (int d = 0; d < b.ndev; ++d) open += b.h_scal[d].active ? 1 : 0;
    if (open == 0) break;
    // inner solve with the FP32-stored matrix on the normalised residual
    rc = bicgstab(b, true, inner_rtol, inner_maxit, true, 8);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ir_update, dim3(b.nblk), dim3(BS), 0, b.stream, R);
    hipLaunchKernelGGL(k_ir_count, dim3((b.ndev + 63) / 64), dim3(64), 0, b.stream, b.d_scal,
                       b.ndev);
    b.launches += 2;
  }
  hipLaunchKernelGGL(k_active_copy, dim3((b.ndev + 63) / 64), dim3(64), 0, b.stream, b.d_scal,
                     b.ndev, save, 1);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

// du = Cc * y : (psi, n, p) column scale (1, n, p)
__global__ void __launch_bounds__(BS) k_unscale(const DevInfo *info, const int *blk2dev,
                                                 long long N, const double *y, const double *n,
                                                 const double *p, double *du) {
  const int d = blk2dev[blockIdx.x];
  const DevInfo &I = info[d];
  const NodeHood H = hood(I, blockIdx.x);
  if (!H.valid) return;
  du[H.g] = y[H.g];
  du[N + H.g] = n[H.g] * y[N + H.g];
  du[2 * N + H.g] = p[H.g] * y[2 * N + H.g];
}

int launch_unscale(Batch &b, const double *n, const double *p, double *du) {
  return launch_unscale_state(b, n, p, du);
}

int launch_unscale_state(Batch &b, const double *n, const double *p, double *du) {
  hipLaunchKernelGGL(k_unscale, dim3(b.nblk), dim3(BS), 0, b.stream, b.d_info, b.d_blk2dev,
                     b.ntot, b.y64, n, p, du);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

// ---------------------------------------------------------------------------
// PCG (Jacobi) for the SPD equilibrium Poisson system.
//   vectors veq: 0 x, 1 r, 2 z, 3 p, 4 q
// ---------------------------------------------------------------------------
struct PArgs {
  const DevInfo *info;
  const int *blk2dev;
  long long N;
  int nblk;
  const double *A;
  const double *b;
  double *x, *r, *z, *p, *q;
  double *part;
  Scal *scal;
  double tol2;
  int maxit;
  int parity;
};

__device__ __forceinline__ double pspmv(const double *A, long long N, const NodeHood &H,
                                        const double *x) {
  double y = A[H.g] * x[H.g];
#pragma unroll
  for (int dir = 0; dir < 6; ++dir)
    if (H.nb[dir]) y += A[(1 + dir) * N + H.g] * x[H.nbg[dir]];
  return y;
}

#define PPROLOGUE                                   \
  __shared__ double sh[BS / 64];                    \
  const int d = P.blk2dev[blockIdx.x];              \
  const DevInfo &I = P.info[d];                     \
  Scal &S = P.scal[d];                              \
  if (S.done) return;                               \
  const bool lead = (blockIdx.x == I.blk0) && threadIdx.x == 0;

__global__ void __launch_bounds__(BS) k_pcg_init(PArgs P) {
  PPROLOGUE
  (void)lead;
  const NodeHood H = hood(I, blockIdx.x);
  double rz = 0.0, rr = 0.0;
  if (H.valid) {
    const double bv = P.b[H.g];
    const double zv = bv / P.A[H.g];
    P.x[H.g] = 0.0;
    P.r[H.g] = bv;
    P.z[H.g] = zv;
    P.p[H.g] = 0.0;
    rz = bv * zv;
    rr = bv * bv;
  }
  const double a = block_sum(rz, sh);
  const double c = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    P.part[blockIdx.x] = a;
    P.part[(long long)P.nblk + blockIdx.x] = c;
  }
}

__global__ void __launch_bounds__(BS) k_pcg_a(PArgs P) {
  PPROLOGUE
  const double rz = device_sum(P.part, P.nblk, 0, I, sh);
  const double rr = device_sum(P.part, P.nblk, 1, I, sh);
  const double bb = (S.it == 0) ? rr : S.bb;
  if (rr <= P.tol2 * bb || S.it >= P.maxit || !isfinite(rr)) {
    if (lead) {
      S.rr = rr;
      if (S.it == 0) S.bb = bb;
      if (!isfinite(rr) || rr > P.tol2 * bb) S.status = 1;
      S.done = 1;
    }
    return;
  }
  const double beta = rz / S.rho[P.parity ^ 1];
  const NodeHood H = hood(I, blockIdx.x);
  if (H.valid) P.p[H.g] = P.z[H.g] + beta * P.p[H.g];
  if (lead) {
    S.rho[P.parity] = rz;
    S.rr = rr;
    if (S.it == 0) S.bb = bb;
  }
}

__global__ void __launch_bounds__(BS) k_pcg_b(PArgs P) {
  PPROLOGUE
  (void)lead;
  const NodeHood H = hood(I, blockIdx.x);
  double pq = 0.0;
  if (H.valid) {
    const double qv = pspmv(P.A, P.N, H, P.p);
    P.q[H.g] = qv;
    pq = P.p[H.g] * qv;
  }
  const double a = block_sum(pq, sh);
  if (threadIdx.x == 0) P.part[2 * (long long)P.nblk + blockIdx.x] = a;
}

__global__ void __launch_bounds__(BS) k_pcg_c(PArgs P) {
  PPROLOGUE
  const double pq = device_sum(P.part, P.nblk, 2, I, sh);
  const double alpha = S.rho[P.parity] / pq;
  const NodeHood H = hood(I, blockIdx.x);
  double rz = 0.0, rr = 0.0;
  if (H.valid) {
    P.x[H.g] += alpha * P.p[H.g];
    const double r = P.r[H.g] - alpha * P.q[H.g];
    P.r[H.g] = r;
    const double zv = r / P.A[H.g];
    P.z[H.g] = zv;
    rz = r * zv;
    rr = r * r;
  }
  const double a = block_sum(rz, sh);
  const double c = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    P.part[blockIdx.x] = a;
    P.part[(long long)P.nblk + blockIdx.x] = c;
  }
  if (lead) S.it = S.it + 1;
}

int pcg_poisson(Batch &b, double rtol, int maxit) {
  PArgs P{};
  P.info = b.d_info; P.blk2dev = b.d_blk2dev; P.N = b.ntot; P.nblk = b.nblk;
  P.A = b.Aeq; P.b = b.beq;
  P.x = b.veq[0]; P.r = b.veq[1]; P.z = b.veq[2]; P.p = b.veq[3]; P.q = b.veq[4];
  P.part = b.part; P.scal = b.d_scal; P.tol2 = rtol * rtol; P.maxit = maxit;
  hipLaunchKernelGGL(k_krylov_reset, dim3((b.ndev + 63) / 64), dim3(64), 0, b.stream, b.d_scal,
                     b.ndev);
  hipLaunchKernelGGL(k_pcg_init, dim3(b.nblk), dim3(BS), 0, b.stream, P);
  b.launches += 2;
  int it = 0, chunk = 16;
  while (true) {
    for (int c = 0; c < chunk; ++c, ++it) {
      P.parity = it & 1;
      hipLaunchKernelGGL(k_pcg_a, dim3(b.nblk), dim3(BS), 0, b.stream, P);
      hipLaunchKernelGGL(k_pcg_b, dim3(b.nblk), dim3(BS), 0, b.stream, P);
      hipLaunchKernelGGL(k_pcg_c, dim3(b.nblk), dim3(BS), 0, b.stream, P);
      b.launches += 3;
    }
    TCAD_HIP_CHECK(hipGetLastError());
    if (it > maxit) {
      P.parity = it & 1;
      hipLaunchKernelGGL(k_pcg_a, dim3(b.nblk), dim3(BS), 0, b.stream, P);
      b.launches++;
    }
    int rc = sync_scalars(b);
    if (rc) return rc;
    int open = 0;
    for (int d = 0; d < b.ndev; ++d) open += b.h_scal[d].done ? 0 : 1;
    if (open == 0 || it > maxit) break;
    if (chunk < 128) chunk *= 2;
  }
  return TCAD_OK;
}

}  // namespace tcad
