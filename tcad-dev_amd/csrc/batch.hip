// Newton drivers, per-node update kernels and the extern "C" entry points.
//
// The Newton loops run on the host side of this library; every per-node
// operation (contact stamping, damping/clipping, the update measure, the
// convergence test and masking of converged devices) runs on the GPU.
// The host reads back one small per-device scalar record per Newton
// iteration (and per Krylov chunk) to decide when every device is done.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "tcad_internal.h"

namespace tcad {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

__device__ __forceinline__ double clip_np(double a, double lo, double hi) {
  // numpy.clip = minimum(maximum(a, lo), hi), NaN propagating
  if (a != a) return a;
  return fmin(fmax(a, lo), hi);
}

struct UArgs {
  const DevInfo *info;
  const int *blk2dev;
  long long N;
  int nblk;
  double *part;
  Scal *scal;
};

__device__ __forceinline__ double ublock_max(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s = fmax(s, sh[w]);
  return s;
}

// Damped Newton update of solve_bias (device2d.py:2079-2100,
// device3d.py:2170-2185): du = Cc*y; psi += clip(du_psi, +-max_dpsi);
// n, p clipped to [0.1x, 10x]; err = max(|dpsi|, rel dn, rel dp) with the
// density floor 1e-10.  NaN in the update propagates to err.  With
// apply == 0 only the measure err is formed (the state is untouched), so
// a step that passes the test can have its linear solve refined before it
// is applied; apply == 1 writes the new state to (psi_o, n_o, p_o), which
// is the state itself or a candidate buffer.
__global__ void __launch_bounds__(BS) k_newton_update(UArgs U, const double *y, const double *psi,
                                                      const double *n, const double *p,
                                                      double max_dpsi, int apply, double *psi_o,
                                                      double *n_o, double *p_o) {
  __shared__ double sh[BS / 64];
  const int d = U.blk2dev[blockIdx.x];
  const DevInfo &I = U.info[d];
  if (!U.scal[d].active) return;
  const NodeAt P = node_at(I, blockIdx.x);
  const long long k = P.k;
  double err = 0.0;
  if (P.valid) {
    const long long g = I.off + k, N = U.N;
    const double n0 = n[g], p0 = p[g];
    const double dpsi = clip_np(y[g], -max_dpsi, max_dpsi);
    const double dn = n0 * y[N + g];
    const double dp = p0 * y[2 * N + g];
    const double n1 = clip_np(n0 + dn, 0.1 * n0, 10.0 * n0);
    const double p1 = clip_np(p0 + dp, 0.1 * p0, 10.0 * p0);
    const double rn = fabs(n1 - n0) / fmax(n0, 1e-10);
    const double rp = fabs(p1 - p0) / fmax(p0, 1e-10);
    err = fmax(fmax(fabs(dpsi), rn), rp);
    if (!(err == err) || !isfinite(n1) || !isfinite(p1) || !isfinite(dpsi)) err = INFINITY;
    if (apply) {
      psi_o[g] = psi[g] + dpsi;
      n_o[g] = n1;
      p_o[g] = p1;
    }
  }
  if (apply) return;
  const double m = ublock_max(err, sh);
  if (threadIdx.x == 0) U.part[blockIdx.x] = m;
}

// One workgroup per device: the update measure of this step and whether it
// passes the convergence test (the step is applied afterwards).  A device
// whose linear solve ended above the acceptance residual stops here, before
// its step is applied (linsolve.py:973-982 refuses such an iterate).
__global__ void k_newton_check(UArgs U, double tol, double accept2, double rtol, double eta_max,
                               int forcing, double eta_c, int eta_pow, double tile_lag) {
  __shared__ double sh[BS / 64];
  const int d = blockIdx.x;
  const DevInfo &I = U.info[d];
  Scal &S = U.scal[d];
  if (!S.active) return;
  double m = 0.0;
  for (int t = threadIdx.x; t < I.nblk; t += blockDim.x) m = fmax(m, U.part[I.blk0 + t]);
  m = ublock_max(m, sh);
  if (threadIdx.x == 0) {
    S.err = m;
    S.conv_now = m < tol ? 1 : 0;
    if (!(S.rr <= fmax(S.eta2, accept2) * S.bb)) {
      S.conv_now = 0;
      S.active = 0;
      S.result = TCAD_NEWTON_LINEAR_FAIL;
    }
    // forcing term of the next solve: far from the solution the step only
    // needs a few digits; near it the target tightens with the update,
    // linearly (eta_pow 1) or with its square (eta_pow 2, which keeps
    // Newton's quadratic convergence) (the converging step is refined to
    // the rounding floor before it is applied)
    // the next Newton step keeps this step's tile inverses when the state
    // moved by less than tile_lag (a preconditioner from a nearby Jacobian;
    // the FP64 refinement still decides the solve)
    S.tile_keep = (tile_lag > 0.0 && m < tile_lag) ? 1 : 0;
    if (forcing && S.forced) {
      const double eta = fmin(fmax(eta_c * (eta_pow == 2 ? m * m : m), rtol), eta_max);
      S.eta2 = eta * eta;
    }
  }
}

// Partial sums over the non-contact nodes of the continuity rows F_n + F_p
// (at a candidate state; storage terms included in a transient step):
// minus the imbalance of the contact currents (plus, in a transient step,
// the displacement of mobile charge, which the rows already contain).
__global__ void __launch_bounds__(BS) k_interior_sum(UArgs U, const int *cidx, const double *F) {
  __shared__ double sh[BS / 64];
  const int d = U.blk2dev[blockIdx.x];
  const DevInfo &I = U.info[d];
  if (!U.scal[d].active) return;
  const NodeAt P = node_at(I, blockIdx.x);
  double v = 0.0;
  if (P.valid) {
    const long long g = I.off + P.k;
    if (cidx[g] < 0) v = F[U.N + g] + F[2 * U.N + g];
  }
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < BS / 64; ++w) s += sh[w];
    U.part[blockIdx.x] = s;
  }
}

// Final-step refinement test, one workgroup per device: the candidate
// state is good enough when the contact currents balance to 1e-8 of the
// largest (the imbalance is the interior residual seen by the currents),
// or when the last refinement sweep no longer halved the true residual
// norm (the FP64 rounding floor).
__global__ void k_refine_check(UArgs U, const double *I_c, const double *cscale, int first) {
  __shared__ double sh[BS / 64];
  const int d = blockIdx.x;
  const DevInfo &I = U.info[d];
  Scal &S = U.scal[d];
  if (!S.active) return;
  double m = 0.0;
  for (int t = threadIdx.x; t < I.nblk; t += blockDim.x) m += U.part[I.blk0 + t];
  for (int o = 32; o > 0; o >>= 1) m += __shfl_down(m, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double delta = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) delta += sh[w];
  double imax = 0.0, scale = 0.0;
  for (int c = I.c0; c < I.c0 + I.nc; ++c) {
    imax = fmax(imax, fabs(I_c[c]));
    scale = cscale[2 * c] * cscale[2 * c + 1];
  }
  const bool balanced = I.nc == 0 || fabs(delta) * scale <= 1e-8 * imax;
  const bool stalled = !first && S.rr > 0.25 * S.rr_ref;
  if (balanced || stalled || !isfinite(delta)) S.active = 0;
  S.rr_ref = S.rr;
}

// ---------------------------------------------------------------------------
// Transient step (transient2d.py:138-183, transient3d.py:152-219): Newton on
// the step residual with Armijo backtracking on the merit 0.5 |F|^2.
// ---------------------------------------------------------------------------
// Partial sums of |F|^2 (all three rows) per workgroup.
__global__ void __launch_bounds__(BS) k_merit_part(UArgs U, const double *F) {
  __shared__ double sh[BS / 64];
  const int d = U.blk2dev[blockIdx.x];
  const DevInfo &I = U.info[d];
  if (!U.scal[d].active) return;
  const NodeAt P = node_at(I, blockIdx.x);
  double v = 0.0;
  if (P.valid) {
    const long long g = I.off + P.k;
    const double f0 = F[g], f1 = F[U.N + g], f2 = F[2 * U.N + g];
    v = f0 * f0 + f1 * f1 + f2 * f2;
  }
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < BS / 64; ++w) s += sh[w];
    U.part[blockIdx.x] = s;
  }
}

// One workgroup per device.  phase 0: the merit at the iterate starts the
// line search at lambda = 1.  phase 1: the trial at lambda is accepted if
// finite and merit <= merit0 (1 - 1e-4 lambda), else lambda halves; after
// 40 rejections lambda = 0.  A settled device leaves the active set.
__global__ void k_ls_merit(UArgs U, int phase) {
  __shared__ double sh[BS / 64];
  const int d = blockIdx.x;
  const DevInfo &I = U.info[d];
  Scal &S = U.scal[d];
  if (!S.active) return;
  double m = 0.0;
  for (int t = threadIdx.x; t < I.nblk; t += blockDim.x) m += U.part[I.blk0 + t];
  for (int o = 32; o > 0; o >>= 1) m += __shfl_down(m, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
  const double merit = 0.5 * s;
  if (phase == 0) {
    S.merit0 = merit;
    S.lam = 1.0;
    S.ls_done = 0;
    S.ls_it = 0;
    return;
  }
  if (isfinite(merit) && merit <= S.merit0 * (1.0 - 1e-4 * S.lam)) {
    S.ls_done = 1;
    S.active = 0;
  } else {
    S.lam = S.lam * 0.5;
    S.ls_it += 1;
    if (S.ls_it >= 40) {
      S.lam = 0.0;
      S.ls_done = 1;
      S.active = 0;
    }
  }
}

__global__ void k_ls_begin(Scal *scal, int ndev) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d < ndev) scal[d].newton_active = scal[d].active;
}

// Step at the device's lambda: psi + lambda clip(dpsi), densities clipped
// to [0.1x, 10x].  apply == 0: into the candidate buffers; apply == 1: into
// the state, with the update measure err = max(|clip(dpsi)|, rel dn,
// rel dp) (psi undamped, densities damped, transient2d.py:172-181).
__global__ void __launch_bounds__(BS) k_tr_step(UArgs U, const double *y, double *psi, double *n,
                                                double *p, double max_dpsi, int apply,
                                                double *cp, double *cn, double *cq) {
  __shared__ double sh[BS / 64];
  const int d = U.blk2dev[blockIdx.x];
  const DevInfo &I = U.info[d];
  const Scal &S = U.scal[d];
  if (!S.active) return;
  const NodeAt P = node_at(I, blockIdx.x);
  double err = 0.0;
  if (P.valid) {
    const long long g = I.off + P.k, N = U.N;
    const double lam = S.lam;
    const double n0 = n[g], p0 = p[g];
    const double dpsi = clip_np(y[g], -max_dpsi, max_dpsi);
    const double psi1 = psi[g] + lam * dpsi;
    const double n1 = clip_np(n0 + lam * (n0 * y[N + g]), 0.1 * n0, 10.0 * n0);
    const double p1 = clip_np(p0 + lam * (p0 * y[2 * N + g]), 0.1 * p0, 10.0 * p0);
    if (apply) {
      err = fmax(fmax(fabs(dpsi), fabs(n1 - n0) / fmax(n0, 1e-10)), fabs(p1 - p0) / fmax(p0, 1e-10));
      if (!(err == err) || !isfinite(psi1) || !isfinite(n1) || !isfinite(p1)) err = INFINITY;
      psi[g] = psi1;
      n[g] = n1;
      p[g] = p1;
    } else {
      cp[g] = psi1;
      cn[g] = n1;
      cq[g] = p1;
    }
  }
  if (!apply) return;
  const double m = ublock_max(err, sh);
  if (threadIdx.x == 0) U.part[blockIdx.x] = m;
}

// After the step: count it; converged, failed (3D: line search found no
// step, transient3d.py:216) or non-finite devices retire.
__global__ void k_tr_finish(UArgs U, double tol) {
  __shared__ double sh[BS / 64];
  const int d = blockIdx.x;
  const DevInfo &I = U.info[d];
  Scal &S = U.scal[d];
  if (!S.active) return;
  double m = 0.0;
  for (int t = threadIdx.x; t < I.nblk; t += blockDim.x) m = fmax(m, U.part[I.blk0 + t]);
  m = ublock_max(m, sh);
  if (threadIdx.x != 0) return;
  S.err = m;
  S.newton_it += 1;
  if (m < tol) {
    S.newton_done = 1;
    S.active = 0;
  } else if (!isfinite(m)) {
    S.result = TCAD_NEWTON_NONFINITE;
    S.active = 0;
  } else if (I.dim == 3 && S.lam == 0.0) {
    S.result = TCAD_NEWTON_MAX_ITER;
    S.active = 0;
  }
}

// Per-device assembly error flags -> the device stops as non-finite.
__global__ void k_flags_to_status(Scal *scal, int *flag, int ndev) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndev) return;
  if (flag[1 + d]) {
    flag[1 + d] = 0;
    scal[d].active = 0;
    scal[d].result = TCAD_NEWTON_NONFINITE;
  }
  if (d == 0) flag[0] = 0;
}

// Refinement mask: only the devices whose step just converged stay active.
__global__ void k_mask_converging(Scal *scal, int ndev) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndev) return;
  Scal &S = scal[d];
  S.newton_active = S.active;
  S.active = S.active && S.conv_now;
}

__global__ void k_unmask(Scal *scal, int ndev) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d < ndev) scal[d].active = scal[d].newton_active;
}

// After the step is applied: count it, retire converged or failed devices.
__global__ void k_newton_finish(Scal *scal, int ndev, int max_iter) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndev) return;
  Scal &S = scal[d];
  if (!S.active) return;
  S.newton_it += 1;
  if (S.conv_now) {
    S.newton_done = 1;
    S.active = 0;
  } else if (!isfinite(S.err)) {
    S.result = TCAD_NEWTON_NONFINITE;
    S.active = 0;
  } else if (S.newton_it >= max_iter) {
    S.result = TCAD_NEWTON_MAX_ITER;
    S.active = 0;
  }
}

// Equilibrium update (device2d.py:1335-1341): psi += clip(d), err = max|d|
__global__ void __launch_bounds__(BS) k_poisson_update(UArgs U, const double *x, double *psi,
                                                       double max_dpsi) {
  __shared__ double sh[BS / 64];
  const int d = U.blk2dev[blockIdx.x];
  const DevInfo &I = U.info[d];
  if (!U.scal[d].active) return;
  const NodeAt P = node_at(I, blockIdx.x);
  const long long k = P.k;
  double err = 0.0;
  if (P.valid) {
    const long long g = I.off + k;
    const double dd = clip_np(x[g], -max_dpsi, max_dpsi);
    err = fabs(dd);
    if (!(err == err)) err = INFINITY;
    psi[g] = psi[g] + dd;
  }
  const double m = ublock_max(err, sh);
  if (threadIdx.x == 0) U.part[blockIdx.x] = m;
}

// One workgroup per device: err = max of its partials; convergence test
// err < tol; Krylov iterations of this Newton step accumulated.
__global__ void k_newton_finalize(UArgs U, double tol, int add_krylov) {
  __shared__ double sh[BS / 64];
  const int d = blockIdx.x;
  const DevInfo &I = U.info[d];
  Scal &S = U.scal[d];
  if (!S.active) return;
  double m = 0.0;
  for (int t = threadIdx.x; t < I.nblk; t += blockDim.x) m = fmax(m, U.part[I.blk0 + t]);
  m = ublock_max(m, sh);
  if (threadIdx.x == 0) {
    S.err = m;
    S.newton_it += 1;
    if (add_krylov) S.kry_total += S.it;
    if (m < tol) {
      S.newton_done = 1;
      S.active = 0;
    } else if (!isfinite(m)) {
      S.result = TCAD_NEWTON_NONFINITE;
      S.active = 0;
    }
  }
}

// Select the devices of the next Newton solve (mask NULL = all) and reset
// their counters.  eta_exact < 0: every device starts at the target eta0 and
// may use forcing terms; eta_exact >= 0 (forcing terms for predicted starts
// only): a device whose start tcad_predict_state did not move solves every
// Newton step to eta_exact.  The predicted flag is consumed here.
__global__ void k_set_active(Scal *scal, int ndev, const int *mask, double eta0,
                             double eta_exact) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndev) return;
  Scal &S = scal[d];
  S.forced = eta_exact < 0.0 || S.predicted;
  S.predicted = 0;
  S.eta2 = S.forced ? eta0 * eta0 : eta_exact * eta_exact;
  S.sel = mask ? (mask[d] != 0) : 1;
  S.active = S.sel;
  S.result = 0;
  S.newton_it = 0;
  S.newton_done = 0;
  S.kry_total = 0;
  S.ir_it = 0;
  S.status = 0;
  S.err = 0.0;
  S.tile_keep = 0;
}

__global__ void k_ir_reset(Scal *scal, int ndev) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndev) return;
  scal[d].ir_it = 0;
}

// Stamp the contact values into the warm start (device2d.py:1949-1953).
__global__ void __launch_bounds__(BS) k_stamp(const DevInfo *info, const int *blk2dev,
                                              const int *cidx, const double *lnr,
                                              const double *n0, const double *p0,
                                              const int *ccontact, const double *V,
                                              const Scal *scal, double *psi, double *n,
                                              double *p) {
  const int d = blk2dev[blockIdx.x];
  if (!scal[d].sel) return;
  const DevInfo &I = info[d];
  const NodeAt P = node_at(I, blockIdx.x);
  if (!P.valid) return;
  const long long k = P.k;
  const long long g = I.off + k;
  const int t = cidx[g];
  if (t < 0) return;
  psi[g] = V[ccontact[t]] / I.VT + lnr[t];
  n[g] = n0[t];
  p[g] = p0[t];
}

// Secant predictor of a sweep's next Newton start from its last two
// converged points (the tangent step of continuation.py:368-404, there in
// psi only; here psi linearly and the densities geometrically, i.e.
// linearly in log n and log p, which tracks the quasi-Fermi levels).  Per
// device ratio f = step to the next point / step from the previous one
// (f <= 0: the device is not touched): psi += f (psi - psi_prev),
// n *= clamp((n / n_prev)^f, 0.1, 10), p likewise -- the Newton update's own
// density clip (device2d.py:2083-2100), so a prediction moves no density
// by more than one Newton step could.
__global__ void __launch_bounds__(BS) k_predict(const DevInfo *info, const int *blk2dev,
                                                const double *ratio, const double *psi0,
                                                const double *n0, const double *p0, double *psi,
                                                double *n, double *p) {
  const int d = blk2dev[blockIdx.x];
  const double f = ratio[d];
  if (!(f > 0.0)) return;
  const DevInfo &I = info[d];
  const NodeAt P = node_at(I, blockIdx.x);
  if (!P.valid) return;
  const long long g = I.off + P.k;
  const double ps = psi[g], nn = n[g], pp = p[g];
  const double rn = f == 1.0 ? nn / n0[g] : pow(nn / n0[g], f);
  const double rp = f == 1.0 ? pp / p0[g] : pow(pp / p0[g], f);
  psi[g] = ps + f * (ps - psi0[g]);
  n[g] = nn * fmin(fmax(rn, 0.1), 10.0);
  p[g] = pp * fmin(fmax(rp, 0.1), 10.0);
}

// Flag the devices whose start k_predict moved.
__global__ void k_mark_predicted(Scal *scal, int ndev, const double *ratio) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d < ndev) scal[d].predicted = ratio[d] > 0.0;
}

// Bulk-neutral guess arcsinh(C / (2 nie_s)) (device2d.py:744-759).
__global__ void __launch_bounds__(BS) k_bulk_guess(const DevInfo *info, const int *blk2dev,
                                                   const double *C, const double *nie_s,
                                                   const Scal *scal, double *psi) {
  const int d = blk2dev[blockIdx.x];
  if (!scal[d].sel) return;
  const DevInfo &I = info[d];
  const NodeAt P = node_at(I, blockIdx.x);
  if (!P.valid) return;
  const long long k = P.k;
  const long long g = I.off + k;
  psi[g] = asinh(C[g] / (2.0 * nie_s[g]));
}

// Slaved equilibrium densities (device2d.py:1365-1367).
__global__ void __launch_bounds__(BS) k_slaved(const DevInfo *info, const int *blk2dev,
                                               const double *nie_s, const Scal *scal,
                                               const double *psi, double *n, double *p) {
  const int d = blk2dev[blockIdx.x];
  if (!scal[d].sel) return;
  const DevInfo &I = info[d];
  const NodeAt P = node_at(I, blockIdx.x);
  if (!P.valid) return;
  const long long k = P.k;
  const long long g = I.off + k;
  const double ps = psi[g];
  n[g] = nie_s[g] * exp(fmin(fmax(ps, -700.0), 700.0));
  p[g] = nie_s[g] * exp(fmin(fmax(-ps, -700.0), 700.0));
}

static UArgs uargs(Batch &b) {
  UArgs U{};
  U.info = b.d_info;
  U.blk2dev = b.d_blk2dev;
  U.N = b.ntot;
  U.nblk = b.nblk;
  U.part = b.part;
  U.scal = b.d_scal;
  return U;
}

int sync_scalars(Batch &b) {
  TCAD_HIP_CHECK(hipMemcpyAsync(b.h_scal, b.d_scal, sizeof(Scal) * b.ndev,
                                hipMemcpyDeviceToHost, b.stream));
  TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
  return TCAD_OK;
}

int launch_set_active(Batch &b, const int32_t *h_mask, double eta0, double eta_exact) {
  if (h_mask)
    TCAD_HIP_CHECK(hipMemcpyAsync(b.d_mask, h_mask, sizeof(int) * b.ndev, hipMemcpyHostToDevice,
                                  b.stream));
  hipLaunchKernelGGL(k_set_active, dim3((b.ndev + 63) / 64), dim3(64), 0, b.stream, b.d_scal,
                     b.ndev, h_mask ? (const int *)b.d_mask : nullptr, eta0, eta_exact);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

static int launch_flags_to_status(Batch &b) {
  hipLaunchKernelGGL(k_flags_to_status, dim3((b.ndev + 63) / 64), dim3(64), 0, b.stream,
                     b.d_scal, b.d_flag, b.ndev);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

int check_flag(Batch &b) {
  int flag = 0;
  TCAD_HIP_CHECK(hipMemcpyAsync(&flag, b.d_flag, sizeof(int), hipMemcpyDeviceToHost, b.stream));
  TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
  if (flag) {
    TCAD_HIP_CHECK(hipMemsetAsync(b.d_flag, 0, sizeof(int) * (b.ndev + 1), b.stream));
    set_error("assembly produced a non-finite or singular node block");
    return TCAD_E_NONFINITE;
  }
  return TCAD_OK;
}

// ---------------------------------------------------------------------------
// Newton drivers
// ---------------------------------------------------------------------------
// ---- per-iteration linear fallback ------------------------------------------
// The CSR form of device d's equilibrated, Dirichlet-eliminated Newton
// system (the stencil blocks of A64), unknowns interleaved u = 3 k + c
// (k the device-local node), columns sorted: z-, y-, x- neighbour, the
// node, x+, y+, z+ neighbour.  Row r of node k holds the node block's
// three entries and, per neighbour that exists, psi<-psi (r = 0),
// n<-psi, n<-n (r = 1) or p<-psi, p<-p (r = 2).
__global__ void k_fb_gather(long long nnz, const long long *src, const double *A, double *val) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) val[k] = A[src[k]];
}

// interleaved <-> plane layout of one device's three unknowns per node
__global__ void k_fb_rhs(long long nd, long long N, long long off, const double *bs,
                         double *bv) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nd) return;
  for (int c = 0; c < 3; ++c) bv[3 * k + c] = bs[c * N + off + k];
}

__global__ void k_fb_sol(long long nd, long long N, long long off, const double *xv, double *y) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nd) return;
  for (int c = 0; c < 3; ++c) y[c * N + off + k] = xv[3 * k + c];
}

static int build_fallback_csr(Batch &b, int d) {
  const DevInfo &I = b.info[d];
  const long long N = b.ntot, nd = I.n;
  const int nx = I.nx, ny = I.ny, nz = I.dim == 3 ? I.nz : 1;
  if (3 * nd >= (1LL << 31) / 16) {
    set_error("linear fallback: device too large for the int32 CSR form");
    return TCAD_E_INVALID;
  }
  std::vector<int> ptr(3 * nd + 1, 0), idx;
  std::vector<long long> src;
  idx.reserve((size_t)nd * (I.dim == 3 ? 39 : 29));
  src.reserve(idx.capacity());
  const long long sxy = (long long)nx * ny;
  for (long long k = 0; k < nd; ++k) {
    const int i = (int)(k % nx), j = (int)((k / nx) % ny), kk = (int)(k / sxy);
    const long long g = I.off + k;
    for (int r = 0; r < 3; ++r) {
      auto nbr = [&](int dir, long long kn) {
        const int e = 9 + 5 * dir;
        if (r == 0) {
          idx.push_back((int)(3 * kn));
          src.push_back((long long)e * N + g);
        } else if (r == 1) {
          idx.push_back((int)(3 * kn));
          src.push_back((long long)(e + 1) * N + g);
          idx.push_back((int)(3 * kn + 1));
          src.push_back((long long)(e + 2) * N + g);
        } else {
          idx.push_back((int)(3 * kn));
          src.push_back((long long)(e + 3) * N + g);
          idx.push_back((int)(3 * kn + 2));
          src.push_back((long long)(e + 4) * N + g);
        }
      };
      if (I.dim == 3 && kk > 0) nbr(4, k - sxy);
      if (j > 0) nbr(2, k - nx);
      if (i > 0) nbr(0, k - 1);
      for (int c = 0; c < 3; ++c) {
        idx.push_back((int)(3 * k + c));
        src.push_back((long long)(3 * r + c) * N + g);
      }
      if (i < nx - 1) nbr(1, k + 1);
      if (j < ny - 1) nbr(3, k + nx);
      if (I.dim == 3 && kk < nz - 1) nbr(5, k + sxy);
      ptr[3 * k + r + 1] = (int)idx.size();
    }
  }
  Batch::FallbackCsr &F = b.fb[d];
  F.n = (int)(3 * nd);
  F.nnz = (long long)idx.size();
  TCAD_HIP_CHECK(hipMalloc(&F.ptr, sizeof(int) * ptr.size()));
  TCAD_HIP_CHECK(hipMalloc(&F.idx, sizeof(int) * idx.size()));
  TCAD_HIP_CHECK(hipMalloc(&F.src, sizeof(long long) * src.size()));
  TCAD_HIP_CHECK(hipMalloc(&F.val, sizeof(double) * idx.size()));
  TCAD_HIP_CHECK(hipMalloc(&F.b, sizeof(double) * F.n));
  TCAD_HIP_CHECK(hipMalloc(&F.x, sizeof(double) * F.n));
  TCAD_HIP_CHECK(hipMemcpy(F.ptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice));
  TCAD_HIP_CHECK(hipMemcpy(F.idx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice));
  TCAD_HIP_CHECK(hipMemcpy(F.src, src.data(), sizeof(long long) * src.size(),
                           hipMemcpyHostToDevice));
  return TCAD_OK;
}

int linear_fallback(Batch &b, int d, double rtol, int32_t *h_iters, double *h_resid) {
  if ((int)b.fb.size() != b.ndev) b.fb.resize(b.ndev);
  int rc;
  if (!b.fb[d].ptr && (rc = build_fallback_csr(b, d))) return rc;
  const Batch::FallbackCsr &F = b.fb[d];
  const DevInfo &I = b.info[d];
  const long long nd = I.n;
  hipLaunchKernelGGL(k_fb_gather, dim3((unsigned)((F.nnz + 255) / 256)), dim3(256), 0, b.stream,
                     F.nnz, F.src, b.A64, F.val);
  hipLaunchKernelGGL(k_fb_rhs, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, b.stream, nd,
                     b.ntot, I.off, b.bs, F.b);
  TCAD_HIP_CHECK(hipMemsetAsync(F.x, 0, sizeof(double) * F.n, b.stream));
  b.launches += 2;
  TCAD_HIP_CHECK(hipGetLastError());
  rc = tcad_csr_solve(F.n, F.ptr, F.idx, F.val, F.b, F.x, TCAD_CSR_GMRES, TCAD_CSR_PREC_ILU0, 1,
                      rtol, 4000, 50, (void *)b.stream, h_iters, h_resid);
  b.fallbacks++;
  if (rc != TCAD_OK) return rc;
  hipLaunchKernelGGL(k_fb_sol, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, b.stream, nd,
                     b.ntot, I.off, F.x, b.y64);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

// Refine the solves of the devices whose step just passed the convergence
// test (Scal.conv_now), one 1e-6 inner sweep at a time, until the contact
// currents at the candidate state balance to 1e-8 of the largest (toward
// the FP64 rounding floor, 1e-16, if need be).  A reverse-bias terminal
// current is a cancellation of fluxes many decades larger: with the
// residual at 1e-14 it still moved by percents with the summation order
// (measured on a 24x16 diode at -2 V), while a forward-biased or on-state
// device usually needs no sweep at all.  Runs on the FP64 matrix in both
// modes: where the Jacobian's conditioning times FP32 rounding approaches
// 1, refinement around the FP32-stored matrix stalls early.
static int refine_converging(Batch &b, const UArgs &U, const double *psi, const double *n,
                             const double *p, double max_dpsi, int maxit,
                             const TransientArgs *tr) {
  int rc;
  const double inner_ref = 1e-6, final_tol = 1e-16;
  const dim3 gd((b.ndev + 63) / 64), td(64);
  hipLaunchKernelGGL(k_mask_converging, gd, td, 0, b.stream, b.d_scal, b.ndev);
  b.launches++;
  double *cp = b.cand, *cn = b.cand + b.ntot, *cq = b.cand + 2 * b.ntot;
  for (int sweep = 0; sweep < 16; ++sweep) {
    hipLaunchKernelGGL(k_newton_update, dim3(b.nblk), dim3(BS), 0, b.stream, U, b.y64, psi, n, p,
                       max_dpsi, 1, cp, cn, cq);
    b.launches++;
    if ((rc = launch_assemble(b, cp, cn, cq, b.F, b.Fnp, false, true, 2, tr))) return rc;
    hipLaunchKernelGGL(k_interior_sum, dim3(b.nblk), dim3(BS), 0, b.stream, U, b.d_cidx, b.F);
    b.launches++;
    if ((rc = launch_contact_sums(b, b.Fnp, b.d_Icand))) return rc;
    hipLaunchKernelGGL(k_refine_check, dim3(b.ndev), dim3(BS), 0, b.stream, U, b.d_Icand,
                       b.d_cscale, sweep == 0 ? 1 : 0);
    b.launches++;
    TCAD_HIP_CHECK(hipGetLastError());
    if ((rc = sync_scalars(b))) return rc;
    int open = 0;
    for (int d = 0; d < b.ndev; ++d) open += b.h_scal[d].active;
    if (open == 0) break;
    if ((rc = ir_solve(b, false, false, final_tol, 1, maxit, inner_ref))) return rc;
  }
  hipLaunchKernelGGL(k_unmask, gd, td, 0, b.stream, b.d_scal, b.ndev);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

// Per-call settings of the coupled Newton loop shared by solve_bias and the
// point queue.
struct NewtonCfg {
  bool forcing, mixed, loose32;
  double tol, max_dpsi, rtol, eta_max, accept;
  int maxit;
};

static NewtonCfg newton_cfg(Batch &b, int &mode, double tol, double max_dpsi, double rtol,
                            int maxit) {
  NewtonCfg c{};
  c.forcing = (mode & TCAD_LIN_FORCING) != 0;
  b.tile = (mode & TCAD_PRECOND_NODE) ? 0 : 1;
  mode &= ~(TCAD_LIN_FORCING | TCAD_PRECOND_NODE);
  c.mixed = mode == TCAD_LIN_BICGSTAB_MIXED;
  // FP32-stored matrix for loose forcing solves
  c.loose32 = c.forcing && !c.mixed && b.forcing_fp32;
  c.tol = tol;
  c.max_dpsi = max_dpsi;
  c.rtol = rtol;
  c.eta_max = rtol > b.forcing_max ? rtol : b.forcing_max;
  c.accept = rtol > 1e-6 ? rtol : 1e-6;
  c.maxit = maxit;
  return c;
}

// One Newton iteration of every active device: assemble, solve (FP64
// refinement; the inner BiCGStab on the FP32-stored matrix in the mixed mode
// and, with loose32, for loose forcing targets), measure, refine a
// converging step to the FP64 floor, apply, and retire devices that
// converged, failed or reached max_iter.  eta_next: the loosest target of
// the next solve over the active devices (read back with the scalars).
static int bias_newton_step(Batch &b, const UArgs &U, const NewtonCfg &c, int max_iter,
                            double &eta_next, double *psi, double *n, double *p) {
  int rc;
  const dim3 gd((b.ndev + 63) / 64), td(64);
  const bool fp32_now = c.mixed || (c.loose32 && eta_next >= 1e-5);
  const double inner = fp32_now ? b.fp32_inner : c.rtol;
  if ((rc = launch_surface_mobility(b, psi, true))) return rc;
  if ((rc = launch_assemble(b, psi, n, p, b.F, nullptr, c.mixed || c.loose32, true))) return rc;
  if (b.tile && (rc = launch_tile_setup(b, true))) return rc;
  if ((rc = launch_flags_to_status(b))) return rc;
  if ((rc = ir_solve(b, fp32_now, true, c.rtol, 16, c.maxit, inner, true, true))) return rc;
  // measure the step, and before applying a step that passes the test
  // refine its solve to the FP64 rounding floor: the converged state is
  // then the reference's (exact Newton steps converge quadratically, so
  // its final state is exact to rounding), which the terminal current
  // needs where it is a cancellation of much larger fluxes (reverse bias)
  hipLaunchKernelGGL(k_newton_update, dim3(b.nblk), dim3(BS), 0, b.stream, U, b.y64, psi, n, p,
                     c.max_dpsi, 0, psi, n, p);
  hipLaunchKernelGGL(k_newton_check, dim3(b.ndev), dim3(BS), 0, b.stream, U, c.tol,
                     c.accept * c.accept, c.rtol, c.eta_max, c.forcing ? 1 : 0, b.forcing_c,
                     b.forcing_pow, b.tile_lag);
  b.launches += 2;
  TCAD_HIP_CHECK(hipGetLastError());
  if ((rc = sync_scalars(b))) return rc;
  int conv = 0;
  eta_next = 1.0;
  for (int d = 0; d < b.ndev; ++d) {
    conv += b.h_scal[d].active && b.h_scal[d].conv_now;
    if (b.h_scal[d].active) eta_next = fmin(eta_next, sqrt(b.h_scal[d].eta2));
  }
  if (!c.forcing) eta_next = c.rtol;
  if (conv && c.rtol > 1e-16) {
    if ((rc = refine_converging(b, U, psi, n, p, c.max_dpsi, c.maxit, nullptr))) return rc;
  }
  hipLaunchKernelGGL(k_newton_update, dim3(b.nblk), dim3(BS), 0, b.stream, U, b.y64, psi, n, p,
                     c.max_dpsi, 1, psi, n, p);
  hipLaunchKernelGGL(k_newton_finish, gd, td, 0, b.stream, b.d_scal, b.ndev, max_iter);
  b.launches += 2;
  TCAD_HIP_CHECK(hipGetLastError());
  return sync_scalars(b);
}

static int solve_bias_impl(Batch &b, int max_iter, double tol, double max_dpsi, int mode,
                           double rtol, int maxit, const int32_t *mask, double *psi, double *n,
                           double *p) {
  int rc;
  const NewtonCfg c = newton_cfg(b, mode, tol, max_dpsi, rtol, maxit);
  if ((rc = launch_set_active(b, mask, c.forcing ? c.eta_max : rtol,
                              c.forcing && b.forcing_pred_only ? rtol : -1.0)))
    return rc;
  hipLaunchKernelGGL(k_stamp, dim3(b.nblk), dim3(BS), 0, b.stream, b.d_info, b.d_blk2dev, b.d_cidx,
                     b.d_cn_lnr, b.d_cn_n0, b.d_cn_p0, b.d_cn_contact, b.d_contact_V, b.d_scal,
                     psi, n, p);
  b.launches++;
  UArgs U = uargs(b);
  double eta_next = c.forcing ? c.eta_max : rtol;
  for (int it = 0; it < max_iter; ++it) {
    b.kry_slot = it;
    rc = bias_newton_step(b, U, c, max_iter, eta_next, psi, n, p);
    b.kry_slot = -1;
    if (rc) return rc;
    int open = 0;
    for (int d = 0; d < b.ndev; ++d) open += b.h_scal[d].active;
    if (open == 0) break;
  }
  return TCAD_OK;
}

// Re-arm the devices of `mask` for their next queued point (the other
// devices' selection flags are cleared, so the stamp that follows touches
// only the re-armed devices).
__global__ void k_rearm(Scal *scal, int ndev, const int *mask, double eta0) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndev) return;
  Scal &S = scal[d];
  S.sel = mask[d] != 0;
  if (!S.sel) return;
  S.active = 1;
  S.tile_keep = 0;
  S.eta2 = eta0 * eta0;
  S.result = 0;
  S.newton_it = 0;
  S.newton_done = 0;
  S.kry_total = 0;
  S.ir_it = 0;
  S.status = 0;
  S.err = 0.0;
  S.conv_now = 0;
}

// Copy the state of the devices of `mask` to (dir 0) or from (dir 1) the
// backup planes.
__global__ void __launch_bounds__(BS) k_backup(const DevInfo *info, const int *blk2dev,
                                               const int *mask, int dir, long long N,
                                               double *psi, double *n, double *p, double *bk) {
  const int d = blk2dev[blockIdx.x];
  if (!mask[d]) return;
  const DevInfo &I = info[d];
  const NodeAt P = node_at(I, blockIdx.x);
  if (!P.valid) return;
  const long long g = I.off + P.k;
  if (dir == 0) {
    bk[g] = psi[g];
    bk[N + g] = n[g];
    bk[2 * N + g] = p[g];
  } else {
    psi[g] = bk[g];
    n[g] = bk[N + g];
    p[g] = bk[2 * N + g];
  }
}

// A queue of bias points per device, solved in one call: each device goes
// on to its next point as soon as it converges at the present one (no
// barrier between devices), warm-started from it, and stops at its first
// failure with the state of its last converged point restored.
static int queue_impl(Batch &b, int max_iter, double tol, double max_dpsi, int mode, double rtol,
                      int maxit, int npmax, const int32_t *h_np, const double *h_qV,
                      const double *h_qG, double *psi, double *n, double *p, int32_t *h_done,
                      int32_t *h_iters, int32_t *h_kry, int32_t *h_status, double *h_I) {
  int rc;
  const NewtonCfg c = newton_cfg(b, mode, tol, max_dpsi, rtol, maxit);
  const int nd = b.ndev;
  const double eta0 = c.forcing ? c.eta_max : rtol;
  if (!b.bk) {
    hipError_t e = hipMalloc(&b.bk, sizeof(double) * 3 * (size_t)b.ntot);
    if (e != hipSuccess) {
      set_error(std::string("tcad_solve_queue: hipMalloc failed: ") + hipGetErrorString(e));
      return TCAD_E_HIP;
    }
  }
  if (!b.d_qmask) {
    TCAD_HIP_CHECK(hipMalloc(&b.d_qmask, sizeof(int) * nd));
  }
  // host copies of the bias tables, overwritten per device as it advances
  std::vector<double> hV(std::max(b.ncontact, 1)), hG(std::max(b.ngate, 1));
  if (b.ncontact)
    TCAD_HIP_CHECK(hipMemcpy(hV.data(), b.d_contact_V, sizeof(double) * b.ncontact,
                             hipMemcpyDeviceToHost));
  if (b.ngate)
    TCAD_HIP_CHECK(hipMemcpy(hG.data(), b.d_gate_Vg, sizeof(double) * b.ngate,
                             hipMemcpyDeviceToHost));
  std::vector<int> pidx(nd, 0), state(nd, 0), qm(nd, 0);   // state: 0 idle, 1 running, 2 done
  std::vector<double> hI(std::max(b.ncontact, 1));
  auto load_bias = [&](int d) {
    const DevInfo &I = b.info[d];
    const int k = pidx[d];
    for (int t = 0; t < I.nc; ++t) hV[I.c0 + t] = h_qV[(size_t)k * b.ncontact + I.c0 + t];
    for (int t = 0; t < I.ng; ++t) hG[I.g0 + t] = h_qG[(size_t)k * b.ngate + I.g0 + t];
  };
  auto upload = [&]() -> int {
    if (b.ncontact)
      TCAD_HIP_CHECK(hipMemcpyAsync(b.d_contact_V, hV.data(), sizeof(double) * b.ncontact,
                                    hipMemcpyHostToDevice, b.stream));
    if (b.ngate)
      TCAD_HIP_CHECK(hipMemcpyAsync(b.d_gate_Vg, hG.data(), sizeof(double) * b.ngate,
                                    hipMemcpyHostToDevice, b.stream));
    TCAD_HIP_CHECK(hipMemcpyAsync(b.d_qmask, qm.data(), sizeof(int) * nd, hipMemcpyHostToDevice,
                                  b.stream));
    // the host tables change again right after: wait for the copies
    TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
    return TCAD_OK;
  };
  auto backup = [&](int dir) {
    hipLaunchKernelGGL(k_backup, dim3(b.nblk), dim3(BS), 0, b.stream, b.d_info, b.d_blk2dev,
                       b.d_qmask, dir, b.ntot, psi, n, p, b.bk);
    b.launches++;
  };
  auto stamp = [&]() {
    hipLaunchKernelGGL(k_stamp, dim3(b.nblk), dim3(BS), 0, b.stream, b.d_info, b.d_blk2dev,
                       b.d_cidx, b.d_cn_lnr, b.d_cn_n0, b.d_cn_p0, b.d_cn_contact,
                       b.d_contact_V, b.d_scal, psi, n, p);
    b.launches++;
  };
  int total_pts = 0;
  for (int d = 0; d < nd; ++d) {
    h_done[d] = 0;
    h_status[d] = TCAD_NEWTON_OK;
    qm[d] = h_np[d] > 0 ? 1 : 0;
    total_pts += std::max(0, h_np[d]);
    if (qm[d]) {
      state[d] = 1;
      load_bias(d);
    }
  }
  if (total_pts == 0) return TCAD_OK;
  if ((rc = upload())) return rc;
  backup(0);                                      // the start state, restored on a first-point failure
  if ((rc = launch_set_active(b, qm.data(), eta0))) return rc;
  stamp();
  UArgs U = uargs(b);
  double eta_next = eta0;
  const long long guard = (long long)max_iter * total_pts + 1;
  int local_it = 0;                               // Newton index at the current point (one device)
  for (long long it = 0; it < guard; ++it) {
    b.kry_slot = nd == 1 ? local_it++ : -1;
    rc = bias_newton_step(b, U, c, max_iter, eta_next, psi, n, p);
    b.kry_slot = -1;
    if (rc) return rc;
    std::vector<int> conv, fail;
    int open = 0;
    for (int d = 0; d < nd; ++d) {
      if (state[d] != 1) continue;
      if (b.h_scal[d].active) {
        ++open;
        continue;
      }
      (b.h_scal[d].newton_done ? conv : fail).push_back(d);
    }
    if (!conv.empty()) {
      // currents at the converged states (every contact; only the converged
      // devices' are kept)
      if ((rc = launch_assemble(b, psi, n, p, b.F, b.Fnp, false, false, 1))) return rc;
      if ((rc = launch_contact_sums(b, b.Fnp, b.v64[0]))) return rc;
      if (b.ncontact)
        TCAD_HIP_CHECK(hipMemcpyAsync(hI.data(), b.v64[0], sizeof(double) * b.ncontact,
                                      hipMemcpyDeviceToHost, b.stream));
      TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
      std::fill(qm.begin(), qm.end(), 0);
      for (int d : conv) {
        const DevInfo &I = b.info[d];
        const int k = pidx[d];
        for (int t = 0; t < I.nc; ++t)
          h_I[(size_t)k * b.ncontact + I.c0 + t] = hI[I.c0 + t];
        h_iters[(size_t)d * npmax + k] = b.h_scal[d].newton_it;
        h_kry[(size_t)d * npmax + k] = b.h_scal[d].kry_total;
        h_done[d] = k + 1;
        qm[d] = 1;
      }
      if ((rc = upload())) return rc;
      backup(0);                                  // their converged states
      std::fill(qm.begin(), qm.end(), 0);
      int rearm = 0;
      for (int d : conv) {
        if (++pidx[d] < h_np[d]) {
          load_bias(d);
          qm[d] = 1;
          ++rearm;
          ++open;
        } else {
          state[d] = 2;
        }
      }
      if (rearm) {
        local_it = 0;
        if ((rc = upload())) return rc;
        hipLaunchKernelGGL(k_rearm, dim3((nd + 63) / 64), dim3(64), 0, b.stream, b.d_scal, nd,
                           b.d_qmask, eta0);
        b.launches++;
        stamp();
        eta_next = c.forcing ? std::min(eta_next, eta0) : rtol;
      }
    }
    if (!fail.empty()) {
      std::fill(qm.begin(), qm.end(), 0);
      for (int d : fail) {
        const Scal &S = b.h_scal[d];
        h_status[d] = S.result ? S.result : TCAD_NEWTON_MAX_ITER;
        h_iters[(size_t)d * npmax + pidx[d]] = b.h_scal[d].newton_it;
        h_kry[(size_t)d * npmax + pidx[d]] = b.h_scal[d].kry_total;
        qm[d] = 1;
        state[d] = 2;
      }
      if ((rc = upload())) return rc;
      backup(1);                                  // back to their last converged state
    }
    TCAD_HIP_CHECK(hipGetLastError());
    if (open == 0) break;
  }
  TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
  return TCAD_OK;
}

static int transient_impl(Batch &b, int max_iter, double tol, double max_dpsi, int mode,
                          double rtol, int maxit, const int32_t *mask, const TransientArgs &tr,
                          double *psi, double *n, double *p) {
  int rc;
  b.tile = (mode & TCAD_PRECOND_NODE) ? 0 : 1;
  mode &= ~(TCAD_LIN_FORCING | TCAD_PRECOND_NODE);
  const bool mixed = mode == TCAD_LIN_BICGSTAB_MIXED;
  if ((rc = launch_set_active(b, mask, rtol))) return rc;
  UArgs U = uargs(b);
  const double inner = mixed ? 1e-6 : rtol;
  const double accept = rtol > 1e-6 ? rtol : 1e-6;
  const dim3 gd((b.ndev + 63) / 64), td(64);
  double *cp = b.cand, *cn = b.cand + b.ntot, *cq = b.cand + 2 * b.ntot;
  for (int it = 0; it < max_iter; ++it) {
    // F (all rows, step terms included) and the eliminated, scaled matrix;
    // the contact rows move the contacts to the new time's values
    if ((rc = launch_assemble(b, psi, n, p, b.F, nullptr, mixed, true, 0, &tr))) return rc;
    if (b.tile && (rc = launch_tile_setup(b, true))) return rc;
    if ((rc = launch_flags_to_status(b))) return rc;
    hipLaunchKernelGGL(k_merit_part, dim3(b.nblk), dim3(BS), 0, b.stream, U, b.F);
    hipLaunchKernelGGL(k_ls_merit, dim3(b.ndev), dim3(BS), 0, b.stream, U, 0);
    b.launches += 2;
    if ((rc = ir_solve(b, mixed, true, rtol, 16, maxit, inner, true))) return rc;
    // a full step that would pass the test is refined first (as in
    // solve_bias), so the state it lands on is exact to rounding
    hipLaunchKernelGGL(k_newton_update, dim3(b.nblk), dim3(BS), 0, b.stream, U, b.y64, psi, n, p,
                       max_dpsi, 0, psi, n, p);
    hipLaunchKernelGGL(k_newton_check, dim3(b.ndev), dim3(BS), 0, b.stream, U, tol,
                       accept * accept, rtol, 1.0, 0, b.forcing_c, b.forcing_pow, 0.0);
    b.launches += 2;
    TCAD_HIP_CHECK(hipGetLastError());
    if ((rc = sync_scalars(b))) return rc;
    int conv = 0;
    for (int d = 0; d < b.ndev; ++d) conv += b.h_scal[d].active && b.h_scal[d].conv_now;
    if (conv && (rc = refine_converging(b, U, psi, n, p, max_dpsi, maxit, &tr))) return rc;
    // Armijo backtracking, every device at its own lambda
    hipLaunchKernelGGL(k_ls_begin, gd, td, 0, b.stream, b.d_scal, b.ndev);
    b.launches++;
    for (int ls = 0; ls < 40; ++ls) {
      hipLaunchKernelGGL(k_tr_step, dim3(b.nblk), dim3(BS), 0, b.stream, U, b.y64, psi, n, p,
                         max_dpsi, 0, cp, cn, cq);
      b.launches++;
      if ((rc = launch_assemble(b, cp, cn, cq, b.F, nullptr, false, true, 2, &tr))) return rc;
      hipLaunchKernelGGL(k_merit_part, dim3(b.nblk), dim3(BS), 0, b.stream, U, b.F);
      hipLaunchKernelGGL(k_ls_merit, dim3(b.ndev), dim3(BS), 0, b.stream, U, 1);
      b.launches += 2;
      TCAD_HIP_CHECK(hipGetLastError());
      if ((rc = sync_scalars(b))) return rc;
      int open = 0;
      for (int d = 0; d < b.ndev; ++d) open += b.h_scal[d].active;
      if (open == 0) break;
    }
    hipLaunchKernelGGL(k_unmask, gd, td, 0, b.stream, b.d_scal, b.ndev);
    hipLaunchKernelGGL(k_tr_step, dim3(b.nblk), dim3(BS), 0, b.stream, U, b.y64, psi, n, p,
                       max_dpsi, 1, cp, cn, cq);
    hipLaunchKernelGGL(k_tr_finish, dim3(b.ndev), dim3(BS), 0, b.stream, U, tol);
    b.launches += 3;
    TCAD_HIP_CHECK(hipGetLastError());
    if ((rc = sync_scalars(b))) return rc;
    int open = 0;
    for (int d = 0; d < b.ndev; ++d) open += b.h_scal[d].active;
    if (open == 0) break;
  }
  return TCAD_OK;
}

static int solve_equilibrium_impl(Batch &b, int max_iter, double tol, double max_dpsi,
                                  int from_guess, const int32_t *mask, double *psi, double *n,
                                  double *p) {
  int rc;
  if ((rc = launch_set_active(b, mask))) return rc;
  if (from_guess) {
    hipLaunchKernelGGL(k_bulk_guess, dim3(b.nblk), dim3(BS), 0, b.stream, b.d_info, b.d_blk2dev,
                       b.f.d_C, b.f.d_nie_s, b.d_scal, psi);
    b.launches++;
  }
  UArgs U = uargs(b);
  for (int it = 0; it < max_iter; ++it) {
    if ((rc = launch_assemble_poisson(b, psi, b.F))) return rc;
    if (b.pcg_tile && (rc = launch_pcg_tile_setup(b))) return rc;
    if ((rc = launch_flags_to_status(b))) return rc;
    if ((rc = pcg_poisson(b, 1e-12, 20000))) return rc;
    hipLaunchKernelGGL(k_poisson_update, dim3(b.nblk), dim3(BS), 0, b.stream, U, b.veq[0], psi,
                       max_dpsi);
    hipLaunchKernelGGL(k_newton_finalize, dim3(b.ndev), dim3(BS), 0, b.stream, U, tol, 1);
    b.launches += 2;
    TCAD_HIP_CHECK(hipGetLastError());
    if ((rc = sync_scalars(b))) return rc;
    int open = 0;
    for (int d = 0; d < b.ndev; ++d) open += b.h_scal[d].active;
    if (open == 0) break;
  }
  hipLaunchKernelGGL(k_slaved, dim3(b.nblk), dim3(BS), 0, b.stream, b.d_info, b.d_blk2dev,
                     b.f.d_nie_s, b.d_scal, psi, n, p);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

// ---------------------------------------------------------------------------
// Batch construction
// ---------------------------------------------------------------------------
template <typename T>
static int dalloc(T **ptr, size_t count) {
  *ptr = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void **)ptr, count * sizeof(T));
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return TCAD_E_HIP;
  }
  return TCAD_OK;
}

template <typename T>
static int upload(T **ptr, const std::vector<T> &v, hipStream_t s) {
  int rc = dalloc(ptr, v.size());
  if (rc) return rc;
  if (!v.empty())
    TCAD_HIP_CHECK(hipMemcpyAsync(*ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return TCAD_OK;
}

static void free_batch(Batch *b) {
  if (!b) return;
  void *ptrs[] = {b->d_info, b->d_blk2dev, b->d_h[0], b->d_h[1], b->d_h[2], b->d_dv[0],
                  b->d_dv[1], b->d_dv[2], b->d_cidx, b->d_gidx, b->d_cn_lnr, b->d_cn_n0,
                  b->d_cn_p0, b->d_cn_contact, b->d_gn_kw, b->d_gn_psib, b->d_gn_gate,
                  b->d_contact_V, b->d_gate_Vg, b->d_gate_Vfb, b->d_cptr, b->d_cnodes,
                  b->d_cscale, b->d_c2dev, b->d_gptr, b->d_gnode, b->d_surf[0], b->d_surf[1], b->d_surf[2],
                  b->d_surf[3], b->A64, b->A32, b->Dinv64, b->Dinv32, b->bs, b->rowscale,
                  b->colscale, b->y64, b->F, b->Fnp, b->part, b->d_scal, b->Aeq, b->beq,
                  b->d_flag, b->d_mask, b->d_ratio, b->cand, b->d_Icand, b->T4, b->T4s, b->CPL,
                  b->P16, b->d_dts};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  for (hipEvent_t e : b->probe_ev) (void)hipEventDestroy(e);
  if (b->bk) (void)hipFree(b->bk);
  if (b->d_qmask) (void)hipFree(b->d_qmask);
  if (b->d_bmap) (void)hipFree(b->d_bmap);
  for (auto &F : b->fb) {
    void *fp[] = {F.ptr, F.idx, F.src, F.val, F.b, F.x};
    for (void *q : fp)
      if (q) (void)hipFree(q);
  }
  if (b->h_bmap) (void)hipHostFree(b->h_bmap);
  for (auto &g : b->graphs) {
    (void)hipGraphExecDestroy(g.exec);
    (void)hipGraphDestroy(g.graph);
  }
  if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
  for (int i = 0; i < 9; ++i)
    if (b->v64[i]) (void)hipFree(b->v64[i]);
  if (b->ir_r) (void)hipFree(b->ir_r);
  if (b->ir_rhs) (void)hipFree(b->ir_rhs);
  for (int i = 0; i < 6; ++i)
    if (b->veq[i]) (void)hipFree(b->veq[i]);
  if (b->h_scal) (void)hipHostFree(b->h_scal);
  delete b;
}

static int create_impl(const tcad_device_desc *devs, int ndev, const tcad_static_fields *fields,
                       void *stream, Batch **out) {
  if (!devs || ndev <= 0 || !fields || !out) {
    set_error("tcad_batch_create: null argument or ndev <= 0");
    return TCAD_E_INVALID;
  }
  int ndevices = 0;
  if (hipGetDeviceCount(&ndevices) != hipSuccess || ndevices == 0) {
    set_error("no HIP device available");
    return TCAD_E_NO_DEVICE;
  }
  Batch *b = new (std::nothrow) Batch();
  if (!b) {
    set_error("out of host memory");
    return TCAD_E_INVALID;
  }
  b->ndev = ndev;
  b->stream = (hipStream_t)stream;
  b->f = *fields;
  std::vector<double> hx[3], dvw[3];
  std::vector<int> blk2dev;
  std::vector<int> cidx, gidx, cn_contact, gn_gate, cptr{0}, cnodes, gptr{0}, gate_node;
  std::vector<double> cn_lnr, cn_n0, cn_p0, gn_kw, gn_psib, cscale;
  std::vector<double> surf[4];
  long long off = 0, end_max = 0;
  std::vector<std::pair<long long, long long>> ranges;
  int blk = 0, cglobal = 0, gglobal = 0;
  b->info.resize(ndev);
  for (int d = 0; d < ndev; ++d) {
    const tcad_device_desc &D = devs[d];
    if (D.dim != 2 && D.dim != 3) {
      set_error("device " + std::to_string(d) + ": dim must be 2 or 3");
      free_batch(b);
      return TCAD_E_INVALID;
    }
    const int nz = D.dim == 3 ? D.nz : 1;
    if (D.nx < 2 || D.ny < 2 || (D.dim == 3 && nz < 2)) {
      set_error("device " + std::to_string(d) + ": every axis needs >= 2 nodes");
      free_batch(b);
      return TCAD_E_INVALID;
    }
    DevInfo &I = b->info[d];
    I.dim = D.dim;
    I.nx = D.nx;
    I.ny = D.ny;
    I.nz = nz;
    if (D.node_offset >= 0) off = D.node_offset;
    I.off = off;
    I.n = (long long)D.nx * D.ny * nz;
    ranges.emplace_back(I.off, I.off + I.n);
    end_max = std::max(end_max, I.off + I.n);
    I.blk0 = blk;
    if (D.dim == 3) {
      I.px = (D.nx + PX3 - 1) / PX3;
      I.py = (D.ny + PY3 - 1) / PY3;
      I.nblk = I.px * I.py * ((nz + PZ3 - 1) / PZ3);
    } else {
      I.px = (D.nx + PX - 1) / PX;
      I.py = (D.ny + PY - 1) / PY;
      I.nblk = I.px * I.py * nz;
    }
    const int n_ax[3] = {D.nx, D.ny, nz};
    const double *hs[3] = {D.h_hx, D.h_hy, D.h_hz};
    const double *dvs[3] = {D.h_dvx, D.h_dvy, D.h_dvz};
    for (int a = 0; a < 3; ++a) {
      I.h_off[a] = (int)hx[a].size();
      I.dv_off[a] = (int)dvw[a].size();
      if (a < D.dim) {
        if (!hs[a] || !dvs[a]) {
          set_error("device " + std::to_string(d) + ": missing geometry array");
          free_batch(b);
          return TCAD_E_INVALID;
        }
        hx[a].insert(hx[a].end(), hs[a], hs[a] + n_ax[a] - 1);
        dvw[a].insert(dvw[a].end(), dvs[a], dvs[a] + n_ax[a]);
      }
    }
    I.VT = D.VT; I.Ns = D.Ns; I.LD = D.LD; I.J0 = D.J0; I.R0 = D.R0;
    I.Cn = D.Cn_auger; I.Cp = D.Cp_auger;
    I.srh = D.srh; I.auger = D.auger;
    I.surf = (D.surface_mobility && D.dim == 2) ? 1 : 0;
    I.T = D.T;
    I.surf_off = (int)surf[0].size();
    if (I.surf) {
      if (!D.h_mun_row0 || !D.h_mun_row1 || !D.h_mup_row0 || !D.h_mup_row1) {
        set_error("device " + std::to_string(d) + ": surface mobility rows missing");
        free_batch(b);
        return TCAD_E_INVALID;
      }
      surf[0].insert(surf[0].end(), D.h_mun_row0, D.h_mun_row0 + D.nx);
      surf[1].insert(surf[1].end(), D.h_mun_row1, D.h_mun_row1 + D.nx);
      surf[2].insert(surf[2].end(), D.h_mup_row0, D.h_mup_row0 + D.nx);
      surf[3].insert(surf[3].end(), D.h_mup_row1, D.h_mup_row1 + D.nx);
      b->any_surf = 1;
    }
    for (int t = 0; t < I.nblk; ++t) blk2dev.push_back(d);
    // boundary tables
    if ((long long)cidx.size() < off + I.n) {
      cidx.resize(off + I.n, -1);
      gidx.resize(off + I.n, -1);
    }
    I.c0 = cglobal;
    I.nc = D.ncontacts;
    for (int c = 0; c < D.ncontacts; ++c) {
      for (int t = D.h_contact_ptr[c]; t < D.h_contact_ptr[c + 1]; ++t) {
        const int node = D.h_contact_nodes[t];
        if (node < 0 || node >= I.n) {
          set_error("device " + std::to_string(d) + ": contact node index out of range");
          free_batch(b);
          return TCAD_E_INDEX;
        }
        const long long gnode = off + node;
        // A node listed by several contacts takes the Dirichlet value of
        // the last one (the reference overwrites F3[kk] contact by contact,
        // device2d.py:1801-1806), while every contact's current still sums
        // its whole node list (terminal_current, device2d.py:2219-2226).
        cidx[gnode] = (int)cn_lnr.size();
        cn_lnr.push_back(D.h_contact_lnr[t]);
        cn_n0.push_back(D.h_contact_n0[t]);
        cn_p0.push_back(D.h_contact_p0[t]);
        cn_contact.push_back(cglobal + c);
        cnodes.push_back((int)gnode);
      }
      cptr.push_back((int)cnodes.size());
      cscale.push_back(D.J0);
      cscale.push_back(D.dim == 3 ? D.LD * D.LD : D.LD);
    }
    cglobal += D.ncontacts;
    I.g0 = gglobal;
    I.ng = D.ngates;
    for (int gg = 0; gg < D.ngates; ++gg) {
      for (int t = D.h_gate_ptr[gg]; t < D.h_gate_ptr[gg + 1]; ++t) {
        const int node = D.h_gate_nodes[t];
        if (node < 0 || node >= I.n) {
          set_error("device " + std::to_string(d) + ": gate node index out of range");
          free_batch(b);
          return TCAD_E_INDEX;
        }
        const long long gnode = off + node;
        if (gidx[gnode] >= 0) {
          set_error("device " + std::to_string(d) + ": node " + std::to_string(node) +
                    " belongs to two gates");
          free_batch(b);
          return TCAD_E_INVALID;
        }
        gidx[gnode] = (int)gn_kw.size();
        gn_kw.push_back(D.h_gate_kw[t]);
        gn_psib.push_back(D.h_gate_psib[t]);
        gn_gate.push_back(gglobal + gg);
        gate_node.push_back((int)gnode);
      }
      gptr.push_back((int)gn_kw.size());
    }
    gglobal += D.ngates;
    b->dim_max = std::max(b->dim_max, D.dim);
    off += I.n;
    blk += I.nblk;
  }
  std::sort(ranges.begin(), ranges.end());
  for (size_t r = 1; r < ranges.size(); ++r)
    if (ranges[r].first < ranges[r - 1].second) {
      set_error("tcad_batch_create: device node ranges overlap");
      free_batch(b);
      return TCAD_E_INVALID;
    }
  b->ntot = end_max;
  b->nblk = blk;
  b->ne = nentries(b->dim_max);
  b->ncontact = cglobal;
  b->ngate = gglobal;
  const hipStream_t s = b->stream;
  const size_t N = (size_t)b->ntot;
  int rc = TCAD_OK;
#define TRY(x)                   \
  do {                           \
    if ((rc = (x))) {            \
      free_batch(b);             \
      return rc;                 \
    }                            \
  } while (0)
  TRY(upload(&b->d_info, b->info, s));
  TRY(upload(&b->d_blk2dev, blk2dev, s));
  for (int a = 0; a < 3; ++a) {
    TRY(upload(&b->d_h[a], hx[a], s));
    TRY(upload(&b->d_dv[a], dvw[a], s));
  }
  TRY(upload(&b->d_cidx, cidx, s));
  TRY(upload(&b->d_gidx, gidx, s));
  TRY(upload(&b->d_cn_lnr, cn_lnr, s));
  TRY(upload(&b->d_cn_n0, cn_n0, s));
  TRY(upload(&b->d_cn_p0, cn_p0, s));
  TRY(upload(&b->d_cn_contact, cn_contact, s));
  TRY(upload(&b->d_gn_kw, gn_kw, s));
  TRY(upload(&b->d_gn_psib, gn_psib, s));
  TRY(upload(&b->d_gn_gate, gn_gate, s));
  TRY(upload(&b->d_cptr, cptr, s));
  TRY(upload(&b->d_cnodes, cnodes, s));
  TRY(upload(&b->d_gptr, gptr, s));
  TRY(upload(&b->d_gnode, gate_node.empty() ? std::vector<int>{0} : gate_node, s));
  TRY(upload(&b->d_cscale, cscale, s));
  for (int i = 0; i < 4; ++i) TRY(upload(&b->d_surf[i], surf[i], s));
  TRY(upload(&b->d_contact_V, std::vector<double>(std::max(1, cglobal), 0.0), s));
  TRY(upload(&b->d_gate_Vg, std::vector<double>(std::max(1, gglobal), 0.0), s));
  TRY(upload(&b->d_gate_Vfb, std::vector<double>(std::max(1, gglobal), 0.0), s));
  TRY(dalloc(&b->A64, (size_t)b->ne * N));
  TRY(dalloc(&b->A32, (size_t)b->ne * N));
  TRY(dalloc(&b->Dinv64, 9 * N));
  TRY(dalloc(&b->Dinv32, 9 * N));
  TRY(dalloc(&b->bs, 3 * N));
  TRY(dalloc(&b->rowscale, 3 * N));
  TRY(dalloc(&b->y64, 3 * N));
  TRY(dalloc(&b->F, 3 * N));
  TRY(dalloc(&b->Fnp, N));
  for (int i = 0; i < 9; ++i) TRY(dalloc(&b->v64[i], 3 * N));
  TRY(dalloc(&b->ir_r, 3 * N));
  TRY(dalloc(&b->ir_rhs, 3 * N));
  TRY(dalloc(&b->part, 8 * (size_t)b->nblk));
  TRY(dalloc(&b->d_scal, (size_t)ndev));
  TRY(dalloc(&b->Aeq, 7 * N));
  TRY(dalloc(&b->beq, N));
  for (int i = 0; i < 6; ++i) TRY(dalloc(&b->veq[i], N));
  TRY(dalloc(&b->d_flag, (size_t)ndev + 1));
  TRY(dalloc(&b->d_mask, (size_t)ndev));
  TRY(dalloc(&b->d_ratio, (size_t)ndev));
  TRY(dalloc(&b->d_dts, (size_t)ndev));
  // tile-inverse rows by patch slot (workgroups x BS; tile_row_slot)
  TRY(dalloc(&b->T4, (size_t)TILE_PLANES * b->nblk * BS));
  TRY(dalloc(&b->T4s, (size_t)3 * b->nblk * BS));
  TRY(dalloc(&b->CPL, (size_t)CPL_GROUPS * 5 * b->nblk * (BS / 2)));
  TRY(dalloc(&b->P16, (size_t)TILE_NODES * N));
  {
    // equilibrium PCG: tile block-Jacobi (default: half the PCG
    // iterations of Jacobi on the MOSFET, its sweep 5% and the tri-gate
    // headline 2% faster), or Jacobi with TCADHIP_PCG_PRECOND=jacobi (the
    // two equilibrium states differ in the last bits)
    const char *e = std::getenv("TCADHIP_PCG_PRECOND");
    b->pcg_tile = (e && std::strcmp(e, "jacobi") == 0) ? 0 : 1;
  }
  TRY(dalloc(&b->cand, 3 * N));
  TRY(dalloc(&b->d_Icand, (size_t)std::max(1, cglobal)));
#undef TRY
  TCAD_HIP_CHECK(hipMemsetAsync(b->d_flag, 0, sizeof(int) * (ndev + 1), s));
  TCAD_HIP_CHECK(hipMemsetAsync(b->d_scal, 0, sizeof(Scal) * ndev, s));
  TCAD_HIP_CHECK(hipMemsetAsync(b->A32, 0, sizeof(float) * b->ne * N, s));
  TCAD_HIP_CHECK(hipMemsetAsync(b->A64, 0, sizeof(double) * b->ne * N, s));
  if (hipHostMalloc((void **)&b->h_scal, sizeof(Scal) * ndev, hipHostMallocDefault) != hipSuccess) {
    free_batch(b);
    set_error("hipHostMalloc failed");
    return TCAD_E_HIP;
  }
  TCAD_HIP_CHECK(hipStreamSynchronize(s));
  *out = b;
  return TCAD_OK;
}

}  // namespace tcad

// ---------------------------------------------------------------------------
// extern "C" boundary
// ---------------------------------------------------------------------------
using tcad::Batch;

extern "C" {

const char *tcad_last_error(void) { return tcad::g_last_error.c_str(); }

const char *tcad_version(void) { return "tcadhip 0.1.0 (gfx950)"; }

int tcad_batch_create(const tcad_device_desc *h_devs, int32_t ndev,
                      const tcad_static_fields *fields, void *stream, tcad_batch **out) {
  if (!out) {
    tcad::set_error("tcad_batch_create: out is NULL");
    return TCAD_E_INVALID;
  }
  *out = nullptr;
  Batch *b = nullptr;
  int rc = tcad::create_impl(h_devs, ndev, fields, stream, &b);
  if (rc) return rc;
  *out = new tcad_batch{b};
  return TCAD_OK;
}

int tcad_batch_destroy(tcad_batch *b) {
  if (!b) return TCAD_OK;
  if (b->impl) (void)hipStreamSynchronize(b->impl->stream);
  tcad::free_batch(b->impl);
  delete b;
  return TCAD_OK;
}

int64_t tcad_batch_nodes(const tcad_batch *b) { return (b && b->impl) ? b->impl->ntot : -1; }

int64_t tcad_launch_count(const tcad_batch *b) { return (b && b->impl) ? b->impl->launches : -1; }

int tcad_set_bias(tcad_batch *bp, const double *h_contact_V, const double *h_gate_Vg,
                  const double *h_gate_Vfb) {
  GUARD(bp);
  Batch &b = *bp->impl;
  if (b.ncontact && h_contact_V)
    TCAD_HIP_CHECK(hipMemcpyAsync(b.d_contact_V, h_contact_V, sizeof(double) * b.ncontact,
                                  hipMemcpyHostToDevice, b.stream));
  if (b.ngate && h_gate_Vg)
    TCAD_HIP_CHECK(hipMemcpyAsync(b.d_gate_Vg, h_gate_Vg, sizeof(double) * b.ngate,
                                  hipMemcpyHostToDevice, b.stream));
  if (b.ngate && h_gate_Vfb)
    TCAD_HIP_CHECK(hipMemcpyAsync(b.d_gate_Vfb, h_gate_Vfb, sizeof(double) * b.ngate,
                                  hipMemcpyHostToDevice, b.stream));
  TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
  return TCAD_OK;
}

int tcad_predict_state(tcad_batch *bp, const double *psi_prev, const double *n_prev,
                       const double *p_prev, double *psi, double *n, double *p,
                       const double *h_ratio) {
  GUARD(bp);
  Batch &b = *bp->impl;
  if (!psi_prev || !n_prev || !p_prev || !psi || !n || !p || !h_ratio) {
    tcad::set_error("tcad_predict_state: null pointer");
    return TCAD_E_INVALID;
  }
  for (int d = 0; d < b.ndev; ++d)
    if (!std::isfinite(h_ratio[d]) || h_ratio[d] > 8.0) {
      tcad::set_error("tcad_predict_state: ratio must be finite and at most 8");
      return TCAD_E_INVALID;
    }
  TCAD_HIP_CHECK(hipMemcpyAsync(b.d_ratio, h_ratio, sizeof(double) * b.ndev,
                                hipMemcpyHostToDevice, b.stream));
  hipLaunchKernelGGL(tcad::k_predict, dim3(b.nblk), dim3(tcad::BS), 0, b.stream, b.d_info,
                     b.d_blk2dev, b.d_ratio, psi_prev, n_prev, p_prev, psi, n, p);
  hipLaunchKernelGGL(tcad::k_mark_predicted, dim3((b.ndev + 63) / 64), dim3(64), 0, b.stream,
                     b.d_scal, b.ndev, b.d_ratio);
  b.launches += 2;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

int tcad_assemble(tcad_batch *bp, const double *psi, const double *n, const double *p, double *F,
                  double *Fnp) {
  GUARD(bp);
  Batch &b = *bp->impl;
  if (!psi || !n || !p || !F) {
    tcad::set_error("tcad_assemble: null state or output pointer");
    return TCAD_E_INVALID;
  }
  int rc = tcad::launch_assemble(b, psi, n, p, F, Fnp, true);
  if (rc) return rc;
  if ((rc = tcad::launch_tile_setup(b, false))) return rc;
  return tcad::check_flag(b);
}

int tcad_export_jacobian(tcad_batch *bp, double *h_blocks, double *h_rowscale, double *h_rhs,
                         int32_t *h_nentries) {
  GUARD(bp);
  Batch &b = *bp->impl;
  const size_t N = (size_t)b.ntot;
  if (h_nentries) *h_nentries = b.ne;
  if (h_blocks)
    TCAD_HIP_CHECK(hipMemcpy(h_blocks, b.A64, sizeof(double) * b.ne * N, hipMemcpyDeviceToHost));
  if (h_rowscale)
    TCAD_HIP_CHECK(hipMemcpy(h_rowscale, b.rowscale, sizeof(double) * 3 * N, hipMemcpyDeviceToHost));
  if (h_rhs) TCAD_HIP_CHECK(hipMemcpy(h_rhs, b.bs, sizeof(double) * 3 * N, hipMemcpyDeviceToHost));
  return TCAD_OK;
}

int tcad_solve_linear(tcad_batch *bp, int32_t mode, double rtol, int32_t maxiter, double *d_du,
                      int32_t *h_iters, double *h_resid) {
  GUARD(bp);
  Batch &b = *bp->impl;
  b.tile = (mode & TCAD_PRECOND_NODE) ? 0 : 1;
  mode &= ~TCAD_PRECOND_NODE;
  if (mode != TCAD_LIN_BICGSTAB_F64 && mode != TCAD_LIN_BICGSTAB_MIXED) {
    tcad::set_error("tcad_solve_linear: unknown mode");
    return TCAD_E_INVALID;
  }
  if (!(rtol > 0.0) || maxiter <= 0) {
    tcad::set_error("tcad_solve_linear: rtol must be > 0 and maxiter > 0");
    return TCAD_E_INVALID;
  }
  if (!b.last_n || !b.last_p) {
    tcad::set_error("tcad_solve_linear: no assembled system");
    return TCAD_E_INVALID;
  }
  int rc = tcad::launch_set_active(b, nullptr);
  if (rc) return rc;
  const bool mixed = mode == TCAD_LIN_BICGSTAB_MIXED;
  rc = tcad::ir_solve(b, mixed, true, rtol, 16, maxiter, mixed ? 1e-6 : rtol);
  if (rc) return rc;
  if ((rc = tcad::sync_scalars(b))) return rc;
  int worst = 0;
  for (int d = 0; d < b.ndev; ++d) {
    const tcad::Scal &S = b.h_scal[d];
    if (h_iters) h_iters[d] = S.kry_total;
    if (h_resid) h_resid[d] = S.bb > 0 ? std::sqrt(S.rr / S.bb) : 0.0;
    if (S.status) worst = 1;
  }
  if (d_du && (rc = tcad::launch_unscale(b, b.last_n, b.last_p, d_du))) return rc;
  TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
  if (worst) {
    tcad::set_error("Krylov solve did not reach the requested tolerance");
    return TCAD_E_NOT_CONVERGED;
  }
  return TCAD_OK;
}

// Per-device TCAD_NEWTON_* outcome after a Newton solve.
static int newton_outcome(const tcad::Scal &S) {
  if (!S.sel || S.newton_done) return TCAD_NEWTON_OK;
  return S.result ? S.result : TCAD_NEWTON_MAX_ITER;
}

int tcad_solve_equilibrium(tcad_batch *bp, int32_t max_iter, double tol_update, double max_dpsi,
                           int32_t from_guess, const int32_t *h_active, double *psi, double *n,
                           double *p, int32_t *h_iters, int32_t *h_converged,
                           int32_t *h_status) {
  GUARD(bp);
  Batch &b = *bp->impl;
  if (!psi || !n || !p || max_iter <= 0) {
    tcad::set_error("tcad_solve_equilibrium: invalid argument");
    return TCAD_E_INVALID;
  }
  int rc = tcad::solve_equilibrium_impl(b, max_iter, tol_update, max_dpsi, from_guess, h_active,
                                        psi, n, p);
  if (rc) return rc;
  if ((rc = tcad::sync_scalars(b))) return rc;
  for (int d = 0; d < b.ndev; ++d) {
    if (h_iters) h_iters[d] = b.h_scal[d].newton_it;
    if (h_converged) h_converged[d] = b.h_scal[d].newton_done;
    if (h_status) h_status[d] = newton_outcome(b.h_scal[d]);
  }
  return TCAD_OK;
}

int tcad_solve_bias(tcad_batch *bp, int32_t max_iter, double tol_update, double max_dpsi,
                    int32_t lin_mode, double lin_rtol, int32_t lin_maxiter,
                    const int32_t *h_active, double *psi, double *n, double *p,
                    int32_t *h_iters, int32_t *h_converged, int32_t *h_krylov,
                    int32_t *h_status) {
  GUARD(bp);
  Batch &b = *bp->impl;
  if (!psi || !n || !p || max_iter <= 0 || lin_maxiter <= 0 || !(lin_rtol > 0.0)) {
    tcad::set_error("tcad_solve_bias: invalid argument");
    return TCAD_E_INVALID;
  }
  const int base_mode = lin_mode & ~(TCAD_LIN_FORCING | TCAD_PRECOND_NODE);
  if (base_mode != TCAD_LIN_BICGSTAB_F64 && base_mode != TCAD_LIN_BICGSTAB_MIXED) {
    tcad::set_error("tcad_solve_bias: unknown linear-solve mode");
    return TCAD_E_INVALID;
  }
  int rc = tcad::solve_bias_impl(b, max_iter, tol_update, max_dpsi, lin_mode, lin_rtol,
                                 lin_maxiter, h_active, psi, n, p);
  if (rc) return rc;
  if ((rc = tcad::sync_scalars(b))) return rc;
  for (int d = 0; d < b.ndev; ++d) {
    if (h_iters) h_iters[d] = b.h_scal[d].newton_it;
    if (h_converged) h_converged[d] = b.h_scal[d].newton_done;
    if (h_krylov) h_krylov[d] = b.h_scal[d].kry_total;
    if (h_status) h_status[d] = newton_outcome(b.h_scal[d]);
  }
  return TCAD_OK;
}

int tcad_solve_queue(tcad_batch *bp, int32_t max_iter, double tol_update, double max_dpsi,
                     int32_t lin_mode, double lin_rtol, int32_t lin_maxiter, int32_t npmax,
                     const int32_t *h_npoints, const double *h_contact_V, const double *h_gate_Vg,
                     double *psi, double *n, double *p, int32_t *h_done, int32_t *h_iters,
                     int32_t *h_krylov, int32_t *h_status, double *h_I) {
  GUARD(bp);
  Batch &b = *bp->impl;
  if (!psi || !n || !p || max_iter <= 0 || lin_maxiter <= 0 || !(lin_rtol > 0.0) ||
      npmax <= 0 || !h_npoints || !h_done || !h_iters || !h_krylov || !h_status || !h_I ||
      (b.ncontact && !h_contact_V) || (b.ngate && !h_gate_Vg)) {
    tcad::set_error("tcad_solve_queue: invalid argument");
    return TCAD_E_INVALID;
  }
  for (int d = 0; d < b.ndev; ++d)
    if (h_npoints[d] < 0 || h_npoints[d] > npmax) {
      tcad::set_error("tcad_solve_queue: npoints[d] must be in [0, npmax]");
      return TCAD_E_INVALID;
    }
  const int base_mode = lin_mode & ~(TCAD_LIN_FORCING | TCAD_PRECOND_NODE);
  if (base_mode != TCAD_LIN_BICGSTAB_F64 && base_mode != TCAD_LIN_BICGSTAB_MIXED) {
    tcad::set_error("tcad_solve_queue: unknown linear-solve mode");
    return TCAD_E_INVALID;
  }
  return tcad::queue_impl(b, max_iter, tol_update, max_dpsi, lin_mode, lin_rtol, lin_maxiter,
                          npmax, h_npoints, h_contact_V, h_gate_Vg, psi, n, p, h_done, h_iters,
                          h_krylov, h_status, h_I);
}

int tcad_transient_step(tcad_batch *bp, int32_t max_iter, double tol_update, double max_dpsi,
                        int32_t lin_mode, double lin_rtol, int32_t lin_maxiter,
                        const int32_t *h_active, const double *h_dt_s, double theta,
                        const double *n_old, const double *p_old, const double *Fn_old,
                        const double *Fp_old, double *psi, double *n, double *p,
                        int32_t *h_iters, int32_t *h_converged, int32_t *h_status) {
  GUARD(bp);
  Batch &b = *bp->impl;
  if (!psi || !n || !p || !n_old || !p_old || !h_dt_s || max_iter <= 0 || lin_maxiter <= 0 ||
      !(lin_rtol > 0.0) || !(theta > 0.0 && theta <= 1.0) ||
      (theta != 1.0 && (!Fn_old || !Fp_old))) {
    tcad::set_error("tcad_transient_step: invalid argument");
    return TCAD_E_INVALID;
  }
  const int base_mode = lin_mode & ~(TCAD_LIN_FORCING | TCAD_PRECOND_NODE);
  if (base_mode != TCAD_LIN_BICGSTAB_F64) {
    // the FP32-matrix path took a different Newton path on a reverse-biased
    // Crank-Nicolson ramp and ended 5e-5 away from the reference's state
    tcad::set_error("tcad_transient_step: only the FP64 linear path is supported");
    return TCAD_E_INVALID;
  }
  for (int d = 0; d < b.ndev; ++d)
    if (!(h_dt_s[d] > 0.0) && (!h_active || h_active[d])) {
      tcad::set_error("tcad_transient_step: every selected device needs dt_s > 0");
      return TCAD_E_INVALID;
    }
  TCAD_HIP_CHECK(hipMemcpyAsync(b.d_dts, h_dt_s, sizeof(double) * b.ndev, hipMemcpyHostToDevice,
                                b.stream));
  tcad::TransientArgs tr{};
  tr.active = 1;
  tr.n_old = n_old;
  tr.p_old = p_old;
  tr.Fn_old = Fn_old;
  tr.Fp_old = Fp_old;
  tr.dt_s = b.d_dts;
  tr.theta = theta;
  int rc = tcad::transient_impl(b, max_iter, tol_update, max_dpsi, lin_mode, lin_rtol,
                                lin_maxiter, h_active, tr, psi, n, p);
  if (rc) return rc;
  if ((rc = tcad::sync_scalars(b))) return rc;
  for (int d = 0; d < b.ndev; ++d) {
    if (h_iters) h_iters[d] = b.h_scal[d].newton_it;
    if (h_converged) h_converged[d] = b.h_scal[d].newton_done;
    if (h_status) h_status[d] = newton_outcome(b.h_scal[d]);
  }
  return TCAD_OK;
}

int tcad_continuity_residuals(tcad_batch *bp, const double *psi, const double *n, const double *p,
                              double *Fn, double *Fp) {
  GUARD(bp);
  Batch &b = *bp->impl;
  if (!psi || !n || !p || !Fn || !Fp) {
    tcad::set_error("tcad_continuity_residuals: null argument");
    return TCAD_E_INVALID;
  }
  tcad::TransientArgs tr{};
  tr.Fn_raw = Fn;
  tr.Fp_raw = Fp;
  int rc = tcad::launch_assemble(b, psi, n, p, b.F, nullptr, false, false, 1, &tr);
  if (rc) return rc;
  TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
  return TCAD_OK;
}

int tcad_residual_jacobian(tcad_batch *bp, const double *psi, const double *n, const double *p,
                           double *F, double *Fn, double *Fp, double *Jraw) {
  GUARD(bp);
  Batch &b = *bp->impl;
  if (!psi || !n || !p || !F || !Fn || !Fp || !Jraw) {
    tcad::set_error("tcad_residual_jacobian: null argument");
    return TCAD_E_INVALID;
  }
  tcad::TransientArgs tr{};
  tr.Fn_raw = Fn;
  tr.Fp_raw = Fp;
  int rc = tcad::launch_assemble(b, psi, n, p, F, nullptr, false, false, 0, &tr, nullptr, Jraw);
  if (rc) return rc;
  TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
  return tcad::check_flag(b);
}

int tcad_terminal_currents(tcad_batch *bp, const double *psi, const double *n, const double *p,
                           double *h_I) {
  GUARD(bp);
  Batch &b = *bp->impl;
  if (!psi || !n || !p || !h_I) {
    tcad::set_error("tcad_terminal_currents: null argument");
    return TCAD_E_INVALID;
  }
  int rc = tcad::launch_assemble(b, psi, n, p, b.F, b.Fnp, false);
  if (rc) return rc;
  double *d_I = b.v64[0];
  if ((rc = tcad::launch_contact_sums(b, b.Fnp, d_I))) return rc;
  if (b.ncontact)
    TCAD_HIP_CHECK(hipMemcpyAsync(h_I, d_I, sizeof(double) * b.ncontact, hipMemcpyDeviceToHost,
                                  b.stream));
  TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
  return TCAD_OK;
}

int tcad_time_kernel(tcad_batch *bp, int32_t which, int32_t reps, double *h_avg_us) {
  GUARD(bp);
  Batch &b = *bp->impl;
  if (which < 0 || which > 4 || reps <= 0 || !h_avg_us) {
    tcad::set_error("tcad_time_kernel: invalid argument");
    return TCAD_E_INVALID;
  }
  if (!b.last_psi) {
    tcad::set_error("tcad_time_kernel: no assembled system");
    return TCAD_E_INVALID;
  }
  int rc = tcad::clear_done(b);
  if (rc) return rc;
  // one untimed launch, then the timed burst
  auto launch = [&]() -> int {
    if (which == 2) return tcad::launch_assemble(b, b.last_psi, b.last_n, b.last_p, b.F, nullptr, true);
    if (which >= 3) return tcad::launch_vector_kernel_only(b, which);
    return tcad::launch_spmv_only(b, which == 1);
  };
  // an untimed burst first (clocks and caches settle), then the timed one:
  // back-to-back dependent launches, as inside the solver
  for (int r = 0; r < reps; ++r)
    if ((rc = launch())) return rc;
  hipEvent_t e0, e1;
  TCAD_HIP_CHECK(hipEventCreate(&e0));
  TCAD_HIP_CHECK(hipEventCreate(&e1));
  TCAD_HIP_CHECK(hipEventRecord(e0, b.stream));
  for (int r = 0; r < reps; ++r)
    if ((rc = launch())) return rc;
  TCAD_HIP_CHECK(hipEventRecord(e1, b.stream));
  TCAD_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  TCAD_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *h_avg_us = 1000.0 * ms / reps;
  return TCAD_OK;
}

int tcad_set_tuning(tcad_batch *bp, int32_t key, double value) {
  GUARD(bp);
  Batch &b = *bp->impl;
  switch (key) {
    case TCAD_TUNE_TILE_ROWS:
      if (value != 0.0 && value != 1.0 && value != 2.0) break;
      b.tile_rows = (int)value;
      return TCAD_OK;
    case TCAD_TUNE_FORCING_SLOPE:
      if (!(value > 0.0 && value <= 1e3)) break;
      b.forcing_c = value;
      return TCAD_OK;
    case TCAD_TUNE_FORCING_MAX:
      if (!(value > 0.0 && value < 1.0)) break;
      b.forcing_max = value;
      return TCAD_OK;
    case TCAD_TUNE_FORCING_FP32:
      if (value != 0.0 && value != 1.0) break;
      b.forcing_fp32 = (int)value;
      return TCAD_OK;
    case TCAD_TUNE_GRAPH:
      if (value != 0.0 && value != 1.0) break;
      b.use_graph = (int)value;
      return TCAD_OK;
    case TCAD_TUNE_IR_CONTINUE:
      if (value != 0.0 && value != 1.0 && value != 2.0) break;
      b.ir_cont = (int)value;
      return TCAD_OK;
    case TCAD_TUNE_FORCING_POW:
      if (value != 1.0 && value != 2.0) break;
      b.forcing_pow = (int)value;
      return TCAD_OK;
    case TCAD_TUNE_LINEAR_FALLBACK:
      if (value != 0.0 && value != 1.0) break;
      b.lin_fallback = (int)value;
      return TCAD_OK;
    case TCAD_TUNE_TILE_LAG:
      if (!(value >= 0.0 && value < 1.0)) break;
      b.tile_lag = value;
      return TCAD_OK;
    case TCAD_TUNE_TILE_RB:
      if (value != 0.0 && value != 1.0) break;
      b.tile_rb = (int)value;
      return TCAD_OK;
    case TCAD_TUNE_FORCING_PREDICTED:
      if (value != 0.0 && value != 1.0) break;
      b.forcing_pred_only = (int)value;
      return TCAD_OK;
    case TCAD_TUNE_FP32_INNER:
      if (!(value >= 1e-7 && value <= 1e-2)) break;
      b.fp32_inner = value;
      return TCAD_OK;
    default:
      break;
  }
  tcad::set_error("tcad_set_tuning: unknown key or value out of range");
  return TCAD_E_INVALID;
}

int tcad_probe_kernel(tcad_batch *bp, int32_t which) {
  GUARD(bp);
  Batch &b = *bp->impl;
  if (which < -1 || which > 4) {
    tcad::set_error("tcad_probe_kernel: which must be -1..4");
    return TCAD_E_INVALID;
  }
  TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
  b.probe = which;
  b.probe_used = 0;
  return TCAD_OK;
}

int tcad_probe_read(tcad_batch *bp, double min_us, int64_t *h_launches, double *h_avg_us) {
  GUARD(bp);
  Batch &b = *bp->impl;
  if (!h_launches || !h_avg_us) {
    tcad::set_error("tcad_probe_read: null output");
    return TCAD_E_INVALID;
  }
  TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
  double total = 0.0;
  int64_t cnt = 0;
  const size_t pairs = b.probe_used / 2;
  for (size_t i = 0; i < pairs; ++i) {
    float ms = 0.f;
    TCAD_HIP_CHECK(hipEventElapsedTime(&ms, b.probe_ev[2 * i], b.probe_ev[2 * i + 1]));
    if (1000.0 * ms >= min_us) {
      total += ms;
      ++cnt;
    }
  }
  *h_launches = cnt;
  *h_avg_us = cnt ? 1000.0 * total / (double)cnt : 0.0;
  b.probe_used = 0;
  return TCAD_OK;
}

int tcad_surface_mobility(tcad_batch *bp, const double *psi) {
  GUARD(bp);
  Batch &b = *bp->impl;
  if (!psi) {
    tcad::set_error("tcad_surface_mobility: null psi");
    return TCAD_E_INVALID;
  }
  int rc = tcad::launch_surface_mobility(b, psi);
  if (rc) return rc;
  TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
  return TCAD_OK;
}

int tcad_edge_currents(tcad_batch *bp, const double *psi, const double *n, const double *p,
                       double *d_J) {
  GUARD(bp);
  Batch &b = *bp->impl;
  if (!psi || !n || !p || !d_J) {
    tcad::set_error("tcad_edge_currents: null argument");
    return TCAD_E_INVALID;
  }
  int rc = tcad::launch_edge_currents(b, psi, n, p, d_J);
  if (rc) return rc;
  TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
  return TCAD_OK;
}

}  // extern "C"
