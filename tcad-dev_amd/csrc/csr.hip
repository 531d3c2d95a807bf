// Iterative solve of a caller-supplied sparse system A x = b in CSR form:
// the GPU counterpart of linsolve.solve_linear's Krylov methods
// (pytcad/pytcad/linsolve.py:725-985) for any matrix, not only the
// stencil-block Newton systems the batch handle assembles itself.
//
//   methods      BiCGStab (right preconditioned) and restarted GMRES(m)
//                (right preconditioned, classical Gram-Schmidt applied
//                twice, Givens rotations on the host), FP64 throughout;
//   precond      node block-Jacobi with blocks of 1 or 3 unknowns
//                (linsolve.py:363-427: the exact inverse of each node's
//                diagonal block), or ILU(0) with level scheduling (rows of
//                one level depend only on rows of earlier levels; one
//                launch per level in the factorisation and in each
//                triangular solve);
//   acceptance   converged when the true relative residual
//                ||b - A x|| / ||b|| is <= max(rtol, 1e-6) at the end, as
//                the reference accepts (linsolve.py:971-982); otherwise
//                TCAD_E_NOT_CONVERGED and x holds the last iterate.
//
// Every reduction runs over a fixed partition in a fixed order, so a
// solve is deterministic.  The Krylov scalars live in device memory and
// are formed by the one-block reduction kernels; the host reads them
// only every `chunk` BiCGStab iterations (and once per GMRES step for
// the small Hessenberg problem).
//
// Two entry points share every kernel: tcad_csr_solve allocates its
// workspace per call; tcad_csr_solve_noalloc runs in `scratch`, a
// file-scope device array of fixed size, for callers that cannot
// allocate.  Roofline: HBM bound; per BiCGStab iteration two SpMVs
// (12 B per stored entry + 16 B per row) and about 20 vector passes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "tcad_internal.h"

namespace tcad {
namespace csr {

constexpr int TB = 256;           // threads per block of the vector kernels
constexpr int NPART = 512;        // blocks (partial sums) of every reduction
constexpr int MAXM = 100;         // GMRES restart length cap of the scratch variant
constexpr int SUB = 8;            // lanes per row in the SpMV

struct Mat {
  int n;
  const int *ptr, *idx;
  const double *val;
};

// device scalars of one solve
struct Sc {
  double bb;                       // ||b||^2
  double rho, rho_old, alpha, omega;
  double rr;                       // ||r||^2 of the current iterate
  double ss;                       // ||s||^2 (early exit of BiCGStab)
  double dots[4];                  // last reductions
  int it, done, status;            // status: 0 ok, 1 breakdown / non-finite
};

__device__ __forceinline__ double block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + o];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// y = A x; SUB lanes per row, a fixed butterfly over them
__global__ void __launch_bounds__(TB) k_spmv(Mat A, const double *__restrict__ x,
                                             double *__restrict__ y) {
  const int lane = threadIdx.x % SUB;
  const int row = blockIdx.x * (TB / SUB) + threadIdx.x / SUB;
  double acc = 0.0;
  if (row < A.n) {
    const int e = A.ptr[row + 1];
    for (int j = A.ptr[row] + lane; j < e; j += SUB) acc = acc + A.val[j] * x[A.idx[j]];
  }
  acc = acc + __shfl_xor(acc, 4, SUB);
  acc = acc + __shfl_xor(acc, 2, SUB);
  acc = acc + __shfl_xor(acc, 1, SUB);
  if (row < A.n && lane == 0) y[row] = acc;
}

// partial sums of up to 4 dot products (a_q, b_q), grid-stride over a
// fixed NPART-block grid; part[q * NPART + block]
struct Dots {
  const double *a[4];
  const double *b[4];
  int nq;
};

__global__ void __launch_bounds__(TB) k_dots(Dots D, int n, double *part) {
  __shared__ double sh[TB];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = blockIdx.x * TB + threadIdx.x; i < n; i += NPART * TB)
    for (int q = 0; q < D.nq; ++q) acc[q] = acc[q] + D.a[q][i] * D.b[q][i];
  for (int q = 0; q < D.nq; ++q) {
    const double s = block_sum(acc[q], sh);
    if (threadIdx.x == 0) part[q * NPART + blockIdx.x] = s;
  }
}

// one block: reduce the partials of nq dots into S.dots[0..nq)
__device__ void reduce_parts(const double *part, int nq, Sc &S, double *sh) {
  for (int q = 0; q < nq; ++q) {
    double v = 0.0;
    for (int i = threadIdx.x; i < NPART; i += TB) v = v + part[q * NPART + i];
    const double s = block_sum(v, sh);
    if (threadIdx.x == 0) S.dots[q] = s;
  }
  __syncthreads();
}

// ---- preconditioners ------------------------------------------------------
// node block-Jacobi setup: Dinv[9 * k + 3 r + c] (bs = 3) or Dinv[k] (bs = 1)
__global__ void k_bj_setup(Mat A, int bs, double *Dinv, int *flag) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k * bs >= A.n) return;
  if (bs == 1) {
    double d = 0.0;
    for (int j = A.ptr[k]; j < A.ptr[k + 1]; ++j)
      if (A.idx[j] == k) d = d + A.val[j];
    if (!(d != 0.0) || !isfinite(d)) atomicOr(flag, 1);
    Dinv[k] = 1.0 / d;
    return;
  }
  double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int r0 = 3 * k;
  for (int r = 0; r < 3; ++r)
    for (int j = A.ptr[r0 + r]; j < A.ptr[r0 + r + 1]; ++j) {
      const int c = A.idx[j] - r0;
      if (c >= 0 && c < 3) a[3 * r + c] = a[3 * r + c] + A.val[j];
    }
  const double c0 = a[4] * a[8] - a[5] * a[7], c1 = a[5] * a[6] - a[3] * a[8],
               c2 = a[3] * a[7] - a[4] * a[6];
  const double det = a[0] * c0 + a[1] * c1 + a[2] * c2;
  if (!(fabs(det) > 0.0) || !isfinite(det)) atomicOr(flag, 1);
  const double id = 1.0 / det;
  double *o = Dinv + 9 * (size_t)k;
  o[0] = c0 * id; o[1] = (a[2] * a[7] - a[1] * a[8]) * id; o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  o[3] = c1 * id; o[4] = (a[0] * a[8] - a[2] * a[6]) * id; o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  o[6] = c2 * id; o[7] = (a[1] * a[6] - a[0] * a[7]) * id; o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

__global__ void k_bj_apply(int n, int bs, const double *Dinv, const double *x, double *y) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k * bs >= n) return;
  if (bs == 1) {
    y[k] = Dinv[k] * x[k];
    return;
  }
  const double *D = Dinv + 9 * (size_t)k;
  const double x0 = x[3 * k], x1 = x[3 * k + 1], x2 = x[3 * k + 2];
  y[3 * k] = D[0] * x0 + D[1] * x1 + D[2] * x2;
  y[3 * k + 1] = D[3] * x0 + D[4] * x1 + D[5] * x2;
  y[3 * k + 2] = D[6] * x0 + D[7] * x1 + D[8] * x2;
}

// ILU(0), IKJ order on the pattern of A (sorted column indices): the rows
// of one level, each by one thread
__global__ void k_ilu_factor_level(int n, const int *ptr, const int *idx, double *lu,
                                   const int *diag, const int *rows, int cnt, int *flag) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  const int e = ptr[i + 1];
  for (int jk = ptr[i]; jk < e && idx[jk] < i; ++jk) {
    const int k = idx[jk];
    const double ukk = lu[diag[k]];
    const double lik = lu[jk] / ukk;
    lu[jk] = lik;
    // a_ij -= l_ik u_kj for j > k present in both rows (both sorted)
    int p = diag[k] + 1;
    const int pe = ptr[k + 1];
    for (int jj = jk + 1; jj < e && p < pe; ++jj) {
      const int j = idx[jj];
      while (p < pe && idx[p] < j) ++p;
      if (p < pe && idx[p] == j) lu[jj] = lu[jj] - lik * lu[p];
    }
  }
  const double d = lu[diag[i]];
  if (!(d != 0.0) || !isfinite(d)) atomicOr(flag, 1);
}

// forward solve with the unit lower factor, one level
__global__ void k_ilu_lower_level(const int *ptr, const int *idx, const double *lu,
                                  const int *rows, int cnt, const double *x, double *y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  double s = x[i];
  for (int j = ptr[i]; j < ptr[i + 1] && idx[j] < i; ++j) s = s - lu[j] * y[idx[j]];
  y[i] = s;
}

// backward solve with the upper factor, one level
__global__ void k_ilu_upper_level(const int *ptr, const int *idx, const double *lu,
                                  const int *diag, const int *rows, int cnt, double *y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  double s = y[i];
  for (int j = diag[i] + 1; j < ptr[i + 1]; ++j) s = s - lu[j] * y[idx[j]];
  y[i] = s / lu[diag[i]];
}

// ---- vector kernels ---------------------------------------------------------
__global__ void k_copy(int n, const double *x, double *y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i];
}

__global__ void k_fill(int n, double v, double *y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

// r = b - r (r holds A x)
__global__ void k_residual(int n, const double *b, double *r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - r[i];
}

// ---- BiCGStab (right preconditioned), scalars on the device ----------------
// setup: S.bb = (b,b), S.rr = (r,r), S.rho = (rhat, r) with rhat = r
__global__ void __launch_bounds__(TB) k_bc_start(const double *part, Sc *S, double tol2, int maxit) {
  __shared__ double sh[TB];
  reduce_parts(part, 2, *S, sh);
  if (threadIdx.x == 0) {
    S->bb = S->dots[0];
    S->rr = S->dots[1];
    S->rho = S->rr;
    S->rho_old = 1.0;
    S->alpha = 1.0;
    S->omega = 1.0;
    S->it = 0;
    S->status = 0;
    const double bb = S->bb > 0.0 ? S->bb : 1e-300;
    S->done = (S->rr <= tol2 * bb || !(S->bb > 0.0)) ? 1 : 0;
    if (!isfinite(S->rr)) { S->status = 1; S->done = 1; }
    (void)maxit;
  }
}

// p = r + beta (p - omega v), first iteration p = r
__global__ void k_bc_p(int n, const Sc *S, const double *r, const double *v, double *p) {
  if (S->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (S->it == 0) {
    p[i] = r[i];
  } else {
    const double beta = (S->rho / S->rho_old) * (S->alpha / S->omega);
    p[i] = r[i] + beta * (p[i] - S->omega * v[i]);
  }
}

// alpha = rho / (rhat, v)
__global__ void __launch_bounds__(TB) k_bc_alpha(const double *part, Sc *S) {
  __shared__ double sh[TB];
  if (S->done) return;
  reduce_parts(part, 1, *S, sh);
  if (threadIdx.x == 0) {
    const double rv = S->dots[0];
    const double a = S->rho / rv;
    if (!(rv != 0.0) || !isfinite(a)) { S->status = 1; S->done = 1; }
    else S->alpha = a;
  }
}

// s = r - alpha v (stored in r)
__global__ void k_bc_s(int n, const Sc *S, const double *v, double *r) {
  if (S->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = r[i] - S->alpha * v[i];
}

// omega = (t, s) / (t, t); also (s, s)
__global__ void __launch_bounds__(TB) k_bc_omega(const double *part, Sc *S, double tol2) {
  __shared__ double sh[TB];
  if (S->done) return;
  reduce_parts(part, 3, *S, sh);
  if (threadIdx.x == 0) {
    const double ts = S->dots[0], tt = S->dots[1], ss = S->dots[2];
    S->ss = ss;
    if (ss <= tol2 * S->bb) {
      S->omega = 0.0;            // x += alpha phat only; r = s
    } else {
      const double w = ts / tt;
      if (!(tt != 0.0) || !isfinite(w) || w == 0.0) { S->status = 1; S->done = 1; }
      else S->omega = w;
    }
  }
}

// x += alpha phat + omega shat; r = s - omega t
__global__ void k_bc_update(int n, const Sc *S, const double *ph, const double *sh_,
                            const double *t, double *x, double *r) {
  if (S->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = x[i] + S->alpha * ph[i] + S->omega * sh_[i];
  r[i] = r[i] - S->omega * t[i];
}

// rr = (r, r), rho_new = (rhat, r): stop test, next iteration's scalars
__global__ void __launch_bounds__(TB) k_bc_next(const double *part, Sc *S, double tol2, int maxit) {
  __shared__ double sh[TB];
  if (S->done) return;
  reduce_parts(part, 2, *S, sh);
  if (threadIdx.x == 0) {
    S->rr = S->dots[0];
    S->rho_old = S->rho;
    S->rho = S->dots[1];
    S->it = S->it + 1;
    if (!isfinite(S->rr)) { S->status = 1; S->done = 1; return; }
    if (S->rr <= tol2 * S->bb || S->omega == 0.0) { S->done = 1; return; }
    if (S->it >= maxit) { S->done = 1; return; }
    if (S->rho == 0.0) { S->status = 1; S->done = 1; }
  }
}

// ---- GMRES helpers ------------------------------------------------------------
// partial sums of (V_i, w) for i < k: grid (NPART_G, k)
constexpr int NPART_G = 64;
__global__ void __launch_bounds__(TB) k_mdot(const double *V, size_t ld, int n, const double *w,
                                             double *part) {
  __shared__ double sh[TB];
  const double *v = V + (size_t)blockIdx.y * ld;
  double acc = 0.0;
  for (int i = blockIdx.x * TB + threadIdx.x; i < n; i += NPART_G * TB) acc = acc + v[i] * w[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.y * NPART_G + blockIdx.x] = s;
}

__global__ void __launch_bounds__(TB) k_mdot_reduce(const double *part, double *h) {
  __shared__ double sh[TB];
  double v = 0.0;
  for (int i = threadIdx.x; i < NPART_G; i += TB) v = v + part[blockIdx.x * NPART_G + i];
  const double s = block_sum(v, sh);
  if (threadIdx.x == 0) h[blockIdx.x] = s;
}

// w -= sum_i h_i V_i (i < k, in order)
__global__ void k_msub(const double *V, size_t ld, int k, int n, const double *h, double *w) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = w[r];
  for (int i = 0; i < k; ++i) s = s - h[i] * V[(size_t)i * ld + r];
  w[r] = s;
}

// u = sum_i y_i V_i (i < k)
__global__ void k_mcomb(const double *V, size_t ld, int k, int n, const double *y, double *u) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int i = 0; i < k; ++i) s = s + y[i] * V[(size_t)i * ld + r];
  u[r] = s;
}

// out = w / sqrt(nrm2[0])
__global__ void k_scale(int n, const double *nrm2, double *w, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = w[i] / sqrt(nrm2[0]);
}

__global__ void k_axpy_dev(int n, double a, const double *x, double *y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = y[i] + a * x[i];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
inline dim3 grid(long long n, int t = TB) { return dim3((unsigned)((n + t - 1) / t)); }

struct Work {
  // vectors of n (BiCGStab: r rhat p v phat shat t x0(unused); GMRES uses V)
  double *r, *rhat, *p, *v, *ph, *shat, *t, *tmp;
  double *Dinv;                    // node block inverses
  double *part;                    // NPART * 4 partials (or NPART_G * (m + 1))
  double *h;                       // m + 2 reduced dots (GMRES)
  double *V;                       // (m + 1) vectors of n (GMRES)
  Sc *S;
  int *flag;
  // ILU(0)
  double *lu;
  int *diag, *lrows, *urows;
  std::vector<int> lofs, uofs;     // level offsets (host)
};

// bytes of scratch a solve needs (8-byte words), BiCGStab or GMRES(m)
static size_t words_needed(int n, int bs, int method, int m) {
  size_t w = 0;
  if (method == TCAD_CSR_BICGSTAB) w += 8 * (size_t)n;
  else w += (size_t)(m + 1) * n + 3 * (size_t)n + (size_t)(m + 2);
  w += (bs == 3 ? 3 * (size_t)n : (size_t)n);
  w += (size_t)std::max(4 * NPART, NPART_G * (m + 2));
  w += 64;                          // Sc + flag, padded
  return w;
}

// carve a workspace out of one base pointer (256-byte aligned pieces)
static void carve(double *base, int n, int bs, int method, int m, Work &W) {
  auto take = [&](size_t words) {
    double *q = base;
    base += (words + 31) / 32 * 32;
    return q;
  };
  W.r = take(n);
  W.tmp = take(n);
  W.ph = take(n);
  if (method == TCAD_CSR_BICGSTAB) {
    W.rhat = take(n); W.p = take(n); W.v = take(n); W.shat = take(n); W.t = take(n);
    W.V = nullptr; W.h = nullptr;
  } else {
    W.V = take((size_t)(m + 1) * n);
    W.h = take(m + 2);
    W.rhat = W.p = W.v = W.shat = W.t = nullptr;
  }
  W.Dinv = take(bs == 3 ? 3 * (size_t)n : (size_t)n);
  W.part = take((size_t)std::max(4 * NPART, NPART_G * (m + 2)));
  W.S = (Sc *)take(sizeof(Sc) / 8 + 1);
  W.flag = (int *)take(1);
}

struct Solve {
  Mat A;
  int n, bs, method, prec;         // prec: 0 block-Jacobi, 1 ILU(0)
  hipStream_t st;
  Work W;
};

static int check_flag_clear(Solve &s, const char *what) {
  int f = 0;
  TCAD_HIP_CHECK(hipMemcpyAsync(&f, s.W.flag, sizeof(int), hipMemcpyDeviceToHost, s.st));
  TCAD_HIP_CHECK(hipStreamSynchronize(s.st));
  if (f) {
    set_error(std::string(what) + ": singular or non-finite diagonal block / pivot");
    return TCAD_E_NONFINITE;
  }
  return TCAD_OK;
}

static void apply_prec(Solve &s, const double *x, double *y) {
  if (s.prec == 0) {
    const int nb = s.n / s.bs;
    hipLaunchKernelGGL(k_bj_apply, grid(nb), dim3(TB), 0, s.st, s.n, s.bs, s.W.Dinv, x, y);
    return;
  }
  const Work &W = s.W;
  for (size_t l = 0; l + 1 < W.lofs.size(); ++l) {
    const int cnt = W.lofs[l + 1] - W.lofs[l];
    hipLaunchKernelGGL(k_ilu_lower_level, grid(cnt), dim3(TB), 0, s.st, s.A.ptr, s.A.idx, W.lu,
                       W.lrows + W.lofs[l], cnt, x, y);
  }
  for (size_t l = 0; l + 1 < W.uofs.size(); ++l) {
    const int cnt = W.uofs[l + 1] - W.uofs[l];
    hipLaunchKernelGGL(k_ilu_upper_level, grid(cnt), dim3(TB), 0, s.st, s.A.ptr, s.A.idx, W.lu,
                       W.diag, W.urows + W.uofs[l], cnt, y);
  }
}

static void dots(Solve &s, int nq, const double *a0, const double *b0, const double *a1 = nullptr,
                 const double *b1 = nullptr, const double *a2 = nullptr,
                 const double *b2 = nullptr) {
  Dots D{};
  D.nq = nq;
  D.a[0] = a0; D.b[0] = b0; D.a[1] = a1; D.b[1] = b1; D.a[2] = a2; D.b[2] = b2;
  hipLaunchKernelGGL(k_dots, dim3(NPART), dim3(TB), 0, s.st, D, s.n, s.W.part);
}

__global__ void k_zero_int(int *f) { *f = 0; }

static int setup_prec(Solve &s) {
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s.st, s.W.flag);
  if (s.prec == 0) {
    const int nb = s.n / s.bs;
    hipLaunchKernelGGL(k_bj_setup, grid(nb), dim3(TB), 0, s.st, s.A, s.bs, s.W.Dinv, s.W.flag);
    TCAD_HIP_CHECK(hipGetLastError());
    return check_flag_clear(s, "block-Jacobi setup");
  }
  const Work &W = s.W;
  for (size_t l = 0; l + 1 < W.lofs.size(); ++l) {
    const int cnt = W.lofs[l + 1] - W.lofs[l];
    hipLaunchKernelGGL(k_ilu_factor_level, grid(cnt), dim3(TB), 0, s.st, s.n, s.A.ptr, s.A.idx,
                       W.lu, W.diag, W.lrows + W.lofs[l], cnt, W.flag);
  }
  TCAD_HIP_CHECK(hipGetLastError());
  return check_flag_clear(s, "ILU(0) factorisation");
}

static int bicgstab(Solve &s, const double *b, double *x, double rtol, int maxit, int32_t *iters,
                    double *resid) {
  Work &W = s.W;
  const int n = s.n;
  const double tol2 = rtol * rtol;
  // r = b - A x, rhat = r
  hipLaunchKernelGGL(k_spmv, grid((long long)n * SUB), dim3(TB), 0, s.st, s.A, x, W.r);
  hipLaunchKernelGGL(k_residual, grid(n), dim3(TB), 0, s.st, n, b, W.r);
  hipLaunchKernelGGL(k_copy, grid(n), dim3(TB), 0, s.st, n, W.r, W.rhat);
  hipLaunchKernelGGL(k_fill, grid(n), dim3(TB), 0, s.st, n, 0.0, W.v);
  dots(s, 2, b, b, W.r, W.r);
  hipLaunchKernelGGL(k_bc_start, dim3(1), dim3(TB), 0, s.st, W.part, W.S, tol2, maxit);
  Sc hs{};
  int it = 0;
  const int chunk = 16;
  while (true) {
    for (int c = 0; c < chunk && it < maxit; ++c, ++it) {
      hipLaunchKernelGGL(k_bc_p, grid(n), dim3(TB), 0, s.st, n, W.S, W.r, W.v, W.p);
      apply_prec(s, W.p, W.ph);
      hipLaunchKernelGGL(k_spmv, grid((long long)n * SUB), dim3(TB), 0, s.st, s.A, W.ph, W.v);
      dots(s, 1, W.rhat, W.v);
      hipLaunchKernelGGL(k_bc_alpha, dim3(1), dim3(TB), 0, s.st, W.part, W.S);
      hipLaunchKernelGGL(k_bc_s, grid(n), dim3(TB), 0, s.st, n, W.S, W.v, W.r);   // r := s
      apply_prec(s, W.r, W.shat);
      hipLaunchKernelGGL(k_spmv, grid((long long)n * SUB), dim3(TB), 0, s.st, s.A, W.shat, W.t);
      dots(s, 3, W.t, W.r, W.t, W.t, W.r, W.r);
      hipLaunchKernelGGL(k_bc_omega, dim3(1), dim3(TB), 0, s.st, W.part, W.S, tol2);
      hipLaunchKernelGGL(k_bc_update, grid(n), dim3(TB), 0, s.st, n, W.S, W.ph, W.shat, W.t, x,
                         W.r);
      dots(s, 2, W.r, W.r, W.rhat, W.r);
      hipLaunchKernelGGL(k_bc_next, dim3(1), dim3(TB), 0, s.st, W.part, W.S, tol2, maxit);
    }
    TCAD_HIP_CHECK(hipGetLastError());
    TCAD_HIP_CHECK(hipMemcpyAsync(&hs, W.S, sizeof(Sc), hipMemcpyDeviceToHost, s.st));
    TCAD_HIP_CHECK(hipStreamSynchronize(s.st));
    if (hs.done || it >= maxit) break;
  }
  *iters = hs.it;
  // the true residual decides acceptance
  hipLaunchKernelGGL(k_spmv, grid((long long)n * SUB), dim3(TB), 0, s.st, s.A, x, W.tmp);
  hipLaunchKernelGGL(k_residual, grid(n), dim3(TB), 0, s.st, n, b, W.tmp);
  dots(s, 2, b, b, W.tmp, W.tmp);                 // S.bb = (b, b), S.rr = (r, r)
  hipLaunchKernelGGL(k_bc_start, dim3(1), dim3(TB), 0, s.st, W.part, W.S, tol2, maxit);
  TCAD_HIP_CHECK(hipMemcpyAsync(&hs, W.S, sizeof(Sc), hipMemcpyDeviceToHost, s.st));
  TCAD_HIP_CHECK(hipStreamSynchronize(s.st));
  *resid = hs.bb > 0.0 ? std::sqrt(hs.rr) / std::sqrt(hs.bb) : std::sqrt(hs.rr);
  return TCAD_OK;
}

// Givens rotation (c, s) zeroing b in (a, b)
static void givens(double a, double b, double &c, double &s) {
  if (b == 0.0) { c = 1.0; s = 0.0; return; }
  const double r = std::hypot(a, b);
  c = a / r;
  s = b / r;
}

static int gmres(Solve &s, const double *b, double *x, double rtol, int maxcycles, int m,
                 int32_t *iters, double *resid) {
  Work &W = s.W;
  const int n = s.n;
  const size_t ld = (size_t)n;
  double H[(MAXM + 1) * MAXM];     // column-major, H[i + j (m + 1)]
  double cs[MAXM], sn[MAXM], g[MAXM + 1], y[MAXM], hv[MAXM + 2];
  int total = 0;
  double bnorm = 0.0;
  {
    hipLaunchKernelGGL(k_mdot, dim3(NPART_G, 1), dim3(TB), 0, s.st, b, ld, n, b, W.part);
    hipLaunchKernelGGL(k_mdot_reduce, dim3(1), dim3(TB), 0, s.st, W.part, W.h);
    TCAD_HIP_CHECK(hipMemcpyAsync(hv, W.h, sizeof(double), hipMemcpyDeviceToHost, s.st));
    TCAD_HIP_CHECK(hipStreamSynchronize(s.st));
    bnorm = std::sqrt(hv[0]);
  }
  const double tol = rtol * (bnorm > 0.0 ? bnorm : 1.0);
  double rn = 0.0;
  for (int cyc = 0; cyc < maxcycles; ++cyc) {
    // v0 = (b - A x) / ||.||
    hipLaunchKernelGGL(k_spmv, grid((long long)n * SUB), dim3(TB), 0, s.st, s.A, x, W.r);
    hipLaunchKernelGGL(k_residual, grid(n), dim3(TB), 0, s.st, n, b, W.r);
    hipLaunchKernelGGL(k_mdot, dim3(NPART_G, 1), dim3(TB), 0, s.st, W.r, ld, n, W.r, W.part);
    hipLaunchKernelGGL(k_mdot_reduce, dim3(1), dim3(TB), 0, s.st, W.part, W.h);
    TCAD_HIP_CHECK(hipMemcpyAsync(hv, W.h, sizeof(double), hipMemcpyDeviceToHost, s.st));
    TCAD_HIP_CHECK(hipStreamSynchronize(s.st));
    rn = std::sqrt(hv[0]);
    if (!std::isfinite(rn)) break;
    if (rn <= tol) break;
    hipLaunchKernelGGL(k_scale, grid(n), dim3(TB), 0, s.st, n, W.h, W.r, W.V);
    g[0] = rn;
    for (int i = 1; i <= m; ++i) g[i] = 0.0;
    int j = 0;
    bool stop = false;
    for (; j < m; ++j) {
      double *vj = W.V + (size_t)j * ld;
      double *w = W.V + (size_t)(j + 1) * ld;
      apply_prec(s, vj, W.ph);
      hipLaunchKernelGGL(k_spmv, grid((long long)n * SUB), dim3(TB), 0, s.st, s.A, W.ph, w);
      // classical Gram-Schmidt, twice
      double h[MAXM + 1];
      for (int i = 0; i <= j; ++i) h[i] = 0.0;
      for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_mdot, dim3(NPART_G, j + 1), dim3(TB), 0, s.st, W.V, ld, n, w, W.part);
        hipLaunchKernelGGL(k_mdot_reduce, dim3(j + 1), dim3(TB), 0, s.st, W.part, W.h);
        hipLaunchKernelGGL(k_msub, grid(n), dim3(TB), 0, s.st, W.V, ld, j + 1, n, W.h, w);
        TCAD_HIP_CHECK(hipMemcpyAsync(hv, W.h, sizeof(double) * (j + 1), hipMemcpyDeviceToHost,
                                      s.st));
        TCAD_HIP_CHECK(hipStreamSynchronize(s.st));
        for (int i = 0; i <= j; ++i) h[i] = h[i] + hv[i];
      }
      hipLaunchKernelGGL(k_mdot, dim3(NPART_G, 1), dim3(TB), 0, s.st, w, ld, n, w, W.part);
      hipLaunchKernelGGL(k_mdot_reduce, dim3(1), dim3(TB), 0, s.st, W.part, W.h);
      TCAD_HIP_CHECK(hipMemcpyAsync(hv, W.h, sizeof(double), hipMemcpyDeviceToHost, s.st));
      TCAD_HIP_CHECK(hipStreamSynchronize(s.st));
      const double hn = std::sqrt(hv[0]);
      h[j + 1] = hn;
      if (hn > 0.0 && std::isfinite(hn))
        hipLaunchKernelGGL(k_scale, grid(n), dim3(TB), 0, s.st, n, W.h, w, w);
      for (int i = 0; i < j; ++i) {
        const double a = cs[i] * h[i] + sn[i] * h[i + 1];
        h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
        h[i] = a;
      }
      givens(h[j], h[j + 1], cs[j], sn[j]);
      h[j] = cs[j] * h[j] + sn[j] * h[j + 1];
      h[j + 1] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      for (int i = 0; i <= j; ++i) H[i + j * (m + 1)] = h[i];
      ++total;
      if (!std::isfinite(g[j + 1]) || std::fabs(g[j + 1]) <= tol || !(hn > 0.0)) {
        stop = true;
        ++j;
        break;
      }
    }
    // y = H^-1 g on the j columns; x += M (V y)
    const int k = j;
    for (int i = k - 1; i >= 0; --i) {
      double t = g[i];
      for (int l = i + 1; l < k; ++l) t -= H[i + l * (m + 1)] * y[l];
      y[i] = t / H[i + i * (m + 1)];
    }
    TCAD_HIP_CHECK(hipMemcpyAsync(W.h, y, sizeof(double) * k, hipMemcpyHostToDevice, s.st));
    hipLaunchKernelGGL(k_mcomb, grid(n), dim3(TB), 0, s.st, W.V, ld, k, n, W.h, W.tmp);
    apply_prec(s, W.tmp, W.ph);
    hipLaunchKernelGGL(k_axpy_dev, grid(n), dim3(TB), 0, s.st, n, 1.0, W.ph, x);
    TCAD_HIP_CHECK(hipStreamSynchronize(s.st));
    if (stop && !(std::fabs(g[k]) > tol)) {
      // the estimate reached the target: confirm on the true residual
      // at the top of the next cycle (which then ends the loop)
    }
    if (!std::isfinite(g[k])) break;
  }
  *iters = total;
  hipLaunchKernelGGL(k_spmv, grid((long long)n * SUB), dim3(TB), 0, s.st, s.A, x, W.r);
  hipLaunchKernelGGL(k_residual, grid(n), dim3(TB), 0, s.st, n, b, W.r);
  hipLaunchKernelGGL(k_mdot, dim3(NPART_G, 1), dim3(TB), 0, s.st, W.r, ld, n, W.r, W.part);
  hipLaunchKernelGGL(k_mdot_reduce, dim3(1), dim3(TB), 0, s.st, W.part, W.h);
  TCAD_HIP_CHECK(hipMemcpyAsync(hv, W.h, sizeof(double), hipMemcpyDeviceToHost, s.st));
  TCAD_HIP_CHECK(hipStreamSynchronize(s.st));
  *resid = std::sqrt(hv[0]) / (bnorm > 0.0 ? bnorm : 1.0);
  return TCAD_OK;
}


// ILU(0) structure on the host: diagonal positions, level sets of the
// lower (forward) and upper (backward) dependency graphs
static int ilu_levels(int n, const std::vector<int> &ptr, const std::vector<int> &idx,
                      std::vector<int> &diag, std::vector<int> &lrows, std::vector<int> &lofs,
                      std::vector<int> &urows, std::vector<int> &uofs) {
  diag.assign(n, -1);
  std::vector<int> lev(n, 0);
  int lmax = 0;
  for (int i = 0; i < n; ++i) {
    int l = 0;
    for (int j = ptr[i]; j < ptr[i + 1]; ++j) {
      if (j > ptr[i] && idx[j] <= idx[j - 1]) {
        set_error("tcad_csr_solve: ILU(0) needs sorted, duplicate-free column indices");
        return TCAD_E_INVALID;
      }
      if (idx[j] < 0 || idx[j] >= n) {
        set_error("tcad_csr_solve: column index out of range");
        return TCAD_E_INDEX;
      }
      if (idx[j] < i) l = std::max(l, lev[idx[j]] + 1);
      if (idx[j] == i) diag[i] = j;
    }
    if (diag[i] < 0) {
      set_error("tcad_csr_solve: ILU(0) needs every diagonal entry stored (row " +
                std::to_string(i) + ")");
      return TCAD_E_INVALID;
    }
    lev[i] = l;
    lmax = std::max(lmax, l);
  }
  auto bucket = [&](const std::vector<int> &lv, int nl, std::vector<int> &rows,
                    std::vector<int> &ofs) {
    ofs.assign(nl + 2, 0);
    for (int i = 0; i < n; ++i) ofs[lv[i] + 1]++;
    for (int l = 0; l <= nl; ++l) ofs[l + 1] += ofs[l];
    rows.assign(n, 0);
    std::vector<int> pos(ofs.begin(), ofs.end() - 1);
    for (int i = 0; i < n; ++i) rows[pos[lv[i]]++] = i;
  };
  bucket(lev, lmax, lrows, lofs);
  std::vector<int> uev(n, 0);
  int umax = 0;
  for (int i = n - 1; i >= 0; --i) {
    int l = 0;
    for (int j = diag[i] + 1; j < ptr[i + 1]; ++j) l = std::max(l, uev[idx[j]] + 1);
    uev[i] = l;
    umax = std::max(umax, l);
  }
  bucket(uev, umax, urows, uofs);
  return TCAD_OK;
}

static int check_args(int n, const int *ptr, const int *idx, const double *val, const double *b,
                      double *x, int method, int prec, int bs, double rtol, int maxiter,
                      const char *who) {
  if (n <= 0 || !ptr || !idx || !val || !b || !x) {
    set_error(std::string(who) + ": n must be > 0 and every array non-null");
    return TCAD_E_INVALID;
  }
  if (method != TCAD_CSR_BICGSTAB && method != TCAD_CSR_GMRES) {
    set_error(std::string(who) + ": method must be TCAD_CSR_BICGSTAB or TCAD_CSR_GMRES");
    return TCAD_E_INVALID;
  }
  if (prec != TCAD_CSR_PREC_BLOCK_JACOBI && prec != TCAD_CSR_PREC_ILU0) {
    set_error(std::string(who) + ": unknown preconditioner");
    return TCAD_E_INVALID;
  }
  if (prec == TCAD_CSR_PREC_BLOCK_JACOBI && ((bs != 1 && bs != 3) || n % bs != 0)) {
    set_error(std::string(who) + ": block_size must be 1 or 3 and divide n");
    return TCAD_E_INVALID;
  }
  if (!(rtol > 0.0) || maxiter <= 0) {
    set_error(std::string(who) + ": rtol must be > 0 and maxiter > 0");
    return TCAD_E_INVALID;
  }
  return TCAD_OK;
}

static int run(Solve &s, const double *b, double *x, double rtol, int maxiter, int m,
               int32_t *h_iters, double *h_resid) {
  int rc = setup_prec(s);
  if (rc) return rc;
  int32_t its = 0;
  double res = 0.0;
  if (s.method == TCAD_CSR_BICGSTAB) rc = bicgstab(s, b, x, rtol, maxiter, &its, &res);
  else rc = gmres(s, b, x, rtol, maxiter, m, &its, &res);
  if (rc) return rc;
  if (h_iters) *h_iters = its;
  if (h_resid) *h_resid = res;
  if (!std::isfinite(res) || res > std::max(rtol, 1e-6)) {
    set_error("Krylov solve did not converge: relative residual " + std::to_string(res) +
              " after " + std::to_string(its) + " iterations (target " + std::to_string(rtol) +
              ")");
    return TCAD_E_NOT_CONVERGED;
  }
  return TCAD_OK;
}

// ---- the no-allocation variant's device scratch ----------------------------
// 64 MiB of FP64 words on every GPU the library runs on, held until the
// process ends.  Calls on one GPU share it: overlapping calls race.
constexpr size_t SCRATCH_WORDS = size_t(8) << 20;
static __device__ __align__(256) double scratch[SCRATCH_WORDS];
constexpr int MAX_DEVICES = 64;
static double *g_scratch[MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// the device address of `scratch` on the current device, looked up once per
// device (the host-side symbol is no device address)
static int scratch_ptr(double **out) {
  int dev = 0;
  TCAD_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) {
    set_error("tcad_csr_solve_noalloc: device index out of range");
    return TCAD_E_INVALID;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("tcad_csr_solve_noalloc: hipGetSymbolAddress failed: ") +
                hipGetErrorString(e));
      return TCAD_E_HIP;
    }
    g_scratch[dev] = (double *)p;
  }
  *out = g_scratch[dev];
  return TCAD_OK;
}

}  // namespace csr
}  // namespace tcad

extern "C" {

int tcad_csr_solve(int32_t n, const int32_t *d_indptr, const int32_t *d_indices,
                   const double *d_data, const double *d_b, double *d_x, int32_t method,
                   int32_t precond, int32_t block_size, double rtol, int32_t maxiter,
                   int32_t restart, void *stream, int32_t *h_iters, double *h_resid) {
  using namespace tcad;
  using namespace tcad::csr;
  int rc = check_args(n, d_indptr, d_indices, d_data, d_b, d_x, method, precond, block_size,
                      rtol, maxiter, "tcad_csr_solve");
  if (rc) return rc;
  const int m = method == TCAD_CSR_GMRES ? std::min(restart > 0 ? restart : 100, n) : 0;
  if (method == TCAD_CSR_GMRES && m > MAXM) {
    set_error("tcad_csr_solve: GMRES restart length is limited to 100");
    return TCAD_E_INVALID;
  }
  Solve s{};
  s.A = Mat{n, d_indptr, d_indices, d_data};
  s.n = n;
  s.bs = precond == TCAD_CSR_PREC_BLOCK_JACOBI ? block_size : 1;
  s.method = method;
  s.prec = precond == TCAD_CSR_PREC_ILU0 ? 1 : 0;
  s.st = (hipStream_t)stream;
  const size_t words = words_needed(n, s.bs, method, m) + 32 * 16;
  int nnz = 0;
  std::vector<int> hptr, hidx, diag, lrows, urows;
  if (s.prec == 1) {
    hptr.resize(n + 1);
    TCAD_HIP_CHECK(hipMemcpy(hptr.data(), d_indptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    nnz = hptr[n];
    hidx.resize(nnz);
    TCAD_HIP_CHECK(hipMemcpy(hidx.data(), d_indices, sizeof(int) * nnz, hipMemcpyDeviceToHost));
    if ((rc = ilu_levels(n, hptr, hidx, diag, lrows, s.W.lofs, urows, s.W.uofs))) return rc;
  }
  double *base = nullptr;
  const size_t ilu_words = s.prec == 1 ? (size_t)nnz + 32 + (3 * (size_t)n + 1) / 2 + 64 : 0;
  hipError_t e = hipMalloc(&base, sizeof(double) * (words + ilu_words));
  if (e != hipSuccess) {
    set_error(std::string("tcad_csr_solve: hipMalloc failed: ") + hipGetErrorString(e));
    return TCAD_E_HIP;
  }
  carve(base, n, s.bs, method, m, s.W);
  if (s.prec == 1) {
    double *q = base + words;
    s.W.lu = q;
    int *iq = (int *)(q + nnz + 32);
    s.W.diag = iq;
    s.W.lrows = iq + n;
    s.W.urows = iq + 2 * n;
    e = hipMemcpyAsync(s.W.lu, d_data, sizeof(double) * nnz, hipMemcpyDeviceToDevice, s.st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(s.W.diag, diag.data(), sizeof(int) * n, hipMemcpyHostToDevice, s.st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(s.W.lrows, lrows.data(), sizeof(int) * n, hipMemcpyHostToDevice, s.st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(s.W.urows, urows.data(), sizeof(int) * n, hipMemcpyHostToDevice, s.st);
    if (e == hipSuccess) e = hipStreamSynchronize(s.st);
    if (e != hipSuccess) {
      (void)hipFree(base);
      set_error(std::string("tcad_csr_solve: ") + hipGetErrorString(e));
      return TCAD_E_HIP;
    }
  }
  rc = run(s, d_b, d_x, rtol, maxiter, m, h_iters, h_resid);
  (void)hipStreamSynchronize(s.st);
  (void)hipFree(base);
  return rc;
}

int tcad_csr_solve_noalloc(int32_t n, const int32_t *d_indptr, const int32_t *d_indices,
                           const double *d_data, const double *d_b, double *d_x, int32_t method,
                           int32_t precond, int32_t block_size, double rtol, int32_t maxiter,
                           int32_t restart, void *stream, int32_t *h_iters, double *h_resid) {
  using namespace tcad;
  using namespace tcad::csr;
  int rc = check_args(n, d_indptr, d_indices, d_data, d_b, d_x, method, precond, block_size,
                      rtol, maxiter, "tcad_csr_solve_noalloc");
  if (rc) return rc;
  if (precond != TCAD_CSR_PREC_BLOCK_JACOBI) {
    set_error("tcad_csr_solve_noalloc: only block-Jacobi runs without allocation (ILU(0) "
              "needs tcad_csr_solve)");
    return TCAD_E_INVALID;
  }
  const int m = method == TCAD_CSR_GMRES ? std::min(restart > 0 ? restart : 100, n) : 0;
  if (method == TCAD_CSR_GMRES && m > MAXM) {
    set_error("tcad_csr_solve_noalloc: GMRES restart length is limited to 100");
    return TCAD_E_INVALID;
  }
  const size_t words = words_needed(n, block_size, method, m) + 32 * 16;
  if (words > SCRATCH_WORDS) {
    set_error("tcad_csr_solve_noalloc: the system needs " + std::to_string(words * 8) +
              " bytes of workspace; the fixed scratch holds " +
              std::to_string(SCRATCH_WORDS * 8) + " (use tcad_csr_solve)");
    return TCAD_E_INVALID;
  }
  double *base = nullptr;
  if ((rc = scratch_ptr(&base))) return rc;
  Solve s{};
  s.A = Mat{n, d_indptr, d_indices, d_data};
  s.n = n;
  s.bs = block_size;
  s.method = method;
  s.prec = 0;
  s.st = (hipStream_t)stream;
  carve(base, n, s.bs, method, m, s.W);
  return run(s, d_b, d_x, rtol, maxiter, m, h_iters, h_resid);
}

}  // extern "C"
