// Small-signal AC admittance: the N-port Y(f) of a converged DC operating
// point, pytcad/pytcad/ac2d.py:265-342 and ac3d.py:187-277.
//
// Every contact and gate is a port.  For port k and frequency f the
// complex linear system (J + j w_s C) du = b_k is solved, J the Newton
// Jacobian at the operating point (Dirichlet rows at ohmic contact nodes),
// C the storage matrix (-dV on electron rows, +dV on hole rows, zero at
// contact nodes; ac2d.py:107-124), w_s = 2 pi f Ns / R0 (ac2d.py:79-83).
// An ohmic port drives du_psi = 1 on its nodes; a gate port forces
// -kappa w on its nodes' potential rows (ac2d.py:296-313).  Row i of Y:
// an ohmic contact observes the raw continuity rows of its nodes,
// S_i . du (ac2d.py:168-194); a gate observes its displacement current
// j w_s sum kappa w (delta_ik - du_psi) (ac2d.py:320-333).  Y is returned
// as Y J0 LD^(dim-1) / VT (A/V; per cm in 2D).
//
// The reference factors J + j w_s C once per frequency (SuperLU) and
// back-solves every port.  Here the system is the equilibrated stencil
// matrix of the Newton solver (R J Cc, Cc = (1, n, p) per node) plus the
// imaginary storage diagonal in the same scaled frame, solved by complex
// BiCGStab preconditioned with the exact inverse of that matrix on every
// 16-node tile (complex tile block-Jacobi, rebuilt per frequency, FP16
// rows; the complex node blocks with TCADHIP_AC_PRECOND=node); FP64
// vectors and reductions.  On the 200x120 MOSFET's 4-port Y at 7
// frequencies: 39,202 complex iterations with node blocks, 15,808 with
// tiles (5.5 -> 9.3 frequencies/s).  Restarted sweeps recompute the true residual r = b - A x
// from the current x (iterative refinement in working precision) until
// ||r|| <= rtol ||b||, or the residual stops falling.  All devices of a
// batch are solved together, one port index at a time; reductions are
// per workgroup partials re-reduced in a fixed order per device, so Y is
// bit-reproducible and identical for a device alone or in a batch.
//
// Roofline: HBM/L2 bound like the real solver.  Algorithmic bytes per node
// and iteration, 2D: 2 SpMV x (29 x 8 matrix + 48 x + 48 y + 16 storage
// diagonal) + 2 preconditioner applications x (576 + 12 tile rows + 48 +
// 48) + ~12 complex vector passes x 48.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "node_helpers.h"

namespace tcad {
namespace {

enum { CX = 0, CR, CRH, CP, CV, CS, CT, CPH, CSH, NCV };

struct Cx {
  double re, im;
};

__device__ __forceinline__ Cx cmul(Cx a, Cx b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

__device__ __forceinline__ Cx cdiv(Cx a, Cx b) {
  // Smith's algorithm: no overflow for |b| near the double range
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, den = b.re + b.im * r;
    return {(a.re + a.im * r) / den, (a.im - a.re * r) / den};
  }
  const double r = b.re / b.im, den = b.re * r + b.im;
  return {(a.re * r + a.im) / den, (a.im * r - a.re) / den};
}

__device__ __forceinline__ Cx csub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Cx cadd(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }

// Per-device scalars of the complex solve, device resident.
struct ACScal {
  double rho[2][2];                 // (rhat, r) by iteration parity: re, im
  double alpha[2], omega[2];        // complex
  double bb;                        // ||r||^2 at the sweep start (true residual)
  double b2;                        // ||b||^2
  double rr;                        // last ||r||^2
  double tol2;                      // sweep stop: rr <= tol2 * bb
  double abs2;                      // overall stop: rr <= abs2 * b2
  int it, maxit, done, status, sweep, pad;
};

struct CArgs {
  Where W;
  long long N;
  int nblk;
  const double *A;                  // scaled real Newton matrix, FP64
  const double *Mi;                 // 18 planes: complex node block inverses (re 0..8, im 9..17)
  const tile4 *Tr, *Ti;             // complex tile inverse rows (tile != 0)
  const float *Ts;                  // their per-row scales
  int tile;                         // 1: complex tile block-Jacobi, 0: node blocks
  const double *dn, *dp;            // scaled storage diagonal (n and p rows)
  const double *omega;              // scaled angular frequency per device
  double *v[NCV];                   // complex vectors: re c*N + g, im (3 + c)*N + g
  const double *b;                  // scaled right-hand side (real), 3 planes
  double *part;                     // 8 partial sums per workgroup
  ACScal *acs;
  int parity;
  int first;
  // red-black tile ordering (tile != 0): k_cb_init / k_cb_s / k_cb_u apply
  // the red tiles' inverses only, k_cb_black the black tiles' after them
  // (v[rb_src] -> v[rb_dst])
  int rb, rb_src, rb_dst;
};

#define CPROLOGUE                                   \
  __shared__ double sh[128];                        \
  const int d = dev_index(K.W);                     \
  ACScal &S = K.acs[d];                             \
  const int done0 = S.done;                         \
  const DevInfo I = dev_info(K.W, d);               \
  KARGS_EARLY(I.off, I.nx, I.ny, I.px, I.blk0)      \
  KARGS_EARLY(done0, I.dim, I.py)                   \
  if (done0) return;                                \
  const bool lead = (blockIdx.x == I.blk0) && threadIdx.x == 0;

__device__ __forceinline__ void cload(const double *v, long long N, long long g, double x[6]) {
#pragma unroll
  for (int c = 0; c < 6; ++c) x[c] = v[c * N + g];
}

__device__ __forceinline__ void cstore(double *v, long long N, long long g, const double x[6]) {
#pragma unroll
  for (int c = 0; c < 6; ++c) v[c * N + g] = x[c];
}

// y = (A + j w diag(0, dn, dp)) x at one node; the real stencil is read
// once for both halves of x.
__device__ __forceinline__ void cspmv(const CArgs &K, const NodeHood &H, double w,
                                      const double *x, double y[6]) {
  const long long N = K.N, g = H.g;
  double xr[6];
  cload(x, N, g, xr);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double a0 = K.A[(3 * r) * N + g], a1 = K.A[(3 * r + 1) * N + g],
                 a2 = K.A[(3 * r + 2) * N + g];
    y[r] = a0 * xr[0] + a1 * xr[1] + a2 * xr[2];
    y[3 + r] = a0 * xr[3] + a1 * xr[4] + a2 * xr[5];
  }
#pragma unroll
  for (int dir = 0; dir < 6; ++dir) {
    if (!H.nb[dir]) continue;
    const long long gn = H.nbg[dir];
    const double *o = K.A + (long long)(9 + 5 * dir) * N + g;
    const double o0 = o[0], o1 = o[N], o2 = o[2 * N], o3 = o[3 * N], o4 = o[4 * N];
    const double pr = x[gn], nr = x[N + gn], qr = x[2 * N + gn];
    const double pi = x[3 * N + gn], ni = x[4 * N + gn], qi = x[5 * N + gn];
    y[0] += o0 * pr;
    y[3] += o0 * pi;
    y[1] += o1 * pr + o2 * nr;
    y[4] += o1 * pi + o2 * ni;
    y[2] += o3 * pr + o4 * qr;
    y[5] += o3 * pi + o4 * qi;
  }
  const double wn = w * K.dn[g], wp = w * K.dp[g];
  y[1] -= wn * xr[4];
  y[4] += wn * xr[1];
  y[2] -= wp * xr[5];
  y[5] += wp * xr[2];
}

// z = M^-1 x with the node's complex block inverse
__device__ __forceinline__ void cprec(const CArgs &K, long long g, const double x[6],
                                      double z[6]) {
  const long long N = K.N;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double zr = 0.0, zi = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double mr = K.Mi[(3 * r + c) * N + g], mi = K.Mi[(9 + 3 * r + c) * N + g];
      zr += mr * x[c] - mi * x[3 + c];
      zi += mr * x[3 + c] + mi * x[c];
    }
    z[r] = zr;
    z[3 + r] = zi;
  }
}

// Complex inverse of each node block of A + j w diag(0, dn, dp), by
// cofactors (the complex counterpart of the node block-Jacobi).
__global__ void __launch_bounds__(BS) k_ac_prec(CArgs K, const Scal *scal, double *Mi,
                                                int *flag) {
  const int d = dev_index(K.W);
  if (!scal[d].active) return;            // not assembled for this call
  const DevInfo I = dev_info(K.W, d);
  const NodeHood H = hood(I, blockIdx.x);
  if (!H.valid) return;
  const long long N = K.N, g = H.g;
  const double w = K.omega[d];
  Cx a[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) a[e] = {K.A[e * N + g], 0.0};
  a[4].im = w * K.dn[g];
  a[8].im = w * K.dp[g];
  const Cx c0 = csub(cmul(a[4], a[8]), cmul(a[5], a[7]));
  const Cx c1 = csub(cmul(a[5], a[6]), cmul(a[3], a[8]));
  const Cx c2 = csub(cmul(a[3], a[7]), cmul(a[4], a[6]));
  const Cx det = cadd(cadd(cmul(a[0], c0), cmul(a[1], c1)), cmul(a[2], c2));
  const Cx adj[9] = {c0, csub(cmul(a[2], a[7]), cmul(a[1], a[8])),
                     csub(cmul(a[1], a[5]), cmul(a[2], a[4])),
                     c1, csub(cmul(a[0], a[8]), cmul(a[2], a[6])),
                     csub(cmul(a[2], a[3]), cmul(a[0], a[5])),
                     c2, csub(cmul(a[1], a[6]), cmul(a[0], a[7])),
                     csub(cmul(a[0], a[4]), cmul(a[1], a[3]))};
  const double dm = hypot(det.re, det.im);
  if (!(dm > 1e-300) || !isfinite(dm)) {
    atomicOr(flag, 1);
    atomicOr(flag + 1 + d, 1);
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const Cx m = cdiv(adj[e], det);
    Mi[e * N + g] = m.re;
    Mi[(9 + e) * N + g] = m.im;
  }
}

// ---------------------------------------------------------------------------
// Complex tile block-Jacobi: the exact inverse of (A + j w D) restricted to
// every 16-node tile (the tiles of the real solver: 4 x 4 in 2D, T3X x T3Y
// x T3Z in 3D), rebuilt per frequency.  Node k keeps its 3 inverse rows as
// real and imaginary planes of four FP16 entries (Tr, Ti: TILE_PLANES
// planes each) with one FP32 scale per row (Ts), the largest |re| or |im|
// of the row.
// ---------------------------------------------------------------------------
constexpr int CTU = TILE_U;                         // 48 unknowns per tile
constexpr int CHC = CTU / 2;                        // columns per setup thread
constexpr int CTPW = 2;                             // tiles per setup workgroup
constexpr int CNODES = CTPW * TILE_NODES;           // 32 nodes per setup workgroup
constexpr int CSCOLS = CTPW * TILE_X, CSCOLS3 = CTPW * T3X;
constexpr int CSPP = (PX / CSCOLS) * (PY / TILE_Y); // setup workgroups per patch
constexpr int CCH = 6;                              // elimination chunk (columns)
static_assert(PX % CSCOLS == 0 && PX3 % CSCOLS3 == 0 && PX3 / CSCOLS3 == CSPP,
              "complex setup groups must tile 2D and 3D patches alike");
static_assert(CHC % 3 == 0 && CHC % 4 == 0 && CHC % CCH == 0, "halves hold whole blocks");

// lane (0..31) of the complex setup workgroup that holds node l of tile tq
__device__ __forceinline__ int csetup_lane(int dim, int tq, int l) {
  if (dim == 3) {
    const int il = l % T3X, ly = (l / T3X) % T3Y, lz = l / (T3X * T3Y);
    return lz * (CSCOLS3 * T3Y) + ly * CSCOLS3 + tq * T3X + il;
  }
  return (l / TILE_X) * CSCOLS + tq * TILE_X + l % TILE_X;
}

// Complex inverse of a 3x3 block (cofactors); returns |det|
__device__ __forceinline__ double cinv3(const Cx a[9], Cx P[9]) {
  const Cx c0 = csub(cmul(a[4], a[8]), cmul(a[5], a[7]));
  const Cx c1 = csub(cmul(a[5], a[6]), cmul(a[3], a[8]));
  const Cx c2 = csub(cmul(a[3], a[7]), cmul(a[4], a[6]));
  const Cx det = cadd(cadd(cmul(a[0], c0), cmul(a[1], c1)), cmul(a[2], c2));
  const Cx adj[9] = {c0, csub(cmul(a[2], a[7]), cmul(a[1], a[8])),
                     csub(cmul(a[1], a[5]), cmul(a[2], a[4])),
                     c1, csub(cmul(a[0], a[8]), cmul(a[2], a[6])),
                     csub(cmul(a[2], a[3]), cmul(a[0], a[5])),
                     c2, csub(cmul(a[1], a[6]), cmul(a[0], a[7])),
                     csub(cmul(a[0], a[4]), cmul(a[1], a[3]))};
#pragma unroll
  for (int e = 0; e < 9; ++e) P[e] = cdiv(adj[e], det);
  return hypot(det.re, det.im);
}

#define CMR(lane, r, col) Mre[((r) * CTU + (col)) * CNODES + (lane)]
#define CMI(lane, r, col) Mim[((r) * CTU + (col)) * CNODES + (lane)]

// Setup: 64 threads per CTPW tiles side by side in x (32 nodes), two
// threads per node over the two halves of the 48 columns; complex block
// Gauss-Jordan in FP64 with the 3x3 node blocks as pivots, in place in
// LDS (column-major, lane fastest, as the real tile setup).
__global__ void __launch_bounds__(64) k_ac_tile_setup(CArgs K, const Scal *scal, tile4 *Tr,
                                                      tile4 *Ti, float *Ts, int *flag) {
  __shared__ double Mre[3 * CTU * CNODES], Mim[3 * CTU * CNODES];
  __shared__ float amx[2 * CNODES * 3];
  const int nt = threadIdx.x % CNODES, half = threadIdx.x / CNODES, c0 = half * CHC;
  const int patch = blockIdx.x / CSPP, sub = blockIdx.x % CSPP;
  const int d = K.W.single ? 0 : K.W.blk2dev[patch];
  if (!scal[d].active) return;            // not assembled for this call
  const DevInfo I = dev_info(K.W, d);
  const long long N = K.N;
  const double w = K.omega[d];
  const unsigned local = (unsigned)(patch - I.blk0);
  const unsigned bx = local % (unsigned)I.px, t2 = local / (unsigned)I.px;
  const unsigned by = t2 % (unsigned)I.py, bz = t2 / (unsigned)I.py;
  int tq, l, i, j, kk;
  bool has[6];
  int ln[6];
  if (I.dim == 3) {
    const int lc = nt % CSCOLS3, ly = (nt / CSCOLS3) % T3Y, lz = nt / (CSCOLS3 * T3Y);
    const int il = lc % T3X;
    tq = lc / T3X;
    l = (lz * T3Y + ly) * T3X + il;
    i = (int)bx * PX3 + sub * CSCOLS3 + lc;
    j = (int)by * PY3 + ly;
    kk = (int)bz * PZ3 + lz;
    has[0] = il > 0 && i > 0;
    has[1] = il < T3X - 1 && i < I.nx - 1;
    has[2] = ly > 0 && j > 0;
    has[3] = ly < T3Y - 1 && j < I.ny - 1;
    has[4] = lz > 0 && kk > 0;
    has[5] = lz < T3Z - 1 && kk < I.nz - 1;
    ln[2] = l - T3X; ln[3] = l + T3X;
    ln[4] = l - T3X * T3Y; ln[5] = l + T3X * T3Y;
  } else {
    const int lc = nt % CSCOLS, lr = nt / CSCOLS;
    const int il = lc % TILE_X;
    tq = lc / TILE_X;
    l = lr * TILE_X + il;
    i = (int)bx * PX + (sub % (PX / CSCOLS)) * CSCOLS + lc;
    j = (int)by * PY + (sub / (PX / CSCOLS)) * TILE_Y + lr;
    kk = 0;
    has[0] = il > 0 && i > 0;
    has[1] = il < TILE_X - 1 && i < I.nx - 1;
    has[2] = lr > 0 && j > 0;
    has[3] = lr < TILE_Y - 1 && j < I.ny - 1;
    has[4] = has[5] = false;
    ln[2] = l - TILE_X; ln[3] = l + TILE_X;
    ln[4] = ln[5] = l;
  }
  ln[0] = l - 1; ln[1] = l + 1;
  const bool valid = i < I.nx && j < I.ny && kk < I.nz;
  const long long g = I.off + ((long long)kk * I.ny + j) * I.nx + i;
#pragma unroll
  for (int e = 0; e < CHC; ++e)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      CMR(nt, r, c0 + e) = 0.0;
      CMI(nt, r, c0 + e) = 0.0;
    }
  __syncthreads();
  if (half == 0) {
    if (valid) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) CMR(nt, r, 3 * l + c) = K.A[(3 * r + c) * N + g];
      CMI(nt, 1, 3 * l + 1) = w * K.dn[g];
      CMI(nt, 2, 3 * l + 2) = w * K.dp[g];
#pragma unroll
      for (int dir = 0; dir < 6; ++dir) {
        if (!has[dir]) continue;
        const double *o = K.A + (long long)(9 + 5 * dir) * N + g;
        const int cb = 3 * ln[dir];
        CMR(nt, 0, cb) = o[0];
        CMR(nt, 1, cb) = o[N];
        CMR(nt, 1, cb + 1) = o[2 * N];
        CMR(nt, 2, cb) = o[3 * N];
        CMR(nt, 2, cb + 2) = o[4 * N];
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) CMR(nt, c, 3 * l + c) = 1.0;
    }
  }
  bool bad = false;
  for (int sp = 0; sp < TILE_NODES; ++sp) {
    const int pc = 3 * sp;
    const int pl = csetup_lane(I.dim, tq, sp);
    __syncthreads();
    Cx a[9], P[9], F[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        a[3 * r + c] = {CMR(pl, r, pc + c), CMI(pl, r, pc + c)};
        F[3 * r + c] = {CMR(nt, r, pc + c), CMI(nt, r, pc + c)};
      }
    const double dm = cinv3(a, P);
    if (l == sp && half == 0 && (!(dm > 1e-300) || !isfinite(dm))) bad = true;
    __syncthreads();                            // pivot block read by all
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int cc = l + TILE_NODES * k;
      if (cc < CHC) {
        const int col = c0 + cc;
        const Cx m[3] = {{CMR(pl, 0, col), CMI(pl, 0, col)}, {CMR(pl, 1, col), CMI(pl, 1, col)},
                         {CMR(pl, 2, col), CMI(pl, 2, col)}};
        const int c = col - pc;
        const bool pb = c >= 0 && c < 3;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          Cx nv = cadd(cadd(cmul(P[3 * r], m[0]), cmul(P[3 * r + 1], m[1])),
                       cmul(P[3 * r + 2], m[2]));
          if (pb) nv = c == 0 ? P[3 * r] : (c == 1 ? P[3 * r + 1] : P[3 * r + 2]);
          CMR(pl, r, col) = nv.re;
          CMI(pl, r, col) = nv.im;
        }
      }
    }
    __syncthreads();
    if (l != sp) {
#pragma unroll
      for (int ch = 0; ch < CHC; ch += CCH) {
        Cx sv[3][CCH], mv[3][CCH];
#pragma unroll
        for (int e = 0; e < CCH; ++e)
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            sv[r][e] = {CMR(pl, r, c0 + ch + e), CMI(pl, r, c0 + ch + e)};
            mv[r][e] = {CMR(nt, r, c0 + ch + e), CMI(nt, r, c0 + ch + e)};
          }
#pragma unroll
        for (int e = 0; e < CCH; ++e)
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const Cx v = csub(mv[r][e], cadd(cadd(cmul(F[3 * r], sv[0][e]),
                                                  cmul(F[3 * r + 1], sv[1][e])),
                                             cmul(F[3 * r + 2], sv[2][e])));
            CMR(nt, r, c0 + ch + e) = v.re;
            CMI(nt, r, c0 + ch + e) = v.im;
          }
      }
      if (pc / CHC == half) {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const Cx v = cadd(cadd(cmul(F[3 * r], P[c]), cmul(F[3 * r + 1], P[3 + c])),
                              cmul(F[3 * r + 2], P[6 + c]));
            CMR(nt, r, pc + c) = -v.re;
            CMI(nt, r, pc + c) = -v.im;
          }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double m = 0.0;
#pragma unroll
    for (int e = 0; e < CHC; ++e)
      m = fmax(m, fmax(fabs(CMR(nt, r, c0 + e)), fabs(CMI(nt, r, c0 + e))));
    amx[(half * CNODES + nt) * 3 + r] = (float)m;
  }
  __syncthreads();
  if (valid) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const float amax = fmaxf(amx[nt * 3 + r], amx[(CNODES + nt) * 3 + r]);
      const float sc = amax > 0.0f && isfinite(amax) ? amax : 1.0f;
      const double inv = 1.0 / (double)sc;
      if (half == 0) Ts[(long long)r * N + g] = sc;
#pragma unroll
      for (int q = c0 / 4; q < (c0 + CHC) / 4; ++q) {
        tile4 hr, hi;
        hr.x = (_Float16)(CMR(nt, r, 4 * q) * inv);
        hr.y = (_Float16)(CMR(nt, r, 4 * q + 1) * inv);
        hr.z = (_Float16)(CMR(nt, r, 4 * q + 2) * inv);
        hr.w = (_Float16)(CMR(nt, r, 4 * q + 3) * inv);
        hi.x = (_Float16)(CMI(nt, r, 4 * q) * inv);
        hi.y = (_Float16)(CMI(nt, r, 4 * q + 1) * inv);
        hi.z = (_Float16)(CMI(nt, r, 4 * q + 2) * inv);
        hi.w = (_Float16)(CMI(nt, r, 4 * q + 3) * inv);
        Tr[(long long)(r * (CTU / 4) + q) * N + g] = hr;
        Ti[(long long)(r * (CTU / 4) + q) * N + g] = hi;
      }
    }
  }
  if (bad) {                              // error flags only; never part of a sum
    atomicOr(flag, 1);
    atomicOr(flag + 1 + d, 1);
  }
}
#undef CMR
#undef CMI

// z = M^-1 x at this thread's node with the complex tile inverse rows read
// at the point of use (every thread of the workgroup calls it: it holds
// barriers; tv is LDS of BS * 6 doubles)
__device__ __forceinline__ void ctile_apply(const CArgs &K, int dim, bool valid, long long g,
                                            const double x[6], double z[6], double *tv) {
  int tile, l;
  tile_slot(dim, tile, l);
  double *vr = tv + tile * CTU, *vi = tv + 3 * BS + tile * CTU;
  __syncthreads();                          // tv free (a previous use is done)
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    vr[3 * l + c] = valid ? x[c] : 0.0;
    vi[3 * l + c] = valid ? x[3 + c] : 0.0;
  }
  __syncthreads();
  if (!valid) return;
  const long long N = K.N;
  double ar[3] = {0.0, 0.0, 0.0}, ai[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < CTU / 4; ++q) {
    const double r0 = vr[4 * q], r1 = vr[4 * q + 1], r2 = vr[4 * q + 2], r3 = vr[4 * q + 3];
    const double i0 = vi[4 * q], i1 = vi[4 * q + 1], i2 = vi[4 * q + 2], i3 = vi[4 * q + 3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const tile4 tr = K.Tr[(long long)(r * (CTU / 4) + q) * N + g];
      const tile4 ti = K.Ti[(long long)(r * (CTU / 4) + q) * N + g];
      ar[r] += ((double)tr.x * r0 + (double)tr.y * r1 + (double)tr.z * r2 + (double)tr.w * r3) -
               ((double)ti.x * i0 + (double)ti.y * i1 + (double)ti.z * i2 + (double)ti.w * i3);
      ai[r] += ((double)tr.x * i0 + (double)tr.y * i1 + (double)tr.z * i2 + (double)tr.w * i3) +
               ((double)ti.x * r0 + (double)ti.y * r1 + (double)ti.z * r2 + (double)ti.w * r3);
    }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double sc = (double)K.Ts[(long long)r * N + g];
    z[r] = ar[r] * sc;
    z[3 + r] = ai[r] * sc;
  }
}

// z = M^-1 x: complex tiles (every thread of the workgroup calls it) or the
// complex node blocks
__device__ __forceinline__ void capply(const CArgs &K, int dim, bool valid, long long g,
                                       const double x[6], double z[6]) {
  if (K.tile) {
    __shared__ double tv[BS * 6];
    ctile_apply(K, dim, valid, g, x, z, tv);
  } else if (valid) {
    cprec(K, g, x, z);
  }
}

// Port k's forcing on every node of every device: -kappa w on the nodes of
// a gate port, 0 elsewhere; an ohmic port's 1 on its nodes is scattered
// by k_ac_contact_forcing afterwards, from the contact's own node list
// (contacts may share nodes, as a tri-gate's body and source do, and the
// reference drives every node of the port's list, ac2d.py:296-304).
__global__ void __launch_bounds__(BS) k_ac_forcing(Where W, const int *gidx,
                                                   const int *gn_gate, const double *gn_kw,
                                                   int k, double *acg) {
  const int d = dev_index(W);
  const DevInfo I = dev_info(W, d);
  const NodeHood H = hood(I, blockIdx.x);
  if (!H.valid) return;
  const long long g = H.g;
  double v = 0.0;
  if (k >= I.nc && k < I.nc + I.ng) {
    const int gt = gidx[g];
    if (gt >= 0 && gn_gate[gt] == I.g0 + (k - I.nc)) v = -gn_kw[gt];
  }
  acg[g] = v;
}

// One workgroup per device: 1 on every node of its ohmic port k's list.
__global__ void __launch_bounds__(BS) k_ac_contact_forcing(const DevInfo *info, const int *cptr,
                                                           const int *cnodes, int k,
                                                           double *acg) {
  const DevInfo &I = info[blockIdx.x];
  if (k >= I.nc) return;
  const int c = I.c0 + k;
  for (int t = cptr[c] + threadIdx.x; t < cptr[c + 1]; t += blockDim.x) acg[cnodes[t]] = 1.0;
}

__global__ void k_ac_reset(ACScal *acs, const int *act, int ndev, double tol2, double abs2,
                           int maxit, int sweep) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndev) return;
  ACScal &S = acs[d];
  S.rho[0][0] = S.rho[1][0] = 1.0;
  S.rho[0][1] = S.rho[1][1] = 0.0;
  S.alpha[0] = 1.0;
  S.alpha[1] = 0.0;
  S.omega[0] = 1.0;
  S.omega[1] = 0.0;
  S.it = 0;
  S.maxit = maxit;
  S.tol2 = tol2;
  S.abs2 = abs2;
  S.sweep = sweep;
  if (sweep == 0) S.b2 = 0.0;
  if (act[d]) {
    S.status = 0;
    S.rr = 0.0;
    S.done = 0;
  } else {
    S.done = 1;
  }
}

// x := xbak on the devices flagged in mask (their sweep diverged)
__global__ void __launch_bounds__(BS) k_ac_restore(Where W, long long N, const int *mask,
                                                   const double *xbak, double *x) {
  const int d = dev_index(W);
  if (!mask[d]) return;
  const DevInfo I = dev_info(W, d);
  const NodeHood H = hood(I, blockIdx.x);
  if (!H.valid) return;
#pragma unroll
  for (int c = 0; c < 6; ++c) x[c * N + H.g] = xbak[c * N + H.g];
}

// r = b - A x (x = 0 when K.first), rhat = p = r, phat = M^-1 p; partials
// ||r||^2 (quantity 0; rho_0 = (rhat, r) is the same sum) and ||b||^2
// (quantity 1, read by the first k_cb_s of a solve: the reference of its
// relative residual)
__global__ void __launch_bounds__(BS) k_cb_init(CArgs K) {
  CPROLOGUE
  (void)lead;
  const NodeHood H = hood(I, blockIdx.x);
  const long long N = K.N;
  double rr = 0.0, bb = 0.0;
  double r[6] = {0, 0, 0, 0, 0, 0}, z[6];
  if (H.valid) {
    if (K.first) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        r[c] = K.b[c * N + H.g];
        r[3 + c] = 0.0;
      }
      const double z[6] = {0, 0, 0, 0, 0, 0};
      cstore(K.v[CX], N, H.g, z);
    } else {
      double ax[6];
      cspmv(K, H, K.omega[d], K.v[CX], ax);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        r[c] = K.b[c * N + H.g] - ax[c];
        r[3 + c] = -ax[3 + c];
      }
    }
    cstore(K.v[CR], N, H.g, r);
    cstore(K.v[CRH], N, H.g, r);
    cstore(K.v[CP], N, H.g, r);
#pragma unroll
    for (int c = 0; c < 6; ++c) rr += r[c] * r[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) bb += K.b[c * N + H.g] * K.b[c * N + H.g];
  }
  const bool on = H.valid && !(K.tile && K.rb && tile_black_at(I.dim, H.i, H.j, H.kk));
  capply(K, I.dim, on, H.g, r, z);
  if (on) cstore(K.v[CPH], N, H.g, z);
  double in[2] = {rr, bb}, s[2];
  block_sums<2>(in, sh, s);
  if (threadIdx.x == 0) {
    K.part[blockIdx.x] = s[0];
    K.part[(long long)K.nblk + blockIdx.x] = s[1];
  }
}

// The black tiles of the red-black tile ordering, after a kernel that
// applied the red ones: z = T^-1 (x - C z_red) on every black tile, x =
// v[rb_src], z = v[rb_dst], C the (real) stencil couplings of the tile's
// nodes to the neighbouring tiles, all red (the storage diagonal jwD has no
// couplings).  Forward block Gauss-Seidel over the two colours, as in the
// real solver.
__global__ void __launch_bounds__(BS) k_cb_black(CArgs K) {
  CPROLOGUE
  (void)lead;
  (void)sh;
  const NodeHood H = hood(I, blockIdx.x);
  const long long N = K.N;
  const bool on = H.valid && tile_black_at(I.dim, H.i, H.j, H.kk);
  const double *src = K.v[K.rb_src];
  double *dst = K.v[K.rb_dst];
  double x[6] = {0, 0, 0, 0, 0, 0};
  if (on) {
    cload(src, N, H.g, x);
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
      if (!H.nb[dir] || !leaves_tile_at(I.dim, H.i, H.j, H.kk, dir)) continue;
      const long long gn = H.nbg[dir];
      const double *o = K.A + (long long)(9 + 5 * dir) * N + H.g;
#pragma unroll
      for (int h = 0; h < 2; ++h) {              // real, imaginary halves
        const double zp = dst[(3 * h) * N + gn], zn = dst[(3 * h + 1) * N + gn],
                     zq = dst[(3 * h + 2) * N + gn];
        x[3 * h] -= o[0] * zp;
        x[3 * h + 1] -= o[N] * zp + o[2 * N] * zn;
        x[3 * h + 2] -= o[3 * N] * zp + o[4 * N] * zq;
      }
    }
  }
  double z[6];
  capply(K, I.dim, on, H.g, x, z);
  if (on) cstore(dst, N, H.g, z);
}

static void launch_cblack(Batch &b, CArgs K, int src, int dst) {
  if (!K.tile || !K.rb) return;
  K.rb_src = src;
  K.rb_dst = dst;
  hipLaunchKernelGGL(k_cb_black, dim3(b.nblk), dim3(BS), 0, b.stream, K);
  b.launches++;
}

// dst = (A + j w D) src; partial sum conj(a) dst at q0, q0 + 1
__global__ void __launch_bounds__(BS) k_cb_spmv(CArgs K, int src, int dst, int dot_a, int q0) {
  CPROLOGUE
  (void)lead;
  const NodeHood H = hood(I, blockIdx.x);
  const long long N = K.N;
  double q[2] = {0.0, 0.0};
  if (H.valid) {
    double y[6], a[6];
    cspmv(K, H, K.omega[d], K.v[src], y);
    cstore(K.v[dst], N, H.g, y);
    cload(K.v[dot_a], N, H.g, a);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      q[0] += a[c] * y[c] + a[3 + c] * y[3 + c];
      q[1] += a[c] * y[3 + c] - a[3 + c] * y[c];
    }
  }
  double o[2];
  block_sums<2>(q, sh, o);
  if (threadIdx.x == 0) {
    K.part[(long long)q0 * K.nblk + blockIdx.x] = o[0];
    K.part[(long long)(q0 + 1) * K.nblk + blockIdx.x] = o[1];
  }
}

// t = (A + j w D) shat; partials t^H s (1, 2), t^H t (3), rhat^H s (4, 5),
// rhat^H t (6, 7): omega and the next rho = rhat^H (s - omega t) without
// another reduction
__global__ void __launch_bounds__(BS) k_cb_spmv_t(CArgs K) {
  CPROLOGUE
  (void)lead;
  const NodeHood H = hood(I, blockIdx.x);
  const long long N = K.N;
  double q[7] = {0, 0, 0, 0, 0, 0, 0};
  if (H.valid) {
    double t[6], s6[6], rh[6];
    cspmv(K, H, K.omega[d], K.v[CSH], t);
    cstore(K.v[CT], N, H.g, t);
    cload(K.v[CS], N, H.g, s6);
    cload(K.v[CRH], N, H.g, rh);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double tr = t[c], ti = t[3 + c], sr = s6[c], si = s6[3 + c];
      const double hr = rh[c], hi = rh[3 + c];
      q[0] += tr * sr + ti * si;
      q[1] += tr * si - ti * sr;
      q[2] += tr * tr + ti * ti;
      q[3] += hr * sr + hi * si;
      q[4] += hr * si - hi * sr;
      q[5] += hr * tr + hi * ti;
      q[6] += hr * ti - hi * tr;
    }
  }
  double o[7];
  block_sums<7>(q, sh, o);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 7; ++k) K.part[(long long)(1 + k) * K.nblk + blockIdx.x] = o[k];
  }
}

// Stop test on ||r||^2 (quantity 0); alpha = rho / rhat^H v (2, 3);
// s = r - alpha v; shat = M^-1 s
__global__ void __launch_bounds__(BS) k_cb_s(CArgs K) {
  CPROLOGUE
  const NodeHood H = hood(I, blockIdx.x);
  const long long N = K.N;
  double r[6] = {0, 0, 0, 0, 0, 0}, v[6] = {0, 0, 0, 0, 0, 0};
  if (H.valid) {
    cload(K.v[CR], N, H.g, r);
    cload(K.v[CV], N, H.g, v);
  }
  // device scalars (written by earlier launches, or rewritten here with the
  // same values): read with the vectors, before the reduction
  const int s_it = S.it, s_sweep = S.sweep, s_maxit = S.maxit;
  const double s_bb = S.bb, s_b2 = S.b2, s_tol2 = S.tol2, s_abs2 = S.abs2;
  const Cx s_rho = {S.rho[K.parity][0], S.rho[K.parity][1]};
  double red[4];
  device_sums<4>(K.part, K.nblk, 0, I, sh, red);
  const bool first_it = s_it == 0;
  const double rr = red[0];
  const double bb = first_it ? rr : s_bb;
  const double b2 = (first_it && s_sweep == 0) ? red[1] : s_b2;
  const Cx rv = {red[2], red[3]};
  const Cx rho = first_it ? Cx{rr, 0.0} : s_rho;
  const bool conv = rr <= s_tol2 * bb || rr <= s_abs2 * b2;
  if (conv || s_it >= s_maxit || !isfinite(rr) || (rho.re == 0.0 && rho.im == 0.0)) {
    if (lead) {
      S.rr = rr;
      if (first_it) {
        S.bb = bb;
        S.b2 = b2;
      }
      if (!conv) S.status = 1;
      S.done = 1;
    }
    return;
  }
  const Cx alpha = cdiv(rho, rv);
  if ((rv.re == 0.0 && rv.im == 0.0) || !isfinite(alpha.re) || !isfinite(alpha.im)) {
    if (lead) {
      S.status = 1;
      S.done = 1;
    }
    return;
  }
  double s6[6] = {0, 0, 0, 0, 0, 0}, z[6];
  if (H.valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      s6[c] = r[c] - (alpha.re * v[c] - alpha.im * v[3 + c]);
      s6[3 + c] = r[3 + c] - (alpha.re * v[3 + c] + alpha.im * v[c]);
    }
    cstore(K.v[CS], N, H.g, s6);
  }
  const bool on = H.valid && !(K.tile && K.rb && tile_black_at(I.dim, H.i, H.j, H.kk));
  capply(K, I.dim, on, H.g, s6, z);
  if (on) cstore(K.v[CSH], N, H.g, z);
  if (lead) {
    S.alpha[0] = alpha.re;
    S.alpha[1] = alpha.im;
    S.rr = rr;
    if (first_it) {
      S.bb = bb;
      S.b2 = b2;
      S.rho[K.parity][0] = rho.re;
      S.rho[K.parity][1] = rho.im;
    }
  }
}

// omega = t^H s / t^H t; x += alpha phat + omega shat; r = s - omega t;
// rho' = rhat^H s - omega rhat^H t; p = r + beta (p - omega v);
// phat = M^-1 p (this iteration's end fused with the next one's start);
// partial ||r||^2 at quantity 0
__global__ void __launch_bounds__(BS) k_cb_u(CArgs K) {
  CPROLOGUE
  const NodeHood H = hood(I, blockIdx.x);
  const long long N = K.N;
  double x[6] = {0, 0, 0, 0, 0, 0}, ph[6] = {0, 0, 0, 0, 0, 0}, shh[6] = {0, 0, 0, 0, 0, 0},
         s6[6] = {0, 0, 0, 0, 0, 0}, t[6] = {0, 0, 0, 0, 0, 0}, p[6] = {0, 0, 0, 0, 0, 0},
         v[6] = {0, 0, 0, 0, 0, 0};
  if (H.valid) {
    cload(K.v[CX], N, H.g, x);
    cload(K.v[CPH], N, H.g, ph);
    cload(K.v[CSH], N, H.g, shh);
    cload(K.v[CS], N, H.g, s6);
    cload(K.v[CT], N, H.g, t);
    cload(K.v[CP], N, H.g, p);
    cload(K.v[CV], N, H.g, v);
  }
  // written by earlier launches only: read with the vectors
  const Cx alpha = {S.alpha[0], S.alpha[1]};
  const Cx rho_old = {S.rho[K.parity][0], S.rho[K.parity][1]};
  const int s_it = S.it;
  double red[7];
  device_sums<7>(K.part, K.nblk, 1, I, sh, red);
  const Cx ts = {red[0], red[1]};
  const double tt = red[2];
  const Cx rs = {red[3], red[4]}, rt = {red[5], red[6]};
  const Cx omega = tt > 0.0 ? Cx{ts.re / tt, ts.im / tt} : Cx{0.0, 0.0};
  // omega == 0: stagnation; a tiny value keeps beta finite and the next
  // stop test (or the next sweep) decides
  const bool zero = omega.re == 0.0 && omega.im == 0.0;
  const Cx om_b = zero ? Cx{1e-300, 0.0} : omega;
  const Cx rho_new = csub(rs, cmul(omega, rt));
  const Cx beta = cmul(cdiv(rho_new, rho_old), cdiv(alpha, om_b));
  double rr = 0.0;
  double pn[6] = {0, 0, 0, 0, 0, 0}, z[6];
  if (H.valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      x[c] = x[c] + (alpha.re * ph[c] - alpha.im * ph[3 + c]) +
             (omega.re * shh[c] - omega.im * shh[3 + c]);
      x[3 + c] = x[3 + c] + (alpha.re * ph[3 + c] + alpha.im * ph[c]) +
                 (omega.re * shh[3 + c] + omega.im * shh[c]);
      const double rre = s6[c] - (omega.re * t[c] - omega.im * t[3 + c]);
      const double rim = s6[3 + c] - (omega.re * t[3 + c] + omega.im * t[c]);
      s6[c] = rre;
      s6[3 + c] = rim;
      rr += rre * rre + rim * rim;
      const Cx pv = {p[c] - (om_b.re * v[c] - om_b.im * v[3 + c]),
                     p[3 + c] - (om_b.re * v[3 + c] + om_b.im * v[c])};
      const Cx bp = cmul(beta, pv);
      pn[c] = rre + bp.re;
      pn[3 + c] = rim + bp.im;
    }
    cstore(K.v[CX], N, H.g, x);
    cstore(K.v[CR], N, H.g, s6);
    cstore(K.v[CP], N, H.g, pn);
  }
  const bool on = H.valid && !(K.tile && K.rb && tile_black_at(I.dim, H.i, H.j, H.kk));
  capply(K, I.dim, on, H.g, pn, z);
  if (on) cstore(K.v[CPH], N, H.g, z);
  const double a = block_sum(rr, sh);
  if (threadIdx.x == 0) K.part[blockIdx.x] = a;
  if (lead) {
    S.rho[K.parity ^ 1][0] = rho_new.re;
    S.rho[K.parity ^ 1][1] = rho_new.im;
    S.omega[0] = om_b.re;
    S.omega[1] = om_b.im;
    S.it = s_it + 1;
  }
}

// Column k of Y for every device: one workgroup per row (contacts, then
// gates, batch-global ids); y is the scaled solution (du = Cc y).
struct ObsArgs {
  const DevInfo *info;
  const int *row_dev;               // row -> device
  const int *cptr, *cnodes;         // contact node lists (global nodes)
  const double *cscale;             // per contact: J0, LD^(dim-1)
  const int *gptr, *gnode;          // gate table rows and their nodes
  const double *gn_kw;
  const double *raw;                // 21 planes: raw contact current rows
  const double *y;                  // complex, 6 planes
  const double *n, *p;
  const double *omega;              // per device
  long long N;
  int ncontact;
  int k;
  double *ycol;                     // 2 per row: re, im
};

__global__ void __launch_bounds__(BS) k_ac_observe(ObsArgs O) {
  __shared__ double sh[BS / 64];
  const int row = blockIdx.x;
  const int d = O.row_dev[row];
  const DevInfo &I = O.info[d];
  const long long N = O.N;
  double sre = 0.0, sim = 0.0;
  if (row < O.ncontact) {
    const long long sy = I.nx, sz = (long long)I.nx * I.ny;
    for (int t = O.cptr[row] + threadIdx.x; t < O.cptr[row + 1]; t += blockDim.x) {
      const long long g = O.cnodes[t];
      const long long kl = g - I.off;
      const int i = (int)(kl % I.nx), j = (int)((kl / I.nx) % I.ny), kz = (int)(kl / sz);
      const bool nb[6] = {i > 0, i < I.nx - 1, j > 0, j < I.ny - 1, I.dim == 3 && kz > 0,
                          I.dim == 3 && kz < I.nz - 1};
      const long long gn[6] = {g - 1, g + 1, g - sy, g + sy, g - sz, g + sz};
      // du at the node: (y_psi, n y_n, p y_p)
      const double nn = O.n[g], pp = O.p[g];
      double re = O.raw[0 * N + g] * O.y[g] + O.raw[1 * N + g] * (nn * O.y[N + g]) +
                  O.raw[2 * N + g] * (pp * O.y[2 * N + g]);
      double im = O.raw[0 * N + g] * O.y[3 * N + g] +
                  O.raw[1 * N + g] * (nn * O.y[4 * N + g]) +
                  O.raw[2 * N + g] * (pp * O.y[5 * N + g]);
#pragma unroll
      for (int dir = 0; dir < 6; ++dir) {
        if (!nb[dir]) continue;
        const long long q = gn[dir];
        const double c0 = O.raw[(3 + 3 * dir) * N + g], c1 = O.raw[(4 + 3 * dir) * N + g],
                     c2 = O.raw[(5 + 3 * dir) * N + g];
        const double nq = O.n[q], pq = O.p[q];
        re += c0 * O.y[q] + c1 * (nq * O.y[N + q]) + c2 * (pq * O.y[2 * N + q]);
        im += c0 * O.y[3 * N + q] + c1 * (nq * O.y[4 * N + q]) + c2 * (pq * O.y[5 * N + q]);
      }
      sre += re;
      sim += im;
    }
    const double a = block_sum(sre, sh);
    const double b = block_sum(sim, sh);
    if (threadIdx.x == 0) {
      const double J0 = O.cscale[2 * row], ld = O.cscale[2 * row + 1];
      O.ycol[2 * row] = a * J0 * ld / I.VT;
      O.ycol[2 * row + 1] = b * J0 * ld / I.VT;
    }
  } else {
    const int gate = row - O.ncontact;
    const bool self = O.k >= I.nc && O.k < I.nc + I.ng && gate == I.g0 + (O.k - I.nc);
    const double delta = self ? 1.0 : 0.0;
    for (int t = O.gptr[gate] + threadIdx.x; t < O.gptr[gate + 1]; t += blockDim.x) {
      const long long g = O.gnode[t];
      const double kw = O.gn_kw[t];
      sre += kw * (delta - O.y[g]);
      sim += kw * (0.0 - O.y[3 * N + g]);
    }
    const double a = block_sum(sre, sh);
    const double b = block_sum(sim, sh);
    if (threadIdx.x == 0) {
      // j w_s (a + j b)
      const double w = O.omega[d];
      const double ld = I.dim == 3 ? I.LD * I.LD : I.LD;
      O.ycol[2 * row] = (-w * b) * I.J0 * ld / I.VT;
      O.ycol[2 * row + 1] = (w * a) * I.J0 * ld / I.VT;
    }
  }
}

// Device workspace of one tcad_y_parameters call.
struct AcWork {
  double *base = nullptr;
  double *v[NCV] = {};
  double *Mi = nullptr, *dn = nullptr, *dp = nullptr, *raw = nullptr, *acg = nullptr;
  double *part = nullptr, *omega = nullptr, *ycol = nullptr;
  double *xbak = nullptr;           // x at the current sweep's start (6 planes)
  ACScal *acs = nullptr;
  int *act = nullptr, *row_dev = nullptr;
  void *tbase = nullptr;            // complex tile rows
  tile4 *Tr = nullptr, *Ti = nullptr;
  float *Ts = nullptr;
  int tile = 1;
  int hint[16] = {};                // iterations of the last sweep of each index (slowest device)
  ~AcWork() {
    if (base) (void)hipFree(base);
    if (tbase) (void)hipFree(tbase);
  }
};

int alloc_work(Batch &b, int nrows, AcWork &w) {
  const size_t N = (size_t)b.ntot;
  const size_t nd = NCV * 6 * N + 6 * N + 18 * N + 2 * N + 21 * N + N + 8 * (size_t)b.nblk + b.ndev +
                    2 * (size_t)nrows;
  const size_t bytes = nd * sizeof(double) + sizeof(ACScal) * b.ndev +
                       sizeof(int) * ((size_t)b.ndev + nrows) + 256;
  TCAD_HIP_CHECK(hipMalloc((void **)&w.base, bytes));
  double *q = w.base;
  for (int i = 0; i < NCV; ++i, q += 6 * N) w.v[i] = q;
  w.xbak = q; q += 6 * N;
  w.Mi = q; q += 18 * N;
  w.dn = q; q += N;
  w.dp = q; q += N;
  w.raw = q; q += 21 * N;
  w.acg = q; q += N;
  w.part = q; q += 8 * (size_t)b.nblk;
  w.omega = q; q += b.ndev;
  w.ycol = q; q += 2 * (size_t)nrows;
  w.acs = reinterpret_cast<ACScal *>(q);
  int *ip = reinterpret_cast<int *>(w.acs + b.ndev);
  w.act = ip;
  w.row_dev = ip + b.ndev;
  // complex tiles unless TCADHIP_AC_PRECOND=node (the complex node blocks)
  const char *e = std::getenv("TCADHIP_AC_PRECOND");
  w.tile = (e && std::strcmp(e, "node") == 0) ? 0 : 1;
  if (w.tile) {
    const size_t tb = 2 * (size_t)TILE_PLANES * N * sizeof(tile4) + 3 * N * sizeof(float);
    TCAD_HIP_CHECK(hipMalloc(&w.tbase, tb));
    w.Tr = reinterpret_cast<tile4 *>(w.tbase);
    w.Ti = w.Tr + (size_t)TILE_PLANES * N;
    w.Ts = reinterpret_cast<float *>(w.Ti + (size_t)TILE_PLANES * N);
  }
  return TCAD_OK;
}

CArgs make_cargs(Batch &b, const AcWork &w) {
  CArgs K{};
  K.W = make_where(b);
  K.N = b.ntot;
  K.nblk = b.nblk;
  K.A = b.A64;
  K.Mi = w.Mi;
  K.Tr = w.Tr;
  K.Ti = w.Ti;
  K.Ts = w.Ts;
  K.tile = w.tile;
  K.dn = w.dn;
  K.dp = w.dp;
  K.omega = w.omega;
  for (int i = 0; i < NCV; ++i) K.v[i] = w.v[i];
  K.b = b.bs;
  K.part = w.part;
  K.acs = w.acs;
  K.rb = b.tile_rb;
  return K;
}

void launch_citeration(Batch &b, CArgs K, int parity) {
  const dim3 g(b.nblk), t(BS);
  K.parity = parity;
  hipLaunchKernelGGL(k_cb_spmv, g, t, 0, b.stream, K, (int)CPH, (int)CV, (int)CRH, 2);
  hipLaunchKernelGGL(k_cb_s, g, t, 0, b.stream, K);
  launch_cblack(b, K, CS, CSH);
  hipLaunchKernelGGL(k_cb_spmv_t, g, t, 0, b.stream, K);
  hipLaunchKernelGGL(k_cb_u, g, t, 0, b.stream, K);
  launch_cblack(b, K, CP, CPH);
  b.launches += 4;
}

// Refinement sweeps of the complex solve of the current right-hand side
// (b.bs) at the current frequency, for the devices in act.  Per device:
// worst outcome so far in status (0 ok, TCAD_NEWTON_LINEAR_FAIL), Krylov
// iterations added to iters.
constexpr int kMaxSweeps = 12;
constexpr double kSweepRtol = 1e-8;     // residual reduction per sweep
constexpr double kAcceptRel = 1e-6;     // stalled above this: a failed solve

// warm: start from the solution left in x (the previous frequency's, same
// port) instead of zero.
int complex_solve(Batch &b, AcWork &w, std::vector<int> act, double rtol, int maxit,
                  bool warm, std::vector<int> &status, std::vector<long long> &iters) {
  CArgs K = make_cargs(b, w);
  std::vector<ACScal> hs(b.ndev);
  std::vector<double> prev(b.ndev, INFINITY);
  std::vector<int> conv(b.ndev, 0), stalled_ok(b.ndev, 0), restored(b.ndev, 0);
  const std::vector<int> act0 = act;
  for (int sweep = 0; sweep < kMaxSweeps; ++sweep) {
    int open = 0;
    for (int d = 0; d < b.ndev; ++d) open += act[d];
    if (!open) break;
    TCAD_HIP_CHECK(hipMemcpyAsync(w.act, act.data(), sizeof(int) * b.ndev,
                                  hipMemcpyHostToDevice, b.stream));
    hipLaunchKernelGGL(k_ac_reset, dim3((b.ndev + 63) / 64), dim3(64), 0, b.stream, w.acs,
                       (const int *)w.act, b.ndev, kSweepRtol * kSweepRtol, rtol * rtol, maxit,
                       sweep);
    K.first = sweep == 0 && !warm ? 1 : 0;
    hipLaunchKernelGGL(k_cb_init, dim3(b.nblk), dim3(BS), 0, b.stream, K);
    b.launches += 2;
    launch_cblack(b, K, CP, CPH);
    // x as this sweep starts (k_cb_init has zeroed it on a cold start)
    TCAD_HIP_CHECK(hipMemcpyAsync(w.xbak, w.v[CX], sizeof(double) * 6 * (size_t)b.ntot,
                                  hipMemcpyDeviceToDevice, b.stream));
    // chunks between host checks as in the real solver: 3/4 of the
    // iterations of the last solve's sweep of the same index (a solve's
    // first sweep is long, its restarts short), then 16, 32, 64
    static_assert(kMaxSweeps <= 16, "one hint per sweep index");
    const int hint = w.hint[sweep];
    int it = 0;
    int chunk = hint > 0 ? std::max(8, (3 * hint) / 4) : 8;
    int step = hint > 0 ? 16 : 8;
    while (true) {
      for (int c = 0; c < chunk && it < maxit; ++c, ++it) launch_citeration(b, K, it & 1);
      if (it >= maxit) {
        CArgs K1 = K;
        K1.parity = it & 1;
        hipLaunchKernelGGL(k_cb_s, dim3(b.nblk), dim3(BS), 0, b.stream, K1);
        b.launches++;
      }
      TCAD_HIP_CHECK(hipGetLastError());
      TCAD_HIP_CHECK(hipMemcpyAsync(hs.data(), w.acs, sizeof(ACScal) * b.ndev,
                                    hipMemcpyDeviceToHost, b.stream));
      TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
      int left = 0;
      for (int d = 0; d < b.ndev; ++d) left += hs[d].done ? 0 : 1;
      if (left == 0 || it >= maxit) break;
      chunk = step;
      if (step < 64) step *= 2;
    }
    int most = 0, failed = it >= maxit ? 1 : 0;
    for (int d = 0; d < b.ndev; ++d) {
      most = std::max(most, hs[d].it);
      if (act[d]) failed |= hs[d].status;
    }
    w.hint[sweep] = failed ? 0 : most;
    // A sweep whose residual ended above where it started (FP16 complex tile
    // inverses of a nearly singular tile can do that) leaves x as the sweep
    // found it, and the rest of the solve uses the complex node blocks; a
    // device that diverges with the node blocks too is a stalled solve.
    std::vector<int> back(b.ndev, 0);
    int nback = 0;
    for (int d = 0; d < b.ndev; ++d) {
      const ACScal &S = hs[d];
      if (act[d] && S.status != 0 && !(S.rr <= S.bb)) {
        back[d] = 1;
        ++nback;
      }
    }
    const bool tiles_failed = nback > 0 && K.tile;
    if (nback) {
      TCAD_HIP_CHECK(hipMemcpyAsync(w.act, back.data(), sizeof(int) * b.ndev,
                                    hipMemcpyHostToDevice, b.stream));
      hipLaunchKernelGGL(k_ac_restore, dim3(b.nblk), dim3(BS), 0, b.stream, K.W, K.N,
                         (const int *)w.act, (const double *)w.xbak, w.v[CX]);
      b.launches++;
      if (tiles_failed) {
        K.tile = 0;
        hipLaunchKernelGGL(k_ac_prec, dim3(b.nblk), dim3(BS), 0, b.stream, K,
                           (const Scal *)b.d_scal, w.Mi, b.d_flag);
        b.launches++;
        TCAD_HIP_CHECK(hipGetLastError());
        int rc = check_flag(b);
        if (rc) return rc;
      }
      TCAD_HIP_CHECK(hipGetLastError());
    }
    for (int d = 0; d < b.ndev; ++d) {
      if (!act[d]) continue;
      const ACScal &S = hs[d];
      iters[d] += S.it;
      if (back[d]) {
        if (tiles_failed) {
          restored[d] = 1;                  // retried from the same x: no stall test next
          continue;
        }
        stalled_ok[d] = std::isfinite(prev[d]) && prev[d] <= kAcceptRel * kAcceptRel * S.b2;
        act[d] = 0;
        continue;
      }
      // a sweep retried from the restored x starts where the diverged one
      // did: no stall test against it
      const bool retried = restored[d] != 0;
      restored[d] = 0;
      // S.bb: true residual at this sweep's start; converged there if the
      // sweep stopped before its first iteration
      const bool done_at_start = S.it == 0 && S.bb <= rtol * rtol * S.b2;
      if (done_at_start || !(S.b2 > 0.0)) {
        conv[d] = 1;
        act[d] = 0;
      } else if (!std::isfinite(S.bb) || (!retried && sweep > 0 && S.bb > 0.25 * prev[d])) {
        stalled_ok[d] = std::isfinite(S.bb) && S.bb <= kAcceptRel * kAcceptRel * S.b2;
        act[d] = 0;
      }
      prev[d] = S.bb;
    }
  }
  for (int d = 0; d < b.ndev; ++d) {
    if (!act0[d]) continue;
    bool ok = conv[d] || stalled_ok[d];
    if (!conv[d] && !stalled_ok[d] && act[d])
      ok = std::isfinite(prev[d]) && prev[d] <= kAcceptRel * kAcceptRel * hs[d].b2;
    if (!ok) status[d] = TCAD_NEWTON_LINEAR_FAIL;
  }
  return TCAD_OK;
}

int y_parameters_impl(Batch &b, const int32_t *h_active, int nfreq, const double *h_freqs,
                      double rtol, int maxit, const double *psi, const double *n,
                      const double *p, double *h_Y, int32_t *h_status, int64_t *h_iters) {
  const int nrows = b.ncontact + b.ngate;
  int P = 0;
  for (const DevInfo &I : b.info) P = std::max(P, I.nc + I.ng);
  std::vector<int> act(b.ndev, 1);
  if (h_active)
    for (int d = 0; d < b.ndev; ++d) act[d] = h_active[d] ? 1 : 0;
  std::vector<int> status(b.ndev, TCAD_NEWTON_OK);
  std::vector<long long> iters(b.ndev, 0);
  std::memset(h_Y, 0, sizeof(double) * 2 * (size_t)b.ndev * nfreq * P * P);
  if (nrows == 0 || P == 0) {
    for (int d = 0; d < b.ndev; ++d) {
      h_status[d] = TCAD_NEWTON_OK;
      if (h_iters) h_iters[d] = 0;
    }
    return TCAD_OK;
  }
  AcWork w;
  int rc = alloc_work(b, nrows, w);
  if (rc) return rc;
  std::vector<int> row_dev(nrows, 0);
  for (int d = 0; d < b.ndev; ++d) {
    const DevInfo &I = b.info[d];
    for (int c = 0; c < I.nc; ++c) row_dev[I.c0 + c] = d;
    for (int g = 0; g < I.ng; ++g) row_dev[b.ncontact + I.g0 + g] = d;
  }
  TCAD_HIP_CHECK(hipMemcpyAsync(w.row_dev, row_dev.data(), sizeof(int) * nrows,
                                hipMemcpyHostToDevice, b.stream));
  TCAD_HIP_CHECK(hipMemsetAsync(w.raw, 0, sizeof(double) * 21 * (size_t)b.ntot, b.stream));
  if ((rc = launch_set_active(b, h_active))) return rc;

  // operating-point matrix, storage diagonal and raw contact rows
  AcAssembly prep{};
  prep.mode = 1;
  prep.dn = w.dn;
  prep.dp = w.dp;
  prep.raw = w.raw;
  if ((rc = launch_assemble(b, psi, n, p, b.F, nullptr, false, true, 0, nullptr, &prep)))
    return rc;
  if ((rc = check_flag(b))) return rc;

  ObsArgs O{};
  O.info = b.d_info;
  O.row_dev = w.row_dev;
  O.cptr = b.d_cptr;
  O.cnodes = b.d_cnodes;
  O.cscale = b.d_cscale;
  O.gptr = b.d_gptr;
  O.gnode = b.d_gnode;
  O.gn_kw = b.d_gn_kw;
  O.raw = w.raw;
  O.y = w.v[CX];
  O.n = n;
  O.p = p;
  O.omega = w.omega;
  O.N = b.ntot;
  O.ncontact = b.ncontact;
  O.ycol = w.ycol;
  std::vector<double> ycol(2 * (size_t)nrows), om(b.ndev);
  CArgs K = make_cargs(b, w);
  for (int k = 0; k < P; ++k) {
    hipLaunchKernelGGL(k_ac_forcing, dim3(b.nblk), dim3(BS), 0, b.stream, K.W,
                       (const int *)b.d_gidx, (const int *)b.d_gn_gate,
                       (const double *)b.d_gn_kw, k, w.acg);
    hipLaunchKernelGGL(k_ac_contact_forcing, dim3(b.ndev), dim3(BS), 0, b.stream,
                       (const DevInfo *)b.d_info, (const int *)b.d_cptr, (const int *)b.d_cnodes,
                       k, w.acg);
    b.launches += 2;
    AcAssembly rhs{};
    rhs.mode = 2;
    rhs.acg = w.acg;
    if ((rc = launch_assemble(b, psi, n, p, b.F, nullptr, false, true, 0, nullptr, &rhs)))
      return rc;
    std::vector<int> act_k(b.ndev);
    for (int d = 0; d < b.ndev; ++d)
      act_k[d] = act[d] && k < b.info[d].nc + b.info[d].ng ? 1 : 0;
    for (int f = 0; f < nfreq; ++f) {
      for (int d = 0; d < b.ndev; ++d)
        om[d] = 2.0 * M_PI * h_freqs[f] * (b.info[d].Ns / b.info[d].R0);
      TCAD_HIP_CHECK(hipMemcpyAsync(w.omega, om.data(), sizeof(double) * b.ndev,
                                    hipMemcpyHostToDevice, b.stream));
      if (w.tile)
        hipLaunchKernelGGL(k_ac_tile_setup, dim3(CSPP * b.nblk), dim3(64), 0, b.stream, K,
                           (const Scal *)b.d_scal, w.Tr, w.Ti, w.Ts, b.d_flag);
      else
        hipLaunchKernelGGL(k_ac_prec, dim3(b.nblk), dim3(BS), 0, b.stream, K,
                           (const Scal *)b.d_scal, w.Mi, b.d_flag);
      b.launches++;
      TCAD_HIP_CHECK(hipGetLastError());
      if ((rc = check_flag(b))) return rc;
      if ((rc = complex_solve(b, w, act_k, rtol, maxit, f > 0, status, iters))) return rc;
      O.k = k;
      hipLaunchKernelGGL(k_ac_observe, dim3(nrows), dim3(BS), 0, b.stream, O);
      b.launches++;
      TCAD_HIP_CHECK(hipGetLastError());
      TCAD_HIP_CHECK(hipMemcpyAsync(ycol.data(), w.ycol, sizeof(double) * 2 * nrows,
                                    hipMemcpyDeviceToHost, b.stream));
      TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
      for (int d = 0; d < b.ndev; ++d) {
        if (!act_k[d]) continue;
        const DevInfo &I = b.info[d];
        double *Yd = h_Y + 2 * (((size_t)d * nfreq + f) * P * P);
        for (int i = 0; i < I.nc + I.ng; ++i) {
          const int row = i < I.nc ? I.c0 + i : b.ncontact + I.g0 + (i - I.nc);
          Yd[2 * ((size_t)i * P + k)] = ycol[2 * row];
          Yd[2 * ((size_t)i * P + k) + 1] = ycol[2 * row + 1];
        }
      }
    }
  }
  for (int d = 0; d < b.ndev; ++d) {
    h_status[d] = status[d];
    if (h_iters) h_iters[d] = iters[d];
  }
  return TCAD_OK;
}

}  // namespace
}  // namespace tcad

extern "C" {

int tcad_y_parameters(tcad_batch *bp, const int32_t *h_active, int32_t nfreq,
                      const double *h_freqs, double rtol, int32_t maxiter, const double *psi,
                      const double *n, const double *p, double *h_Y, int32_t *h_status,
                      int64_t *h_iters) {
  GUARD(bp);
  tcad::Batch &b = *bp->impl;
  if (!psi || !n || !p || !h_Y || !h_status || (nfreq > 0 && !h_freqs)) {
    tcad::set_error("tcad_y_parameters: null argument");
    return TCAD_E_INVALID;
  }
  if (nfreq < 0 || maxiter <= 0 || !(rtol > 0.0) || !(rtol < 1.0)) {
    tcad::set_error("tcad_y_parameters: need nfreq >= 0, maxiter > 0 and 0 < rtol < 1");
    return TCAD_E_INVALID;
  }
  for (int f = 0; f < nfreq; ++f)
    if (!(h_freqs[f] >= 0.0) || !std::isfinite(h_freqs[f])) {
      tcad::set_error("tcad_y_parameters: frequencies must be finite and >= 0");
      return TCAD_E_INVALID;
    }
  return tcad::y_parameters_impl(b, h_active, nfreq, h_freqs, rtol, maxiter, psi, n, p, h_Y,
                                 h_status, h_iters);
}

}  // extern "C"
