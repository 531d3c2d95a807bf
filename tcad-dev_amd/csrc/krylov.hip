// Krylov solvers on the stencil-block Newton matrix.
//
// * BiCGStab (right block-Jacobi preconditioning: 3x3 node blocks, or the
//   16-node tiles, 4 x 4 in 2D and 2 x 1 x 8 in 3D), templated on the
//   storage precision T of the matrix and block inverses (double or
//   float); vectors and dot products are FP64.  One iteration is four
//   kernels:
//     k_bicg_spmv    v = A phat, partial (rhat, v)
//     k_bicg_s       reduce (r, r) -> stop test; reduce -> alpha;
//                    s = r - alpha v; shat = M^-1 s
//     k_bicg_spmv_t  t = A shat, partials (t,s), (t,t), (rhat,s), (rhat,t)
//     k_bicg_u       reduce -> omega; x += alpha phat + omega shat;
//                    r = s - omega t; rho' = (rhat,s) - omega (rhat,t);
//                    p = r + beta (p - omega v); phat = M^-1 p; partial (r,r)
//   The update of one iteration and the direction of the next share a
//   kernel: rho' comes from the dot products of t, so the textbook's fifth
//   kernel and its reduction point disappear (the stop test on the new r
//   moves to k_bicg_s).  Each workgroup writes one partial sum per
//   quantity; every workgroup of a device re-reduces that device's
//   partials in a fixed order at the top of the next kernel, so the
//   scalars are identical in all workgroups, bit-reproducible run to run
//   and identical for a device alone or in a batch, with no atomics and no
//   cross-workgroup synchronisation inside a kernel.  (A last-arriving-
//   workgroup finalize was measured 1.5x slower on gfx950: its
//   device-scope release/acquire writes back and invalidates the XCD's L2
//   per workgroup.)  The vector loads are issued before the reductions so
//   the two latencies overlap.  At this size every launch is latency: a
//   kernel reaches its first vector load after two rounds of loads (its
//   arguments; then its device's metadata with the stop flag, see
//   KPROLOGUE), reads the device scalars with the vectors, and sums its
//   partials with DPP row shifts where the tree allows (wave_sum0).
// * Every Newton linear solve is FP64 iterative refinement: the true
//   residual r = b - A64 y is formed in FP64 and BiCGStab solves
//   A d = r / ||r|| with the matrix and block inverses stored in FP64
//   (linsolve "bicgstab") or FP32 ("bicgstab_mixed"; FP64 vector
//   arithmetic -- FP32 vector arithmetic was measured unstable on these
//   Jacobians, residual growth by 1e20+ in the second sweep).
// * PCG (Jacobi; tile block-Jacobi as an option) for the SPD equilibrium
//   Poisson system, FP64, same
//   reduction scheme (two kernels per iteration: the direction update is
//   recomputed at the neighbours inside the product kernel).
//
// Replaces linsolve.solve_linear's node-block-Jacobi Krylov path
// (pytcad/pytcad/linsolve.py:363-427, 725-985).
//
// Roofline: HBM/L2 bound.  Algorithmic bytes per node and BiCGStab
// iteration, 2D (29 entries), matrix width w (8 or 4): 2 SpMV x
// (29 w + 24 + 24 + 24) + 2 precond x (9 w + 48) + ~12 vector passes x 24.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include <hip/hip_ext.h>

#include "node_helpers.h"

namespace tcad {

enum { VX = 0, VR, VRH, VP, VV, VS, VT, VPH, VSH };

// ---------------------------------------------------------------------------
// BiCGStab
// ---------------------------------------------------------------------------
template <typename T>
struct KArgs {
  Where W;
  long long N;
  int nblk;
  const T *A;                       // matrix storage precision
  const T *Dinv;                    // preconditioner storage precision
  const tile4 *T4;                  // tile inverses (FP16 rows), tile != 0
  const float *T4s;                 // their per-row scales
  long long NS;                     // their plane stride (workgroups x BS; tile_row_slot)
  const float *CPL;                 // black tiles' couplings to the red ones (cpl_index)
  int tile;                         // 1: tile block-Jacobi, 0: node block-Jacobi
  double *v[9];                     // Krylov vectors, always FP64
  const double *b;
  double *part;
  Scal *scal;
  int scale_b;                      // init: divide b by Scal::nrm (refinement rhs)
  int parity;
  int late_stop;                    // 1: finished devices return at KSTOP (late_stop_ok)
  // red-black tile ordering (tile != 0): the preconditioner kernels apply
  // the red tiles' inverses only, k_bicg_black the black tiles' after them
  // (v[rb_src] -> v[rb_dst]; rb_init: the solve's first application,
  // skipped by a continued solve, which rescales the last one)
  int rb, rb_src, rb_dst, rb_init;
  // launch grid: grid workgroups; with bmap, workgroup i does the work of
  // the batch's workgroup bmap[i] (the workgroups of the devices still
  // iterating), else workgroup i is workgroup i (grid == nblk)
  const int *bmap;
  int grid;
};

// Where a finished device's workgroups return.  Early (before any load)
// in a batch of several devices, most of whose workgroups may belong to
// retired devices, and on large grids, where the iterations issued after
// convergence inside a chunk would otherwise each read the whole system.
// Late (KSTOP, after the vector loads are issued, so that the stop flag's
// latency overlaps theirs) for one device on a small, latency-bound grid.
constexpr int kLateStopBlocks = 512;
inline int late_stop_ok(const Batch &b) { return b.ndev == 1 && b.nblk <= kLateStopBlocks; }

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
__device__ __forceinline__ void prec3(const T *Di, long long N, long long g, const double x[3],
                                      double z[3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r)
    z[r] = (double)Di[(3 * r) * N + g] * x[0] + (double)Di[(3 * r + 1) * N + g] * x[1] +
           (double)Di[(3 * r + 2) * N + g] * x[2];
}

// ---------------------------------------------------------------------------
// Tile block-Jacobi.  Every 16-node tile (4 x 4 in 2D, T3X x T3Y x T3Z in
// 3D) is one diagonal block of the preconditioner: the scaled Newton matrix
// restricted to the tile (couplings leaving the tile dropped), inverted
// exactly.  The inverse rows of a node are TILE_PLANES planes of four FP16
// entries, T4[(r * TILE_U / 4 + q) * NS + slot] (row r of the node, columns
// 4q..4q+3 of the tile, tile unknown u = 3 l + c with l the node's place in
// the tile, x fastest), each row scaled by its largest magnitude, kept in
// T4s[r * NS + slot]; NS = workgroups x BS, and the slot of a node is
// tile_row_slot's (patch-major, each colour of the red-black ordering
// contiguous in its patch).  Measured BiCGStab iterations to 1e-10 on the
// equilibrated systems of the 200x120 MOSFET (Vg = 0.5 V): node
// block-Jacobi ~630, 4 x 4 tiles ~255 (4 x 8 tiles ~185 for twice the
// bytes: slower on the GPU).  On a 32x24x24 tri-gate: node block-Jacobi
// 145-217, 4 x 4 x 1 x-y tiles 114-190 (the z couplings of the finely
// meshed fin stay outside them), 2 x 2 x 4 tiles 72-86, 2 x 1 x 8 tiles
// 68-89.  On the 64x48x48 tri-gate sweep (GPU, 34 points): node
// block-Jacobi 74,304 Krylov iterations, 9.3 points/s; 2 x 2 x 4 tiles
// 25,969, 18.1; 2 x 1 x 8 tiles (16 x 1 x 8 patches) 23,922, 20.3 (1 x 2 x 8
// tiles: 19,540 iterations, but 8-node x rows, 16.0 points/s).
// ---------------------------------------------------------------------------
constexpr int TU = TILE_U;

// Storage slot of the tile-inverse rows of the node at lane t of workgroup
// (patch) blk, whose tile has colour `black`: the patch's BS slots hold its
// red tiles' nodes, then its black tiles' (each in lane order: every x row
// of a patch alternates tile colours, so each colour fills half of every
// row).  A launch over the tiles of one colour then reads whole cache
// lines, where the node order interleaves the colours every tile width
// (2 nodes, 16 B of FP16 rows, in 3D).
__device__ __forceinline__ long long tile_row_slot(int dim, long long blk, int t, bool black) {
  const int px = dim == 3 ? PX3 : PX, tw = dim == 3 ? T3X : TILE_X;
  const int il = t % px, row = t / px;
  const int rank = row * (px / 2) + ((il / tw) >> 1) * tw + il % tw;
  return blk * BS + (black ? BS / 2 : 0) + rank;
}

// The black tiles' couplings to the red tiles, FP32, for the black-tile
// kernel: for every black node, its out-of-tile directions in direction
// order (at most CPL_GROUPS: one side per axis, both along an axis the tile
// is one node wide), 5 entries each (the stencil coupling entries psi<-psi,
// n<-psi, n<-n, p<-psi, p<-p), zero where the neighbour is outside the
// device.  Entry (grp * 5 + e) of black slot s sits at
// (grp * 5 + e) * NC + cpl_index(s), NC = workgroups x BS / 2: contiguous
// per patch, like the tile rows.
static_assert((TILE_X > 1) + (TILE_Y > 1) <= 2 &&
                  (T3X > 1 ? 1 : 2) + (T3Y > 1 ? 1 : 2) + (T3Z > 1 ? 1 : 2) <= CPL_GROUPS,
              "at most CPL_GROUPS out-of-tile directions per node");
__device__ __forceinline__ long long cpl_index(long long slot) {
  return (slot / BS) * (BS / 2) + (slot % BS) - BS / 2;
}
// tile setup: a 64-thread workgroup inverts STPW tiles side by side in x
constexpr int STPW = 64 / TILE_NODES;
constexpr int SCOLS = STPW * TILE_X;        // node columns of one 2D setup workgroup
constexpr int SCOLS3 = STPW * T3X;          // node columns of one 3D setup workgroup
static_assert(64 % TILE_NODES == 0 && PX % SCOLS == 0 && PX3 % SCOLS3 == 0,
              "setup groups must tile the patch");
constexpr int SPP = (PX / SCOLS) * (PY / TILE_Y);   // setup workgroups per patch
static_assert(PX3 / SCOLS3 == SPP, "2D and 3D patches need the same setup grid");

// The inverse rows of this thread's node, loaded into registers at the top
// of a kernel, before its reductions: they depend on nothing the kernel
// computes, so their memory latency overlaps the reductions' (left to the
// compiler after the barriers, one kernel issued them one at a time,
// waiting for each).
struct TileRows {
  tile4 t[TILE_PLANES];
  float sc[3];
};

__device__ __forceinline__ void tile_load(const tile4 *T4, const float *T4s, long long N,
                                          long long g, bool load, TileRows &R) {
  if (!load) return;
#pragma unroll
  for (int e = 0; e < TILE_PLANES; ++e) R.t[e] = T4[(long long)e * N + g];
#pragma unroll
  for (int r = 0; r < 3; ++r) R.sc[r] = T4s[(long long)r * N + g];
}

// The same rows read from memory at the point of use.
struct TileMem {
  const tile4 *T4;
  const float *T4s;
  long long N, g;
};

__device__ __forceinline__ tile4 tile_row(const TileRows &R, int e) { return R.t[e]; }
__device__ __forceinline__ tile4 tile_row(const TileMem &M, int e) {
  return M.T4[(long long)e * M.N + M.g];
}
__device__ __forceinline__ float tile_scale(const TileRows &R, int r) { return R.sc[r]; }
__device__ __forceinline__ float tile_scale(const TileMem &M, int r) {
  return M.T4s[(long long)r * M.N + M.g];
}

// z = M^-1 x for this thread's node; every thread of the workgroup calls it
// (it holds barriers); tv is LDS of BS * 3 doubles.
template <typename Rows>
__device__ __forceinline__ void tile_apply(const Rows &R, int dim, bool valid, const double x[3],
                                           double z[3], double *tv) {
  int tile, l;
  tile_slot(dim, tile, l);
  double *v = tv + tile * TU;
  __syncthreads();                          // tv free (a previous use is done)
#pragma unroll
  for (int c = 0; c < 3; ++c) v[3 * l + c] = valid ? x[c] : 0.0;
  __syncthreads();
  if (!valid) return;
  // each tile value is read from LDS once and used for all three rows (the
  // LDS traffic of the row-by-row order was 3x, and bound the kernels on
  // large batches); every row still sums in the same order
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < TU / 4; ++q) {
    const double v0 = v[4 * q], v1 = v[4 * q + 1], v2 = v[4 * q + 2], v3 = v[4 * q + 3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const tile4 t = tile_row(R, r * (TU / 4) + q);
      acc[r] += (double)t.x * v0 + (double)t.y * v1 + (double)t.z * v2 + (double)t.w * v3;
    }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) z[r] = acc[r] * (double)tile_scale(R, r);
}

// Node of lane nt (0..63) of a tile-setup workgroup (setup workgroup sub of
// patch (bx, by, bz) of device I): its coordinates (i, j, kk), its tile in
// the group (tq), its place in the tile (l), and which of its neighbours
// x-, x+, y-, y+, z-, z+ lie in the tile (has) at which places (ln).
struct SetupNode {
  int tq, l, i, j, kk;
  bool has[6];
  int ln[6];
};

__device__ __forceinline__ SetupNode setup_node(const DevInfo &I, unsigned bx, unsigned by,
                                                unsigned bz, int sub, int nt) {
  SetupNode S;
  int &tq = S.tq, &l = S.l, &i = S.i, &j = S.j, &kk = S.kk;
  bool *has = S.has;
  int *ln = S.ln;
  if (I.dim == 3) {
    const int lc = nt % SCOLS3, ly = (nt / SCOLS3) % T3Y, lz = nt / (SCOLS3 * T3Y);
    const int il = lc % T3X;
    tq = lc / T3X;
    l = (lz * T3Y + ly) * T3X + il;
    i = (int)bx * PX3 + sub * SCOLS3 + lc;
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
    const int lc = nt % SCOLS, lr = nt / SCOLS;
    const int il = lc % TILE_X;
    tq = lc / TILE_X;
    l = lr * TILE_X + il;
    i = (int)bx * PX + (sub % (PX / SCOLS)) * SCOLS + lc;
    j = (int)by * PY + (sub / (PX / SCOLS)) * TILE_Y + lr;
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
  return S;
}

// Setup: one workgroup of 128 threads per STPW tiles side by side in x (64
// nodes), two threads per node, each owning half of the 48 columns; block
// Gauss-Jordan in FP64 with the 3x3 node blocks as pivots (no row exchange,
// like the node block-Jacobi inverse), in place in LDS.  The tile matrices
// are stored column-major with the node's lane as the fastest index,
// M[(r * 48 + col) * 64 + lane]: a wavefront's own-row accesses are 64
// consecutive doubles (no bank conflicts; rows of a stride of 145 doubles
// put the four tiles of a half-wavefront on the same banks), its pivot-row
// reads are broadcasts.  Each pivot step: every thread inverts the pivot
// block, the tile's threads scale the pivot rows column by column, then
// every other row eliminates over its half of the columns without a branch
// (the pivot block, updated like the rest, is then overwritten with -F P),
// in two 12-column chunks whose reads all come before their writes (the
// pivot rows and the own rows share one array, so a read after a write
// could not be hoisted and each column waited a full LDS round trip).
// Measured per call: 222 -> 73 us on the 200x120 MOSFET, 652 -> 310 us on
// the 64x48x48 tri-gate (one thread per node with row-major rows).
constexpr int HCOL = TU / 2;                // columns per thread (24)
static_assert(HCOL % 3 == 0 && HCOL % 4 == 0, "a pivot block or a quad never straddles halves");
#define TMAT(lane, r, col) M[((r) * TU + (col)) * 64 + (lane)]

// lane (0..63) of the setup workgroup that holds node l of tile tq
__device__ __forceinline__ int setup_lane(int dim, int tq, int l) {
  if (dim == 3) {
    const int il = l % T3X, ly = (l / T3X) % T3Y, lz = l / (T3X * T3Y);
    return lz * (SCOLS3 * T3Y) + ly * SCOLS3 + tq * T3X + il;
  }
  return (l / TILE_X) * SCOLS + tq * TILE_X + l % TILE_X;
}

__global__ void __launch_bounds__(128) k_tile_setup(Where W, long long N, long long NS,
                                                    const double *A, const Scal *scal,
                                                    tile4 *T4, float *T4s, float *CPL,
                                                    int *flag) {
  __shared__ double M[3 * TU * 64];
  __shared__ float amx[2 * 64 * 3];           // row maxima of each half
  const int nt = threadIdx.x % 64, half = threadIdx.x / 64, c0 = half * HCOL;
  const int patch = blockIdx.x / SPP, sub = blockIdx.x % SPP;
  const int d = W.single ? 0 : W.blk2dev[patch];
  if (scal && (!scal[d].active || scal[d].tile_keep)) return;
  const DevInfo I = dev_info(W, d);
  const unsigned local = (unsigned)(patch - I.blk0);
  const unsigned bx = local % (unsigned)I.px, t2 = local / (unsigned)I.px;
  const unsigned by = t2 % (unsigned)I.py, bz = t2 / (unsigned)I.py;
  const SetupNode SN = setup_node(I, bx, by, bz, sub, nt);
  const int tq = SN.tq, l = SN.l, i = SN.i, j = SN.j, kk = SN.kk;
  const bool *has = SN.has;
  const int *ln = SN.ln;
  const bool valid = i < I.nx && j < I.ny && kk < I.nz;
  const long long g = I.off + ((long long)kk * I.ny + j) * I.nx + i;
  // the node's lane in its patch's node workgroups, and its row slot
  const int lane = I.dim == 3
      ? ((kk - (int)bz * PZ3) * PY3 + (j - (int)by * PY3)) * PX3 + (i - (int)bx * PX3)
      : (j - (int)by * PY) * PX + (i - (int)bx * PX);
  const bool blk_tile = tile_black_at(I.dim, i, j, kk);
  const long long ts = tile_row_slot(I.dim, patch, lane, blk_tile);
  if (valid && blk_tile && half == 1 && CPL) {
    // the black node's couplings to the red tiles around its tile
    const long long NC = NS / 2, ci = cpl_index(ts);
    const bool exists[6] = {i > 0, i < I.nx - 1, j > 0, j < I.ny - 1,
                            I.dim == 3 && kk > 0, I.dim == 3 && kk < I.nz - 1};
    int grp = 0;
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
      if (!leaves_tile_at(I.dim, i, j, kk, dir)) continue;
#pragma unroll
      for (int e = 0; e < 5; ++e)
        CPL[(long long)(grp * 5 + e) * NC + ci] =
            exists[dir] ? (float)A[(long long)(9 + 5 * dir + e) * N + g] : 0.0f;
      ++grp;
    }
  }
#pragma unroll
  for (int e = 0; e < HCOL; ++e)
#pragma unroll
    for (int r = 0; r < 3; ++r) TMAT(nt, r, c0 + e) = 0.0;
  __syncthreads();
  if (half == 0) {
    if (valid) {
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) TMAT(nt, r, 3 * l + c) = A[(3 * r + c) * N + g];
#pragma unroll
      for (int dir = 0; dir < 6; ++dir) {
        if (!has[dir]) continue;
        const double *o = A + (long long)(9 + 5 * dir) * N + g;
        const int cb = 3 * ln[dir];
        TMAT(nt, 0, cb) = o[0];
        TMAT(nt, 1, cb) = o[N];
        TMAT(nt, 1, cb + 1) = o[2 * N];
        TMAT(nt, 2, cb) = o[3 * N];
        TMAT(nt, 2, cb + 2) = o[4 * N];
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) TMAT(nt, c, 3 * l + c) = 1.0;
    }
  }
  bool bad = false;
  for (int sp = 0; sp < TILE_NODES; ++sp) {
    const int pc = 3 * sp;
    const int pl = setup_lane(I.dim, tq, sp);                       // pivot node's lane
    __syncthreads();
    const double a0 = TMAT(pl, 0, pc), a1 = TMAT(pl, 0, pc + 1), a2 = TMAT(pl, 0, pc + 2);
    const double a3 = TMAT(pl, 1, pc), a4 = TMAT(pl, 1, pc + 1), a5 = TMAT(pl, 1, pc + 2);
    const double a6 = TMAT(pl, 2, pc), a7 = TMAT(pl, 2, pc + 1), a8 = TMAT(pl, 2, pc + 2);
    const double c0d = a4 * a8 - a5 * a7, c1d = a5 * a6 - a3 * a8, c2d = a3 * a7 - a4 * a6;
    const double det = a0 * c0d + a1 * c1d + a2 * c2d;
    if (l == sp && half == 0 && (!(fabs(det) > 1e-300) || !isfinite(det))) bad = true;
    const double id = 1.0 / det;
    const double P[9] = {c0d * id, (a2 * a7 - a1 * a8) * id, (a1 * a5 - a2 * a4) * id,
                         c1d * id, (a0 * a8 - a2 * a6) * id, (a2 * a3 - a0 * a5) * id,
                         c2d * id, (a1 * a6 - a0 * a7) * id, (a0 * a4 - a1 * a3) * id};
    double F[9];                                // this row's pivot-column block
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) F[3 * r + c] = TMAT(nt, r, pc + c);
    __syncthreads();                            // pivot block read by all
    // pivot rows := P * pivot rows (pivot block := P): the tile's 16 x 2
    // threads take the 48 columns, 24 per half
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int cc = l + TILE_NODES * k;
      if (cc < HCOL) {
        const int col = c0 + cc;
        const double m0 = TMAT(pl, 0, col), m1 = TMAT(pl, 1, col), m2 = TMAT(pl, 2, col);
        double n0 = P[0] * m0 + P[1] * m1 + P[2] * m2;
        double n1 = P[3] * m0 + P[4] * m1 + P[5] * m2;
        double n2 = P[6] * m0 + P[7] * m1 + P[8] * m2;
        const int c = col - pc;
        if (c >= 0 && c < 3) {
          n0 = c == 0 ? P[0] : (c == 1 ? P[1] : P[2]);
          n1 = c == 0 ? P[3] : (c == 1 ? P[4] : P[5]);
          n2 = c == 0 ? P[6] : (c == 1 ? P[7] : P[8]);
        }
        TMAT(pl, 0, col) = n0;
        TMAT(pl, 1, col) = n1;
        TMAT(pl, 2, col) = n2;
      }
    }
    __syncthreads();
    if (l != sp) {
      // two chunks of 12 columns, all reads of a chunk before its writes
#pragma unroll
      for (int ch = 0; ch < HCOL; ch += HCOL / 2) {
        double sv[3][HCOL / 2], mv[3][HCOL / 2];
#pragma unroll
        for (int e = 0; e < HCOL / 2; ++e)
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            sv[r][e] = TMAT(pl, r, c0 + ch + e);
            mv[r][e] = TMAT(nt, r, c0 + ch + e);
          }
#pragma unroll
        for (int e = 0; e < HCOL / 2; ++e)
#pragma unroll
          for (int r = 0; r < 3; ++r)
            TMAT(nt, r, c0 + ch + e) = mv[r][e] - (F[3 * r] * sv[0][e] + F[3 * r + 1] * sv[1][e] +
                                                   F[3 * r + 2] * sv[2][e]);
      }
      if (pc / HCOL == half) {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c)
            TMAT(nt, r, pc + c) = -(F[3 * r] * P[c] + F[3 * r + 1] * P[3 + c] +
                                    F[3 * r + 2] * P[6 + c]);
      }
    }
  }
  __syncthreads();
  // FP16 entries scaled by the row's largest magnitude (entries below ~6e-8
  // of it, far under the preconditioner's accuracy, flush); each half
  // stores its own columns
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double m = 0.0;
#pragma unroll
    for (int e = 0; e < HCOL; ++e) m = fmax(m, fabs(TMAT(nt, r, c0 + e)));
    amx[(half * 64 + nt) * 3 + r] = (float)m;
  }
  __syncthreads();
  if (valid) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const float amax = fmaxf(amx[nt * 3 + r], amx[(64 + nt) * 3 + r]);
      const float sc = amax > 0.0f && isfinite(amax) ? amax : 1.0f;
      const double inv = 1.0 / (double)sc;
      if (half == 0) T4s[(long long)r * NS + ts] = sc;
#pragma unroll
      for (int q = c0 / 4; q < (c0 + HCOL) / 4; ++q) {
        tile4 h;
        h.x = (_Float16)(TMAT(nt, r, 4 * q) * inv);
        h.y = (_Float16)(TMAT(nt, r, 4 * q + 1) * inv);
        h.z = (_Float16)(TMAT(nt, r, 4 * q + 2) * inv);
        h.w = (_Float16)(TMAT(nt, r, 4 * q + 3) * inv);
        T4[(long long)(r * (TU / 4) + q) * NS + ts] = h;
      }
    }
  }
  if (bad) {                              // error flags only; never part of a sum
    atomicOr(flag, 1);
    atomicOr(flag + 1 + d, 1);
  }
}
#undef TMAT

// KSTOP: the late return of a finished device's workgroups (late_stop_ok;
// the loads before it stay in bounds).
#define KPROLOGUE                                   \
  __shared__ double sh[64];                         \
  const int blk = K.bmap ? K.bmap[blockIdx.x] : (int)blockIdx.x; \
  const int d = K.W.single ? 0 : K.W.blk2dev[blk];  \
  Scal &S = K.scal[d];                              \
  const int done0 = S.done;                         \
  const DevInfo I = dev_info(K.W, d);               \
  KARGS_EARLY(I.off, I.nx, I.ny, I.px, I.blk0)      \
  KARGS_EARLY(done0, I.dim, I.py)                   \
  if (!K.late_stop && done0) return;                \
  const bool lead = (blk == I.blk0) && threadIdx.x == 0;
#define KSTOP \
  if (done0) return;

// The kernels that apply the preconditioner come in three variants (PV):
// node block-Jacobi only (no tile code: full occupancy, for batches
// without tile-preconditioned devices); tiles with the 144 inverse entries
// loaded into registers before the reductions (their latency overlaps
// the reductions', at 2 waves per SIMD: best when the grid is small and
// latency-bound); tiles read at the point of use (full occupancy: large
// grids, where more waves in flight hide the latency instead).
enum { PV_NODE = 0, PV_TILE_REG = 1, PV_TILE_MEM = 2 };

// z = M^-1 x at this thread's node where `on` (every thread of the
// workgroup calls it: the tile variant holds barriers)
template <typename T, int PV>
__device__ __forceinline__ void apply_prec(const KArgs<T> &K, const NodeHood &H, int dim,
                                           bool use_tile, bool on, long long ts,
                                           const TileRows &TR, const double x[3], double z[3]) {
  if (!use_tile) {
    if (on) prec3<T>(K.Dinv, K.N, H.g, x, z);
    return;
  }
  __shared__ double tv[BS * 3];
  if (PV == PV_TILE_REG) {
    tile_apply(TR, dim, on, x, z, tv);
  } else {
    const TileMem TM{K.T4, K.T4s, K.NS, ts};
    tile_apply(TM, dim, on, x, z, tv);
  }
}

__device__ __forceinline__ bool tile_black(int dim, const NodeHood &H) {
  return tile_black_at(dim, H.i, H.j, H.kk);
}

// Whether the neighbour in direction dir (x-, x+, y-, y+, z-, z+) lies in
// another tile.
__device__ __forceinline__ bool leaves_tile(int dim, const NodeHood &H, int dir) {
  return leaves_tile_at(dim, H.i, H.j, H.kk, dir);
}

// With red-black ordering one launch applies the tiles of one colour, and
// two lanes share each of their nodes: the lane at the same place in the
// x-adjacent tile (the other colour, lane ^ TILE_X in 2D, ^ T3X in 3D)
// takes the second half of the node's 48 tile columns, so every lane of
// the workgroup works.  The halves are summed with one exchange.
constexpr int QH = TU / 8;                  // quads (4 columns) per half row
static_assert((TILE_X & (TILE_X - 1)) == 0 && (T3X & (T3X - 1)) == 0 &&
                  (PX / TILE_X) % 2 == 0 && (PX3 / T3X) % 2 == 0 && TU % 8 == 0,
              "a tile's x-neighbour in the patch is lane ^ tile width");
__device__ __forceinline__ int mate_mask(int dim) { return dim == 3 ? T3X : TILE_X; }

// The node this lane works for, the columns it takes and its tile's LDS
// slot: its own node (primary: its tile has the colour being applied) or
// its mate's.
struct PairSlot {
  long long slot;                   // the node's tile-row slot
  bool act;
  int q0, vt;
};

__device__ __forceinline__ PairSlot pair_slot(int dim, long long blk, const NodeHood &H,
                                              bool primary, bool on) {
  int tile, l;
  tile_slot(dim, tile, l);
  const int m = mate_mask(dim);
  const int onm = __shfl_xor(on ? 1 : 0, m);
  // the mate's tile has the other colour
  const bool black = tile_black(dim, H);
  PairSlot P;
  P.slot = primary ? tile_row_slot(dim, blk, threadIdx.x, black)
                   : tile_row_slot(dim, blk, threadIdx.x ^ m, !black);
  P.act = primary ? on : onm != 0;
  P.q0 = primary ? 0 : QH;
  P.vt = primary ? tile : (tile ^ 1);
  return P;
}

// The half rows of a lane, loaded into registers before the reductions
// (PV_TILE_REG: the first 3 QH entries of TileRows, whose registers the
// whole rows would take without the red-black ordering), or read at the
// point of use.
__device__ __forceinline__ void half_load(const tile4 *T4, const float *T4s, long long N,
                                          const PairSlot &P, TileRows &R) {
  if (!P.act) return;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int k = 0; k < QH; ++k)
      R.t[r * QH + k] = T4[(long long)(r * (TU / 4) + P.q0 + k) * N + P.slot];
#pragma unroll
  for (int r = 0; r < 3; ++r) R.sc[r] = T4s[(long long)r * N + P.slot];
}

struct HalfMem {
  const tile4 *T4;
  const float *T4s;
  long long N, g;
  int q0;
};

__device__ __forceinline__ tile4 half_row(const TileRows &R, int r, int k) { return R.t[r * QH + k]; }
__device__ __forceinline__ tile4 half_row(const HalfMem &M, int r, int k) {
  return M.T4[(long long)(r * (TU / 4) + M.q0 + k) * M.N + M.g];
}
__device__ __forceinline__ float half_scale(const TileRows &R, int r) { return R.sc[r]; }
__device__ __forceinline__ float half_scale(const HalfMem &M, int r) {
  return M.T4s[(long long)r * M.N + M.g];
}

// z = T^-1 x on the tiles of one colour, two lanes per node: lanes with
// `on` stage their x (their tile has the colour); z is valid on those lanes.
// Every lane calls it (barriers and the exchange).
template <typename Rows>
__device__ __forceinline__ void pair_apply(const Rows &R, const PairSlot &P, int dim, bool on,
                                           const double x[3], double z[3], double *tv) {
  int tile, l;
  tile_slot(dim, tile, l);
  __syncthreads();                          // tv free (a previous use is done)
#pragma unroll
  for (int c = 0; c < 3; ++c) tv[tile * TU + 3 * l + c] = on ? x[c] : 0.0;
  __syncthreads();
  double acc[3] = {0.0, 0.0, 0.0};
  if (P.act) {
    const double *v = tv + P.vt * TU;
#pragma unroll
    for (int k = 0; k < QH; ++k) {
      const int q = P.q0 + k;
      const double v0 = v[4 * q], v1 = v[4 * q + 1], v2 = v[4 * q + 2], v3 = v[4 * q + 3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const tile4 t = half_row(R, r, k);
        acc[r] += (double)t.x * v0 + (double)t.y * v1 + (double)t.z * v2 + (double)t.w * v3;
      }
    }
  }
  const int m = mate_mask(dim);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double other = __shfl_xor(acc[r], m);
    // the node's own half first, then its mate's (the same order on
    // either lane of the pair)
    z[r] = (P.q0 == 0 ? acc[r] + other : other + acc[r]) * (double)half_scale(R, r);
  }
}

// The same for one colour of the red-black ordering (pair_apply).
template <typename T, int PV>
__device__ __forceinline__ void apply_prec_rb(const KArgs<T> &K, int dim, const PairSlot &P,
                                              bool on, const TileRows &HR, const double x[3],
                                              double z[3]) {
  __shared__ double tv[BS * 3];
  if (PV == PV_TILE_REG) {
    pair_apply(HR, P, dim, on, x, z, tv);
  } else {
    const HalfMem HM{K.T4, K.T4s, K.NS, P.slot, P.q0};
    pair_apply(HM, P, dim, on, x, z, tv);
  }
}

// x = 0, r = rhat = p = b (optionally b / nrm), phat = M^-1 p; partial
// (r, r) at quantity 1 (rho_0 = (rhat, r) is the same sum)
template <typename T, int PV>
__device__ __forceinline__ void bicg_init_body(const KArgs<T> &K) {
  KPROLOGUE
  KSTOP
  (void)lead;
  const NodeHood H = hood(I, blk);
  const long long N = K.N;
  if (S.cont) {
    // continue the previous sweep's process on this sweep's rhs: x = 0,
    // r = b / nrm (the FP64 true residual replaces the iterated one), p and
    // phat = M^-1 p rescaled from the previous rhs scale to this one
    // (f = nrm_prev / nrm), rhat kept; partials (r, r) at quantity 1 and
    // (rhat, r), the first rho, at quantity 3.  One device per workgroup:
    // the branch is uniform.
    const double sc = 1.0 / S.nrm, f = S.nrm_prev * sc;
    double q[2] = {0.0, 0.0}, o[2];
    if (H.valid) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const long long ix = c * N + H.g;
        const double bv = K.b[ix] * sc, rh = K.v[VRH][ix];
        K.v[VX][ix] = 0.0;
        K.v[VR][ix] = bv;
        K.v[VP][ix] = f * K.v[VP][ix];
        K.v[VPH][ix] = f * K.v[VPH][ix];
        q[0] += bv * bv;
        q[1] += rh * bv;
      }
    }
    block_sums<2>(q, sh, o);
    if (threadIdx.x == 0) {
      K.part[(long long)K.nblk + blk] = o[0];
      K.part[3 * (long long)K.nblk + blk] = o[1];
    }
    return;
  }
  const bool use_tile = PV != PV_NODE && K.tile;
  // red-black ordering: this launch applies the red tiles only, two lanes
  // per node
  const bool rb = use_tile && K.rb;
  const bool black = tile_black(I.dim, H);
  const bool on = H.valid && !(rb && black);
  const long long tslot = tile_row_slot(I.dim, blk, threadIdx.x, black);
  TileRows TR;
  PairSlot PS{};
  if (rb) {
    PS = pair_slot(I.dim, blk, H, !black, on);
    if (PV == PV_TILE_REG) half_load(K.T4, K.T4s, K.NS, PS, TR);
  } else if (PV == PV_TILE_REG) {
    tile_load(K.T4, K.T4s, K.NS, tslot, use_tile && on, TR);
  }
  const double sc = K.scale_b ? 1.0 / S.nrm : 1.0;
  double b3[3] = {0, 0, 0}, rr = 0.0;
  if (H.valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long ix = c * N + H.g;
      const double bv = K.scale_b ? K.b[ix] * sc : K.b[ix];
      b3[c] = bv;
      K.v[VX][ix] = 0.0;
      K.v[VR][ix] = bv;
      K.v[VRH][ix] = bv;
      K.v[VP][ix] = bv;
      rr += bv * bv;
    }
  }
  double z[3];
  if (rb) apply_prec_rb<T, PV>(K, I.dim, PS, on, TR, b3, z);
  else apply_prec<T, PV>(K, H, I.dim, use_tile, on, tslot, TR, b3, z);
  if (on) {
#pragma unroll
    for (int c = 0; c < 3; ++c) K.v[VPH][c * N + H.g] = z[c];
  }
  const double s = block_sum(rr, sh);
  if (threadIdx.x == 0) K.part[(long long)K.nblk + blk] = s;
}

// v = A phat; partial (rhat, v) at quantity 2 (the vectors are fixed, so
// their addresses arrive with the first kernel-argument loads)
template <typename T>
__global__ void __launch_bounds__(BS) k_bicg_spmv(KArgs<T> K) {
  KARGS_EARLY(K.A, K.v[VPH], K.v[VRH], K.v[VV], K.N)
  KPROLOGUE
  (void)lead;
  const NodeHood H = hood(I, blk);
  const long long N = K.N;
  double s0 = 0.0, y[3] = {0, 0, 0}, a3[3] = {0, 0, 0};
  if (H.valid) {
    spmv_node<T>(K.A, N, H, K.v[VPH], y);
#pragma unroll
    for (int c = 0; c < 3; ++c) a3[c] = K.v[VRH][c * N + H.g];
  }
  KSTOP
  if (H.valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      K.v[VV][c * N + H.g] = y[c];
      s0 += a3[c] * y[c];
    }
  }
  const double a = block_sum(s0, sh);
  if (threadIdx.x == 0) K.part[2 * (long long)K.nblk + blk] = a;
}

// t = A shat; partials (t, s), (t, t), (rhat, s), (rhat, t) at quantities
// 3..6: with them the update kernel forms omega and the next rho
// (rhat, s - omega t) without another reduction
template <typename T>
__global__ void __launch_bounds__(BS) k_bicg_spmv_t(KArgs<T> K) {
  KARGS_EARLY(K.A, K.v[VSH], K.v[VS], K.v[VRH], K.N)
  KPROLOGUE
  (void)lead;
  const NodeHood H = hood(I, blk);
  const long long N = K.N;
  double q[4] = {0.0, 0.0, 0.0, 0.0}, y[3] = {0, 0, 0}, s3[3] = {0, 0, 0}, r3[3] = {0, 0, 0};
  if (H.valid) {
    spmv_node<T>(K.A, N, H, K.v[VSH], y);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      s3[c] = K.v[VS][c * N + H.g];
      r3[c] = K.v[VRH][c * N + H.g];
    }
  }
  KSTOP
  if (H.valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long ix = c * N + H.g;
      K.v[VT][ix] = y[c];
      const double sv = s3[c], rh = r3[c];
      q[0] += y[c] * sv;
      q[1] += y[c] * y[c];
      q[2] += rh * sv;
      q[3] += rh * y[c];
    }
  }
  double o[4];
  block_sums<4>(q, sh, o);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) K.part[(long long)(3 + k) * K.nblk + blk] = o[k];
  }
}

// Stop test on (r, r) (quantity 1); alpha = rho / (rhat, v) (quantity 2);
// s = r - alpha v; shat = M^-1 s
template <typename T, int PV>
__device__ __forceinline__ void bicg_s_body(const KArgs<T> &K) {
  KARGS_EARLY(K.v[VR], K.v[VV], K.T4, K.T4s, K.N)
  KPROLOGUE
  const NodeHood H = hood(I, blk);
  const long long N = K.N;
  const bool use_tile = PV != PV_NODE && K.tile;
  const bool rb = use_tile && K.rb;
  const bool black = tile_black(I.dim, H);
  const bool on = H.valid && !(rb && black);
  const long long tslot = tile_row_slot(I.dim, blk, threadIdx.x, black);
  TileRows TR;
  PairSlot PS{};
  if (rb) {
    PS = pair_slot(I.dim, blk, H, !black, on);
    if (PV == PV_TILE_REG) half_load(K.T4, K.T4s, K.NS, PS, TR);
  } else if (PV == PV_TILE_REG) {
    tile_load(K.T4, K.T4s, K.NS, tslot, use_tile && on, TR);
  }
  double r[3] = {0, 0, 0}, v[3] = {0, 0, 0};
  if (H.valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      r[c] = K.v[VR][c * N + H.g];
      v[c] = K.v[VV][c * N + H.g];
    }
  }
  // the device scalars this kernel reads are written by earlier launches
  // only: read with the vectors, before the reduction
  const int s_it = S.it, s_maxit = S.maxit, s_cont = S.cont;
  const double s_bb = S.bb, s_rho = S.rho2[K.parity], s_tol2 = S.tol2;
  KSTOP
  const bool first = s_it == 0;
  // the first iteration of a continued solve takes rho = (rhat, r) from
  // quantity 3 (after a fresh start rhat = r and rho = (r, r))
  double red[3];
  if (first && s_cont) device_sums<3>(K.part, K.nblk, 1, I, sh, red);
  else device_sums<2>(K.part, K.nblk, 1, I, sh, red);
  const double rr = red[0], rv = red[1];
  const double bb = first ? rr : s_bb;
  const double rho = first ? (s_cont ? red[2] : rr) : s_rho;
  if (rr <= s_tol2 * bb || s_it >= s_maxit || !isfinite(rr) || rho == 0.0) {
    if (lead) {
      S.rr = rr;
      if (first) S.bb = bb;
      if (!(rr <= s_tol2 * bb)) S.status = 1;
      S.done = 1;
    }
    return;
  }
  const double alpha = rho / rv;
  if (!isfinite(alpha) || rv == 0.0) {
    if (lead) {
      S.status = 1;
      S.done = 1;
    }
    return;
  }
  double s3[3] = {0, 0, 0}, z[3];
  if (H.valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      s3[c] = r[c] - alpha * v[c];
      K.v[VS][c * N + H.g] = s3[c];
    }
  }
  if (rb) apply_prec_rb<T, PV>(K, I.dim, PS, on, TR, s3, z);
  else apply_prec<T, PV>(K, H, I.dim, use_tile, on, tslot, TR, s3, z);
  if (on) {
#pragma unroll
    for (int c = 0; c < 3; ++c) K.v[VSH][c * N + H.g] = z[c];
  }
  if (lead) {
    S.alpha = alpha;
    S.rr = rr;
    if (first) {
      S.bb = bb;
      S.rho2[K.parity] = rho;
    }
  }
}

// omega = (t, s) / (t, t); x += alpha phat + omega shat; r = s - omega t;
// the next rho = (rhat, s) - omega (rhat, t), beta, p = r + beta (p -
// omega v), phat = M^-1 p: the end of one iteration fused with the start of
// the next (one launch and one reduction point fewer than the textbook
// order); partial (r, r) at quantity 1 for the next stop test.
template <typename T, int PV>
__device__ __forceinline__ void bicg_u_body(const KArgs<T> &K) {
  KARGS_EARLY(K.v[VX], K.v[VPH], K.v[VSH], K.v[VS], K.v[VT])
  KARGS_EARLY(K.v[VP], K.v[VV], K.T4, K.T4s, K.N)
  KPROLOGUE
  const NodeHood H = hood(I, blk);
  const long long N = K.N;
  const bool use_tile = PV != PV_NODE && K.tile;
  const bool rb = use_tile && K.rb;
  const bool black = tile_black(I.dim, H);
  const bool on = H.valid && !(rb && black);
  const long long tslot = tile_row_slot(I.dim, blk, threadIdx.x, black);
  TileRows TR;
  PairSlot PS{};
  if (rb) {
    PS = pair_slot(I.dim, blk, H, !black, on);
    if (PV == PV_TILE_REG) half_load(K.T4, K.T4s, K.NS, PS, TR);
  } else if (PV == PV_TILE_REG) {
    tile_load(K.T4, K.T4s, K.NS, tslot, use_tile && on, TR);
  }
  double x[3] = {0, 0, 0}, ph[3] = {0, 0, 0}, shh[3] = {0, 0, 0}, s[3] = {0, 0, 0},
         t[3] = {0, 0, 0}, p[3] = {0, 0, 0}, v[3] = {0, 0, 0};
  if (H.valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long ix = c * N + H.g;
      x[c] = K.v[VX][ix];
      ph[c] = K.v[VPH][ix];
      shh[c] = K.v[VSH][ix];
      s[c] = K.v[VS][ix];
      t[c] = K.v[VT][ix];
      p[c] = K.v[VP][ix];
      v[c] = K.v[VV][ix];
    }
  }
  // written by earlier launches only: read with the vectors
  const double alpha = S.alpha, rho_old = S.rho2[K.parity];
  const int s_it = S.it;
  KSTOP
  double red[4];
  device_sums<4>(K.part, K.nblk, 3, I, sh, red);
  const double ts = red[0], tt = red[1], rs = red[2], rt = red[3];
  const double omega = (tt > 0.0) ? ts / tt : 0.0;
  // omega == 0: stagnation; the tiny value keeps beta finite and the next
  // stop test decides between converged and failed
  const double om_b = omega != 0.0 ? omega : 1e-300;
  const double rho_new = rs - omega * rt;
  const double beta = (rho_new / rho_old) * (alpha / om_b);
  double rr = 0.0, pn[3] = {0, 0, 0};
  if (H.valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long ix = c * N + H.g;
      K.v[VX][ix] = x[c] + alpha * ph[c] + omega * shh[c];
      const double r = s[c] - omega * t[c];
      K.v[VR][ix] = r;
      rr += r * r;
      pn[c] = r + beta * (p[c] - om_b * v[c]);
      K.v[VP][ix] = pn[c];
    }
  }
  double z[3];
  if (rb) apply_prec_rb<T, PV>(K, I.dim, PS, on, TR, pn, z);
  else apply_prec<T, PV>(K, H, I.dim, use_tile, on, tslot, TR, pn, z);
  if (on) {
#pragma unroll
    for (int c = 0; c < 3; ++c) K.v[VPH][c * N + H.g] = z[c];
  }
  const double a = block_sum(rr, sh);
  if (threadIdx.x == 0) K.part[(long long)K.nblk + blk] = a;
  if (lead) {
    S.rho2[K.parity ^ 1] = rho_new;
    S.omega = om_b;
    S.it = s_it + 1;
  }
}

// The black tiles of the red-black tile preconditioner (forward block
// Gauss-Seidel over the two colours): z = T^-1 (x - C z_red) on every
// black tile, x = v[rb_src], z = v[rb_dst] (its red tiles written by the
// launch before), C the couplings of the tile's nodes to the neighbouring
// tiles, all red.  Halves the BiCGStab iterations of tile block-Jacobi on
// the equilibrated Newton systems (CPU emulation, 200x120 MOSFET: 152 ->
// 82, 300 -> 138; 32x24x24 tri-gate: 89 -> 47, 100 -> 48) for two more
// launches per iteration.  Each black node's product is shared with the
// red lane beside it (pair_apply).
template <typename T, int PV>
__device__ __forceinline__ void bicg_black_body(const KArgs<T> &K) {
  KPROLOGUE
  KSTOP
  (void)lead;
  (void)sh;
  // a continued solve's first direction is the last solve's, rescaled by
  // the init kernel (uniform per device: one device per workgroup)
  if (K.rb_init && S.cont) return;
  const NodeHood H = hood(I, blk);
  const long long N = K.N;
  const bool black = tile_black(I.dim, H);
  const bool on = H.valid && black;
  const PairSlot PS = pair_slot(I.dim, blk, H, black, on);
  TileRows HR;
  if (PV == PV_TILE_REG) half_load(K.T4, K.T4s, K.NS, PS, HR);
  const double *src = K.v[K.rb_src];
  double *dst = K.v[K.rb_dst];
  double x[3] = {0, 0, 0};
  if (on) {
#pragma unroll
    for (int c = 0; c < 3; ++c) x[c] = src[c * N + H.g];
    // the couplings from the compact FP32 store (k_tile_setup), by
    // out-of-tile direction in direction order
    const long long NC = K.NS / 2;
    const float *o = K.CPL + cpl_index(PS.slot);
    int grp = 0;
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
      if (!leaves_tile(I.dim, H, dir)) continue;
      const int gq = grp++;
      if (!H.nb[dir]) continue;
      const long long gn = H.nbg[dir];
      const float *og = o + (long long)(gq * 5) * NC;
      const double zp = dst[gn], zn = dst[N + gn], zq = dst[2 * N + gn];
      x[0] -= (double)og[0] * zp;
      x[1] -= (double)og[NC] * zp + (double)og[2 * NC] * zn;
      x[2] -= (double)og[3 * NC] * zp + (double)og[4 * NC] * zq;
    }
  }
  double z[3];
  apply_prec_rb<T, PV>(K, I.dim, PS, on, HR, x, z);
  if (on) {
#pragma unroll
    for (int c = 0; c < 3; ++c) dst[c * N + H.g] = z[c];
  }
}

#define TCAD_PV_KERNELS(name, body)                                                        \
  template <typename T>                                                                    \
  __global__ void __launch_bounds__(BS) name##_node(KArgs<T> K) { body<T, PV_NODE>(K); }   \
  template <typename T>                                                                    \
  __global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(1, 2)))         \
  name##_treg(KArgs<T> K) { body<T, PV_TILE_REG>(K); }                                     \
  template <typename T>                                                                    \
  __global__ void __launch_bounds__(BS) name##_tmem(KArgs<T> K) { body<T, PV_TILE_MEM>(K); }

TCAD_PV_KERNELS(k_bicg_init, bicg_init_body)
TCAD_PV_KERNELS(k_bicg_s, bicg_s_body)
TCAD_PV_KERNELS(k_bicg_u, bicg_u_body)
TCAD_PV_KERNELS(k_bicg_black, bicg_black_body)

// Variant of the preconditioner kernels for this batch: node-only when no
// device applies tiles; else registers up to kTileRegBlocks workgroups
// (about two resident per CU at 2 waves per SIMD), memory beyond.  The
// environment variable TCADHIP_TILE_ROWS = reg | mem overrides the choice.
constexpr int kTileRegBlocks = 512;

static int precond_variant(const Batch &b) {
  if (!b.tile) return PV_NODE;
  static const int forced = []() -> int {
    const char *e = std::getenv("TCADHIP_TILE_ROWS");
    if (e && std::strcmp(e, "reg") == 0) return PV_TILE_REG;
    if (e && std::strcmp(e, "mem") == 0) return PV_TILE_MEM;
    return -1;
  }();
  if (forced >= 0) return forced;
  if (b.tile_rows == 1) return PV_TILE_REG;
  if (b.tile_rows == 2) return PV_TILE_MEM;
  return b.nblk <= kTileRegBlocks ? PV_TILE_REG : PV_TILE_MEM;
}

#define TCAD_PV_LAUNCH(name, probe_id)                                                     \
  template <typename T>                                                                    \
  static void launch_##name(Batch &b, const KArgs<T> &K) {                                 \
    const dim3 g(K.grid), t(BS);                                                           \
    hipEvent_t e0 = nullptr, e1 = nullptr;                                                 \
    const bool pr = probe_id >= 0 && probe_events(b, probe_id, e0, e1);                    \
    switch (precond_variant(b)) {                                                          \
      case PV_NODE:                                                                        \
        if (pr) hipExtLaunchKernelGGL(name##_node<T>, g, t, 0, b.stream, e0, e1, 0, K);   \
        else hipLaunchKernelGGL(name##_node<T>, g, t, 0, b.stream, K);                     \
        break;                                                                             \
      case PV_TILE_REG:                                                                    \
        if (pr) hipExtLaunchKernelGGL(name##_treg<T>, g, t, 0, b.stream, e0, e1, 0, K);   \
        else hipLaunchKernelGGL(name##_treg<T>, g, t, 0, b.stream, K);                     \
        break;                                                                             \
      default:                                                                             \
        if (pr) hipExtLaunchKernelGGL(name##_tmem<T>, g, t, 0, b.stream, e0, e1, 0, K);   \
        else hipLaunchKernelGGL(name##_tmem<T>, g, t, 0, b.stream, K);                     \
        break;                                                                             \
    }                                                                                      \
    b.launches++;                                                                          \
  }

TCAD_PV_LAUNCH(k_bicg_init, -1)
TCAD_PV_LAUNCH(k_bicg_s, 1)
TCAD_PV_LAUNCH(k_bicg_u, 3)
TCAD_PV_LAUNCH(k_bicg_black, 4)

// The black tiles' application after a kernel that applied the red ones.
template <typename T>
static void launch_black(Batch &b, KArgs<T> K, int src, int dst, int init) {
  if (!K.tile || !K.rb) return;
  K.rb_src = src;
  K.rb_dst = dst;
  K.rb_init = init;
  launch_k_bicg_black<T>(b, K);
}

// Reset the per-device Krylov state for a new solve; devices not active
// are marked done.  per_dev: each device's own target S.eta2 (not below
// floor2) instead of tol2.
__global__ void k_krylov_reset(Scal *scal, int ndev, double tol2, int maxit, int per_dev,
                               double floor2, int cont) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndev) return;
  Scal &S = scal[d];
  S.cont = (cont && S.active) ? S.cont_ok : 0;
  S.rho2[0] = S.rho2[1] = 1.0;
  S.alpha = 1.0;
  S.omega = 1.0;
  S.it = 0;
  // per_dev (a refinement sweep of a Newton solve): the device's target
  // eta; with a floor (inner solves on the FP32-stored matrix, several
  // sweeps per solve) the reduction that brings this sweep's residual
  // (S.rr, from the refinement check) to eta relative to the Newton rhs
  // (S.bb), not below the floor.  FP64 solves keep eta per sweep: a second
  // sweep there follows a breakdown, and its Newton path is the one the
  // fixtures pin.
  const double t2 = (floor2 > 0.0 && S.rr > 0.0) ? S.eta2 * (S.bb / S.rr) : S.eta2;
  S.tol2 = per_dev ? fmax(t2, floor2) : tol2;
  S.maxit = maxit;
  if (S.active) {
    S.status = 0;
    S.rr = 0.0;
    S.bb = 0.0;
    S.done = 0;
  } else {
    S.done = 1;
  }
}


// Kernel arguments of the refinement's inner solve: rhs = ir_r / ||ir_r||,
// solution in v64[VX].
template <typename T>
static KArgs<T> make_kargs(Batch &b, bool fp32) {
  KArgs<T> K{};
  K.W = make_where(b);
  K.N = b.ntot;
  K.nblk = b.nblk;
  K.part = b.part;
  K.scal = b.d_scal;
  for (int i = 0; i < 9; ++i) K.v[i] = b.v64[i];
  K.A = fp32 ? (const T *)b.A32 : (const T *)b.A64;
  K.Dinv = fp32 ? (const T *)b.Dinv32 : (const T *)b.Dinv64;
  K.T4 = b.T4;
  K.T4s = b.T4s;
  K.NS = (long long)b.nblk * BS;
  K.CPL = b.CPL;
  K.tile = b.tile;
  K.b = b.ir_r;
  K.scale_b = 1;
  K.late_stop = late_stop_ok(b);
  K.rb = b.tile_rb;
  K.bmap = nullptr;
  K.grid = b.nblk;
  return K;
}

// Workgroups of a batch's devices still iterating, for the Krylov launches
// after a check: the devices that converged no longer occupy the grid (an
// early-exit workgroup still costs a dispatch and three dependent loads).
// At the start of a solve (initial) the open devices are the active ones
// (the host copy of the scalars is current then, and the reset marks every
// other device done); after a check, those not done.  Done when the open
// devices hold under 3/4 of the current grid; returns false (grid
// unchanged) otherwise or if the map cannot be placed.
template <typename T>
static bool compact_grid(Batch &b, KArgs<T> &K, bool initial = false) {
  if (b.ndev < 2) return false;
  auto is_open = [&b, initial](int d) {
    return initial ? b.h_scal[d].active != 0 : !b.h_scal[d].done;
  };
  long long open = 0;
  for (int d = 0; d < b.ndev; ++d)
    if (is_open(d)) open += b.info[d].nblk;
  if (open == 0 || 4 * open >= 3 * (long long)K.grid) return false;
  if (!b.d_bmap) {
    if (hipMalloc(&b.d_bmap, sizeof(int) * 2 * (size_t)b.nblk) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    if (hipHostMalloc(&b.h_bmap, sizeof(int) * 2 * (size_t)b.nblk) != hipSuccess) {
      (void)hipGetLastError();
      b.h_bmap = nullptr;
      return false;
    }
  }
  // two halves used in turn: the copy into one may still be pending while
  // the host writes the other (a check drains the stream before the host
  // writes again, so two suffice)
  int *h = b.h_bmap + (b.bmap_half ? b.nblk : 0);
  int *dmap = b.d_bmap + (b.bmap_half ? b.nblk : 0);
  b.bmap_half ^= 1;
  int n = 0;
  for (int d = 0; d < b.ndev; ++d) {
    if (!is_open(d)) continue;
    const DevInfo &I = b.info[d];
    for (int k = 0; k < I.nblk; ++k) h[n++] = I.blk0 + k;
  }
  if (hipMemcpyAsync(dmap, h, sizeof(int) * n, hipMemcpyHostToDevice, b.stream) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  K.bmap = dmap;
  K.grid = n;
  return true;
}

int launch_tile_setup(Batch &b, bool only_active) {
  hipLaunchKernelGGL(k_tile_setup, dim3(SPP * b.nblk), dim3(128), 0, b.stream, make_where(b), b.ntot,
                     (long long)b.nblk * BS,
                     b.A64, only_active ? b.d_scal : nullptr, b.T4, b.T4s, b.CPL, b.d_flag);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

// Events of the next probed launch: true (and e0, e1 set) when `which` is
// the probed kernel; the launch then goes through hipExtLaunchKernelGGL,
// whose events take the kernel's own start and end times.
bool probe_events(Batch &b, int which, hipEvent_t &e0, hipEvent_t &e1) {
  if (b.probe != which) return false;
  while (b.probe_ev.size() < b.probe_used + 2) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) { b.probe = -1; return false; }
    b.probe_ev.push_back(e);
  }
  e0 = b.probe_ev[b.probe_used];
  e1 = b.probe_ev[b.probe_used + 1];
  b.probe_used += 2;
  return true;
}

template <typename T>
static void launch_iteration(Batch &b, KArgs<T> K, int parity) {
  const dim3 g(K.grid), t(BS);
  K.parity = parity;
  hipEvent_t e0, e1;
  if (probe_events(b, 0, e0, e1))
    hipExtLaunchKernelGGL(k_bicg_spmv<T>, g, t, 0, b.stream, e0, e1, 0, K);
  else
    hipLaunchKernelGGL(k_bicg_spmv<T>, g, t, 0, b.stream, K);
  launch_k_bicg_s<T>(b, K);
  launch_black<T>(b, K, VS, VSH, 0);
  if (probe_events(b, 2, e0, e1))
    hipExtLaunchKernelGGL(k_bicg_spmv_t<T>, g, t, 0, b.stream, e0, e1, 0, K);
  else
    hipLaunchKernelGGL(k_bicg_spmv_t<T>, g, t, 0, b.stream, K);
  launch_k_bicg_u<T>(b, K);
  launch_black<T>(b, K, VP, VPH, 0);
  b.launches += 2;
}

// Blocks of kGraphIts BiCGStab iterations (an even count, so the iteration
// parity is the same at every block start) captured once per kernel-argument
// set as a HIP graph and replayed on the batch stream: one graph launch
// instead of 4 kGraphIts (6 kGraphIts with red-black tiles) kernel launches
// from the host.
constexpr int kGraphIts = 16;

template <typename T>
static hipGraphExec_t krylov_graph(Batch &b, const KArgs<T> &K) {
  const int variant = precond_variant(b) * 2 + (sizeof(T) == 4 ? 1 : 0);
  // the key: every argument field, written into zeroed storage (no padding
  // bytes of unknown value)
  std::vector<unsigned char> key(256, 0);
  {
    unsigned char *q = key.data();
    auto put = [&q](const void *v, size_t n) {
      std::memcpy(q, v, n);
      q += n;
    };
    put(&K.W.info, sizeof K.W.info); put(&K.W.blk2dev, sizeof K.W.blk2dev);
    put(&K.W.single, sizeof K.W.single); put(&K.N, sizeof K.N); put(&K.nblk, sizeof K.nblk);
    put(&K.A, sizeof K.A); put(&K.Dinv, sizeof K.Dinv); put(&K.T4, sizeof K.T4);
    put(&K.T4s, sizeof K.T4s); put(&K.tile, sizeof K.tile); put(K.v, sizeof K.v);
    put(&K.b, sizeof K.b); put(&K.part, sizeof K.part); put(&K.scal, sizeof K.scal);
    put(&K.scale_b, sizeof K.scale_b); put(&K.late_stop, sizeof K.late_stop);
    put(&K.rb, sizeof K.rb);
  }
  for (const auto &g : b.graphs)
    if (g.variant == variant && g.key == key) return g.exec;
  if (b.graphs.size() >= 8) {             // bounded cache: drop the oldest
    (void)hipGraphExecDestroy(b.graphs.front().exec);
    (void)hipGraphDestroy(b.graphs.front().graph);
    b.graphs.erase(b.graphs.begin());
  }
  if (!b.capture_stream &&
      hipStreamCreateWithFlags(&b.capture_stream, hipStreamNonBlocking) != hipSuccess)
    return nullptr;
  hipStream_t user = b.stream;
  const long long launches = b.launches;
  b.stream = b.capture_stream;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool ok = hipStreamBeginCapture(b.capture_stream, hipStreamCaptureModeRelaxed) == hipSuccess;
  if (ok) {
    for (int i = 0; i < kGraphIts; ++i) launch_iteration<T>(b, K, i & 1);
    ok = hipStreamEndCapture(b.capture_stream, &graph) == hipSuccess && graph;
  }
  b.stream = user;
  b.launches = launches;
  if (ok) ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
  if (!ok) {
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    b.use_graph = 0;                    // capture refused: plain launches from now on
    return nullptr;
  }
  b.graphs.push_back({key, variant, graph, exec});
  return exec;
}

// Host driver: reset, init, then chunks of iterations with one scalar
// read-back per chunk until every device is done or maxit is reached
// (one extra stop-test kernel then records the final state).
template <typename T>
static int run_bicgstab(Batch &b, const KArgs<T> &K, double rtol, int maxit, int chunk,
                        bool per_dev = false, double floor = 0.0, bool cont = false,
                        double target = 0.0, int slot = -1) {
  hipLaunchKernelGGL(k_krylov_reset, dim3((b.ndev + 63) / 64), dim3(64), 0, b.stream, b.d_scal,
                     b.ndev, rtol * rtol, maxit, per_dev ? 1 : 0, floor * floor, cont ? 1 : 0);
  b.launches++;
  // its grid: the workgroups of the devices this solve selects (the host
  // copy of the scalars is current on entry; the reset marks the others
  // done), shrinking to those still iterating as devices finish
  KArgs<T> KL = K;
  compact_grid<T>(b, KL, true);
  launch_k_bicg_init<T>(b, KL);
  launch_black<T>(b, KL, VP, VPH, 1);
  // Iterations are issued in chunks between host checks of the stop flags;
  // a check costs a host round trip with the stream drained, an iteration
  // issued after convergence four (six with red-black tiles) early-exit
  // launches.  Consecutive Newton
  // solves need similar counts per digit of residual reduction, so the
  // first chunk is 3/4 of the last solve's count, scaled down by the ratio
  // of the two targets' logarithms when this one's is looser (a loose
  // forcing solve after a tight one needs a fraction of its iterations),
  // then 16, 32, 64 (without a history: chunk, doubling to 64).
  if (!(target > 0.0)) target = rtol;
  const double lt = std::log(std::min(std::max(target, 1e-300), 0.5));
  if (slot >= Batch::kSlots) slot = -1;
  // the k-th Newton solve of the last bias point when there is one (9/10
  // of its count: a short first chunk costs one more round trip, a long one
  // early-exit launches), else the last solve's (3/4 of it)
  const bool by_slot = slot >= 0 && b.slot_hint[slot] > 0;
  const int hint = by_slot ? b.slot_hint[slot] : b.kry_hint;
  const double hint_lt = by_slot ? b.slot_lt[slot] : b.kry_hint_lt;
  double predicted = (by_slot ? 0.9 : 0.75) * (double)hint;
  // only ever scaled down: a continued refinement sweep needs far fewer
  // iterations than a fresh solve to the same target, and a chunk issued
  // past convergence costs four to six launches per iteration
  if (hint > 0 && hint_lt < 0.0) predicted *= std::min(1.0, lt / hint_lt);
  int it = 0;
  int next = hint > 0 ? std::max(chunk, (int)predicted) : chunk;
  int step = hint > 0 ? 16 : chunk;
  while (true) {
    int c = 0;
    // graph replays cover the whole grid (their kernel arguments are
    // captured): only while the grid is not compacted
    if (b.use_graph && b.probe < 0 && !KL.bmap) {
      if ((it & 1) && c < next && it < maxit) {
        launch_iteration<T>(b, KL, it & 1);
        ++c;
        ++it;
      }
      while (next - c >= kGraphIts && it + kGraphIts <= maxit) {
        hipGraphExec_t g = krylov_graph<T>(b, KL);
        if (!g) break;
        TCAD_HIP_CHECK(hipGraphLaunch(g, b.stream));
        b.launches += (KL.tile && KL.rb ? 6 : 4) * kGraphIts;
        c += kGraphIts;
        it += kGraphIts;
      }
    }
    for (; c < next && it < maxit; ++c, ++it) launch_iteration<T>(b, KL, it & 1);
    if (it >= maxit) {
      KArgs<T> K1 = KL;
      K1.parity = it & 1;
      launch_k_bicg_s<T>(b, K1);
    }
    TCAD_HIP_CHECK(hipGetLastError());
    int rc = sync_scalars(b);
    if (rc) return rc;
    int open = 0;
    // the iterations the open devices still need at their average rate of
    // residual reduction so far: the next chunk is that plus a margin, at
    // most the geometric step (an underestimate costs one more round trip,
    // an overestimate four to six early-exit launches per iteration)
    double need = 0.0;
    bool rated = true;
    for (int d = 0; d < b.ndev; ++d) {
      const Scal &S = b.h_scal[d];
      if (S.done) continue;
      ++open;
      const double q = S.rr / S.bb, goal = S.tol2;
      if (S.it > 0 && q > 0.0 && q < 1.0 && goal > 0.0 && std::isfinite(q)) {
        const double rate = std::log(q) / S.it;
        need = std::max(need, std::log(goal / q) / rate);
      } else {
        rated = false;
      }
    }
    if (open == 0 || it >= maxit) break;
    compact_grid<T>(b, KL);
    next = step;
    if (rated && need > 0.0)
      next = std::min(step, std::max(4, (int)std::ceil(need * 1.125) + 2));
    if (step < 64) step *= 2;
  }
  // a solve stopped by its cap (or failed) says nothing about the next one's
  // count: no hint then
  int most = 0, failed = it >= maxit ? 1 : 0;
  for (int d = 0; d < b.ndev; ++d) {
    most = std::max(most, b.h_scal[d].it);
    if (b.h_scal[d].active) failed |= b.h_scal[d].status;
  }
  b.kry_hint = failed ? 0 : most;
  b.kry_hint_lt = lt;
  if (slot >= 0) {
    b.slot_hint[slot] = failed ? 0 : most;
    b.slot_lt[slot] = lt;
  }
  return TCAD_OK;
}

// Plain BiCGStab on rhs ir_r (scaled by 1/nrm, nrm set by the caller).
int bicgstab(Batch &b, bool fp32, double rtol, int maxit, bool, int chunk) {
  if (fp32) return run_bicgstab<float>(b, make_kargs<float>(b, true), rtol, maxit, chunk);
  return run_bicgstab<double>(b, make_kargs<double>(b, false), rtol, maxit, chunk);
}

// One stencil SpMV launch as issued inside BiCGStab (v = A phat with the
// (rhat, v) partial); used by the bench's event timing.
int launch_spmv_only(Batch &b, bool fp32) {
  if (fp32) {
    KArgs<float> K = make_kargs<float>(b, true);
    hipLaunchKernelGGL(k_bicg_spmv<float>, dim3(b.nblk), dim3(BS), 0, b.stream, K);
  } else {
    KArgs<double> K = make_kargs<double>(b, false);
    hipLaunchKernelGGL(k_bicg_spmv<double>, dim3(b.nblk), dim3(BS), 0, b.stream, K);
  }
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

__global__ void k_clear_done(Scal *scal, int ndev) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d < ndev) scal[d].done = 0;
}

// Timing only: a BiCGStab state whose stop test never fires.
__global__ void k_time_prep(Scal *scal, int ndev) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndev) return;
  scal[d].done = 0;
  scal[d].tol2 = 0.0;
  scal[d].it = 1;
  scal[d].maxit = 1 << 30;
}

int launch_vector_kernel_only(Batch &b, int which) {
  hipLaunchKernelGGL(k_time_prep, dim3((b.ndev + 63) / 64), dim3(64), 0, b.stream, b.d_scal,
                     b.ndev);
  KArgs<double> K = make_kargs<double>(b, false);
  K.parity = 0;
  if (which == 3) launch_k_bicg_u<double>(b, K);
  else launch_k_bicg_s<double>(b, K);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

int clear_done(Batch &b) {
  hipLaunchKernelGGL(k_clear_done, dim3((b.ndev + 63) / 64), dim3(64), 0, b.stream, b.d_scal,
                     b.ndev);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

// ---------------------------------------------------------------------------
// FP64 iterative refinement.
// ---------------------------------------------------------------------------
struct IRArgs {
  Where W;
  long long N;
  int nblk;
  const double *A64;
  const double *bs;
  double *y64;
  double *r64;
  const double *xin;                // inner solution
  double *part;
  Scal *scal;
};

// r64 = bs - A64 y64 (y64 = 0 when `first`), partials (r,r), (b,b)
__global__ void __launch_bounds__(BS) k_ir_residual(IRArgs R, int first) {
  __shared__ double sh[64];
  const int d = dev_index(R.W);
  const DevInfo I = dev_info(R.W, d);
  Scal &S = R.scal[d];
  if (!S.active) return;
  const NodeHood H = hood(I, blockIdx.x);
  const long long N = R.N;
  double rr = 0.0, bb = 0.0;
  if (H.valid) {
    double y[3] = {0.0, 0.0, 0.0};
    if (!first) spmv_node<double>(R.A64, N, H, R.y64, y);
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
    R.part[blockIdx.x] = a;
    R.part[(long long)R.nblk + blockIdx.x] = b2;
  }
}

// One workgroup per device: refinement stop test; nrm for the inner rhs.
__global__ void k_ir_check(IRArgs R, int ndev, double tol2, int per_dev) {
  __shared__ double sh[64];
  const int d = blockIdx.x;
  if (d >= ndev) return;
  const DevInfo I = dev_info(R.W, d);
  Scal &S = R.scal[d];
  if (!S.active) return;
  double red[2];
  device_sums<2>(R.part, R.nblk, 0, I, sh, red);
  const double rr = red[0], bb = red[1];
  if (threadIdx.x == 0) {
    S.nrm_prev = S.nrm;
    S.nrm = sqrt(rr);
    S.rr = rr;
    S.bb = bb;
    const double t2 = per_dev ? S.eta2 : tol2;
    if (rr <= t2 * bb || S.ir_it >= S.ir_maxsweep || !isfinite(rr)) {
      S.active = 0;
      S.status = (rr <= t2 * bb) ? 0 : 2;
    }
  }
}

// y += ||r|| d for the devices whose inner solve did not diverge: a sweep
// that ended with a larger residual than it started with (S.rr > S.bb, the
// inner solve's own figures) would only move y away; the next sweep then
// starts again from the same y
__global__ void __launch_bounds__(BS) k_ir_update(IRArgs R) {
  const int d = dev_index(R.W);
  const DevInfo I = dev_info(R.W, d);
  Scal &S = R.scal[d];
  if (!S.active) return;
  const NodeHood H = hood(I, blockIdx.x);
  const bool diverged = !(S.rr <= S.bb);
  if (H.valid && !diverged) {
    const long long N = R.N;
    const double nrm = S.nrm;
#pragma unroll
    for (int c = 0; c < 3; ++c) R.y64[c * N + H.g] += nrm * R.xin[c * N + H.g];
  }
  if (blockIdx.x == I.blk0 && threadIdx.x == 0) {
    S.kry_total += S.it;
    S.ir_it += 1;
    // only a solve that converged without breakdown is continued
    S.cont_ok = (!diverged && S.status == 0) ? 1 : 0;
  }
}

// restart: a new system (y = 0), whose first sweep starts a fresh process
__global__ void k_ir_setup(Scal *scal, int ndev, int maxsweep, int restart) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndev) return;
  scal[d].saved_active = scal[d].active;
  scal[d].ir_it = 0;
  scal[d].ir_maxsweep = maxsweep;
  if (restart) scal[d].cont_ok = 0;
}

__global__ void k_ir_restore(Scal *scal, int ndev) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d < ndev) scal[d].active = scal[d].saved_active;
}

int ir_solve(Batch &b, bool fp32, bool restart, double rtol, int maxsweep, int inner_maxit,
             double inner_rtol, bool per_dev, bool fallback) {
  IRArgs R{};
  R.W = make_where(b);
  R.N = b.ntot; R.nblk = b.nblk;
  R.A64 = b.A64; R.bs = b.bs; R.y64 = b.y64; R.r64 = b.ir_r;
  R.xin = b.v64[VX];
  R.part = b.part; R.scal = b.d_scal;
  const dim3 gd((b.ndev + 63) / 64), td(64);
  hipLaunchKernelGGL(k_ir_setup, gd, td, 0, b.stream, b.d_scal, b.ndev, maxsweep,
                     restart ? 1 : 0);
  b.launches++;
  const int tile_mode = b.tile;
  // continued sweeps keep phat = M^-1 p: only while the preconditioner is
  // the one it was formed with
  const bool cont_mode = b.ir_cont == 2 || (b.ir_cont == 1 && fp32);
  bool cont = cont_mode;
  struct Restore {
    Batch &b;
    int t;
    ~Restore() { b.tile = t; }
  } restore{b, tile_mode};
  for (int sweep = 0; sweep <= maxsweep; ++sweep) {
    hipLaunchKernelGGL(k_ir_residual, dim3(b.nblk), dim3(BS), 0, b.stream, R,
                       (restart && sweep == 0) ? 1 : 0);
    hipLaunchKernelGGL(k_ir_check, dim3(b.ndev), dim3(BS), 0, b.stream, R, b.ndev, rtol * rtol,
                       per_dev ? 1 : 0);
    b.launches += 2;
    TCAD_HIP_CHECK(hipGetLastError());
    int rc = sync_scalars(b);
    if (rc) return rc;
    int open = 0;
    for (int d = 0; d < b.ndev; ++d) open += b.h_scal[d].active ? 1 : 0;
    if (open == 0) break;
    // per-device targets: the inner solve aims at the device's own target,
    // not below inner_rtol (the FP32-stored matrix's limit); the tightest
    // of them (as k_krylov_reset forms them) sizes the first chunk
    const double floor = fp32 ? inner_rtol : 0.0;
    double target = inner_rtol;
    if (per_dev) {
      target = 1.0;
      for (int d = 0; d < b.ndev; ++d) {
        const Scal &S = b.h_scal[d];
        if (!S.active) continue;
        const double t2 = (floor > 0.0 && S.rr > 0.0) ? S.eta2 * (S.bb / S.rr) : S.eta2;
        target = std::min(target, std::sqrt(std::max(t2, floor * floor)));
      }
    }
    // the first sweep of a Newton solve keeps its own hint per Newton index
    const int slot = (restart && sweep == 0) ? b.kry_slot : -1;
    rc = fp32 ? run_bicgstab<float>(b, make_kargs<float>(b, true), inner_rtol, inner_maxit, 8,
                                    per_dev, inner_rtol, cont, target, slot)
              : run_bicgstab<double>(b, make_kargs<double>(b, false), inner_rtol, inner_maxit, 8,
                                     per_dev, 0.0, cont, target, slot);
    if (rc) return rc;
    // a tile-preconditioned sweep that broke down or diverged (FP16 tile
    // inverses of a nearly singular tile, as some Crank-Nicolson steps at
    // reverse bias assemble) is followed by sweeps with the FP64 node
    // block-Jacobi, the reference's own preconditioner, for the rest of
    // this solve
    // (a continued solve that breaks down is followed by a fresh one first:
    // k_ir_update clears its cont_ok)
    if (b.tile) {
      bool broke = false;
      for (int d = 0; d < b.ndev; ++d)
        broke |= b.h_scal[d].active && b.h_scal[d].status != 0 && !b.h_scal[d].cont;
      if (broke) {
        b.tile = 0;
        cont = false;
      }
    }
    hipLaunchKernelGGL(k_ir_update, dim3(b.nblk), dim3(BS), 0, b.stream, R);
    b.launches++;
  }
  // a device whose refinement stopped unconverged (its last check, read
  // back above, set status 2) above the acceptance residual the Newton
  // check applies (linsolve.py:973-982: max(rtol, 1e-6), or the device's
  // forcing target) is solved once more by the fallback method, for this
  // Newton iteration only (device3d.py:2154-2168 falls back to a direct
  // solve there); below it the step is taken as the reference takes an
  // accepted iterative solve
  if (fallback && b.lin_fallback) {
    const double acc = std::max(rtol, 1e-6);
    for (int d = 0; d < b.ndev; ++d) {
      Scal &S = b.h_scal[d];
      if (!S.saved_active || S.status != 2) continue;
      if (S.rr <= std::max(S.eta2, acc * acc) * S.bb) continue;
      const double target = per_dev ? std::sqrt(std::max(S.eta2, rtol * rtol)) : rtol;
      int32_t its = 0;
      double res = 0.0;
      const int rc = linear_fallback(b, d, target, &its, &res);
      if (rc != TCAD_OK && rc != TCAD_E_NOT_CONVERGED) return rc;
      if (rc == TCAD_OK) {
        // the device's scalars as if its refinement had ended there
        S.rr = res * res * S.bb;
        S.status = 0;
        S.kry_total += its;
        TCAD_HIP_CHECK(hipMemcpyAsync(b.d_scal + d, &S, sizeof(Scal), hipMemcpyHostToDevice,
                                      b.stream));
        TCAD_HIP_CHECK(hipStreamSynchronize(b.stream));
      }
    }
  }
  hipLaunchKernelGGL(k_ir_restore, gd, td, 0, b.stream, b.d_scal, b.ndev);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

// du = Cc * y : (psi, n, p) column scale (1, n, p)
__global__ void __launch_bounds__(BS) k_unscale(Where W, long long N, const double *y,
                                                 const double *n, const double *p, double *du) {
  const int d = dev_index(W);
  const DevInfo I = dev_info(W, d);
  const NodeHood H = hood(I, blockIdx.x);
  if (!H.valid) return;
  du[H.g] = y[H.g];
  du[N + H.g] = n[H.g] * y[N + H.g];
  du[2 * N + H.g] = p[H.g] * y[2 * N + H.g];
}

int launch_unscale_state(Batch &b, const double *n, const double *p, double *du) {
  hipLaunchKernelGGL(k_unscale, dim3(b.nblk), dim3(BS), 0, b.stream, make_where(b), b.ntot,
                     b.y64, n, p, du);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

int launch_unscale(Batch &b, const double *n, const double *p, double *du) {
  return launch_unscale_state(b, n, p, du);
}

// ---------------------------------------------------------------------------
// PCG (tile block-Jacobi, or Jacobi) for the SPD equilibrium Poisson system.
//   vectors veq: 0 x, 1 r, 2 z, 3 p, 4 q
// ---------------------------------------------------------------------------
struct PArgs {
  Where W;
  long long N;
  int nblk;
  const double *A;
  const double *b;
  double *x, *r, *z, *p, *q;
  double *p2;                       // second buffer of p (by iteration parity)
  double *part;
  Scal *scal;
  const double *P16;                // tile-inverse rows (tile != 0)
  int tile;                         // 1: tile block-Jacobi, 0: Jacobi
  int parity;
  int late_stop;                    // as KArgs::late_stop
};

// The equilibrium Poisson matrix is SPD and scalar per node: its tile
// block-Jacobi preconditioner is the exact inverse of the matrix restricted
// to every 16-node tile (the coupled solver's tiles), FP64 (the inverse's
// entries span the equilibrium's dynamic range), node k keeping its row,
// P16[c * N + g] for the tile's node c.  Measured PCG iterations to 1e-12
// on the equilibrium Newton systems (CPU probe): 200x120 MOSFET ~520
// (Jacobi) -> ~260; 32x24x24 tri-gate 35-58 -> 22-34 (2 x 2 x 4 tiles);
// the batch's uniformly meshed diodes 140 -> 160-180 (Jacobi suits them
// slightly better; they spend under 3% of their time here).
__global__ void __launch_bounds__(64) k_pcg_tile_setup(Where W, long long N, const double *A,
                                                       const Scal *scal, double *P16, int *flag) {
  __shared__ double M[TILE_NODES * 64];     // M[col * 64 + lane]
  const int nt = threadIdx.x;
  const int patch = blockIdx.x / SPP, sub = blockIdx.x % SPP;
  const int d = W.single ? 0 : W.blk2dev[patch];
  if (!scal[d].active) return;
  const DevInfo I = dev_info(W, d);
  const unsigned local = (unsigned)(patch - I.blk0);
  const unsigned bx = local % (unsigned)I.px, t2 = local / (unsigned)I.px;
  const unsigned by = t2 % (unsigned)I.py, bz = t2 / (unsigned)I.py;
  const SetupNode SN = setup_node(I, bx, by, bz, sub, nt);
  const int tq = SN.tq, l = SN.l;
  const bool valid = SN.i < I.nx && SN.j < I.ny && SN.kk < I.nz;
  const long long g = I.off + ((long long)SN.kk * I.ny + SN.j) * I.nx + SN.i;
#pragma unroll
  for (int c = 0; c < TILE_NODES; ++c) M[c * 64 + nt] = 0.0;
  if (valid) {
    M[l * 64 + nt] = A[g];
#pragma unroll
    for (int dir = 0; dir < 6; ++dir)
      if (SN.has[dir]) M[SN.ln[dir] * 64 + nt] = A[(1 + dir) * N + g];
  } else {
    M[l * 64 + nt] = 1.0;
  }
  bool bad = false;
  for (int sp = 0; sp < TILE_NODES; ++sp) {
    const int pl = setup_lane(I.dim, tq, sp);
    __syncthreads();
    const double piv = M[sp * 64 + pl];
    const double f = M[sp * 64 + nt];
    if (l == sp && (!(fabs(piv) > 1e-300) || !isfinite(piv))) bad = true;
    const double pinv = 1.0 / piv;
    __syncthreads();
    // pivot row /= pivot (its pivot entry := 1 / pivot), one column per
    // thread of the tile
    M[l * 64 + pl] = l == sp ? pinv : M[l * 64 + pl] * pinv;
    __syncthreads();
    if (l != sp) {
      double sv[TILE_NODES], mv[TILE_NODES];
#pragma unroll
      for (int c = 0; c < TILE_NODES; ++c) {
        sv[c] = M[c * 64 + pl];
        mv[c] = M[c * 64 + nt];
      }
#pragma unroll
      for (int c = 0; c < TILE_NODES; ++c) M[c * 64 + nt] = mv[c] - f * sv[c];
      M[sp * 64 + nt] = -f * pinv;
    }
  }
  __syncthreads();
  if (valid) {
#pragma unroll
    for (int c = 0; c < TILE_NODES; ++c) P16[(long long)c * N + g] = M[c * 64 + nt];
  }
  if (bad) {                                // error flags only; never part of a sum
    atomicOr(flag, 1);
    atomicOr(flag + 1 + d, 1);
  }
}

int launch_pcg_tile_setup(Batch &b) {
  hipLaunchKernelGGL(k_pcg_tile_setup, dim3(SPP * b.nblk), dim3(64), 0, b.stream, make_where(b),
                     b.ntot, (const double *)b.Aeq, (const Scal *)b.d_scal, b.P16, b.d_flag);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

// z = M^-1 x at this thread's node (every thread of the workgroup calls it
// when P.tile: it holds barriers; tv is LDS of BS doubles)
__device__ __forceinline__ double pcg_prec(const PArgs &P, int dim, bool valid, long long g,
                                           double x, double *tv) {
  if (!P.tile) return valid ? x / P.A[g] : 0.0;
  int tile, l;
  tile_slot(dim, tile, l);
  double *v = tv + tile * TILE_NODES;
  __syncthreads();                          // tv free (a previous use is done)
  v[l] = valid ? x : 0.0;
  __syncthreads();
  if (!valid) return 0.0;
  double z = 0.0;
#pragma unroll
  for (int c = 0; c < TILE_NODES; ++c) z += P.P16[(long long)c * P.N + g] * v[c];
  return z;
}

__device__ __forceinline__ double pspmv(const double *A, long long N, const NodeHood &H,
                                        const double *x) {
  double y = A[H.g] * x[H.g];
#pragma unroll
  for (int dir = 0; dir < 6; ++dir)
    if (H.nb[dir]) y += A[(1 + dir) * N + H.g] * x[H.nbg[dir]];
  return y;
}

#define PPROLOGUE                                   \
  __shared__ double sh[64];                    \
  const int d = dev_index(P.W);                     \
  Scal &S = P.scal[d];                              \
  const int done0 = S.done;                         \
  const DevInfo I = dev_info(P.W, d);               \
  KARGS_EARLY(I.off, I.nx, I.ny, I.px, I.blk0)      \
  KARGS_EARLY(done0, I.dim, I.py)                   \
  if (!P.late_stop && done0) return;                \
  const bool lead = (blockIdx.x == I.blk0) && threadIdx.x == 0;

__global__ void __launch_bounds__(BS) k_pcg_init(PArgs P) {
  PPROLOGUE
  KSTOP
  (void)lead;
  const NodeHood H = hood(I, blockIdx.x);
  __shared__ double tv[BS];
  double rz = 0.0, rr = 0.0;
  const double bv = H.valid ? P.b[H.g] : 0.0;
  const double zv = pcg_prec(P, I.dim, H.valid, H.g, bv, tv);
  if (H.valid) {
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

// Reduce (r, z), (r, r) -> stop test, beta; p = z + beta p_old at this
// node and, on the fly, at its neighbours (p is double-buffered by
// parity, so no workgroup overwrites a value another still reads);
// q = A p; partial (p, q).  The direction update and the product share a
// kernel: two launches per iteration.
__global__ void __launch_bounds__(BS) k_pcg_ab(PArgs P) {
  KARGS_EARLY(P.z, P.p, P.p2, P.A, P.N)
  PPROLOGUE
  const NodeHood H = hood(I, blockIdx.x);
  const long long N = P.N;
  const double *pin = P.parity ? P.p2 : P.p;
  double *pout = P.parity ? P.p : P.p2;
  double zc = 0.0, pc = 0.0, zn[6] = {0, 0, 0, 0, 0, 0}, pn[6] = {0, 0, 0, 0, 0, 0};
  if (H.valid) {
    zc = P.z[H.g];
    pc = pin[H.g];
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
      if (!H.nb[dir]) continue;
      zn[dir] = P.z[H.nbg[dir]];
      pn[dir] = pin[H.nbg[dir]];
    }
  }
  // written by earlier launches only: read with the vectors
  const int s_it = S.it, s_maxit = S.maxit;
  const double s_bb = S.bb, s_tol2 = S.tol2, rho_old = S.rho2[P.parity ^ 1];
  KSTOP
  double red[2];
  device_sums<2>(P.part, P.nblk, 0, I, sh, red);
  const double rz = red[0], rr = red[1];
  const double bb = (s_it == 0) ? rr : s_bb;
  if (rr <= s_tol2 * bb || s_it >= s_maxit || !isfinite(rr)) {
    if (lead) {
      S.rr = rr;
      if (s_it == 0) S.bb = bb;
      if (!(rr <= s_tol2 * bb)) S.status = 1;
      S.done = 1;
    }
    return;
  }
  const double beta = rz / rho_old;
  double pq = 0.0;
  if (H.valid) {
    const double p0 = zc + beta * pc;
    double q = P.A[H.g] * p0;
#pragma unroll
    for (int dir = 0; dir < 6; ++dir)
      if (H.nb[dir]) q += P.A[(1 + dir) * N + H.g] * (zn[dir] + beta * pn[dir]);
    pout[H.g] = p0;
    P.q[H.g] = q;
    pq = p0 * q;
  }
  const double a = block_sum(pq, sh);
  if (threadIdx.x == 0) P.part[2 * (long long)P.nblk + blockIdx.x] = a;
  if (lead) {
    S.rho2[P.parity] = rz;
    S.rr = rr;
    if (s_it == 0) S.bb = bb;
  }
}

// Reduce (p, q) -> alpha; x += alpha p; r -= alpha q; z = r / diag;
// partials (r, z), (r, r)
__global__ void __launch_bounds__(BS) k_pcg_c(PArgs P) {
  KARGS_EARLY(P.x, P.r, P.q, P.p, P.p2)
  PPROLOGUE
  const NodeHood H = hood(I, blockIdx.x);
  const double *pcur = P.parity ? P.p : P.p2;
  __shared__ double tv[BS];
  double x = 0, pv = 0, r = 0, q = 0;
  if (H.valid) {
    x = P.x[H.g];
    pv = pcur[H.g];
    r = P.r[H.g];
    q = P.q[H.g];
  }
  const double rho = S.rho2[P.parity];            // written by k_pcg_ab
  const int s_it = S.it;
  KSTOP
  const double pq = device_sum(P.part, P.nblk, 2, I, sh);
  const double alpha = rho / pq;
  double in[2] = {0.0, 0.0};
  const double rn = r - alpha * q;
  const double zv = pcg_prec(P, I.dim, H.valid, H.g, rn, tv);
  if (H.valid) {
    P.x[H.g] = x + alpha * pv;
    P.r[H.g] = rn;
    P.z[H.g] = zv;
    in[0] = rn * zv;
    in[1] = rn * rn;
  }
  double o[2];
  block_sums<2>(in, sh, o);
  if (threadIdx.x == 0) {
    P.part[blockIdx.x] = o[0];
    P.part[(long long)P.nblk + blockIdx.x] = o[1];
  }
  if (lead) S.it = s_it + 1;
}

int pcg_poisson(Batch &b, double rtol, int maxit) {
  PArgs P{};
  P.W = make_where(b);
  P.N = b.ntot; P.nblk = b.nblk;
  P.A = b.Aeq; P.b = b.beq;
  P.x = b.veq[0]; P.r = b.veq[1]; P.z = b.veq[2]; P.p = b.veq[3]; P.q = b.veq[4];
  P.p2 = b.veq[5];
  P.part = b.part; P.scal = b.d_scal;
  P.P16 = b.P16; P.tile = b.pcg_tile;
  P.late_stop = late_stop_ok(b);
  hipLaunchKernelGGL(k_krylov_reset, dim3((b.ndev + 63) / 64), dim3(64), 0, b.stream, b.d_scal,
                     b.ndev, rtol * rtol, maxit, 0, 0.0, 0);
  hipLaunchKernelGGL(k_pcg_init, dim3(b.nblk), dim3(BS), 0, b.stream, P);
  b.launches += 2;
  // chunks between host checks as in BiCGStab: 3/4 of the last solve's
  // iterations first, then 16, 32, 64, 128
  int it = 0;
  int chunk = b.pcg_hint > 0 ? std::max(16, (3 * b.pcg_hint) / 4) : 16;
  int step = 16;
  const dim3 g(b.nblk), t(BS);
  while (true) {
    for (int c = 0; c < chunk && it < maxit; ++c, ++it) {
      P.parity = it & 1;
      hipLaunchKernelGGL(k_pcg_ab, g, t, 0, b.stream, P);
      hipLaunchKernelGGL(k_pcg_c, g, t, 0, b.stream, P);
      b.launches += 2;
    }
    if (it >= maxit) {
      P.parity = it & 1;
      hipLaunchKernelGGL(k_pcg_ab, g, t, 0, b.stream, P);
      b.launches++;
    }
    TCAD_HIP_CHECK(hipGetLastError());
    int rc = sync_scalars(b);
    if (rc) return rc;
    int open = 0;
    for (int d = 0; d < b.ndev; ++d) open += b.h_scal[d].done ? 0 : 1;
    if (open == 0 || it >= maxit) break;
    chunk = step;
    if (step < 128) step *= 2;
  }
  int most = 0, failed = it >= maxit ? 1 : 0;
  for (int d = 0; d < b.ndev; ++d) {
    most = std::max(most, b.h_scal[d].it);
    if (b.h_scal[d].active) failed |= b.h_scal[d].status;
  }
  b.pcg_hint = failed ? 0 : most;
  return TCAD_OK;
}

}  // namespace tcad
