// Internal declarations shared by the libtcadhip translation units.
//
// Data layout in HBM (all planes length N_total = sum of the batch's node
// counts, devices concatenated):
//   state        psi, n, p                 FP64, 3 planes (caller-owned)
//   static       C, nie_s, nie, tau_n, tau_p, dV, an[3], ap[3], dlnnie[3]
//                FP64 (caller-owned), edge arrays node-shaped
//   matrix       A64[e*N + k], A32[e*N + k] stencil-block form, e < ne:
//                0..8 the 3x3 node block (row-major), then per neighbour
//                direction x-,x+,y-,y+,z-,z+ five coupling entries
//                (psi<-psi, n<-psi, n<-n, p<-psi, p<-p); the matrix is the
//                Dirichlet-ELIMINATED Jacobian scaled as R*J*Cc with
//                Cc = (1, n, p) per node and R the inverse row max.
//   precond      Dinv64/Dinv32[e*N + k], e < 9: inverse of the scaled node
//                block (node block-Jacobi)
//   vectors      3 planes of N_total (v[c*N + k])
// Kernels map one thread to one node.  Each 2D device is cut into patches
// of PX x PY nodes (x fastest), each 3D device into PX3 x PY3 x PZ3 boxes,
// one workgroup per patch: thread t is node (i0 + t % PX, j0 + t / PX), or
// (i0 + t % PX3, j0 + t / PX3 % PY3, k0 + t / (PX3 PY3)), so every
// wavefront reads contiguous x-row segments (coalesced), and a patch holds
// whole 16-node tiles (the tile preconditioner inverts them in LDS).
// Workgroups never straddle devices, so per-device reductions run over a
// contiguous block range in a fixed order and are deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "tcad_hip.h"

namespace tcad {

#ifndef TCAD_BS
#define TCAD_BS 128
#endif
#ifndef TCAD_PY
#define TCAD_PY 4
#endif
constexpr int BS = TCAD_BS;         // threads per node workgroup (multiple of 64)
constexpr int PY = TCAD_PY;         // 2D patch height (whole tiles)
constexpr int PX = BS / PY;         // 2D patch width
static_assert(BS % PX == 0, "workgroup must hold whole patch rows");
constexpr int NDIAG = 9;
constexpr int NOFF = 5;

inline int nentries(int dim) { return NDIAG + NOFF * 2 * dim; }

// Tile block-Jacobi: 16-node tiles, 4 x 4 in 2D and T3X x T3Y x T3Z in 3D,
// each inverted exactly; node k keeps its 3 rows of the tile inverse as
// TILE_PLANES planes of four FP16 entries (tile4), each row with its own
// FP32 scale.  3D devices are cut into PX3 x PY3 x PZ3 patches (x fastest)
// that hold whole tiles: the z couplings of the finely meshed fin are the
// strongest, so the 3D tiles extend along z.
#ifndef TCAD_T3X
#define TCAD_T3X 2
#endif
#ifndef TCAD_T3Y
#define TCAD_T3Y 1
#endif
constexpr int TILE_X = 4, TILE_Y = 4;
constexpr int TILE_NODES = TILE_X * TILE_Y;
constexpr int T3X = TCAD_T3X, T3Y = TCAD_T3Y, T3Z = TILE_NODES / (T3X * T3Y);
constexpr int PY3 = T3Y, PZ3 = T3Z, PX3 = BS / (PY3 * PZ3);
constexpr int TILE_U = 3 * TILE_NODES;          // unknowns per tile
constexpr int TILE_PLANES = 3 * TILE_U / 4;     // tile4 planes per node
constexpr int CPL_GROUPS = 4;                   // out-of-tile directions per node, at most
static_assert(PX % TILE_X == 0 && PY % TILE_Y == 0, "2D patches must hold whole tiles");
static_assert(T3X * T3Y * T3Z == TILE_NODES && PX3 * PY3 * PZ3 == BS && PX3 % T3X == 0,
              "3D patches must hold whole 16-node tiles");
typedef _Float16 tile4 __attribute__((ext_vector_type(4)));

// Per-device metadata, device resident.
struct DevInfo {
  int dim, nx, ny, nz;
  long long off;                    // first node in the batch planes
  long long n;                      // node count
  int blk0, nblk;                   // node-kernel block range (patches)
  int px, py;                       // patches along x and y (PX x PY in 2D, PX3 x PY3 in 3D)
  int h_off[3];                     // offsets into the concatenated spacing arrays
  int dv_off[3];                    // offsets into the concatenated width arrays
  double VT, Ns, LD, J0, R0, Cn, Cp;
  int srh, auger, surf;
  double T;
  int surf_off;                     // offset into surface-mobility row arrays
  int c0, nc;                       // global contact id range
  int g0, ng;                       // global gate id range
};

// Per-device Krylov / Newton scalars, device resident.
struct Scal {
  double rho2[2];                   // rho by iteration parity
  double rho, alpha, omega, beta;   // Krylov scalars of the current iteration
  double bb;                        // ||b||^2
  double rr;                        // last ||r||^2
  double rr_ref;                    // final refinement: ||r||^2 before the last sweep
  double nrm;                       // IR residual norm (rhs normalisation)
  double nrm_prev;                  // the previous sweep's nrm (continued inner solves)
  double err;                       // Newton update measure
  double tol2;                      // stop test rr <= tol2 * bb of the current solve
  double eta2;                      // squared forcing term of this device's next Newton solve
  double lam;                       // transient line search: current step fraction
  double merit0;                    // transient line search: 0.5 |F|^2 at the iterate
  int ls_done;                      // transient line search: step fraction settled
  int ls_it;                        // transient line search: rejected trials
  int maxit;                        // iteration cap of the current solve
  int it;                           // Krylov iterations of the current solve
  int done;                         // Krylov finished (converged or stopped)
  int status;                       // 0 ok, 1 breakdown / non-finite / not converged
  int active;                       // device takes part in the current solve
  int newton_it, newton_done, kry_total, ir_it, ir_maxsweep;
  int conv_now;                     // this Newton step's update passed the test
  int saved_active;                 // active flag kept across one refinement
  int tile_keep;                    // this Newton step keeps the last step's tile inverses
  int cont;                         // this inner solve continues the last one's Krylov process
  int cont_ok;                      // the last inner solve may be continued
  int newton_active;                // Newton-level flag kept across the final-step refinement
  int sel;                          // device selected for the current Newton solve
  int predicted;                    // tcad_predict_state moved this device's start (consumed
                                    // by the next solve's selection)
  int forced;                       // the current Newton solve applies forcing terms
  int result;                       // TCAD_NEWTON_* outcome of the current Newton solve
};

// Node of thread threadIdx.x in workgroup blk of device I.
struct NodeAt {
  bool valid;
  long long k;                      // device-local node index
  int i, j, kk;
};

__device__ __forceinline__ NodeAt node_at(const DevInfo &I, int blk) {
  NodeAt P;
  const unsigned local = (unsigned)(blk - I.blk0);
  const unsigned upx = (unsigned)I.px, upy = (unsigned)I.py;
  const unsigned bx = local % upx, t = local / upx;
  const unsigned by = t % upy, bz = t / upy;
  if (I.dim == 3) {
    P.i = (int)(bx * PX3 + threadIdx.x % PX3);
    P.j = (int)(by * PY3 + (threadIdx.x / PX3) % PY3);
    P.kk = (int)(bz * PZ3 + threadIdx.x / (PX3 * PY3));
  } else {
    P.i = (int)(bx * PX + threadIdx.x % PX);
    P.j = (int)(by * PY + threadIdx.x / PX);
    P.kk = 0;
  }
  P.valid = P.i < I.nx && P.j < I.ny && P.kk < I.nz;
  P.k = ((long long)P.kk * I.ny + P.j) * I.nx + P.i;
  return P;
}

struct Batch {
  int ndev = 0;
  long long ntot = 0;
  int nblk = 0;
  int dim_max = 2;
  int ne = 0;                       // stencil entries per node
  hipStream_t stream = nullptr;
  long long launches = 0;
  // iterations of the last BiCGStab / PCG solve (slowest device): the next
  // solve's first chunk is sized from it
  int kry_hint = 0, pcg_hint = 0;
  double kry_hint_lt = 0.0;         // ln of the relative target the kry_hint solve aimed at
  // per Newton-iteration index: the k-th solve of a bias point resembles
  // the k-th solve of the point before far more than the solve just before
  // it (forcing targets loosen and tighten along a point's Newton steps);
  // kry_slot: the index of the coming Newton solve (-1: none)
  static constexpr int kSlots = 16;
  int kry_slot = -1;
  int slot_hint[kSlots] = {};
  double slot_lt[kSlots] = {};
  std::vector<DevInfo> info;        // host copy
  DevInfo *d_info = nullptr;
  int *d_blk2dev = nullptr;
  tcad_static_fields f{};
  double *d_h[3] = {nullptr, nullptr, nullptr};
  double *d_dv[3] = {nullptr, nullptr, nullptr};
  // boundary tables
  int *d_cidx = nullptr, *d_gidx = nullptr;      // per node: table index or -1
  double *d_cn_lnr = nullptr, *d_cn_n0 = nullptr, *d_cn_p0 = nullptr;
  int *d_cn_contact = nullptr;                   // table row -> global contact id
  double *d_gn_kw = nullptr, *d_gn_psib = nullptr;
  int *d_gn_gate = nullptr;
  int ncontact = 0, ngate = 0;
  double *d_contact_V = nullptr, *d_gate_Vg = nullptr, *d_gate_Vfb = nullptr;
  int *d_cptr = nullptr;                         // contact node lists (global nodes)
  int *d_cnodes = nullptr;
  int *d_gptr = nullptr;                         // gate g: table rows gptr[g]..gptr[g+1]
  int *d_gnode = nullptr;                        // gate table row -> global node
  double *d_cscale = nullptr;                    // per contact J0*LD^(dim-1)
  int *d_c2dev = nullptr;
  double *d_surf[4] = {nullptr, nullptr, nullptr, nullptr};
  int any_surf = 0;
  // coupled system
  double *A64 = nullptr;
  float *A32 = nullptr;
  double *Dinv64 = nullptr;
  float *Dinv32 = nullptr;
  tile4 *T4 = nullptr;                           // tile inverses, TILE_PLANES planes (FP16)
  float *CPL = nullptr;                          // black tiles' couplings to the red ones
  float *T4s = nullptr;                          // their per-row scales, 3 planes
  double *P16 = nullptr;                         // equilibrium PCG tile-inverse rows, 16 planes
  int pcg_tile = 0;                              // 1: tile block-Jacobi PCG, 0: Jacobi
  int tile = 1;                                  // preconditioner of the next solves
  double *bs = nullptr;                          // scaled rhs
  double *rowscale = nullptr;
  double *colscale = nullptr;                    // (1, n, p) at assembly
  double *y64 = nullptr;                         // scaled solution
  double *F = nullptr;                           // internal residual buffer
  double *Fnp = nullptr;
  double *cand = nullptr;                        // candidate state psi, n, p (3 planes)
  double *d_dts = nullptr;                       // per-device scaled time step
  double *d_Icand = nullptr;                     // contact currents at the candidate
  // Krylov work vectors (FP64) and the refinement buffers
  double *v64[9] = {};
  double *ir_r = nullptr, *ir_rhs = nullptr;
  double *part = nullptr;                        // nblk * 8 partial sums
  Scal *d_scal = nullptr;
  Scal *h_scal = nullptr;                        // pinned host mirror
  // equilibrium Poisson system
  double *Aeq = nullptr;                         // (1 + 2*dim) entries per node
  double *beq = nullptr;
  double *veq[6] = {};
  int *d_flag = nullptr;                         // [0] any, [1 + d] per-device error flag
  double *bk = nullptr;                          // point queue: last converged states (3 planes)
  int *d_qmask = nullptr;                        // point queue: per-device selection
  int *d_bmap = nullptr, *h_bmap = nullptr;       // Krylov grid of the open devices (2 halves)
  int bmap_half = 0;
  int *d_mask = nullptr;                         // per-device selection of the next solve
  double *d_ratio = nullptr;                     // per-device secant ratio (tcad_predict_state)
  const double *last_n = nullptr, *last_p = nullptr;   // state of the last assembly
  const double *last_psi = nullptr;
  std::vector<int> dev_ncontact, dev_ngate;
  // run-time tuning (tcad_set_tuning): tile-row variant (0 by grid size,
  // 1 registers, 2 memory), forcing-term slope and cap
  int tile_rows = 0;
  double forcing_c = 0.01, forcing_max = 1e-4;
  int forcing_pow = 1;              // eta = c * err^forcing_pow
  double tile_lag = 0.0;            // keep the tile inverses after a Newton update below this
  int tile_rb = 1;                  // red-black tile ordering of the tile preconditioner
  int forcing_pred_only = 0;        // forcing terms only for devices with a predicted start
  // forcing solves whose target is loose (every active device's eta >=
  // 1e-5) run BiCGStab on the FP32-stored matrix inside the FP64 refinement
  int forcing_fp32 = 0;
  // refinement sweeps after the first continue the previous sweep's
  // BiCGStab (search direction and shadow residual kept, the residual
  // replaced by the FP64 true residual) instead of restarting it:
  // 0 never, 1 for inner solves on the FP32-stored matrix, 2 always
  int ir_cont = 1;
  // relative residual reduction an inner solve on the FP32-stored matrix
  // aims at, at least, per sweep of the Newton solves
  double fp32_inner = 1e-6;
  // BiCGStab iterations replayed as captured HIP graphs (tcad_set_tuning
  // TCAD_TUNE_GRAPH): blocks of kGraphIts iterations captured on a private
  // stream, one executable per kernel-argument set
  int use_graph = 0;
  hipStream_t capture_stream = nullptr;
  struct GraphEntry {
    std::vector<unsigned char> key;  // the kernel arguments it was captured with
    int variant;
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;
  // per-iteration linear fallback (tcad_set_tuning TCAD_TUNE_LINEAR_FALLBACK):
  // a Newton solve whose refinement ends unconverged is solved again, for
  // that iteration, by restarted GMRES with ILU(0) on the device's CSR form
  // of the same equilibrated system (the pattern built on first use)
  int lin_fallback = 1;
  long long fallbacks = 0;          // fallback solves run (all devices, whole life)
  struct FallbackCsr {
    int n = 0;                      // 3 x the device's nodes
    long long nnz = 0;
    int *ptr = nullptr, *idx = nullptr;
    long long *src = nullptr;       // value k = A64[src[k]]
    double *val = nullptr, *b = nullptr, *x = nullptr;
  };
  std::vector<FallbackCsr> fb;
  // in-solver timing of one BiCGStab kernel (tcad_probe_kernel): every
  // launch of kernel `probe` (0 SpMV, 1 alpha/s, 2 second SpMV, 3 update;
  // -1 off) is bracketed by a pair of events on the batch stream
  int probe = -1;
  std::vector<hipEvent_t> probe_ev;
  size_t probe_used = 0;
};

// start/stop events of a probed launch (krylov.hip); false unless b.probe == which
bool probe_events(Batch &b, int which, hipEvent_t &e0, hipEvent_t &e1);

// error handling (tcad_capi.cpp)
void set_error(const std::string &msg);
#define TCAD_HIP_CHECK(expr)                                               \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess) {                                                \
      ::tcad::set_error(std::string("HIP error: ") + hipGetErrorString(e_) \
                        + " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
      return TCAD_E_HIP;                                                   \
    }                                                                      \
  } while (0)

// assemble.hip
// Transient-step terms of the assembly (transient2d.py:104-135) and the
// optional raw continuity residual outputs.
struct TransientArgs {
  int active;                       // add the storage / theta terms
  const double *n_old, *p_old;      // state at the start of the step
  const double *Fn_old, *Fp_old;    // raw continuity residuals there (theta != 1)
  const double *dt_s;               // per device, scaled step
  double theta;
  double *Fn_raw, *Fp_raw;          // if set: raw F_n, F_p per node
};

// Small-signal AC terms of the assembly (ac2d.py:107-194, 296-313).
// mode 1: assemble the matrix as usual and also write the scaled storage
//         diagonal (dn, dp) and the raw contact current rows (raw, 21
//         planes); mode 2: write only the scaled right-hand side R b of
//         the port forcing acg (du_psi = acg on contact nodes, psi-row
//         forcing acg elsewhere), using the row scales of mode 1.
struct AcAssembly {
  int mode;
  const double *acg;
  double *dn, *dp;
  double *raw;
};

// only_active: skip devices whose Newton solve is not running (Scal.active);
// res_only: 1 = write only the raw F_n + F_p per node (Fnp); 2 = only F
// (all rows, boundary and transient terms included), no matrix
int launch_assemble(Batch &b, const double *psi, const double *n, const double *p,
                    double *F, double *Fnp, bool want_fp32, bool only_active = false,
                    int res_only = 0, const TransientArgs *tr = nullptr,
                    const AcAssembly *ac = nullptr, double *Jpre = nullptr);
int launch_surface_mobility(Batch &b, const double *psi, bool only_active = false);
int launch_assemble_poisson(Batch &b, const double *psi, double *F);
int launch_edge_currents(Batch &b, const double *psi, const double *n, const double *p,
                         double *J);
int launch_contact_sums(Batch &b, const double *Fnp, double *d_out);

// krylov.hip
int bicgstab(Batch &b, bool fp32, double rtol, int maxit, bool zero_guess, int chunk);
// FP64 iterative refinement of the scaled Newton system A64 y = bs around
// BiCGStab on A32 (fp32) or A64; restart: y = 0 first, else continue from y.
// per_dev: every device's own target S.eta2 replaces rtol (and the inner
// BiCGStab aims at it, not below inner_rtol for fp32)
int ir_solve(Batch &b, bool fp32, bool restart, double rtol, int maxsweep, int inner_maxit,
             double inner_rtol, bool per_dev = false, bool fallback = false);
// the per-iteration fallback solve of device d (batch.hip): rtol relative
// residual target; returns TCAD_OK and the solution in y64, or an error
int linear_fallback(Batch &b, int d, double rtol, int32_t *h_iters, double *h_resid);
int pcg_poisson(Batch &b, double rtol, int maxit);
// inverses of the tile diagonal blocks of the last assembled matrix
int launch_tile_setup(Batch &b, bool only_active);
// inverses of the tile diagonal blocks of the equilibrium Poisson matrix
int launch_pcg_tile_setup(Batch &b);
int launch_unscale(Batch &b, const double *n, const double *p, double *du);
int launch_unscale_state(Batch &b, const double *n, const double *p, double *du);
int launch_spmv_only(Batch &b, bool fp32);
// timing only: one k_bicg_u (which 3) or k_bicg_s (which 4), FP64, with a
// one-thread reset launch before it
int launch_vector_kernel_only(Batch &b, int which);
int clear_done(Batch &b);

// newton.hip
int launch_newton_update(Batch &b, double *psi, double *n, double *p, double max_dpsi,
                         double tol);
int launch_poisson_update(Batch &b, double *psi, double max_dpsi, double tol);
int launch_stamp_contacts(Batch &b, double *psi, double *n, double *p, int zero_bias);
int launch_bulk_guess(Batch &b, double *psi);
int launch_slaved_densities(Batch &b, const double *psi, double *n, double *p);
// select the devices of the next Newton solve (h_mask NULL = all); eta0 is
// the forcing term of each device's first Newton solve
int launch_set_active(Batch &b, const int32_t *h_mask, double eta0 = 0.0,
                      double eta_exact = -1.0);
int sync_scalars(Batch &b);
// TCAD_E_NONFINITE (and clears the flags) if an assembly flagged a
// non-finite or singular node block
int check_flag(Batch &b);

}  // namespace tcad

struct tcad_batch {
  tcad::Batch *impl;
};

#define GUARD(bp)                                 \
  if (!(bp) || !(bp)->impl) {                     \
    tcad::set_error("null tcad_batch handle");    \
    return TCAD_E_INVALID;                        \
  }
