/*
 * tcad_hip.h -- C ABI of libtcadhip.so, the MI355X (gfx950) implementation
 * of the coupled (psi, n, p) drift-diffusion Newton solve on structured
 * 2D/3D tensor-product meshes.
 *
 * One handle ("batch") holds one or more independent devices that share
 * every kernel launch: a single device is a batch of one.  All array
 * arguments named `d_*` are DEVICE pointers (HIP global memory, e.g.
 * torch tensors' data_ptr()); arguments named `h_*` are host pointers.
 * Sizes are explicit.  No entry point frees caller memory; library
 * workspaces live in the handle and are released by tcad_batch_destroy.
 *
 * Unknown layout at this boundary: node k = j*Nx + i (2D),
 * k = kk*Nx*Ny + j*Nx + i (3D) (pytcad/pytcad/mesh2d.py:1-8,
 * mesh3d.py:1-9); state vectors psi, n, p are separate FP64 planes of
 * length N_total; each device occupies the node range
 * [node_offset, node_offset + n) of every plane (packed in order by
 * default, or at explicit offsets so that a resident pool of devices can
 * be re-batched without moving their state); N_total is the end of the
 * last range.  Scaled units as in pytcad/pytcad/device2d.py:319-345.
 *
 * Return codes (pytcad/core/include/tcad/base/errors.hpp:21-67 mapping):
 *   TCAD_OK                0
 *   TCAD_E_INVALID        -1  invalid argument          -> ValueError
 *   TCAD_E_INDEX          -2  index out of range        -> IndexError
 *   TCAD_E_NONFINITE      -3  non-finite / singular     -> LinearSolveError
 *   TCAD_E_NOT_CONVERGED  -4  Krylov did not converge   -> LinearSolveError
 *   TCAD_E_HIP            -5  HIP runtime error         -> RuntimeError
 *   TCAD_E_NO_DEVICE      -6  no HIP device             -> RuntimeError
 * The message of the last failure on the calling thread is returned by
 * tcad_last_error().
 */
#ifndef TCAD_HIP_H
#define TCAD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TCAD_OK 0
#define TCAD_E_INVALID (-1)
#define TCAD_E_INDEX (-2)
#define TCAD_E_NONFINITE (-3)
#define TCAD_E_NOT_CONVERGED (-4)
#define TCAD_E_HIP (-5)
#define TCAD_E_NO_DEVICE (-6)

/* Linear-solve modes (pytcad/pytcad/linsolve.py:338 `_METHODS` analogue). */
#define TCAD_LIN_BICGSTAB_F64 0   /* FP64 BiCGStab + block-Jacobi            */
#define TCAD_LIN_BICGSTAB_MIXED 1 /* FP32 BiCGStab + block-Jacobi in FP64 IR */
/* Flag OR'd into lin_mode of tcad_solve_bias: Newton forcing terms.  Each
 * Newton solve's relative residual target is clamp(0.01 * err, lin_rtol,
 * max(1e-4, lin_rtol)) with err the previous step's update measure (1e-4
 * for the first step); without the flag every solve aims at lin_rtol.
 * Either way the step that passes the convergence test is refined to the
 * FP64 rounding floor (true relative residual 1e-16, or until a sweep no
 * longer halves it) before it is applied. */
#define TCAD_LIN_FORCING 0x100
/* Flag OR'd into lin_mode (tcad_solve_bias, tcad_solve_linear): precondition
 * every device with the reference's node block-Jacobi (3x3 block per node,
 * linsolve.py:363-427).  Without it, every device uses the exact inverse
 * of the matrix restricted to each 16-node tile (4 x 4 in 2D, 2 x 1 x 8 in
 * 3D; 48 unknowns), stored as FP16 rows with an FP32 scale per row, and
 * applied in red-black order over the tile checkerboard (or as tile
 * block-Jacobi: TCAD_TUNE_TILE_RB). */
#define TCAD_PRECOND_NODE 0x200

/* Per-device outcome of a Newton solve (h_status). */
#define TCAD_NEWTON_OK 0             /* converged (or masked out)            */
#define TCAD_NEWTON_MAX_ITER 1       /* not converged within max_iter        */
#define TCAD_NEWTON_LINEAR_FAIL 2    /* a linear solve ended above 1e-6      */
#define TCAD_NEWTON_NONFINITE 3      /* non-finite or singular assembly/step */

typedef struct tcad_batch tcad_batch;

/* Per-device description.  Host arrays are copied at creation. */
typedef struct {
  int32_t dim;              /* 2 or 3 */
  int32_t nx, ny, nz;       /* nz = 1 in 2D */
  /* scaled geometry (x/LD): spacings (n-1) and box widths (n) per axis */
  const double *h_hx, *h_hy, *h_hz;
  const double *h_dvx, *h_dvy, *h_dvz;
  /* scalars, pytcad/pytcad/device2d.py:319-328 */
  double VT, Ns, LD, J0, R0;
  double Cn_auger, Cp_auger;
  int32_t srh, auger;
  /* lagged Lombardi surface mobility on the row-0 y-edges (2D only),
   * pytcad/pytcad/device2d.py:612-642: T and the bulk mobilities of
   * rows 0 and 1 (length nx each), NULL when off */
  int32_t surface_mobility;
  double T;
  const double *h_mun_row0, *h_mun_row1, *h_mup_row0, *h_mup_row1;
  /* contacts: ncontacts groups; node lists (device-local node index),
   * ohmic values ln(n0/nie), n0, p0 per listed node (contacts.py:21-37) */
  int32_t ncontacts;
  const int32_t *h_contact_ptr;    /* ncontacts+1 offsets into the lists */
  const int32_t *h_contact_nodes;
  const double *h_contact_lnr, *h_contact_n0, *h_contact_p0;
  /* gates: ngates groups; node lists, kappa*w and psi_b per listed node
   * (device2d.py:1623-1637, device3d.py:939-947) */
  int32_t ngates;
  const int32_t *h_gate_ptr;
  const int32_t *h_gate_nodes;
  const double *h_gate_kw, *h_gate_psib;
  /* first node of this device in the batch planes; < 0 packs it right
   * after the previous device.  Ranges must not overlap. */
  int64_t node_offset;
} tcad_device_desc;

/* Static per-node / per-edge coefficient arrays of the whole batch, as
 * DEVICE pointers, each of length N_total (edge arrays are node-shaped:
 * entry k holds the edge from node k to its +axis neighbour, 0 where
 * none).  The library keeps the pointers; the caller keeps the memory
 * alive until tcad_batch_destroy. */
typedef struct {
  const double *d_C;        /* doping / Ns */
  const double *d_nie_s;    /* n_ie / Ns */
  const double *d_nie;      /* n_ie [cm^-3] */
  const double *d_tau_n, *d_tau_p;
  const double *d_dV;       /* scaled box volume */
  double *d_an[3];          /* dn_edge / h per axis (x, y, z); row 0 of the y
                               * array is rewritten each Newton iteration when
                               * the surface-mobility model is on */
  double *d_ap[3];          /* dp_edge / h per axis (same) */
  const double *d_dlnnie[3];/* ln(nie_s[k+axis] / nie_s[k]) per axis */
} tcad_static_fields;

/* Create a batch of `ndev` devices; `stream` is a hipStream_t (NULL =
 * default stream).  Replaces: Device2D.__init__ / Device3D.__init__
 * device-resident setup (pytcad/pytcad/device2d.py:168-453,
 * device3d.py:437-724). */
int tcad_batch_create(const tcad_device_desc *h_devs, int32_t ndev,
                      const tcad_static_fields *fields, void *stream,
                      tcad_batch **out);
int tcad_batch_destroy(tcad_batch *b);

/* Plane length N_total of the batch (end of the last device range). */
int64_t tcad_batch_nodes(const tcad_batch *b);

/* Contact voltages [V] (sum of ncontacts over devices, batch order) and
 * gate voltages [V] and flat-band voltages [V] (sum of ngates).
 * Replaces the bc.V / bc.Vg assignment of solve_bias
 * (pytcad/pytcad/device2d.py:1941-1947). */
int tcad_set_bias(tcad_batch *b, const double *h_contact_V,
                  const double *h_gate_Vg, const double *h_gate_Vfb);

/* Coupled residual/Jacobian at (psi, n, p).  Writes F (3 planes of
 * N_total, raw residual after contact/gate rows, as the reference's F)
 * and, if d_Fnp != NULL, the raw pre-Dirichlet F_n + F_p per node.
 * The eliminated, scaled Jacobian stays inside the handle.
 * Replaces: Device2D._residual_jacobian (device2d.py:1402-1920),
 * Device3D._residual_jacobian (device3d.py:1485-1978) and
 * dirichlet.eliminate_csr (dirichlet.py:161-191). */
int tcad_assemble(tcad_batch *b, const double *d_psi, const double *d_n,
                  const double *d_p, double *d_F, double *d_Fnp);

/* Secant predictor of the next bias point's Newton start, from the state at
 * the sweep's previous converged point (d_*_prev) and the present one
 * (d_psi, d_n, d_p, updated in place).  Per device, h_ratio[d] (host,
 * length ndev) is the step to the next point over the step from the
 * previous one (the bias vectors parallel); a device with h_ratio[d] <= 0
 * is not touched, and ratios above 8 or non-finite are refused
 * (TCAD_E_INVALID).  psi moves linearly, n and p geometrically (linearly
 * in log n, log p), each density by at most the Newton update's clip
 * [0.1x, 10x].  Asynchronous on the handle's stream.  Replaces: the secant
 * (tangent) predictor of continuation.py:368-404 (arc_length_sweep, psi
 * only there), applied to the warm start that id_vg_sweep (mosfet.py:90-103)
 * and id_vg_sweep_3d (finfet3d.py:246-263) take from the previous point. */
int tcad_predict_state(tcad_batch *b, const double *d_psi_prev, const double *d_n_prev,
                       const double *d_p_prev, double *d_psi, double *d_n, double *d_p,
                       const double *h_ratio);

/* The reference's raw residual/Jacobian at (psi, n, p) and the voltages
 * set by tcad_set_bias, without Dirichlet elimination or scaling:
 *   d_F     3 planes of N_total: F after the contact and gate rows;
 *   d_Fn,
 *   d_Fp    N_total each: raw F_n, F_p per node before the contact rows;
 *   d_Jraw  ne planes of N_total in the stencil-block layout of
 *           tcad_export_jacobian (ne = 9 + 10*dim_max): every node's
 *           Jacobian row entries BEFORE contact rows become Dirichlet
 *           rows.  At a contact node these are the raw rows the
 *           reference's current_rows_for captures; the reference's J
 *           holds e_k there (the caller stamps them).  Entries coupling
 *           to a contact neighbour are kept (not eliminated).
 * The handle's Newton matrix is overwritten, as by tcad_assemble.
 * Replaces: Device2D._residual_jacobian (device2d.py:1402-1920, its
 * F, J, F_n, F_p outputs and the current_rows_for capture :1751-1758),
 * Device3D._residual_jacobian (device3d.py:1485-1978). */
int tcad_residual_jacobian(tcad_batch *b, const double *d_psi, const double *d_n,
                           const double *d_p, double *d_F, double *d_Fn, double *d_Fp,
                           double *d_Jraw);

/* Copy the last assembled Jacobian to the host in stencil-block form:
 * h_blocks[e * N_total + k] for e < 9 + 5*2*dim_max (entries 0..8: the
 * 3x3 node block, then per neighbour direction x-,x+,y-,y+,z-,z+ the
 * five coupling entries (psi,psi) (n,psi) (n,n) (p,psi) (p,p)), of the
 * ELIMINATED matrix scaled as R * J * Cc; h_rowscale (3*N_total, R) and
 * h_rhs (3*N_total, the scaled eliminated right-hand side R*(-F - J g)).
 * Column scale Cc is (1, n, p) per node.  Test/inspection entry. */
int tcad_export_jacobian(tcad_batch *b, double *h_blocks, double *h_rowscale,
                         double *h_rhs, int32_t *h_nentries);

/* Solve the last assembled Newton system; writes the unscaled update du
 * (3 planes of N_total) to d_du, unscaled with the n and p buffers that
 * tcad_assemble was given (they must still hold that state).  lin_mode
 * selects the precision and the preconditioner (tcad_assemble built the
 * tile inverses).  Without an assembled system: TCAD_E_INVALID.  h_iters /
 * h_resid (length ndev) receive Krylov iterations and final relative
 * residual per device.
 * Replaces: linsolve.solve_linear (pytcad/pytcad/linsolve.py:725-985)
 * for the node-block-Jacobi Krylov methods. */
int tcad_solve_linear(tcad_batch *b, int32_t mode, double rtol, int32_t maxiter,
                      double *d_du, int32_t *h_iters, double *h_resid);

/* Equilibrium Poisson Newton from the bulk-neutral guess (or from the
 * given psi if from_guess == 0).  Writes psi, n, p.  h_active (length
 * ndev, NULL = all) selects the devices solved; the others' state is
 * not touched.  Per device: Newton iterations, converged flag and a
 * TCAD_NEWTON_* outcome (each output may be NULL).  Replaces:
 * Device2D.solve_equilibrium (device2d.py:1308-1373),
 * Device3D.solve_equilibrium (device3d.py:1311-1454). */
int tcad_solve_equilibrium(tcad_batch *b, int32_t max_iter, double tol_update,
                           double max_dpsi, int32_t from_guess,
                           const int32_t *h_active,
                           double *d_psi, double *d_n, double *d_p,
                           int32_t *h_iters, int32_t *h_converged, int32_t *h_status);

/* Coupled Newton at the biases set by tcad_set_bias, warm-started from
 * (psi, n, p) with contact values stamped.  Writes the state in place.
 * h_active (length ndev, NULL = all) selects the devices solved; the
 * others' state is not touched.  Per device (outputs may be NULL):
 * Newton iterations, converged flag, total Krylov iterations and a
 * TCAD_NEWTON_* outcome.  A device whose linear solve ends with a
 * relative residual above 1e-6 stops with TCAD_NEWTON_LINEAR_FAIL
 * before the step is applied (the reference raises LinearSolveError
 * there, linsolve.py:973-982); the other devices go on.  Replaces:
 * Device2D.solve_bias (device2d.py:1923-2191), Device3D.solve_bias
 * (device3d.py:1981-2271). */
int tcad_solve_bias(tcad_batch *b, int32_t max_iter, double tol_update,
                    double max_dpsi, int32_t lin_mode, double lin_rtol,
                    int32_t lin_maxiter, const int32_t *h_active,
                    double *d_psi, double *d_n, double *d_p,
                    int32_t *h_iters, int32_t *h_converged, int32_t *h_krylov,
                    int32_t *h_status);

/* A queue of bias points per device, solved in one call without a barrier
 * between devices: device d solves its points 0 .. h_npoints[d]-1 in order,
 * each warm-started from the previous converged one (the first from the
 * present state) with its contact values stamped, and goes on to its next
 * point as soon as it converges at the present one; a device stops at its
 * first failure, its state restored to its last converged point.  Bias
 * tables, point-major: h_contact_V[k * ncontact_total + c] (contact c's
 * voltage at a device's point k, global contact order as for
 * tcad_set_bias) and h_gate_Vg[k * ngate_total + g]; entries of other
 * devices' contacts are ignored.  Outputs: h_done[d] points converged;
 * h_iters / h_krylov[d * npmax + k] Newton and Krylov iterations of point
 * k (of the failed point too); h_status[d] the TCAD_NEWTON_* outcome of the
 * point a device stopped at (OK when all converged); h_I[k * ncontact_total
 * + c] the terminal currents after point k (as tcad_terminal_currents; only
 * converged points are written).  Solver arguments as for tcad_solve_bias.
 * Replaces: the warm-started sweep loops of id_vg_sweep
 * (pytcad/pytcad/mosfet.py:90-103), id_vg_sweep_3d (finfet3d.py:246-263)
 * and the sweep runner (gui/services/solver_runner.py:360), for many
 * devices at once. */
int tcad_solve_queue(tcad_batch *b, int32_t max_iter, double tol_update, double max_dpsi,
                     int32_t lin_mode, double lin_rtol, int32_t lin_maxiter, int32_t npmax,
                     const int32_t *h_npoints, const double *h_contact_V,
                     const double *h_gate_Vg, double *d_psi, double *d_n, double *d_p,
                     int32_t *h_done, int32_t *h_iters, int32_t *h_krylov, int32_t *h_status,
                     double *h_I);

/* One implicit time step of the theta scheme, from the state at the start
 * of the step to the new time, whose contact voltages are set by
 * tcad_set_bias.  Per non-contact node the continuity rows become
 *   theta F_n + (1 - theta) F_n(old) - dV/dt (n - n_old)   (electrons)
 *   theta F_p + (1 - theta) F_p(old) + dV/dt (p - p_old)   (holes)
 * with the Poisson and contact rows unchanged; Newton with Armijo
 * backtracking on 0.5 |F|^2 (up to 40 halvings), psi clip, density clip
 * and the update test; in 3D a Newton iteration whose line search finds no
 * step ends the step unconverged.  h_dt_s: per device, the scaled step
 * dt * R0 / Ns as the reference forms it; d_Fn_old / d_Fp_old: the raw
 * continuity residuals at the old state and old voltages
 * (tcad_continuity_residuals), needed when theta != 1.  lin_mode must
 * select the FP64 path (TCAD_LIN_BICGSTAB_F64, flags allowed).  Writes the
 * new state in place; per device iterations, converged flag and outcome.
 * Replaces: transient2d._step_residual_jacobian / _newton_step
 * (pytcad/pytcad/transient2d.py:104-183), transient3d.py:118-219. */
int tcad_transient_step(tcad_batch *b, int32_t max_iter, double tol_update,
                        double max_dpsi, int32_t lin_mode, double lin_rtol,
                        int32_t lin_maxiter, const int32_t *h_active,
                        const double *h_dt_s, double theta,
                        const double *d_n_old, const double *d_p_old,
                        const double *d_Fn_old, const double *d_Fp_old,
                        double *d_psi, double *d_n, double *d_p,
                        int32_t *h_iters, int32_t *h_converged, int32_t *h_status);

/* Raw box-integration continuity residuals F_n, F_p per node (before the
 * contact rows), the last two outputs of _residual_jacobian
 * (device2d.py:1920, device3d.py:1978), at the voltages set by
 * tcad_set_bias. */
int tcad_continuity_residuals(tcad_batch *b, const double *d_psi, const double *d_n,
                              const double *d_p, double *d_Fn, double *d_Fp);

/* Small-signal AC N-port admittance at the DC operating point
 * (psi, n, p) and the voltages set by tcad_set_bias, for each of nfreq
 * frequencies [Hz].  Every contact and gate of a device is a port, in the
 * order: its contacts, then its gates (creation order).  h_Y receives,
 * per device d, frequency f, row i and column k, the complex
 * Y[d][f][i][k] = dI_i / dV_k [A/V; per cm in 2D] as (re, im) at
 * h_Y[2 * (((d * nfreq + f) * P + i) * P + k)], P the largest port count
 * of the batch (entries beyond a device's ports are 0).  Each complex
 * system is solved by FP64 complex BiCGStab, preconditioned with the
 * exact inverse of J + j w C on every 16-node tile, with restarted sweeps
 * until the true residual is below rtol relative to the forcing (maxiter
 * iterations per sweep).  h_status[d]: TCAD_NEWTON_OK, or
 * TCAD_NEWTON_LINEAR_FAIL when a solve stalled above 1e-6; h_iters[d]
 * (may be NULL): Krylov iterations spent.  h_active as for
 * tcad_solve_bias.  Replaces: ac2d.y_parameters
 * (pytcad/pytcad/ac2d.py:265-342), ac3d.y_parameters
 * (pytcad/pytcad/ac3d.py:187-277). */
int tcad_y_parameters(tcad_batch *b, const int32_t *h_active, int32_t nfreq,
                      const double *h_freqs, double rtol, int32_t maxiter,
                      const double *d_psi, const double *d_n, const double *d_p,
                      double *h_Y, int32_t *h_status, int64_t *h_iters);

/* Terminal currents of every contact of every device at (psi, n, p):
 * h_I[c] = sum over the contact's nodes of the raw F_n + F_p times
 * J0*LD (2D, A/cm) or J0*LD^2 (3D, A).  Replaces:
 * Device2D.terminal_current (device2d.py:2194-2227),
 * Device3D.terminal_current (device3d.py:2274-2306). */
int tcad_terminal_currents(tcad_batch *b, const double *d_psi, const double *d_n,
                           const double *d_p, double *h_I);

/* Scharfetter-Gummel edge currents (scaled) per axis, node-shaped
 * planes d_J[axis*2*N_total + carrier*N_total + k] (carrier 0 = Jn,
 * 1 = Jp), at (psi, n, p).  Replaces the Jn_x/Jn_y/Jp_x/Jp_y outputs of
 * _residual_jacobian (device2d.py:1463-1479). */
int tcad_edge_currents(tcad_batch *b, const double *d_psi, const double *d_n,
                       const double *d_p, double *d_J);

/* Lagged Lombardi surface-mobility update of the row-0 y-edge
 * coefficients at psi (2D devices with the model on; no-op otherwise).
 * solve_bias applies it at the start of every Newton iteration.
 * Replaces: Device2D._update_surface_mobility (device2d.py:612-642). */
int tcad_surface_mobility(tcad_batch *b, const double *d_psi);

/* Bench instrumentation: average duration [microseconds] of `reps`
 * back-to-back launches of one kernel on the last assembled system, timed
 * with HIP events on the library stream after an untimed burst of the
 * same length (as inside the solver, each launch waits for the previous
 * one; the time per launch includes that dependency).  which: 0 FP64 stencil SpMV,
 * 1 FP32 stencil SpMV, 2 coupled assembly (FP32 copy on), 3 the fused
 * BiCGStab update kernel and 4 the alpha / s kernel, both FP64 and
 * applying the preconditioner (each launch preceded by a one-thread state
 * reset, included in the time).  Has no reference counterpart. */
int tcad_time_kernel(tcad_batch *b, int32_t which, int32_t reps, double *h_avg_us);

/* Iterative solve of a caller-supplied sparse system A x = b in CSR form
 * (n rows, device arrays d_indptr[n+1], d_indices[nnz], d_data[nnz] with
 * int32 indices; d_b[n]; d_x[n] holds the initial guess on entry -- zero
 * it for none -- and the solution on return).
 *   method     TCAD_CSR_BICGSTAB or TCAD_CSR_GMRES (restart length
 *              `restart`, <= 0 for min(100, n); at most 100);
 *   precond    TCAD_CSR_PREC_BLOCK_JACOBI with block_size 1 or 3 (the exact
 *              inverse of each node's diagonal block), or TCAD_CSR_PREC_ILU0
 *              (level-scheduled ILU(0); needs sorted indices and a stored
 *              diagonal in every row);
 *   maxiter    BiCGStab iterations, or GMRES restart cycles;
 *   rtol       relative to ||b||; the result is accepted when the true
 *              residual ||b - A x|| / ||b|| <= max(rtol, 1e-6).
 * Returns TCAD_OK, or TCAD_E_NOT_CONVERGED (x holds the last iterate),
 * TCAD_E_NONFINITE (singular block / pivot), TCAD_E_INVALID.  h_iters and
 * h_resid (may be NULL) receive the Krylov iterations and the final true
 * relative residual.  `stream` is a hipStream_t (NULL = default).  The
 * workspace is allocated and freed inside the call.  Replaces:
 * linsolve.solve_linear for methods "gmres" / "bicgstab" with
 * block_size (node block-Jacobi, linsolve.py:363-427) or without it (an
 * incomplete factorisation, linsolve.py:429-585), linsolve.py:725-985. */
#define TCAD_CSR_BICGSTAB 0
#define TCAD_CSR_GMRES 1
#define TCAD_CSR_PREC_BLOCK_JACOBI 0
#define TCAD_CSR_PREC_ILU0 1
int tcad_csr_solve(int32_t n, const int32_t *d_indptr, const int32_t *d_indices,
                   const double *d_data, const double *d_b, double *d_x, int32_t method,
                   int32_t precond, int32_t block_size, double rtol, int32_t maxiter,
                   int32_t restart, void *stream, int32_t *h_iters, double *h_resid);

/* tcad_csr_solve without any allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and kernels as tcad_csr_solve, with the workspace in
 * a fixed 64 MiB device array of the library (one per GPU, held until the
 * process ends).  Block-Jacobi only (TCAD_CSR_PREC_ILU0 is refused);
 * systems whose workspace exceeds the array are refused with
 * TCAD_E_INVALID (BiCGStab with 3-blocks: about 11 n words, so n up to
 * about 760,000; GMRES(m): about (m + 7) n words).  Every call on one GPU
 * uses the same array: calls that overlap on one GPU race. */
int tcad_csr_solve_noalloc(int32_t n, const int32_t *d_indptr, const int32_t *d_indices,
                           const double *d_data, const double *d_b, double *d_x,
                           int32_t method, int32_t precond, int32_t block_size, double rtol,
                           int32_t maxiter, int32_t restart, void *stream, int32_t *h_iters,
                           double *h_resid);

/* Run-time tuning of one batch (the autotuner's knobs; defaults in
 * brackets):
 *   TCAD_TUNE_TILE_ROWS      0 by grid size [registers up to 512 workgroups,
 *                            memory above], 1 tile-inverse rows in registers,
 *                            2 read at the point of use;
 *   TCAD_TUNE_FORCING_SLOPE  c (0 < c <= 1000) of the forcing term
 *                            eta = clamp(c * err^q, rtol, eta_max) of
 *                            TCAD_LIN_FORCING solves [0.01];
 *   TCAD_TUNE_FORCING_MAX    eta_max [1e-4];
 *   TCAD_TUNE_FORCING_FP32   1: a forcing solve whose target is at least
 *                            1e-5 for every active device runs BiCGStab on
 *                            the FP32-stored matrix inside the FP64
 *                            refinement (the converging step is still
 *                            refined on the FP64 matrix) [0];
 *   TCAD_TUNE_GRAPH          1: BiCGStab iterations are issued as replays of
 *                            a captured HIP graph of 16 iterations (captured
 *                            once per kernel-argument set on a private
 *                            stream, launched on the batch stream) [0];
 *   TCAD_TUNE_IR_CONTINUE    refinement sweeps after the first continue the
 *                            previous sweep's BiCGStab (search direction and
 *                            shadow residual kept, the iterated residual
 *                            replaced by the FP64 true residual) instead of
 *                            restarting it: 0 never, 1 for inner solves on
 *                            the FP32-stored matrix, 2 always [1];
 *   TCAD_TUNE_FP32_INNER     the residual reduction (1e-7..1e-2) an inner
 *                            solve on the FP32-stored matrix stops at, at
 *                            least, in each refinement sweep of a Newton
 *                            solve [1e-6];
 *   TCAD_TUNE_FORCING_POW    q: 1 (eta linear in the last Newton update
 *                            measure err) or 2 (quadratic) [1];
 *   TCAD_TUNE_TILE_LAG       a Newton step of tcad_solve_bias keeps the
 *                            tile inverses of the step before when that
 *                            step's update measure was below this (0: every
 *                            step inverts its tiles) [0];
 *   TCAD_TUNE_LINEAR_FALLBACK 1: a Newton solve of tcad_solve_bias /
 *                            tcad_solve_queue whose refinement ends
 *                            unconverged is solved again, for that iteration
 *                            only, by restarted GMRES(50) with ILU(0) on the
 *                            device's CSR form of the same system (the
 *                            counterpart of the reference's per-iteration
 *                            direct fallback, device3d.py:2154-2168); only
 *                            if that fails too does the device stop with
 *                            TCAD_NEWTON_LINEAR_FAIL; 0: stop at once [1];
 *   TCAD_TUNE_FORCING_PREDICTED 1: in a TCAD_LIN_FORCING solve of
 *                            tcad_solve_bias only the devices whose start
 *                            the last tcad_predict_state moved (ratio > 0)
 *                            use forcing terms; the others solve every
 *                            Newton step to lin_rtol, as without the flag
 *                            [0];
 *   TCAD_TUNE_TILE_RB        1: the tile preconditioner orders its tiles
 *                            red-black (a checkerboard over the tile grid)
 *                            and applies the black tiles' inverses after the
 *                            red ones, to the residual less the couplings to
 *                            the red tiles' results (forward block
 *                            Gauss-Seidel): about half the BiCGStab
 *                            iterations of 0, plain tile block-Jacobi, for
 *                            two more launches per iteration [1].
 * Has no reference counterpart (the reference's Krylov tolerance is fixed,
 * linsolve.py:725-727). */
#define TCAD_TUNE_TILE_ROWS 1
#define TCAD_TUNE_FORCING_SLOPE 2
#define TCAD_TUNE_FORCING_MAX 3
#define TCAD_TUNE_FORCING_FP32 4
#define TCAD_TUNE_GRAPH 5
#define TCAD_TUNE_IR_CONTINUE 6
#define TCAD_TUNE_FP32_INNER 7
#define TCAD_TUNE_FORCING_POW 8
#define TCAD_TUNE_TILE_LAG 9
#define TCAD_TUNE_LINEAR_FALLBACK 10
#define TCAD_TUNE_FORCING_PREDICTED 11
#define TCAD_TUNE_TILE_RB 12
int tcad_set_tuning(tcad_batch *b, int32_t key, double value);

/* Bench instrumentation, in-solver kernel timing: from this call on,
 * every launch of one BiCGStab kernel inside the solves of this batch is
 * bracketed by a pair of HIP events on the batch stream (which: 0 the
 * stencil SpMV, 1 the alpha / s kernel, 2 the second SpMV, 3 the fused
 * update kernel, 4 the black-tile kernel of the red-black tile ordering
 * (both of its launches per iteration), whichever preconditioner variant
 * runs; -1 stops).
 * The events take the kernel's own start and end (hipExtLaunchKernelGGL).
 * tcad_probe_read waits for the stream and returns, of the launches
 * bracketed since the last read, those that ran at least min_us
 * microseconds (launches issued after a solve converged return at once and
 * move no data) and their average duration [microseconds].  Has no
 * reference counterpart. */
int tcad_probe_kernel(tcad_batch *b, int32_t which);
int tcad_probe_read(tcad_batch *b, double min_us, int64_t *h_launches, double *h_avg_us);

/* Wall-clock-free counters of the last call: kernel launches issued. */
int64_t tcad_launch_count(const tcad_batch *b);

const char *tcad_last_error(void);
const char *tcad_version(void);

#ifdef __cplusplus
}
#endif
#endif /* TCAD_HIP_H */
