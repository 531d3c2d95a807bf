// Node-centric gather assembly of the Scharfetter-Gummel box-integration
// residual and Jacobian on the structured 2D/3D stencils.
//
// One thread owns one node.  It evaluates the 2*dim edges touching its
// node (each edge is evaluated by both endpoints with identical
// arithmetic, so the two sides of every flux agree bit for bit), sums its
// own row, applies the gate Robin term and the contact Dirichlet rows,
// eliminates contact columns into the right-hand side, scales the row and
// inverts its 3x3 node block.  No atomics: every output is written by
// exactly one thread, and sums run in the same order as the reference's
// per-edge scatter (hi edge then lo edge, axis x then y then z), so the
// residual F rounds exactly as pytcad/pytcad/device2d.py:1450-1529 /
// device3d.py:1485-1621 does except where expm1/log differ from glibc.
//
// Roofline: HBM bound.  Algorithmic bytes per node (2D, z = 29 entries):
// reads 24 (psi,n,p) + 11*8 static; writes 24 F + 24 rhs + 24 rowscale +
// 8*29 A64 + 4*29 A32 + 8*9 + 4*9 block inverses = ~640 B.
#include "physics_dev.h"
#include "tcad_internal.h"

namespace tcad {

struct AsmArgs {
  const DevInfo *info;
  const int *blk2dev;
  long long N;
  const double *psi, *n, *p;
  const double *C, *nie_s, *nie, *tau_n, *tau_p, *dV;
  const double *an[3], *ap[3], *dln[3];
  const double *h[3], *dv[3];
  const int *cidx, *gidx;
  const double *cn_lnr, *cn_n0, *cn_p0;
  const int *cn_contact;
  const double *contact_V;
  const double *gn_kw, *gn_psib;
  const int *gn_gate;
  const double *gate_Vg, *gate_Vfb;
  int ne;
  int want32;
  int res_only;                     // 1: only the raw F_n + F_p; 2: only F (all rows)
  double *Fn_raw, *Fp_raw;          // optional raw continuity residuals per node
  // transient step (transient2d.py:104-135): storage term on the
  // non-contact continuity rows; theta-scheme with the old residuals
  int transient;
  const double *n_old, *p_old, *Fn_old, *Fp_old;
  const double *dt_s;               // per device, scaled time step
  double theta;
  // small-signal AC (ac2d.py:265-342): 1 = prepare (the matrix as usual,
  // plus the scaled storage diagonal and the raw current rows of the
  // contact nodes); 2 = only the scaled right-hand side of a port's
  // forcing acg (row scales from the prepare call)
  int ac;
  const double *acg;
  double *ac_dn, *ac_dp;            // R * (-dV) * n, R * dV * p (0 at contacts)
  double *ac_raw;                   // 21 planes: raw (n + p)-row sums, node then directions
  double *F, *Fnp, *A64, *Dinv64, *bs, *rowscale;
  float *A32, *Dinv32;
  double *Jpre;                     // optional: raw rows before the contact rows (ne planes)
  int *flag;                        // [0] any, [1 + d] per device
  const Scal *scal;                 // non-NULL: skip devices not active
};

struct EdgeVals {
  double Fpsi, w;                   // potential flux, Poisson weight face/h
  double Jn, Jp;
  double dpsiRn, dnL, dnR;          // electron current derivatives
  double dpsiRp, dpL, dpR;          // hole current derivatives
};

__device__ __forceinline__ EdgeVals eval_edge(double psiL, double psiR, double nL, double nR,
                                              double pL, double pR, double dln, double an,
                                              double ap, double h, double face) {
  EdgeVals e;
  double dpsi = psiR - psiL;
  double d = dpsi + dln + 0.0;
  double dh = dpsi - dln + 0.0;
  BQuad q = bernoulli4(d);
  BQuad qh = bernoulli4(dh);
  e.Fpsi = 1.0 * dpsi / h;
  e.w = face * 1.0 / h;
  e.Jn = an * (nR * q.Bp - nL * q.Bm);
  e.Jp = -ap * (pR * qh.Bm - pL * qh.Bp);
  double Sn = nR * q.dBp + nL * q.dBm;
  e.dpsiRn = an * Sn;
  e.dnL = -an * q.Bm;
  e.dnR = an * q.Bp;
  double Sp = pR * qh.dBm + pL * qh.dBp;
  e.dpsiRp = ap * Sp;
  e.dpL = ap * qh.Bp;
  e.dpR = -ap * qh.Bm;
  return e;
}

__device__ __forceinline__ void node_coords(const DevInfo &I, long long k, int c[3]) {
  c[0] = (int)(k % I.nx);
  c[1] = (int)((k / I.nx) % I.ny);
  c[2] = (int)(k / ((long long)I.nx * I.ny));
}

// Face measure of an axis-a edge at node coordinates c: product of the
// other axes' widths in increasing axis order (device2d.py:875,
// device3d.py:1600-1603).
template <int DIM>
__device__ __forceinline__ double face_of(const AsmArgs &A, const DevInfo &I, int a,
                                          const int c[3]) {
  double f = 0.0;
  bool first = true;
#pragma unroll
  for (int b = 0; b < DIM; ++b) {
    if (b == a) continue;
    double v = A.dv[b][I.dv_off[b] + c[b]];
    f = first ? v : f * v;
    first = false;
  }
  return f;
}

__device__ __forceinline__ void contact_target(const AsmArgs &A, const DevInfo &I, int t,
                                               double &psi0, double &n0, double &p0) {
  double V = A.contact_V[A.cn_contact[t]];
  psi0 = V / I.VT + A.cn_lnr[t];
  n0 = A.cn_n0[t];
  p0 = A.cn_p0[t];
}

template <int DIM>
__device__ void assemble_node(const AsmArgs &A, const DevInfo &I, long long k, int dev) {
  const long long N = A.N;
  const long long g = I.off + k;
  int c[3];
  node_coords(I, k, c);
  const int dims[3] = {I.nx, I.ny, I.nz};
  const long long str[3] = {1, (long long)I.nx, (long long)I.nx * I.ny};

  const double psi = A.psi[g], nn = A.n[g], pp = A.p[g];
  const double dV = A.dV[g];

  double D[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) D[e] = 0.0;
  double off[6][5];
#pragma unroll
  for (int dir = 0; dir < 6; ++dir)
#pragma unroll
    for (int m = 0; m < 5; ++m) off[dir][m] = 0.0;
  bool has_nb[6] = {false, false, false, false, false, false};

  double tpsi = 0.0, tn = 0.0, tp = 0.0;
  // The reference sums scatter contributions in COO order: every Poisson
  // scatter (all axes) precedes every electron scatter, which precedes
  // every hole scatter; within one scatter a node's (L,L) entry (its hi
  // edge) precedes its (R,R) entry (its lo edge).  Each node-block entry
  // only collects terms of one kind, so per-axis hi-then-lo order is the
  // whole story.
  double s00 = 0.0, s10 = 0.0, s11 = 0.0, s20 = 0.0, s22 = 0.0;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    const double face = face_of<DIM>(A, I, a, c);
    double divF = 0.0, divJn = 0.0, divJp = 0.0;
    if (c[a] < dims[a] - 1) {         // hi edge: this node is L
      const long long gR = g + str[a];
      const double h = A.h[a][I.h_off[a] + c[a]];
      EdgeVals e = eval_edge(psi, A.psi[gR], nn, A.n[gR], pp, A.p[gR], A.dln[a][g],
                             A.an[a][g], A.ap[a][g], h, face);
      divF = divF + e.Fpsi;
      divJn = divJn + e.Jn;
      divJp = divJp + e.Jp;
      s00 = s00 + e.w * -1.0;
      s10 = s10 + face * -e.dpsiRn;
      s11 = s11 + face * e.dnL;
      s20 = s20 + face * -e.dpsiRp;
      s22 = s22 + face * e.dpL;
      const int dir = 2 * a + 1;
      has_nb[dir] = true;
      off[dir][0] = e.w * 1.0;
      off[dir][1] = face * e.dpsiRn;
      off[dir][2] = face * e.dnR;
      off[dir][3] = face * e.dpsiRp;
      off[dir][4] = face * e.dpR;
    }
    if (c[a] > 0) {                   // lo edge: this node is R
      const long long gL = g - str[a];
      const double h = A.h[a][I.h_off[a] + c[a] - 1];
      EdgeVals e = eval_edge(A.psi[gL], psi, A.n[gL], nn, A.p[gL], pp, A.dln[a][gL],
                             A.an[a][gL], A.ap[a][gL], h, face);
      divF = divF - e.Fpsi;
      divJn = divJn - e.Jn;
      divJp = divJp - e.Jp;
      const double mw = -e.w;
      const double mf = -face;
      s00 = s00 + mw * 1.0;
      s10 = s10 + mf * e.dpsiRn;
      s11 = s11 + mf * e.dnR;
      s20 = s20 + mf * e.dpsiRp;
      s22 = s22 + mf * e.dpR;
      const int dir = 2 * a;
      has_nb[dir] = true;
      off[dir][0] = mw * -1.0;
      off[dir][1] = mf * -e.dpsiRn;
      off[dir][2] = mf * e.dnL;
      off[dir][3] = mf * -e.dpsiRp;
      off[dir][4] = mf * e.dpL;
    }
    const double tF = face * divF, tJn = face * divJn, tJp = face * divJp;
    if (a == 0) { tpsi = tF; tn = tJn; tp = tJp; }
    else { tpsi = tpsi + tF; tn = tn + tJn; tp = tp + tJp; }
  }

  // recombination, physical units (device2d.py:1481-1508)
  double Rs = 0.0, dRs_dn = 0.0, dRs_dp = 0.0;
  if (I.srh) {
    double R, dRdn, dRdp;
    recombination(nn * I.Ns, pp * I.Ns, A.nie[g], A.tau_n[g], A.tau_p[g], I.Cn, I.Cp,
                  I.auger, R, dRdn, dRdp);
    Rs = R / I.R0;
    dRs_dn = dRdn * I.Ns / I.R0;
    dRs_dp = dRdp * I.Ns / I.R0;
  }
  double Fpsi = tpsi - dV * (nn - pp - A.C[g]);
  double Fn = tn - Rs * dV;
  double Fp = tp + Rs * dV;

  D[0] = s00;
  D[1] = -dV;
  D[2] = dV;
  D[3] = s10;
  D[4] = s11 + -dRs_dn * dV;
  D[5] = -dRs_dp * dV;
  D[6] = s20;
  D[7] = dRs_dn * dV;
  D[8] = s22 + dRs_dp * dV;

  // gate Robin term on the potential row (device2d.py:1623-1637)
  const int gt = A.gidx[g];
  if (gt >= 0) {
    const int gate = A.gn_gate[gt];
    const double kw = A.gn_kw[gt];
    const double Vg_s = A.gate_Vg[gate] / I.VT, Vfb_s = A.gate_Vfb[gate] / I.VT;
    Fpsi = Fpsi + kw * (Vg_s - Vfb_s - (psi - A.gn_psib[gt]));
    D[0] = D[0] + -kw;
  }

  if (A.Jpre) {
    // the node's raw Jacobian row entries before contact rows become
    // Dirichlet rows (device2d.py:1742-1760, device3d.py:1875-1912):
    // node block, then per direction the five coupling entries
#pragma unroll
    for (int e = 0; e < 9; ++e) A.Jpre[e * N + g] = D[e];
    const int ndir = (A.ne - 9) / 5;
    for (int dir = 0; dir < ndir; ++dir)
#pragma unroll
      for (int m = 0; m < 5; ++m) A.Jpre[(9 + 5 * dir + m) * N + g] = off[dir][m];
  }

  const int ct = A.cidx[g];
  if (A.ac >= 2) {
    // AC forcing b (ac2d.py:296-313) on the psi rows: an ohmic port's 1 on
    // its nodes is du_psi there (Dirichlet rows; the interior sees
    // -J[:, contact] b after elimination), a gate port's -kappa w forces
    // its nodes' potential rows
    double r3[3] = {0.0, 0.0, 0.0};
    r3[0] = A.acg[g];
    if (ct < 0) {
#pragma unroll
      for (int dir = 0; dir < 6; ++dir) {
        if (dir / 2 >= DIM || !has_nb[dir]) continue;
        const long long gn = g + ((dir & 1) ? str[dir / 2] : -str[dir / 2]);
        if (A.cidx[gn] < 0) continue;
        const double gv = A.acg[gn];
        r3[0] = r3[0] - off[dir][0] * gv;
        r3[1] = r3[1] - off[dir][1] * gv;
        r3[2] = r3[2] - off[dir][3] * gv;
      }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) A.bs[r * N + g] = r3[r] * A.rowscale[r * N + g];
    return;
  }
  if (A.ac == 1 && ct >= 0) {
    // d(F_n + F_p)/du of a contact node, before its rows become Dirichlet
    // rows (device2d.py:1751-1758): node columns, then per direction
    A.ac_raw[0 * N + g] = D[3] + D[6];
    A.ac_raw[1 * N + g] = D[4] + D[7];
    A.ac_raw[2 * N + g] = D[5] + D[8];
#pragma unroll
    for (int dir = 0; dir < 6; ++dir) {
      const bool on = dir / 2 < DIM && has_nb[dir];
      A.ac_raw[(3 + 3 * dir) * N + g] = on ? off[dir][1] + off[dir][3] : 0.0;
      A.ac_raw[(4 + 3 * dir) * N + g] = on ? off[dir][2] : 0.0;
      A.ac_raw[(5 + 3 * dir) * N + g] = on ? off[dir][4] : 0.0;
    }
  }

  double F3[3] = {Fpsi, Fn, Fp};
  double rhs[3];
  if (A.Fnp) A.Fnp[g] = Fn + Fp;      // raw, pre-Dirichlet (terminal current)
  if (A.Fn_raw) {
    A.Fn_raw[g] = Fn;
    A.Fp_raw[g] = Fp;
  }
  if (A.res_only == 1) return;
  if (A.transient && ct < 0) {
    // F_n <- theta F_n + (1 - theta) F_n(old) - dV/dt (n - n_old), rows
    // scaled by theta, storage on the diagonal; same order of operations
    // as transient2d.py:118-135 (theta == 1 leaves F and J untouched)
    const double th = A.theta, dts = A.dt_s[dev];
    if (th != 1.0) {
      F3[1] = th * F3[1] + (1.0 - th) * A.Fn_old[g];
      F3[2] = th * F3[2] + (1.0 - th) * A.Fp_old[g];
#pragma unroll
      for (int e = 3; e < 9; ++e) D[e] = th * D[e];
#pragma unroll
      for (int dir = 0; dir < 6; ++dir)
#pragma unroll
        for (int m = 1; m < 5; ++m) off[dir][m] = th * off[dir][m];
    }
    F3[1] = F3[1] - dV / dts * (nn - A.n_old[g]);
    F3[2] = F3[2] + dV / dts * (pp - A.p_old[g]);
    D[4] = D[4] + -dV / dts;
    D[8] = D[8] + dV / dts;
  }
  if (ct >= 0) {
    // contact node: Dirichlet rows (device2d.py:1800-1852)
    double psi0, n0, p0;
    contact_target(A, I, ct, psi0, n0, p0);
    F3[0] = psi - psi0;
    F3[1] = nn - n0;
    F3[2] = pp - p0;
#pragma unroll
    for (int e = 0; e < 9; ++e) D[e] = (e % 4 == 0) ? 1.0 : 0.0;
#pragma unroll
    for (int dir = 0; dir < 6; ++dir)
#pragma unroll
      for (int m = 0; m < 5; ++m) off[dir][m] = 0.0;
#pragma unroll
    for (int r = 0; r < 3; ++r) rhs[r] = -F3[r];
    if (A.res_only == 2) {
#pragma unroll
      for (int r = 0; r < 3; ++r) A.F[r * N + g] = F3[r];
      return;
    }
  } else {
    if (A.res_only == 2) {
#pragma unroll
      for (int r = 0; r < 3; ++r) A.F[r * N + g] = F3[r];
      return;
    }
    // symmetric elimination of contact columns (dirichlet.py:161-191):
    // rhs -= J[:, contact] * g with g = -F(contact), in column order
    double corr[3] = {0.0, 0.0, 0.0};
    const int order[6] = {4, 2, 0, 1, 3, 5};   // z-, y-, x-, x+, y+, z+
#pragma unroll
    for (int oi = 0; oi < 6; ++oi) {
      const int dir = order[oi];
      if (dir / 2 >= DIM || !has_nb[dir]) continue;
      const long long gn = g + ((dir & 1) ? str[dir / 2] : -str[dir / 2]);
      const int cn = A.cidx[gn];
      if (cn < 0) continue;
      double psi0, n0, p0;
      contact_target(A, I, cn, psi0, n0, p0);
      const double gp = -(A.psi[gn] - psi0), gnn = -(A.n[gn] - n0), gpp = -(A.p[gn] - p0);
      corr[0] = corr[0] + off[dir][0] * gp;
      corr[1] = corr[1] + off[dir][1] * gp;
      corr[1] = corr[1] + off[dir][2] * gnn;
      corr[2] = corr[2] + off[dir][3] * gp;
      corr[2] = corr[2] + off[dir][4] * gpp;
#pragma unroll
      for (int m = 0; m < 5; ++m) off[dir][m] = 0.0;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) rhs[r] = -F3[r] - corr[r];
  }

  // column scaling Cc = (1, n, p) and row equilibration R = 1/max|J Cc|
  double nb_n[6], nb_p[6];
#pragma unroll
  for (int dir = 0; dir < 6; ++dir) {
    nb_n[dir] = 0.0;
    nb_p[dir] = 0.0;
    if (dir / 2 < DIM && has_nb[dir]) {
      const long long gn = g + ((dir & 1) ? str[dir / 2] : -str[dir / 2]);
      nb_n[dir] = A.n[gn];
      nb_p[dir] = A.p[gn];
    }
  }
  const double cs[3] = {1.0, nn, pp};
  double Ds[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) Ds[3 * r + cc] = D[3 * r + cc] * cs[cc];
  double os[6][5];
  double rmax[3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
    rmax[r] = fmax(fmax(fabs(Ds[3 * r]), fabs(Ds[3 * r + 1])), fabs(Ds[3 * r + 2]));
#pragma unroll
  for (int dir = 0; dir < 6; ++dir) {
    os[dir][0] = off[dir][0];
    os[dir][1] = off[dir][1];
    os[dir][2] = off[dir][2] * nb_n[dir];
    os[dir][3] = off[dir][3];
    os[dir][4] = off[dir][4] * nb_p[dir];
    rmax[0] = fmax(rmax[0], fabs(os[dir][0]));
    rmax[1] = fmax(rmax[1], fmax(fabs(os[dir][1]), fabs(os[dir][2])));
    rmax[2] = fmax(rmax[2], fmax(fabs(os[dir][3]), fabs(os[dir][4])));
  }
  double rs[3];
  bool bad = false;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    if (!(rmax[r] > 0.0) || !isfinite(rmax[r])) { bad = true; rs[r] = 1.0; }
    else rs[r] = 1.0 / rmax[r];
  }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) Ds[3 * r + cc] *= rs[r];
#pragma unroll
  for (int dir = 0; dir < 6; ++dir) {
    os[dir][0] *= rs[0];
    os[dir][1] *= rs[1];
    os[dir][2] *= rs[1];
    os[dir][3] *= rs[2];
    os[dir][4] *= rs[2];
  }

  // 3x3 inverse by cofactors (node block-Jacobi, linsolve.py:363-427)
  const double a0 = Ds[0], a1 = Ds[1], a2 = Ds[2], a3 = Ds[3], a4 = Ds[4], a5 = Ds[5],
               a6 = Ds[6], a7 = Ds[7], a8 = Ds[8];
  const double c0 = a4 * a8 - a5 * a7, c1 = a5 * a6 - a3 * a8, c2 = a3 * a7 - a4 * a6;
  const double det = a0 * c0 + a1 * c1 + a2 * c2;
  if (!(fabs(det) > 1e-300) || !isfinite(det)) bad = true;
  const double id = 1.0 / det;
  double Di[9] = {c0 * id, (a2 * a7 - a1 * a8) * id, (a1 * a5 - a2 * a4) * id,
                  c1 * id, (a0 * a8 - a2 * a6) * id, (a2 * a3 - a0 * a5) * id,
                  c2 * id, (a1 * a6 - a0 * a7) * id, (a0 * a4 - a1 * a3) * id};
  if (bad) {                          // error flags only; never part of a sum
    atomicOr(A.flag, 1);
    atomicOr(A.flag + 1 + dev, 1);
  }

#pragma unroll
  for (int r = 0; r < 3; ++r) {
    A.F[r * N + g] = F3[r];
    A.bs[r * N + g] = rhs[r] * rs[r];
    A.rowscale[r * N + g] = rs[r];
  }
  if (A.ac == 1) {
    // storage term of the AC system (ac2d.py:107-124) in the scaled frame
    A.ac_dn[g] = ct < 0 ? rs[1] * (-dV) * nn : 0.0;
    A.ac_dp[g] = ct < 0 ? rs[2] * dV * pp : 0.0;
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    A.A64[e * N + g] = Ds[e];
    A.Dinv64[e * N + g] = Di[e];
  }
  const int ndir = (A.ne - 9) / 5;
  for (int dir = 0; dir < ndir; ++dir)
#pragma unroll
    for (int m = 0; m < 5; ++m) A.A64[(9 + 5 * dir + m) * N + g] = os[dir][m];
  if (A.want32) {
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      A.A32[e * N + g] = (float)Ds[e];
      A.Dinv32[e * N + g] = (float)Di[e];
    }
    for (int dir = 0; dir < ndir; ++dir)
#pragma unroll
      for (int m = 0; m < 5; ++m) A.A32[(9 + 5 * dir + m) * N + g] = (float)os[dir][m];
  }
}

__global__ void __launch_bounds__(BS) k_assemble(AsmArgs A) {
  const int d = A.blk2dev[blockIdx.x];
  if (A.scal && !A.scal[d].active) return;
  const DevInfo &I = A.info[d];
  const NodeAt P = node_at(I, blockIdx.x);
  if (!P.valid) return;
  const long long k = P.k;
  if (I.dim == 2) assemble_node<2>(A, I, k, d);
  else assemble_node<3>(A, I, k, d);
}

static AsmArgs make_args(Batch &b, const double *psi, const double *n, const double *p) {
  AsmArgs A{};
  A.info = b.d_info;
  A.blk2dev = b.d_blk2dev;
  A.N = b.ntot;
  A.psi = psi; A.n = n; A.p = p;
  A.C = b.f.d_C; A.nie_s = b.f.d_nie_s; A.nie = b.f.d_nie;
  A.tau_n = b.f.d_tau_n; A.tau_p = b.f.d_tau_p; A.dV = b.f.d_dV;
  for (int a = 0; a < 3; ++a) {
    A.an[a] = b.f.d_an[a]; A.ap[a] = b.f.d_ap[a]; A.dln[a] = b.f.d_dlnnie[a];
    A.h[a] = b.d_h[a]; A.dv[a] = b.d_dv[a];
  }
  A.cidx = b.d_cidx; A.gidx = b.d_gidx;
  A.cn_lnr = b.d_cn_lnr; A.cn_n0 = b.d_cn_n0; A.cn_p0 = b.d_cn_p0;
  A.cn_contact = b.d_cn_contact; A.contact_V = b.d_contact_V;
  A.gn_kw = b.d_gn_kw; A.gn_psib = b.d_gn_psib; A.gn_gate = b.d_gn_gate;
  A.gate_Vg = b.d_gate_Vg; A.gate_Vfb = b.d_gate_Vfb;
  A.ne = b.ne;
  A.flag = b.d_flag;
  return A;
}

int launch_assemble(Batch &b, const double *psi, const double *n, const double *p,
                    double *F, double *Fnp, bool want_fp32, bool only_active, int res_only,
                    const TransientArgs *tr, const AcAssembly *ac, double *Jpre) {
  AsmArgs A = make_args(b, psi, n, p);
  A.Jpre = Jpre;
  if (ac) {
    A.ac = ac->mode;
    A.acg = ac->acg;
    A.ac_dn = ac->dn;
    A.ac_dp = ac->dp;
    A.ac_raw = ac->raw;
  }
  A.scal = only_active ? b.d_scal : nullptr;
  A.res_only = res_only;
  if (tr) {
    A.transient = tr->active;
    A.n_old = tr->n_old;
    A.p_old = tr->p_old;
    A.Fn_old = tr->Fn_old;
    A.Fp_old = tr->Fp_old;
    A.dt_s = tr->dt_s;
    A.theta = tr->theta;
    A.Fn_raw = tr->Fn_raw;
    A.Fp_raw = tr->Fp_raw;
  }
  A.want32 = want_fp32 ? 1 : 0;
  A.F = F; A.Fnp = Fnp;
  A.A64 = b.A64; A.A32 = b.A32; A.Dinv64 = b.Dinv64; A.Dinv32 = b.Dinv32;
  A.bs = b.bs; A.rowscale = b.rowscale;
  if (!res_only && !(ac && ac->mode >= 2)) {
    b.last_n = n;
    b.last_p = p;
    b.last_psi = psi;
  }
  hipLaunchKernelGGL(k_assemble, dim3(b.nblk), dim3(BS), 0, b.stream, A);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

// ---------------------------------------------------------------------------
// Equilibrium Poisson (device2d.py:844-942, device3d.py:190-258, 952-1020):
// slaved n = nie_s e^psi, p = nie_s e^-psi.  The stored system is the
// eliminated Jacobian NEGATED on free rows (symmetric positive definite,
// solved by PCG) with contact rows kept as e_k; entries per node:
// centre, x-, x+, y-, y+, z-, z+.
// ---------------------------------------------------------------------------
template <int DIM>
__device__ void assemble_poisson_node(const AsmArgs &A, const DevInfo &I, long long k, int dev,
                                      double *Aeq, double *beq, double *F) {
  const long long N = A.N;
  const long long g = I.off + k;
  int c[3];
  node_coords(I, k, c);
  const int dims[3] = {I.nx, I.ny, I.nz};
  const long long str[3] = {1, (long long)I.nx, (long long)I.nx * I.ny};
  const double psi = A.psi[g];
  const double nie = A.nie_s[g];
  const double nn = nie * exp(fmin(fmax(psi, -700.0), 700.0));
  const double pp = nie * exp(fmin(fmax(-psi, -700.0), 700.0));
  const double dnp = nn + pp;
  const double dV = A.dV[g];
  double diag = 0.0;
  double offv[6] = {0, 0, 0, 0, 0, 0};
  bool has_nb[6] = {false, false, false, false, false, false};
  double tpsi = 0.0;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    const double face = face_of<DIM>(A, I, a, c);
    double divF = 0.0;
    if (c[a] < dims[a] - 1) {
      const double h = A.h[a][I.h_off[a] + c[a]];
      const double Fa = 1.0 * (A.psi[g + str[a]] - psi) / h;
      const double w = face / h * 1.0;
      divF = divF + Fa;
      diag = diag + -w;
      offv[2 * a + 1] = w;
      has_nb[2 * a + 1] = true;
    }
    if (c[a] > 0) {
      const double h = A.h[a][I.h_off[a] + c[a] - 1];
      const double Fa = 1.0 * (psi - A.psi[g - str[a]]) / h;
      const double w = face / h * 1.0;
      divF = divF - Fa;
      diag = diag + -w;
      offv[2 * a] = w;
      has_nb[2 * a] = true;
    }
    const double t = face * divF;
    tpsi = (a == 0) ? t : tpsi + t;
  }
  double Fv = tpsi - dV * (nn - pp - A.C[g]);
  diag = diag + -dV * dnp;
  const int gt = A.gidx[g];
  if (gt >= 0) {
    const int gate = A.gn_gate[gt];
    const double kw = A.gn_kw[gt];
    const double Vfb_s = A.gate_Vfb[gate] / I.VT;
    Fv = Fv + kw * (0.0 - Vfb_s - (psi - A.gn_psib[gt]));
    diag = diag + -kw;
  }
  const int ct = A.cidx[g];
  double rhs;
  double M[7];
  if (ct >= 0) {
    Fv = psi - (0.0 / I.VT + A.cn_lnr[ct]);
    rhs = -Fv;
    M[0] = 1.0;
    for (int e = 1; e < 7; ++e) M[e] = 0.0;
  } else {
    double corr = 0.0;
    const int order[6] = {4, 2, 0, 1, 3, 5};
#pragma unroll
    for (int oi = 0; oi < 6; ++oi) {
      const int dir = order[oi];
      if (dir / 2 >= DIM || !has_nb[dir]) continue;
      const long long gn = g + ((dir & 1) ? str[dir / 2] : -str[dir / 2]);
      const int cn = A.cidx[gn];
      if (cn < 0) continue;
      const double gv = -(A.psi[gn] - (0.0 / I.VT + A.cn_lnr[cn]));
      corr = corr + offv[dir] * gv;
      offv[dir] = 0.0;
    }
    rhs = -Fv - corr;
    // negate the free rows: SPD system -J du = -rhs
    M[0] = -diag;
    for (int dir = 0; dir < 6; ++dir) M[1 + dir] = -offv[dir];
    rhs = -rhs;
  }
  if (!(M[0] > 0.0) || !isfinite(M[0])) {
    atomicOr(A.flag, 1);
    atomicOr(A.flag + 1 + dev, 1);
  }
  F[g] = Fv;
  beq[g] = rhs;
  const int nent = 1 + 2 * DIM;
  for (int e = 0; e < nent; ++e) Aeq[e * N + g] = M[e];
}

__global__ void __launch_bounds__(BS) k_assemble_poisson(AsmArgs A, double *Aeq, double *beq,
                                                         double *F) {
  const int d = A.blk2dev[blockIdx.x];
  if (!A.scal[d].active) return;
  const DevInfo &I = A.info[d];
  const NodeAt P = node_at(I, blockIdx.x);
  if (!P.valid) return;
  const long long k = P.k;
  if (I.dim == 2) assemble_poisson_node<2>(A, I, k, d, Aeq, beq, F);
  else assemble_poisson_node<3>(A, I, k, d, Aeq, beq, F);
}

int launch_assemble_poisson(Batch &b, const double *psi, double *F) {
  AsmArgs A = make_args(b, psi, nullptr, nullptr);
  A.scal = b.d_scal;
  hipLaunchKernelGGL(k_assemble_poisson, dim3(b.nblk), dim3(BS), 0, b.stream, A, b.Aeq,
                     b.beq, F);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

// ---------------------------------------------------------------------------
// Lagged Lombardi surface mobility on the row-0 y-edges
// (device2d.py:612-642): one thread per surface column of each 2D device
// with the model on.  Rewrites an_y / ap_y of row 0.
// ---------------------------------------------------------------------------
__global__ void k_surface_mobility(const DevInfo *info, int ndev, const Scal *scal,
                                   const double *psi, const double *hy, double *an_y,
                                   double *ap_y, const double *mun0, const double *mun1,
                                   const double *mup0, const double *mup1) {
  const int d = blockIdx.y;
  if (d >= ndev) return;
  if (scal && !scal[d].active) return;
  const DevInfo &I = info[d];
  if (!I.surf || I.dim != 2) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I.nx) return;
  const long long g = I.off + i;
  const double hy0 = hy[I.h_off[1]];
  const double E = fabs(psi[g] - psi[g + I.nx]) * I.VT / (hy0 * I.LD);
  const int so = I.surf_off + i;
  const double mns = mobility_cvt(E, mun0[so], 0, I.T);
  const double mps = mobility_cvt(E, mup0[so], 1, I.T);
  const double dn = 2.0 * mns * mun1[so] / (mns + mun1[so]) * I.VT / 1.0;
  const double dp = 2.0 * mps * mup1[so] / (mps + mup1[so]) * I.VT / 1.0;
  an_y[g] = dn / hy0;
  ap_y[g] = dp / hy0;
}

int launch_surface_mobility(Batch &b, const double *psi, bool only_active) {
  if (!b.any_surf) return TCAD_OK;
  int maxnx = 1;
  for (auto &I : b.info) maxnx = I.nx > maxnx ? I.nx : maxnx;
  dim3 grid((maxnx + 127) / 128, b.ndev);
  hipLaunchKernelGGL(k_surface_mobility, grid, dim3(128), 0, b.stream, b.d_info, b.ndev,
                     only_active ? b.d_scal : nullptr, psi,
                     b.d_h[1], const_cast<double *>(b.f.d_an[1]),
                     const_cast<double *>(b.f.d_ap[1]), b.d_surf[0], b.d_surf[1], b.d_surf[2],
                     b.d_surf[3]);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

// ---------------------------------------------------------------------------
// Edge currents Jn, Jp (scaled) per axis, device2d.py:1463-1479.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BS) k_edge_currents(AsmArgs A, double *J) {
  const int d = A.blk2dev[blockIdx.x];
  const DevInfo &I = A.info[d];
  const NodeAt P = node_at(I, blockIdx.x);
  if (!P.valid) return;
  const long long k = P.k;
  const long long N = A.N, g = I.off + k;
  int c[3];
  node_coords(I, k, c);
  const int dims[3] = {I.nx, I.ny, I.nz};
  const long long str[3] = {1, (long long)I.nx, (long long)I.nx * I.ny};
  for (int a = 0; a < I.dim; ++a) {
    double jn = 0.0, jp = 0.0;
    if (c[a] < dims[a] - 1) {
      const long long gR = g + str[a];
      EdgeVals e = eval_edge(A.psi[g], A.psi[gR], A.n[g], A.n[gR], A.p[g], A.p[gR],
                             A.dln[a][g], A.an[a][g], A.ap[a][g],
                             A.h[a][I.h_off[a] + c[a]], 1.0);
      jn = e.Jn;
      jp = e.Jp;
    }
    J[(2 * a) * N + g] = jn;
    J[(2 * a + 1) * N + g] = jp;
  }
}

int launch_edge_currents(Batch &b, const double *psi, const double *n, const double *p,
                         double *J) {
  AsmArgs A = make_args(b, psi, n, p);
  hipLaunchKernelGGL(k_edge_currents, dim3(b.nblk), dim3(BS), 0, b.stream, A, J);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

// ---------------------------------------------------------------------------
// Terminal currents: one workgroup per contact sums the raw F_n + F_p of
// its node list in a fixed order (strided partial sums, then a fixed
// tree), times J0 * LD^(dim-1) (device2d.py:2217-2226).
// ---------------------------------------------------------------------------
__global__ void k_contact_sums(const int *cptr, const int *cnodes, const double *Fnp,
                               const double *cscale, double *out) {
  __shared__ double sh[256];
  const int c = blockIdx.x;
  const int b0 = cptr[c], b1 = cptr[c + 1];
  double s = 0.0;
  for (int t = b0 + threadIdx.x; t < b1; t += blockDim.x) s = s + Fnp[cnodes[t]];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = sh[0] * cscale[2 * c] * cscale[2 * c + 1];
}

int launch_contact_sums(Batch &b, const double *Fnp, double *d_out) {
  if (b.ncontact == 0) return TCAD_OK;
  hipLaunchKernelGGL(k_contact_sums, dim3(b.ncontact), dim3(256), 0, b.stream, b.d_cptr,
                     b.d_cnodes, Fnp, b.d_cscale, d_out);
  b.launches++;
  TCAD_HIP_CHECK(hipGetLastError());
  return TCAD_OK;
}

}  // namespace tcad
