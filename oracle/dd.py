"""Oracle drift-diffusion devices: Device2D and Device3D restated in FP64
NumPy/SciPy on one dimension-generic core.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows:
  Device2D  pytcad/pytcad/device2d.py
            __init__ :168-453, add_contact/add_gate :672-680,
            _bc_contact_values :689-742, _bulk_psi_guess :744-790,
            _update_surface_mobility :612-642,
            _residual_jacobian_poisson :844-942, solve_equilibrium
            :1308-1373, _residual_jacobian :1402-1920, solve_bias
            :1923-2191, terminal_current :2194-2227
  Device3D  pytcad/pytcad/device3d.py
            __init__ :437-724, _gate_face_weight :939-947,
            _residual_jacobian_poisson :952-1020 (stencil oracle
            :190-258), solve_equilibrium :1311-1454, _residual_jacobian
            :1485-1978 (COO oracles :320-408), solve_bias :1981-2271,
            terminal_current :2274-2306
restricted to what the structured Boltzmann path uses: one
semiconductor (edge permittivity ratio 1), band_offset "nie" (zero band
shift), ohmic Dirichlet contacts with S_n = S_p = 0, Robin gates,
optional lagged Lombardi surface mobility (2D), SRH/Auger, BGN,
Caughey-Thomas mobility.

Unknowns are interleaved (psi, n, p) per node, node k = j*Nx + i in 2D
and k = kk*Nx*Ny + j*Nx + i in 3D (mesh2d.py:1-8, mesh3d.py:1-9).  All
quantities are in the reference's scaled units (psi/VT, n/Ns, x/LD).

Pinning: golden fixtures generated from the reference itself, matched
bit for bit (tests/test_oracle_golden.py).  2D: tests/golden/make_golden.py
(reference Device2D).  3D: tests/golden/make_golden_3d.py (reference
Device3D, its five compiled assembly blocks replaced by the reference's
own pure-NumPy bodies, device3d.py:190-258, 320-408).  Each fixture holds
the residual F, the raw Jacobian J, the edge currents and F_n/F_p at a
perturbed state (3D also the equilibrium Poisson F/J), and the
equilibrium psi, bias-point psi/n/p, Newton iteration counts and terminal
currents of a sweep: 3D covers a tri-gate (y- and z-normal gates, body
contact overlapping source and drain) at Vd = 0.05 and 0.7 V and slabs
with x- and z-normal gates.  tests/test_oracle_dd3d.py adds invariants
(z-invariant 3D equals 2D, axis permutation, finite-difference Jacobian,
current conservation).
"""
import time
import warnings

import numpy as np
from scipy.sparse import csr_matrix

from . import physics as ph
from .linsolve import LinearSolveError, eliminate_csr, solve_linear


class DirichletBC:
    """device2d.py:70-76 / device3d.py:81-88."""

    def __init__(self, idx, V=0.0):
        self.idx = idx          # tuple of index arrays in array order
        self.V = float(V)


class GateBC:
    """device2d.py:79-88 / device3d.py:120-136."""

    def __init__(self, idx, kappa, Vfb, Vg, normal_axis):
        self.idx = idx
        self.kappa = float(kappa)
        self.Vfb = float(Vfb)
        self.Vg = float(Vg)
        self.normal_axis = normal_axis


def _hmean(lo, hi):
    return 2.0 * lo * hi / (lo + hi)


class _DeviceND:
    """Shared core; `ndim` is 2 or 3, arrays are shaped (Ny, Nx) or
    (Nz, Ny, Nx).  Axis letters map to array axes x -> ndim-1,
    y -> ndim-2, z -> 0 (3D)."""

    def __init__(self, coords, doping, Ntotal=None, T=300.0,
                 material=ph.SILICON, models=None):
        self.ndim = len(coords)
        self.coords = [np.asarray(c, dtype=float) for c in coords]  # x, y[, z]
        self.shape = tuple(c.size for c in self.coords[::-1])
        self.N = int(np.prod(self.shape))
        self.doping = np.asarray(doping, dtype=float).reshape(self.shape)
        self.Ntot = (np.abs(self.doping) if Ntotal is None
                     else np.asarray(Ntotal, dtype=float).reshape(self.shape))
        self.T = T
        self.mat = material
        self.models = models or ph.Models()
        if self.Ntot.max() > 1e19:
            warnings.warn("Doping exceeds ~1e19 cm^-3: Boltzmann statistics "
                          "used here overestimate the carrier density.")
        # device2d.py:319-345 / device3d.py:537-567
        self.VT = ph.thermal_voltage(T)
        self.eps = material.eps_r * ph.EPS0
        self.ni = material.ni(T)
        self.Ns = max(float(np.abs(self.doping).max()), self.ni)
        self.LD = np.sqrt(self.eps * self.VT / (ph.Q * self.Ns))
        self.J0 = ph.Q * ph.D0_REF * self.Ns / self.LD
        self.R0 = ph.D0_REF * self.Ns / self.LD ** 2
        self.h = [np.diff(c / self.LD) for c in self.coords]      # per axis
        self.dVw = [ph.control_volume_widths(h) for h in self.h]
        if self.ndim == 2:
            self.dV = np.outer(self.dVw[1], self.dVw[0])
        else:
            self.dV = (self.dVw[2][:, None, None] * self.dVw[1][None, :, None]
                       * self.dVw[0][None, None, :])
        self.C = self.doping / self.Ns
        # device2d.py:357-395 (single material)
        nt = self.Ntot
        self.nie = ph.nie_effective(nt, material, T, self.models.bgn)
        self.nie_s = self.nie / self.Ns
        if self.models.doping_mobility:
            self.mu_n0 = ph.mobility_caughey_thomas(nt, material, T, "n")
            self.mu_p0 = ph.mobility_caughey_thomas(nt, material, T, "p")
        else:
            self.mu_n0 = np.full(self.shape, material.mu_n_max)
            self.mu_p0 = np.full(self.shape, material.mu_p_max)
        self.tau_n = ph.lifetime_scharfetter(nt, material.tau_n0, material.tau_Nref)
        self.tau_p = ph.lifetime_scharfetter(nt, material.tau_p0, material.tau_Nref)
        # device2d.py:442-445: harmonic-mean edge diffusivities
        self.dn_edge = [_hmean(*self._pair(self.mu_n0, a)) * self.VT / ph.D0_REF
                        for a in range(self.ndim)]
        self.dp_edge = [_hmean(*self._pair(self.mu_p0, a)) * self.VT / ph.D0_REF
                        for a in range(self.ndim)]
        self.bcs = {}
        self.psi = self.n = self.p = None
        self.last_converged = None
        self.last_iterations = None
        self.last_eq_iterations = None

    # -- geometry helpers -------------------------------------------------
    def _ax(self, a):
        """Array axis of coordinate axis a (0=x, 1=y, 2=z)."""
        return self.ndim - 1 - a

    def _pair(self, f, a):
        """(lo, hi) endpoint views of every edge along coordinate axis a."""
        ax = self._ax(a)
        lo = [slice(None)] * self.ndim
        hi = [slice(None)] * self.ndim
        lo[ax] = slice(None, -1)
        hi[ax] = slice(1, None)
        return f[tuple(lo)], f[tuple(hi)]

    def _h_b(self, a):
        """Spacing along axis a broadcast over the edge array."""
        shp = [1] * self.ndim
        shp[self._ax(a)] = -1
        return self.h[a].reshape(shp)

    def _dv_b(self, a):
        shp = [1] * self.ndim
        shp[self._ax(a)] = -1
        return self.dVw[a].reshape(shp)

    def _face(self, a):
        """Control-volume face measure crossed by an axis-a edge: the
        product of the other axes' widths (device2d.py:875,
        device3d.py:1600-1603)."""
        others = [b for b in range(self.ndim) if b != a]
        f = self._dv_b(others[0])
        for b in others[1:]:
            f = f * self._dv_b(b)
        return f

    def _div(self, flux, a):
        """div[lo] += flux; div[hi] -= flux (device2d.py:872)."""
        d = np.zeros(self.shape)
        lo, hi = self._pair(d, a)
        lo += flux
        hi -= flux
        return d

    def _edge_pairs(self, a):
        k = np.arange(self.N).reshape(self.shape)
        lo, hi = self._pair(k, a)
        return lo.ravel(), hi.ravel()

    def _flat(self, bc):
        return np.ravel_multi_index(bc.idx, self.shape)

    def _gate_face_weight(self, bc):
        """device3d.py:939-947; in 2D the gate face is always y-normal and
        the weight is dVx[i] (device2d.py:904)."""
        a = "xyz".index(bc.normal_axis)
        others = [b for b in range(self.ndim) if b != a]
        w = self.dVw[others[0]][bc.idx[self._ax(others[0])]]
        for b in others[1:]:
            w = w * self.dVw[b][bc.idx[self._ax(b)]]
        return w

    # -- boundary values --------------------------------------------------
    def _bc_contact_values(self, bc, V):
        """device2d.py:689-742 (ohmic, Boltzmann branch)."""
        return ph.ohmic_values(self.C[bc.idx], self.nie_s[bc.idx], V, self.VT)

    def _bulk_psi_guess(self):
        """device2d.py:744-759."""
        return np.arcsinh(self.C / (2.0 * self.nie_s))

    def _psi_b(self, bc):
        return np.arcsinh(self.C[bc.idx] / (2.0 * self.nie_s[bc.idx]))

    # -- equilibrium -------------------------------------------------------
    def _residual_jacobian_poisson(self, psi):
        """device2d.py:844-942 / device3d.py:190-258, 952-1020."""
        N = self.N
        n = self.nie_s * np.exp(np.clip(psi, -700, 700))
        p = self.nie_s * np.exp(np.clip(-psi, -700, 700))
        dnp = n + p
        F = -self.dV * (n - p - self.C)
        acc = None
        for a in range(self.ndim):
            lo, hi = self._pair(psi, a)
            Fa = 1.0 * (hi - lo) / self._h_b(a)
            term = self._face(a) * self._div(Fa, a)
            acc = term if acc is None else acc + term
        F = acc + F
        rows, cols, vals = [], [], []
        for a in range(self.ndim):
            kL, kR = self._edge_pairs(a)
            w = np.broadcast_to(self._face(a) / self._h_b(a),
                                self._pair(psi, a)[0].shape).ravel() * 1.0
            rows += [kL, kR, kL, kR]
            cols += [kL, kR, kR, kL]
            vals += [-w, -w, w, w]
        rows.append(np.arange(N)); cols.append(np.arange(N))
        vals.append((-self.dV * dnp).ravel())
        Ff = F.ravel()
        for bc in self.bcs.values():
            if isinstance(bc, GateBC):
                kk = self._flat(bc)
                w = self._gate_face_weight(bc)
                Vfb_s = bc.Vfb / self.VT
                Ff[kk] += bc.kappa * w * (0.0 - Vfb_s - (psi[bc.idx] - self._psi_b(bc)))
                rows.append(kk); cols.append(kk); vals.append(-bc.kappa * w)
        contact_k = [self._flat(bc) for bc in self.bcs.values()
                     if isinstance(bc, DirichletBC)]
        for bc in self.bcs.values():
            if isinstance(bc, DirichletBC):
                Ff[self._flat(bc)] = psi[bc.idx] - self._bc_contact_values(bc, 0.0)[0]
        rows = np.concatenate(rows); cols = np.concatenate(cols)
        vals = np.concatenate(vals)
        if contact_k:
            contact_k = np.unique(np.concatenate(contact_k))
            keep = ~np.isin(rows, contact_k)
            rows = np.concatenate([rows[keep], contact_k])
            cols = np.concatenate([cols[keep], contact_k])
            vals = np.concatenate([vals[keep], np.ones(contact_k.size)])
        else:
            contact_k = np.zeros(0, dtype=int)
        self._dirichlet_rows_poisson = contact_k
        return F, csr_matrix((vals, (rows, cols)), shape=(N, N))

    def solve_equilibrium(self, opts=None):
        """device2d.py:1308-1373 / device3d.py:1311-1454 (direct solve)."""
        opts = opts or ph.NewtonOptions()
        psi = self._bulk_psi_guess()
        self.last_eq_iterations = 0
        for it in range(opts.max_iter):
            F, J = self._residual_jacobian_poisson(psi)
            Jd, rhs = eliminate_csr(J, -F.ravel(), self._dirichlet_rows_poisson)
            d, _ = solve_linear(Jd.tocsc(), rhs, method="direct")
            self.last_eq_iterations += 1
            d = np.clip(d.reshape(self.shape), -opts.max_dpsi, opts.max_dpsi)
            psi = psi + d
            if np.abs(d).max() < opts.tol_update:
                break
        else:
            warnings.warn("equilibrium Poisson solve did not converge.")
        self.psi = psi
        self.n = self.nie_s * np.exp(np.clip(psi, -700, 700))
        self.p = self.nie_s * np.exp(np.clip(-psi, -700, 700))
        return self

    # -- coupled residual / Jacobian ----------------------------------------
    def _update_surface_mobility(self, psi):
        """device2d.py:612-642 (2D only): lagged Lombardi CVT on the
        row-0 y-edges."""
        E_eff = np.abs(psi[0, :] - psi[1, :]) * self.VT / (self.h[1][0] * self.LD)
        mu_n_s = ph.mobility_cvt(E_eff, self.mu_n0[0, :], "n", self.T)
        mu_p_s = ph.mobility_cvt(E_eff, self.mu_p0[0, :], "p", self.T)
        self.dn_edge[1][0, :] = (2.0 * mu_n_s * self.mu_n0[1, :]
                                 / (mu_n_s + self.mu_n0[1, :]) * self.VT / ph.D0_REF)
        self.dp_edge[1][0, :] = (2.0 * mu_p_s * self.mu_p0[1, :]
                                 / (mu_p_s + self.mu_p0[1, :]) * self.VT / ph.D0_REF)

    def _residual_jacobian(self, psi, n, p, voltages, current_rows_for=None):
        """device2d.py:1402-1920 / device3d.py:1485-1978.  Returns
        (F, J, Jn_edges, Jp_edges, F_n, F_p): F interleaved (3N) after
        contact/gate BCs, J CSR, per-axis SG edge currents (scaled) and
        the raw pre-Dirichlet continuity residuals."""
        N, nd = self.N, self.ndim
        ph_ = ph
        Jn, Jp, edges = [], [], []
        for a in range(nd):
            lo_psi, hi_psi = self._pair(psi, a)
            lo_nie, hi_nie = self._pair(self.nie_s, a)
            dlnnie = np.log(hi_nie / lo_nie)
            ds = 0.0                                    # band_shift == 0
            d = hi_psi - lo_psi + dlnnie + ds
            Bp, Bm = ph_.bernoulli(d), ph_.bernoulli(-d)
            dBp, dBm = ph_.dbernoulli(d), ph_.dbernoulli(-d)
            dp_ = hi_psi - lo_psi - dlnnie + ds
            Bph, Bmh = ph_.bernoulli(dp_), ph_.bernoulli(-dp_)
            dBph, dBmh = ph_.dbernoulli(dp_), ph_.dbernoulli(-dp_)
            an = self.dn_edge[a] / self._h_b(a)
            ap = self.dp_edge[a] / self._h_b(a)
            nL, nR = self._pair(n, a)
            pL, pR = self._pair(p, a)
            Jn.append(an * (nR * Bp - nL * Bm))
            Jp.append(-ap * (pR * Bmh - pL * Bph))
            edges.append(dict(Bp=Bp, Bm=Bm, dBp=dBp, dBm=dBm, Bph=Bph, Bmh=Bmh,
                              dBph=dBph, dBmh=dBmh, an=an, ap=ap,
                              nL=nL, nR=nR, pL=pL, pR=pR))
        if self.models.srh:
            R, dRdn, dRdp = ph_.recombination(n * self.Ns, p * self.Ns, self.nie,
                                              self.tau_n, self.tau_p, self.mat,
                                              auger=self.models.auger)
        else:
            R = dRdn = dRdp = np.zeros(self.shape)
        Rs = R / self.R0
        dRs_dn = dRdn * self.Ns / self.R0
        dRs_dp = dRdp * self.Ns / self.R0

        def divsum(fluxes):
            acc = None
            for a in range(nd):
                t = self._face(a) * self._div(fluxes[a], a)
                acc = t if acc is None else acc + t
            return acc

        F_psi = divsum([1.0 * (self._pair(psi, a)[1] - self._pair(psi, a)[0])
                        / self._h_b(a) for a in range(nd)]) - self.dV * (n - p - self.C)
        F_n = divsum(Jn) - Rs * self.dV
        F_p = divsum(Jp) + Rs * self.dV
        F = np.empty((N, 3))
        F[:, 0] = F_psi.ravel(); F[:, 1] = F_n.ravel(); F[:, 2] = F_p.ravel()
        F = F.ravel()

        rows, cols, vals = [], [], []

        def scatter(kL, kR, w, rc, cL, dL, cR, dR):
            w = np.broadcast_to(w, dL.shape).ravel(); dL = dL.ravel(); dR = dR.ravel()
            rows.extend([3 * kL + rc, 3 * kL + rc, 3 * kR + rc, 3 * kR + rc])
            cols.extend([3 * kL + cL, 3 * kR + cR, 3 * kL + cL, 3 * kR + cR])
            vals.extend([w * dL, w * dR, -w * dL, -w * dR])

        pairs = [self._edge_pairs(a) for a in range(nd)]
        for a in range(nd):           # Poisson rows, x then y then z
            e = edges[a]
            w = np.broadcast_to(self._face(a), e["Bp"].shape) * 1.0 / self._h_b(a)
            one = np.ones(e["Bp"].shape)
            scatter(*pairs[a], w, 0, 0, -one, 0, one)
        for a in range(nd):           # electron rows
            e = edges[a]
            Sn = e["nR"] * e["dBp"] + e["nL"] * e["dBm"]
            dpsiR = e["an"] * Sn
            wA = np.broadcast_to(self._face(a), e["Bp"].shape)
            scatter(*pairs[a], wA, 1, 0, -dpsiR, 0, dpsiR)
            scatter(*pairs[a], wA, 1, 1, -e["an"] * e["Bm"], 1, e["an"] * e["Bp"])
        for a in range(nd):           # hole rows
            e = edges[a]
            Sp = e["pR"] * e["dBmh"] + e["pL"] * e["dBph"]
            dpsiR = e["ap"] * Sp
            wA = np.broadcast_to(self._face(a), e["Bp"].shape)
            scatter(*pairs[a], wA, 2, 0, -dpsiR, 0, dpsiR)
            scatter(*pairs[a], wA, 2, 2, e["ap"] * e["Bph"], 2, -e["ap"] * e["Bmh"])
        k = np.arange(N)
        dVf = self.dV.ravel()
        rows += [3 * k, 3 * k, 3 * k + 1, 3 * k + 1, 3 * k + 2, 3 * k + 2]
        cols += [3 * k + 1, 3 * k + 2, 3 * k + 1, 3 * k + 2, 3 * k + 2, 3 * k + 1]
        vals += [-dVf, dVf, -dRs_dn.ravel() * dVf, -dRs_dp.ravel() * dVf,
                 dRs_dp.ravel() * dVf, dRs_dn.ravel() * dVf]
        F3 = F.reshape(N, 3)
        for bc in self.bcs.values():
            if isinstance(bc, GateBC):
                kk = self._flat(bc)
                w = self._gate_face_weight(bc)
                Vg_s, Vfb_s = bc.Vg / self.VT, bc.Vfb / self.VT
                F3[kk, 0] += bc.kappa * w * (Vg_s - Vfb_s - (psi[bc.idx] - self._psi_b(bc)))
                rows.append(3 * kk); cols.append(3 * kk); vals.append(-bc.kappa * w)
        rows = np.concatenate(rows); cols = np.concatenate(cols)
        vals = np.concatenate(vals)
        if current_rows_for is not None:
            # device2d.py:1751-1758: the raw continuity rows of the requested
            # nodes (the exact derivative of F_n, F_p there), before the
            # contact rows replace them
            k_req = np.asarray(current_rows_for, dtype=int)
            sel = np.isin(rows, np.concatenate([3 * k_req + 1, 3 * k_req + 2]))
            self._raw_current_rows = (rows[sel].copy(), cols[sel].copy(), vals[sel].copy())
        strip = []
        for name, bc in self.bcs.items():
            if isinstance(bc, DirichletBC):
                V = voltages.get(name, bc.V)
                kk = self._flat(bc)
                psi0, n0, p0 = self._bc_contact_values(bc, V)
                F3[kk, 0] = psi[bc.idx] - psi0
                F3[kk, 1] = n[bc.idx] - n0
                F3[kk, 2] = p[bc.idx] - p0
                strip += [3 * kk, 3 * kk + 1, 3 * kk + 2]
        if strip:
            strip = np.unique(np.concatenate(strip))
            if nd == 3:
                # device3d.py:1910-1922: the eliminated rows are listed by
                # component block (every contact node's psi row, then n, then p)
                ck = strip[0::3] // 3
                strip = np.concatenate([3 * ck, 3 * ck + 1, 3 * ck + 2])
            keep = ~np.isin(rows, strip)
            rows = np.concatenate([rows[keep], strip])
            cols = np.concatenate([cols[keep], strip])
            vals = np.concatenate([vals[keep], np.ones(strip.size)])
        else:
            strip = np.zeros(0, dtype=int)
        self._dirichlet_rows = strip
        J = csr_matrix((vals, (rows, cols)), shape=(3 * N, 3 * N))
        return F3.ravel(), J, Jn, Jp, F_n, F_p

    # -- bias Newton ---------------------------------------------------------
    def _set_voltages(self, voltages):
        for name, V in (voltages or {}).items():
            bc = self.bcs[name]
            if isinstance(bc, DirichletBC):
                bc.V = V
            else:
                bc.Vg = V

    def solve_bias(self, voltages=None, opts=None, deadline=None):
        """device2d.py:1923-2191 / device3d.py:1981-2271 (plain path: one
        stage, no line search, density floor 1e-10).  `deadline` (a
        time.perf_counter() value) stops the Newton loop after the first
        iteration that ends past it, unconverged, with last_timed_out set
        (a cap for timing runs; the reference has none)."""
        opts = opts or ph.NewtonOptions()
        self.last_timed_out = False
        if self.psi is None:
            self.solve_equilibrium(opts)
        self._set_voltages(voltages)
        psi, n, p = self.psi.copy(), self.n.copy(), self.p.copy()
        for bc in self.bcs.values():
            if isinstance(bc, DirichletBC):
                psi0, n0, p0 = self._bc_contact_values(bc, bc.V)
                psi[bc.idx], n[bc.idx], p[bc.idx] = psi0, n0, p0
        cur = {nm: b.V for nm, b in self.bcs.items() if isinstance(b, DirichletBC)}
        converged = False
        self.last_iterations = 0
        self.last_krylov_iterations = 0
        for it in range(opts.max_iter):
            if self.models.surface_mobility and self.ndim == 2:
                self._update_surface_mobility(psi)
            F, J, *_ = self._residual_jacobian(psi, n, p, cur)
            Jd, rhs = eliminate_csr(J, -F, self._dirichlet_rows)
            self._cur_np = (n, p)
            du = self._linear_solve(Jd, rhs, opts)
            self.last_iterations += 1
            dpsi = du[0::3].reshape(self.shape)
            dn = du[1::3].reshape(self.shape)
            dp = du[2::3].reshape(self.shape)
            dpsi = np.clip(dpsi, -opts.max_dpsi, opts.max_dpsi)
            n_new = np.clip(n + dn, 0.1 * n, 10.0 * n)
            p_new = np.clip(p + dp, 0.1 * p, 10.0 * p)
            rel_n = (np.abs(n_new - n) / np.maximum(n, 1e-10)).max()
            rel_p = (np.abs(p_new - p) / np.maximum(p, 1e-10)).max()
            err = max(np.abs(dpsi).max(), rel_n, rel_p)
            psi = psi + dpsi
            n, p = n_new, p_new
            if err < opts.tol_update:
                converged = True
                break
            if deadline is not None and time.perf_counter() > deadline:
                self.last_timed_out = True
                break
        else:
            warnings.warn(f"Newton did not converge; last update {err:.2e}")
        self.last_converged = converged
        self.psi, self.n, self.p = psi, n, p
        _, _, Jn, Jp, _, _ = self._residual_jacobian(psi, n, p, cur)
        self.Jn = [j * self.J0 for j in Jn]
        self.Jp = [j * self.J0 for j in Jp]
        return self

    def terminal_current(self, name):
        """device2d.py:2194-2227 (A/cm) / device3d.py:2274-2306 (A)."""
        bc = self.bcs[name]
        if not isinstance(bc, DirichletBC):
            raise ValueError(f"terminal_current: '{name}' is not a Dirichlet contact")
        if self.psi is None:
            raise RuntimeError("terminal_current: solve the device first")
        cur = {nm: b.V for nm, b in self.bcs.items() if isinstance(b, DirichletBC)}
        *_, F_n, F_p = self._residual_jacobian(self.psi, self.n, self.p, cur)
        return float((F_n[bc.idx] + F_p[bc.idx]).sum()) * self.J0 * self.LD ** (self.ndim - 1)


def equilibrated_direct(Jd, rhs, n, p):
    """Direct solve of the same Newton system after column scaling by
    (1, n, p) per node and row equilibration (inverse row max).  Not a
    reference method: the reference's SuperLU on the unscaled Jacobian
    (linsolve.py:860) loses digits when minority densities span tens of
    decades (deep reverse bias), and this variant is the exact-arithmetic
    Newton step the GPU's equilibrated Krylov solve converges to."""
    cs = np.stack([np.ones(n.size), n.ravel(), p.ravel()], 1).ravel()
    A = Jd.tocsr() @ diags_(cs)
    rs = 1.0 / abs(A).max(axis=1).toarray().ravel()
    y = solve_linear((diags_(rs) @ A).tocsc(), rs * rhs, method="direct")[0]
    return cs * y


def diags_(v):
    from scipy.sparse import diags
    return diags(v, format="csr")


class Device2D(_DeviceND):
    """Oracle twin of device2d.Device2D (structured, Boltzmann)."""

    def __init__(self, x, y, doping, Ntotal=None, T=300.0,
                 material=ph.SILICON, models=None):
        super().__init__((x, y), doping, Ntotal, T, material, models)
        self.Ny, self.Nx = self.shape

    def add_contact(self, name, i, j, V=0.0):
        """device2d.py:672-674."""
        i = np.atleast_1d(np.asarray(i, dtype=int)); j = np.atleast_1d(np.asarray(j, dtype=int))
        i, j = np.broadcast_arrays(i, j)
        self.bcs[name] = DirichletBC((j, i), V)
        return self.bcs[name]

    def add_gate(self, name, i, j, tox_cm, Vfb, Vg=0.0):
        """device2d.py:676-680."""
        i = np.atleast_1d(np.asarray(i, dtype=int)); j = np.atleast_1d(np.asarray(j, dtype=int))
        i, j = np.broadcast_arrays(i, j)
        kappa = ph.EPS_OX_R * ph.EPS0 * self.LD / (self.eps * tox_cm)
        self.bcs[name] = GateBC((j, i), kappa, Vfb, Vg, "y")
        return self.bcs[name]

    def _linear_solve(self, Jd, rhs, opts):
        """device2d.py:2057-2071: direct is spsolve(Jd.tocsc(), rhs)."""
        if opts.linsolve == "direct":
            return solve_linear(Jd.tocsc(), rhs, method="direct")[0]
        if opts.linsolve == "direct_equilibrated":
            return equilibrated_direct(Jd, rhs, *self._cur_np)
        x, info = solve_linear(Jd, rhs, method=opts.linsolve, rtol=opts.linsolve_rtol,
                               block_size=opts.block_size)
        self.last_krylov_iterations += info["iterations"]
        return x


class Device3D(_DeviceND):
    """Oracle twin of device3d.Device3D (structured, Boltzmann)."""

    def __init__(self, x, y, z, doping, Ntotal=None, T=300.0,
                 material=ph.SILICON, models=None):
        models = models or ph.Models()
        if models.surface_mobility:
            raise NotImplementedError("surface_mobility is 2D only (device.py:431-434)")
        super().__init__((x, y, z), doping, Ntotal, T, material, models)
        self.Nz, self.Ny, self.Nx = self.shape

    def add_contact(self, name, i, j, k, V=0.0):
        """device3d.py:748-750."""
        i, j, k = np.broadcast_arrays(*(np.atleast_1d(np.asarray(a, dtype=int))
                                        for a in (i, j, k)))
        self.bcs[name] = DirichletBC((k, j, i), V)
        return self.bcs[name]

    def add_gate(self, name, i, j, k, tox_cm, Vfb, Vg=0.0, normal_axis="z"):
        """device3d.py:752-756."""
        i, j, k = np.broadcast_arrays(*(np.atleast_1d(np.asarray(a, dtype=int))
                                        for a in (i, j, k)))
        kappa = ph.EPS_OX_R * ph.EPS0 * self.LD / (self.eps * tox_cm)
        self.bcs[name] = GateBC((k, j, i), kappa, Vfb, Vg, normal_axis)
        return self.bcs[name]

    def _linear_solve(self, Jd, rhs, opts):
        """device3d.py:2137-2168: direct on the CSR matrix; iterative with a
        per-iteration fallback to direct on LinearSolveError."""
        if opts.linsolve == "direct":
            return solve_linear(Jd, rhs, method="direct")[0]
        if opts.linsolve == "direct_equilibrated":
            return equilibrated_direct(Jd, rhs, *self._cur_np)
        try:
            x, info = solve_linear(Jd, rhs, method=opts.linsolve,
                                   rtol=opts.linsolve_rtol, block_size=opts.block_size)
            self.last_krylov_iterations += info["iterations"]
            return x
        except LinearSolveError:
            if not opts.direct_fallback:
                raise
            return solve_linear(Jd, rhs, method="direct")[0]
