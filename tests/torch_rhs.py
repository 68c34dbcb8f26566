"""Test-only PyTorch FP64 restatement of the first-order RHS (AUSM or Rusanov, species, inflow / outflow ghosts),
written from the reference formulas (ausm.cpp:2-43, navierStokesTransport.cpp:99-207,
speciesTransport.cpp:263-341, perfectGas.cpp:67-174, cellInterpolant.cpp:604-629) with torch ops so that
torch.autograd provides an independent reference gradient for the hand-written adjoint."""
from __future__ import annotations

import numpy as np
import torch


class TorchRHS:
    def __init__(self, mesh, gamma=1.4, Rgas=287.0, n_species=0, flux="ausm", bcs=None, device="cpu"):
        a = mesh.arrays
        self.dim, self.N, self.NI, self.ns = a.dim, a.n_cells, a.n_interior, n_species
        self.nc = self.dim + 2 + n_species
        t = lambda x, dt=torch.float64: torch.tensor(np.asarray(x), dtype=dt, device=device)
        self.L = t(a.face_cells[:, 0], torch.long)
        self.R = t(a.face_cells[:, 1], torch.long)
        n = t(a.face_normal[:, : self.dim])
        self.area = torch.sqrt((n * n).sum(1))
        self.nhat = n / self.area[:, None]
        self.vol = t(a.cell_volume)
        self.gamma, self.Rg, self.cv = gamma, Rgas, Rgas / (gamma - 1.0)
        self.flux = flux
        self.bcs = bcs or {}
        self.bface = np.where(a.face_cells[:, 1] >= a.n_interior)[0]
        self.bslot = a.face_bc[self.bface] if self.bface.size else np.zeros(0, dtype=np.int32)
        self.device = device

    def ghost_state(self, U, thetas):
        """Fills ghost rows: inflow from theta = (rho, |u|, T) and the slot's direction / Yi (with "at_face": twice
        that state minus the interior); outflow copies; "ghost" (caller-filled) rows are inputs and stay as given."""
        U = U.clone()
        for slot in np.unique(self.bslot):
            fsel = self.bface[self.bslot == slot]
            g = self.R[fsel]
            i = self.L[fsel]
            b = self.bcs[int(slot)]
            if b["kind"] == "ghost":
                continue
            if b["kind"] == "outflow":
                U[:, g] = U[:, i]
            else:
                th = thetas[int(slot)]
                rho, um, T = th[0], th[1], th[2]
                d = torch.tensor(b["dir"][: self.dim], dtype=torch.float64, device=self.device)
                vals = [rho, rho * (self.cv * T + 0.5 * um * um)] + [rho * um * d[k] for k in range(self.dim)]
                Y = b.get("Yi") or [0.0] * self.ns
                vals += [rho * Y[k] for k in range(self.ns)]
                col = torch.stack(vals)
                U[:, g] = col[:, None].expand(-1, g.numel())
                if b.get("at_face"):
                    U[:, g] = 2.0 * U[:, g] - U[:, i]
        return U

    def decode(self, u):
        dim = self.dim
        rho = u[0]
        vel = u[2 : 2 + dim] / rho
        un = (vel * self.nhat.T).sum(0)
        eint = u[1] / rho - 0.5 * (vel * vel).sum(0)
        T = eint / self.cv
        p = self.Rg * rho * T
        a = torch.sqrt(self.gamma * p / rho)
        H = T * self.cv + 0.5 * (vel * vel).sum(0) + p / rho
        return rho, vel, un, p, a, H

    def __call__(self, U, thetas=None):
        dim, ns = self.dim, self.ns
        U = self.ghost_state(U, thetas or {})
        uL, uR = U[:, self.L], U[:, self.R]
        rL, vL, unL, pL, aL, HL = self.decode(uL)
        rR, vR, unR, pR, aR, HR = self.decode(uR)
        A = self.area
        if self.flux == "ausm":
            Mm, Mp = unR / aR, unL / aL
            sub_m = Mm.abs() <= 1
            sMm = torch.where(sub_m, -0.25 * (Mm - 1) ** 2, 0.5 * (Mm - Mm.abs()))
            # the untaken branch must stay finite: its gradient is multiplied by zero, not dropped
            sPm = torch.where(sub_m, -sMm * (2 + Mm), sMm / torch.where(sub_m, torch.ones_like(Mm), Mm))
            sub_p = Mp.abs() <= 1
            sMp = torch.where(sub_p, 0.25 * (Mp + 1) ** 2, 0.5 * (Mp + Mp.abs()))
            sPp = torch.where(sub_p, sMp * (2 - Mp), sMp / torch.where(sub_p, torch.ones_like(Mp), Mp))
            m = sMm + sMp
            right = m < 0
            mdot = torch.where(right, m * aR * rR, m * aL * rL)
            p12 = pR * sPm + pL * sPp
            H = torch.where(right, HR, HL)
            vS = torch.where(right[None], vR, vL)
            F = [mdot * A, H * mdot * A] + [vS[d] * mdot * A + p12 * self.nhat[:, d] * A for d in range(dim)]
            rS = torch.where(right, rR, rL)
            for k in range(ns):
                yS = torch.where(right, uR[2 + dim + k], uL[2 + dim + k])
                F.append(mdot * yS / rS * A)
        else:  # rusanov
            lam = torch.maximum(unL.abs() + aL, unR.abs() + aR)
            fL = [uL[0] * unL, (uL[1] + pL) * unL] + [uL[2 + d] * unL + pL * self.nhat[:, d] for d in range(dim)]
            fR = [uR[0] * unR, (uR[1] + pR) * unR] + [uR[2 + d] * unR + pR * self.nhat[:, d] for d in range(dim)]
            fL += [uL[2 + dim + k] * unL for k in range(ns)]
            fR += [uR[2 + dim + k] * unR for k in range(ns)]
            F = [A * (0.5 * (fL[c] + fR[c]) - 0.5 * lam * (uR[c] - uL[c])) for c in range(self.nc)]
        F = torch.stack(F)
        interior_R = self.R < self.NI
        out = torch.zeros((self.nc, self.N), dtype=torch.float64, device=self.device)
        out = out.index_add(1, self.L, -F / self.vol[self.L])
        Rsafe = torch.where(interior_R, self.R, self.L)
        out = out.index_add(1, Rsafe, torch.where(interior_R, F / self.vol[self.R], torch.zeros_like(F)))
        return out


class TorchRHS2:
    """Second-order restatement with torch ops: least-squares cell gradients (boundary faces excluded), the
    Barth-Jespersen limiter (or none), reconstructed face states, AUSM or Rusanov advection, the Navier-Stokes
    viscous flux from face-averaged cell stress / heat-flux terms (aux gradients over all faces, ghost cells with
    their one-neighbour stencil), species advection and diffusion.  Written from the reference formulas
    (cellInterpolant.cpp:330-497,890-950, navierStokesTransport.cpp:99-207,454-532,
    stencils/leastSquaresAverage.cpp:5-155, speciesTransport.cpp:196-341) for autograd checks of the adjoint.
    Weights come from the oracle's least-squares pseudo-inverse."""

    def __init__(self, mesh, gamma=1.4, Rgas=287.0, n_species=0, flux="ausm", bcs=None, limiter="barthjespersen",
                 mu=0.0, k=0.0, diff=0.0, device="cpu", recon=True):
        from oracle import binding as orc

        a = mesh.arrays
        self.base = TorchRHS(mesh, gamma, Rgas, n_species, flux, bcs, device)
        self.dim, self.N, self.NI, self.ns = a.dim, a.n_cells, a.n_interior, n_species
        self.nc = self.dim + 2 + n_species
        self.limiter, self.mu, self.k, self.diff, self.recon = limiter, mu, k, diff, recon
        self.visc = bool(mu or k or diff)
        dim, NI, F = self.dim, self.NI, a.n_faces
        fc = a.face_cells
        shift = a.face_shift if a.face_shift is not None else np.zeros((F, 3))
        xc, xf = a.cell_centroid[:, :dim], a.face_centroid[:, :dim]
        sh = shift[:, :dim]
        slots = [[] for _ in range(NI)]
        for f in range(F):
            slots[fc[f, 0]].append((f, 0))
            if fc[f, 1] < NI:
                slots[fc[f, 1]].append((f, 1))
        K = max(len(s) for s in slots)
        nbr = np.zeros((NI, K), dtype=np.int64)
        mask = np.zeros((NI, K))
        dxf = np.zeros((NI, K, dim))
        w = np.zeros((NI, K, dim))
        wa = np.zeros((NI, K, dim))
        has_bnd = bool(np.any(fc[:, 1] >= NI))
        for c in range(NI):
            dxn = []
            for kk, (f, s) in enumerate(sorted(slots[c])):
                L, R = fc[f]
                n = R if s == 0 else L
                nbr[c, kk] = n
                mask[c, kk] = 1.0
                dxf[c, kk] = xf[f] - xc[L] if s == 0 else xf[f] - (xc[R] + sh[f])
                dxn.append((xc[R] + sh[f] - xc[L]) if s == 0 else (xc[L] - (xc[R] + sh[f])))
            dxn = np.asarray(dxn)
            inner = np.array([nbr[c, kk] < NI for kk in range(len(dxn))])
            if inner.any():
                w[c, np.where(inner)[0]] = orc.lsq_weights(dxn[inner])
            wa[c, : len(dxn)] = orc.lsq_weights(dxn) if (has_bnd and self.visc) else w[c, : len(dxn)]
            nbr[c, len(dxn):] = c
        bf = np.where(fc[:, 1] >= NI)[0]
        wg = np.zeros((len(bf), dim))
        for i, f in enumerate(bf):
            dx = (xc[fc[f, 0]] - (xc[fc[f, 1]] + sh[f]))[None, :]
            wg[i] = orc.lsq_weights(dx)[0]
        t = lambda x, dt=torch.float64: torch.tensor(np.asarray(x), dtype=dt, device=device)
        self.nbr, self.mask, self.dxf, self.w, self.wa = t(nbr, torch.long), t(mask), t(dxf), t(w), t(wa)
        self.bf, self.wg = t(bf, torch.long), t(wg)
        self.dxL = t(xf - xc[fc[:, 0]])
        self.dxR = t(xf - (xc[fc[:, 1]] + sh))

    def _limited(self, Uc, Un, G):
        """Barth-Jespersen (face-centroid projections) per component; Uc [P, NI], Un [P, NI, K], G [P, NI, dim]."""
        if self.limiter == "none":
            return G
        m = self.mask[None]
        big = torch.finfo(torch.float64).max
        umax = torch.maximum(Uc, torch.where(m > 0, Un, torch.full_like(Un, -big)).amax(dim=2))
        umin = torch.minimum(Uc, torch.where(m > 0, Un, torch.full_like(Un, big)).amin(dim=2))
        D = torch.einsum("pcd,ckd->pck", G, self.dxf) * m
        dpos = torch.clamp(D.amax(dim=2), min=0.0)
        dneg = torch.clamp(D.amin(dim=2), max=0.0)
        one = torch.ones_like(Uc)
        phi = one
        rp = (umax - Uc) / torch.where(dpos > 0, dpos, one)
        rn = (umin - Uc) / torch.where(dneg < 0, dneg, one)
        phi = torch.where(dpos > 0, torch.minimum(phi, rp), phi)
        phi = torch.where(dneg < 0, torch.minimum(phi, rn), phi)
        return G * phi[..., None]

    def _stress_heat(self, GA):
        """GA [dim + 1, M, dim] (T then velocity gradients) -> tau (dim, dim, M) and q (dim, M)."""
        dim = self.dim
        div = sum(GA[1 + c, :, c] for c in range(dim))
        tau = [[None] * dim for _ in range(dim)]
        for c in range(dim):
            for d in range(dim):
                if c == d:
                    tau[c][d] = 2.0 * self.mu * (GA[1 + c, :, d] - div / 3.0)
                else:
                    tau[c][d] = self.mu * (GA[1 + c, :, d] + GA[1 + d, :, c])
        q = [self.k * GA[0, :, d] for d in range(dim)]
        return tau, q

    def __call__(self, U, thetas=None):
        b = self.base
        dim, NI, ns, N = self.dim, self.NI, self.ns, self.N
        Ug = b.ghost_state(U, thetas or {})
        Uc = Ug[:, :NI]
        Un = Ug[:, self.nbr]
        delta = (Un - Uc[:, :, None]) * self.mask[None]
        G = torch.einsum("pck,ckd->pcd", delta, self.w)
        Glim = self._limited(Uc, Un, G) if self.recon else torch.zeros_like(G)
        GN = torch.cat([Glim, torch.zeros((self.nc, N - NI, dim), dtype=U.dtype, device=U.device)], dim=1)
        uL = Ug[:, b.L] + (GN[:, b.L] * self.dxL[None]).sum(-1)
        uR = Ug[:, b.R] + (GN[:, b.R] * self.dxR[None]).sum(-1)
        F = self._advection(uL, uR)
        if self.visc:
            F = F + self._viscous(Ug)
        interior_R = b.R < NI
        out = torch.zeros((self.nc, N), dtype=torch.float64, device=U.device)
        out = out.index_add(1, b.L, -F / b.vol[b.L])
        Rsafe = torch.where(interior_R, b.R, b.L)
        out = out.index_add(1, Rsafe, torch.where(interior_R, F / b.vol[b.R], torch.zeros_like(F)))
        return out

    def _advection(self, uL, uR):
        b = self
        tb = self.base
        dim, ns = self.dim, self.ns
        rL, vL, unL, pL, aL, HL = tb.decode(uL)
        rR, vR, unR, pR, aR, HR = tb.decode(uR)
        A = tb.area
        if tb.flux == "ausm":
            Mm, Mp = unR / aR, unL / aL
            sub_m = Mm.abs() <= 1
            sMm = torch.where(sub_m, -0.25 * (Mm - 1) ** 2, 0.5 * (Mm - Mm.abs()))
            sPm = torch.where(sub_m, -sMm * (2 + Mm), sMm / torch.where(sub_m, torch.ones_like(Mm), Mm))
            sub_p = Mp.abs() <= 1
            sMp = torch.where(sub_p, 0.25 * (Mp + 1) ** 2, 0.5 * (Mp + Mp.abs()))
            sPp = torch.where(sub_p, sMp * (2 - Mp), sMp / torch.where(sub_p, torch.ones_like(Mp), Mp))
            m = sMm + sMp
            right = m < 0
            mdot = torch.where(right, m * aR * rR, m * aL * rL)
            p12 = pR * sPm + pL * sPp
            H = torch.where(right, HR, HL)
            vS = torch.where(right[None], vR, vL)
            Fl = [mdot * A, H * mdot * A] + [vS[d] * mdot * A + p12 * tb.nhat[:, d] * A for d in range(dim)]
            rS = torch.where(right, rR, rL)
            for k in range(ns):
                yS = torch.where(right, uR[2 + dim + k], uL[2 + dim + k])
                Fl.append(mdot * yS / rS * A)
        else:
            lam = torch.maximum(unL.abs() + aL, unR.abs() + aR)
            fL = [uL[0] * unL, (uL[1] + pL) * unL] + [uL[2 + d] * unL + pL * tb.nhat[:, d] for d in range(dim)]
            fR = [uR[0] * unR, (uR[1] + pR) * unR] + [uR[2 + d] * unR + pR * tb.nhat[:, d] for d in range(dim)]
            fL += [uL[2 + dim + k] * unL for k in range(ns)]
            fR += [uR[2 + dim + k] * unR for k in range(ns)]
            Fl = [A * (0.5 * (fL[c] + fR[c]) - 0.5 * lam * (uR[c] - uL[c])) for c in range(self.nc)]
        return torch.stack(Fl)

    def _viscous(self, Ug):
        tb = self.base
        dim, NI, ns = self.dim, self.NI, self.ns
        rho = Ug[0]
        vel = Ug[2 : 2 + dim] / rho
        T = (Ug[1] / rho - 0.5 * (vel * vel).sum(0)) / tb.cv
        aux = torch.cat([T[None], vel], dim=0)  # [dim + 1, N]
        an = aux[:, self.nbr]
        da = (an - aux[:, :NI, None]) * self.mask[None]
        GA = torch.einsum("qck,ckd->qcd", da, self.wa)  # [dim + 1, NI, dim]
        tau_c, q_c = self._stress_heat(GA)
        # ghost cells: one-neighbour stencil w_g (a_interior - a_ghost), evaluated per boundary face
        L, R = tb.L, tb.R
        nf = L.shape[0]
        tau_f = [[None] * dim for _ in range(dim)]
        q_f = [None] * dim
        is_b = R >= NI
        Lc = L
        Rc = torch.where(is_b, L, R)  # interior stand-in for the ghost row lookups
        if self.bf.numel():
            dag = aux[:, L[self.bf]] - aux[:, R[self.bf]]
            GAg = dag[:, :, None] * self.wg[None]
            tau_g, q_g = self._stress_heat(GAg)
        for c in range(dim):
            for d in range(dim):
                tR = tau_c[c][d][Rc]
                if self.bf.numel():
                    tR = tR.index_put((self.bf,), tau_g[c][d])
                tau_f[c][d] = 0.5 * tau_c[c][d][Lc] + 0.5 * tR
        for d in range(dim):
            qR = q_c[d][Rc]
            if self.bf.numel():
                qR = qR.index_put((self.bf,), q_g[d])
            q_f[d] = 0.5 * q_c[d][Lc] + 0.5 * qR
        vf = 0.5 * vel[:, L] + 0.5 * vel[:, R]
        normal = tb.nhat * tb.area[:, None]
        zeros = torch.zeros(nf, dtype=torch.float64, device=Ug.device)
        Fv = [zeros, zeros.clone()] + [zeros.clone() for _ in range(dim)]
        for c in range(dim):
            Fv[2 + c] = -sum(normal[:, d] * tau_f[c][d] for d in range(dim))
        Fv[1] = -sum(normal[:, d] * (sum(vf[c] * tau_f[d][c] for c in range(dim)) + q_f[d]) for d in range(dim))
        for k in range(ns):
            Y = Ug[2 + dim + k] / rho
            dY = (Y[self.nbr] - Y[:NI, None]) * self.mask
            GY = torch.einsum("ck,ckd->cd", dY, self.wa)  # [NI, dim]
            GYR = GY[Rc]
            if self.bf.numel():
                GYR = GYR.index_put((self.bf,), self.wg * (Y[L[self.bf]] - Y[R[self.bf]])[:, None])
            dot = (tb.nhat * (GY[Lc] + GYR)).sum(1)
            cd = -self.diff * (0.5 * rho[L] + 0.5 * rho[R]) * 0.5 * tb.area
            Fv.append(cd * dot)
        return torch.stack(Fv)
