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
        """Fills ghost rows: inflow from theta = (rho, |u|, T) and the slot's direction / Yi; outflow copies."""
        U = U.clone()
        for slot in np.unique(self.bslot):
            fsel = self.bface[self.bslot == slot]
            g = self.R[fsel]
            i = self.L[fsel]
            b = self.bcs[int(slot)]
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
