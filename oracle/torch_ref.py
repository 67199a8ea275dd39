"""Oracle (test infrastructure only): the reference's CPU step in PyTorch
float32, timed by bench.py as the CPU baseline.

Restates one BaseModel.forward step (core/md_model.py:103-153) as the
reference runs it on the CPU, on a half Verlet list held fixed between
rebuilds:
  core/integrator/integrator.py:91-99    first half: a = F/m, v_half = v + (0.5 a) dt,
                                         x += v_half dt
  io_utils/reader.py:136-163, 364-376    skin trigger (max |mi(x - x_last)| > skin/2)
                                         and the edge distances |mi(x_i - x_j)|
  core/box.py:109-130                    minimum image
  core/force/eam_force_cu.py:461-579     EAM with use_extension=False: rho by
                                         scatter_add_, F(rho), F'(rho), pair energy,
                                         force scalar, forces by scatter_add_, virial
                                         and virial tensor
  core/integrator/integrator.py:102-124  second half: Langevin O with torch.randn
                                         kicks, then B; kinetic energy

Tables are the reference's own float32 tables (oracle.eam.EAMTables with
dtype float32, bit for bit equal to the reference's, tests/test_golden.py).
torch runs the elementwise and scatter operations on all the threads it is
given (torch.set_num_threads).

pinned by: equality with the numpy oracle's float32 EAM (oracle.eam.eam_forward)
and integrator halves (oracle.integrate) on the same inputs
(tests/test_oracle_torch_ref.py).
"""
import numpy as np
import torch

from . import box as obox
from . import integrate as oint


class TorchBox:
    """core/box.py:109-130 in float32 torch."""

    def __init__(self, H):
        H = np.asarray(H, np.float64)
        self.ortho = obox.is_orthogonal(H)
        self.H = torch.tensor(H, dtype=torch.float32)
        self.Hi = torch.tensor(obox.inverse(H), dtype=torch.float32)
        self.L = torch.tensor(np.diag(H), dtype=torch.float32)

    def minimum_image(self, r):
        if self.ortho:
            return r - self.L * torch.round(r / self.L)
        s = r @ self.Hi
        s = s - torch.round(s)
        return s @ self.H


class TorchEAM:
    """core/force/eam_force_cu.py:437-579 (use_extension=False), one element
    type or several (per type-pair tables)."""

    def __init__(self, tab):
        if tab.dtype != np.float32:
            raise ValueError("the reference path uses float32 tables")
        self.ne = tab.ne
        self.cutoff = float(tab.cutoff)
        self.r_max = float(tab.r_max)
        self.inv_dr = float(tab.inv_dr)
        self.n_r = int(tab.n_r)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a))
        self.dens, self.ddens = t(tab.dens), t(tab.ddens)
        self.phi, self.dphi = t(tab.phi), t(tab.dphi)
        self.F, self.dF = t(tab.F), t(tab.dF)
        self.rho_min, self.drho = list(tab.rho_min), list(tab.drho)
        self.inv_drho, self.n_rho = list(tab.inv_drho), list(tab.n_rho)

    def _interp_r(self, r, table):
        rc = torch.clamp(r, 0.0, self.r_max * (1 - 1e-7))
        idx_f = rc * self.inv_dr
        idx = idx_f.long()
        frac = idx_f - idx
        nxt = torch.clamp(idx + 1, max=self.n_r - 1)
        v0 = table[..., idx]
        return v0 + frac * (table[..., nxt] - v0)

    def _interp_rho(self, rho, a, deriv):
        lo, d, inv, n = self.rho_min[a], self.drho[a], self.inv_drho[a], self.n_rho[a]
        rc = torch.clamp(rho, lo, lo + d * (n - 1) * (1 - 1e-7))
        idx_f = (rc - lo) * inv
        idx = idx_f.long()
        frac = idx_f - idx
        nxt = torch.clamp(idx + 1, max=n - 1)
        t = self.dF[a] if deriv else self.F[a]
        return t[idx] + frac * (t[nxt] - t[idx])

    def forward(self, x, row, col, types, box):
        """x float32 [N,3]; row, col int64 half list; types int64 [N]."""
        rij_all = box.minimum_image(x[row] - x[col])
        dist_all = torch.norm(rij_all, dim=1)
        m = dist_all < self.cutoff
        r = dist_all[m]
        row_m, col_m = row[m], col[m]
        rij = box.minimum_image(x[row_m] - x[col_m])
        n = x.shape[0]
        rho = torch.zeros(n, dtype=torch.float32)
        tr, tc = types[row_m], types[col_m]
        for a in range(self.ne):
            for b in range(self.ne):
                sel = (tr == a) & (tc == b)
                if sel.any():
                    rho.scatter_add_(0, row_m[sel], self._interp_r(r[sel], self.dens[a, b]))
                sel = (tc == a) & (tr == b)
                if sel.any():
                    rho.scatter_add_(0, col_m[sel], self._interp_r(r[sel], self.dens[a, b]))
        emb = torch.zeros_like(rho)
        dF = torch.zeros_like(rho)
        for a in range(self.ne):
            sel = types == a
            if sel.any():
                emb[sel] = self._interp_rho(rho[sel], a, False)
                dF[sel] = self._interp_rho(rho[sel], a, True)
        pair_e = torch.zeros_like(r)
        dd_row = torch.zeros_like(r)
        dd_col = torch.zeros_like(r)
        dphi = torch.zeros_like(r)
        for a in range(self.ne):
            for b in range(self.ne):
                sel = (tr == a) & (tc == b)
                if sel.any():
                    rs = r[sel]
                    pair_e[sel] = self._interp_r(rs, self.phi[a, b])
                    dd_row[sel] = self._interp_r(rs, self.ddens[b, a])
                    dd_col[sel] = self._interp_r(rs, self.ddens[a, b])
                    dphi[sel] = self._interp_r(rs, self.dphi[a, b])
        s = dF[row_m] * dd_col + dF[col_m] * dd_row + dphi
        fv = -s.unsqueeze(1) * (rij / (r.unsqueeze(1) + 1e-8))
        forces = torch.zeros_like(x)
        forces.scatter_add_(0, row_m.unsqueeze(1).expand_as(fv), fv)
        forces.scatter_add_(0, col_m.unsqueeze(1).expand_as(fv), -fv)
        return {"energy": emb.sum() + pair_e.sum(), "forces": forces, "virial": (s * r).sum(),
                "virial_tensor": torch.einsum("ei,ej->ij", fv, rij)}


class TorchEAMStep:
    """One reference MD step (BAOB Langevin or NVE) with EAM forces on a
    fixed half list; `rebuilds_due` counts steps whose skin trigger fired
    (the sample keeps its list)."""

    def __init__(self, x, v, mass_amu, H, edge_index, tab, dt, skin, gamma=None, T=None,
                 seed=0):
        self.x = torch.tensor(np.asarray(x), dtype=torch.float32)
        self.v = torch.tensor(np.asarray(v), dtype=torch.float32)
        self.x_last = self.x.clone()
        n = self.x.shape[0]
        self.m = torch.from_numpy(oint.internal_mass(np.full(n, mass_amu)))[:, None]
        self.box = TorchBox(H)
        self.row = torch.as_tensor(np.asarray(edge_index[0]), dtype=torch.int64)
        self.col = torch.as_tensor(np.asarray(edge_index[1]), dtype=torch.int64)
        self.types = torch.zeros(n, dtype=torch.int64)
        self.eam = TorchEAM(tab)
        self.dt = float(dt)
        self.half_skin = torch.tensor(skin, dtype=torch.float32) / 2
        self.gamma, self.T = gamma, T
        if gamma is not None:
            damp, sig = oint.langevin_factors(self.m[:, 0].double().numpy(), dt, gamma, T)
            self.damp = float(damp)
            self.sig = torch.from_numpy(sig.astype(np.float32))[:, None]
        self.gen = torch.Generator().manual_seed(seed)
        self.f = self.eam.forward(self.x, self.row, self.col, self.types, self.box)["forces"]
        self.rebuilds_due = 0
        self.out = None

    def step(self):
        vh = self.v + 0.5 * (self.f / self.m) * self.dt
        self.x = self.x + vh * self.dt
        d = torch.norm(self.box.minimum_image(self.x - self.x_last), dim=1)
        if bool(d.max() > self.half_skin):
            self.rebuilds_due += 1
        self.out = self.eam.forward(self.x, self.row, self.col, self.types, self.box)
        self.f = self.out["forces"]
        v = vh
        if self.gamma is not None:
            xi = torch.randn(v.shape, generator=self.gen, dtype=torch.float32)
            v = v * self.damp + xi * self.sig
        self.v = v + 0.5 * (self.f / self.m) * self.dt
        return (0.5 * self.m * self.v ** 2).sum()
