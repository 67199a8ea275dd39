"""Oracle: EAM potential file, uniform tables and two-pass EAM (test infrastructure only).

Restates
  io_utils/eam_parser.py:23-210     file layout detection, numeric-line flattening,
                                    FS/alloy data order, r*phi -> phi, not-a-knot
                                    cubic splines (scipy.interpolate.CubicSpline)
  core/force/eam_force_cu.py:334-395  uniform tables on linspace(0, rc, n_r) and
                                    linspace(rho_min, rho_max, n_rho)
  core/force/eam_force_cu.py:437-459  clamped linear interpolation
  core/force/eam_force_cu.py:461-579  rho, F(rho), F'(rho), pair energy, force scalar
                                    s = F'_i f'_{ti,tj} + F'_j f'_{tj,ti} + phi',
                                    f = -s rij/(r + 1e-8), virial = sum s r,
                                    virial_tensor = sum f (x) rij

The spline stage calls scipy's CubicSpline, as the reference does (scipy is a
third-party dependency of the reference, present in this image at 1.15.3).
Tables are evaluated in the requested dtype: float32 reproduces the
reference's float32 knots/coefficients (`_spline_eval`, eam_force_cu.py:397-415),
float64 evaluates the same splines in double precision for the FP64 run.

pinned by (tests/test_oracle_physics.py): analytic tables (synthetic file)
within the linear-interpolation bound, a hand-computed two-atom case and a
closed-form analytic EAM evaluated independently, the reference's zero-stress and
elastic-constant regressions on run_data/W/WRe_YC2.eam.fs
(cuda_test/test_eam_fs_parser_regression.py, cuda_test/test_w_elastic_static.py),
and the reference's own tables and outputs (tests/golden/).
"""
from collections import defaultdict

import numpy as np
from scipy.interpolate import CubicSpline

from . import box as obox


class EAMFile:
    """eam_parser.py:11-210 (same detection rules and data placement)."""

    def __init__(self, path):
        self.filepath = str(path)
        with open(self.filepath) as f:
            lines = f.readlines()
        first = 0
        for i, line in enumerate(lines):
            if i > 2 and len(line.strip()) > 0:
                first = i
                break
        head = lines[first:]
        parts = head[0].split()
        if len(parts) > 1 and all(c.isalpha() for c in parts[1:]):
            self._parse_alloy(head)
        else:
            self._parse_setfl(lines)

    def _parse_setfl(self, lines):
        """eam_parser.py:50-102."""
        hs = 0
        for i in range(len(lines)):
            if len(lines[i].strip()) > 0 and lines[i][0] != "#":
                hs = i
                break
        self.elements = [lines[hs + 3].split()[3]]
        p = lines[hs + 4].split()
        self.n_rho, self.d_rho = int(p[0]), float(p[1])
        self.n_r, self.d_r, self.cutoff = int(p[2]), float(p[3]), float(p[4])
        data = [float(v) for line in lines[hs + 5:] for v in line.split()]
        self.is_fs_format = False
        rho = np.arange(self.n_rho) * self.d_rho
        r = np.arange(self.n_r) * self.d_r
        self.F = {0: CubicSpline(rho, np.array(data[:self.n_rho]), bc_type="not-a-knot")}
        dens = CubicSpline(r, np.array(data[self.n_rho:self.n_rho + self.n_r]), bc_type="not-a-knot")
        self.dens = {0: {0: dens}}
        rphi = np.array(data[self.n_rho + self.n_r:self.n_rho + 2 * self.n_r])
        phi = np.divide(rphi, r, out=np.zeros_like(rphi), where=r != 0)
        self.phi = {0: {0: CubicSpline(r, phi, bc_type="not-a-knot")}}

    def _parse_alloy(self, lines):
        """eam_parser.py:104-210."""
        h = lines[0].split()
        ne = int(h[0])
        self.elements = h[1:]
        p = lines[1].split()
        self.n_rho, self.d_rho = int(p[0]), float(p[1])
        self.n_r, self.d_r, self.cutoff = int(p[2]), float(p[3]), float(p[4])
        data = []
        for line in lines[2:]:
            line = line.strip()
            if line:
                try:
                    row = [float(v) for v in line.split()]
                except ValueError:
                    continue
                data.extend(row)
        exp_alloy = ne * (self.n_rho + self.n_r) + (ne * (ne + 1) // 2) * self.n_r
        exp_fs = ne * (self.n_rho + ne * self.n_r) + (ne * (ne + 1) // 2) * self.n_r
        self.is_fs_format = (
            self.filepath.lower().endswith(".eam.fs")
            or any("eam.fs" in line.lower() for line in lines[:3])
            or (len(data) >= exp_fs and len(data) - exp_alloy >= ne * (ne - 1) * self.n_r))
        ptr = 0
        rho = np.arange(self.n_rho) * self.d_rho
        r = np.arange(self.n_r) * self.d_r
        self.F = {}
        self.dens = defaultdict(dict)
        for beta in range(ne):
            if ptr + self.n_rho > len(data):
                raise ValueError("embedding data short")
            self.F[beta] = CubicSpline(rho, np.array(data[ptr:ptr + self.n_rho]), bc_type="not-a-knot")
            ptr += self.n_rho
            if self.is_fs_format:
                for alpha in range(ne):
                    if ptr + self.n_r > len(data):
                        raise ValueError("density data short")
                    self.dens[alpha][beta] = CubicSpline(r, np.array(data[ptr:ptr + self.n_r]),
                                                         bc_type="not-a-knot")
                    ptr += self.n_r
            else:
                if ptr + self.n_r > len(data):
                    raise ValueError("density data short")
                sp = CubicSpline(r, np.array(data[ptr:ptr + self.n_r]), bc_type="not-a-knot")
                ptr += self.n_r
                for alpha in range(ne):
                    self.dens[alpha][beta] = sp
        self.phi = defaultdict(dict)
        for i in range(ne):
            for j in range(i, ne):
                rphi = np.array(data[ptr:ptr + self.n_r])
                ptr += self.n_r
                phi = np.divide(rphi, r, out=np.zeros_like(rphi), where=r != 0)
                sp = CubicSpline(r, phi, bc_type="not-a-knot")
                self.phi[i][j] = sp
                self.phi[j][i] = sp


def spline_eval(sp, x, dtype):
    """eam_force_cu.py:397-415 in `dtype` (float32 = reference semantics)."""
    knots = np.asarray(sp.x).astype(dtype)
    c = np.asarray(sp.c).astype(dtype)
    x = np.clip(np.asarray(x, dtype), knots[0], knots[-1])
    idx = np.searchsorted(knots, x, side="right") - 1
    idx = np.clip(idx, 0, len(knots) - 2)
    dx = x - knots[idx]
    if c.shape[0] == 4:
        return c[3, idx] + dx * (c[2, idx] + dx * (c[1, idx] + dx * c[0, idx]))
    if c.shape[0] == 3:
        return c[2, idx] + dx * (c[1, idx] + dx * c[0, idx])
    return c[1, idx] + dx * c[0, idx]


class EAMTables:
    """Uniform tables, eam_force_cu.py:334-395. dtype float32 or float64."""

    def __init__(self, f: EAMFile, n_r=8192, n_rho=4096, dtype=np.float64):
        self.dtype = dtype
        self.ne = len(f.elements)
        self.cutoff = float(f.cutoff)
        self.n_r = n_r
        self.r_max = self.cutoff
        self.dr = self.r_max / (n_r - 1)
        self.inv_dr = 1.0 / self.dr
        if dtype == np.float32:
            import torch  # torch.linspace fixes the float32 grid points of the reference
            rg = torch.linspace(0.0, self.r_max, n_r, dtype=torch.float32).numpy()
        else:
            rg = np.linspace(0.0, self.r_max, n_r)
        ne = self.ne
        self.dens = np.zeros((ne, ne, n_r), dtype)
        self.ddens = np.zeros((ne, ne, n_r), dtype)
        self.phi = np.zeros((ne, ne, n_r), dtype)
        self.dphi = np.zeros((ne, ne, n_r), dtype)
        for a in range(ne):
            for b in range(ne):
                sp = f.dens[a][b]
                self.dens[a, b] = spline_eval(sp, rg, dtype)
                self.ddens[a, b] = spline_eval(sp.derivative(1), rg, dtype)
                if b in f.phi[a]:
                    sp = f.phi[a][b]
                    self.phi[a, b] = spline_eval(sp, rg, dtype)
                    self.dphi[a, b] = spline_eval(sp.derivative(1), rg, dtype)
        self.rho_min, self.drho, self.inv_drho, self.n_rho = [], [], [], []
        self.F = np.zeros((ne, n_rho), dtype)
        self.dF = np.zeros((ne, n_rho), dtype)
        for a in range(ne):
            x = np.asarray(f.F[a].x).astype(np.float32 if dtype == np.float32 else np.float64)
            lo, hi = float(x[0]), float(x[-1])
            d = (hi - lo) / (n_rho - 1)
            if dtype == np.float32:
                import torch
                g = torch.linspace(lo, hi, n_rho, dtype=torch.float32).numpy()
            else:
                g = np.linspace(lo, hi, n_rho)
            self.F[a] = spline_eval(f.F[a], g, dtype)
            self.dF[a] = spline_eval(f.F[a].derivative(1), g, dtype)
            self.rho_min.append(lo)
            self.drho.append(d)
            self.inv_drho.append(1.0 / d)
            self.n_rho.append(n_rho)

    def interp_r(self, r, table):
        """eam_force_cu.py:437-446."""
        dt = self.dtype
        r = np.asarray(r, dt)
        rcl = np.clip(r, dt(0.0), dt(self.r_max * (1 - 1e-7)))
        idx_f = rcl * dt(self.inv_dr)
        idx = idx_f.astype(np.int64)
        frac = idx_f - idx.astype(dt)
        nxt = np.minimum(idx + 1, self.n_r - 1)
        v0 = table[..., idx]
        v1 = table[..., nxt]
        return v0 + frac * (v1 - v0)

    def interp_rho(self, rho, a, deriv=False):
        """eam_force_cu.py:448-459."""
        dt = self.dtype
        lo, d, inv, n = self.rho_min[a], self.drho[a], self.inv_drho[a], self.n_rho[a]
        rc = np.clip(np.asarray(rho, dt), dt(lo), dt(lo + d * (n - 1) * (1 - 1e-7)))
        idx_f = (rc - dt(lo)) * dt(inv)
        idx = idx_f.astype(np.int64)
        frac = idx_f - idx.astype(dt)
        nxt = np.minimum(idx + 1, n - 1)
        t = self.dF[a] if deriv else self.F[a]
        return t[idx] + frac * (t[nxt] - t[idx])


def eam_forward(pos, edge_index, types, tab: EAMTables, H, Hinv=None, ortho=None):
    """eam_force_cu.py:461-579 on a half list, in tab.dtype arithmetic.

    Returns dict(energy, forces, virial, virial_tensor, rho, dF).
    """
    dt = tab.dtype
    pos = np.asarray(pos)
    types = np.asarray(types)
    n = pos.shape[0]
    row, col = np.asarray(edge_index[0]), np.asarray(edge_index[1])
    rij_all = obox.minimum_image(pos[row] - pos[col], H, Hinv, ortho).astype(dt)
    dist_all = np.sqrt((rij_all[:, 0] * rij_all[:, 0] + rij_all[:, 1] * rij_all[:, 1])
                       + rij_all[:, 2] * rij_all[:, 2])
    m = dist_all < dt(tab.cutoff)
    r = dist_all[m]
    rij = rij_all[m]
    row, col = row[m], col[m]
    tr, tc = types[row], types[col]
    rho = np.zeros(n, dt)
    np.add.at(rho, row, _interp_pairs(tab, r, tab.dens, tr, tc))
    np.add.at(rho, col, _interp_pairs(tab, r, tab.dens, tc, tr))
    Femb = np.zeros(n, dt)
    dF = np.zeros(n, dt)
    for a in range(tab.ne):
        sel = types == a
        Femb[sel] = tab.interp_rho(rho[sel], a)
        dF[sel] = tab.interp_rho(rho[sel], a, deriv=True)
    pair_e = _interp_pairs(tab, r, tab.phi, tr, tc)
    dd_row = _interp_pairs(tab, r, tab.ddens, tc, tr)
    dd_col = _interp_pairs(tab, r, tab.ddens, tr, tc)
    dphi = _interp_pairs(tab, r, tab.dphi, tr, tc)
    s = dF[row] * dd_col + dF[col] * dd_row + dphi
    fv = -s[:, None] * (rij / (r[:, None] + dt(1e-8)))
    forces = np.zeros((n, 3), dt)
    np.add.at(forces, row, fv)
    np.add.at(forces, col, -fv)
    return {
        "energy": Femb.astype(np.float64).sum() + pair_e.astype(np.float64).sum(),
        "forces": forces,
        "virial": (s * r).astype(np.float64).sum(),
        "virial_tensor": np.einsum("ei,ej->ij", fv.astype(np.float64), rij.astype(np.float64)),
        "rho": rho,
        "dF": dF,
        "embedding_energy": Femb.astype(np.float64).sum(),
        "pair_potential": pair_e.astype(np.float64).sum(),
    }


def _interp_pairs(tab, r, table, ta, tb):
    """table[ta, tb] interpolated at r, per edge (eam_force_cu.py:437-446)."""
    dt = tab.dtype
    rcl = np.clip(r, dt(0.0), dt(tab.r_max * (1 - 1e-7)))
    idx_f = rcl * dt(tab.inv_dr)
    idx = idx_f.astype(np.int64)
    frac = idx_f - idx.astype(dt)
    nxt = np.minimum(idx + 1, tab.n_r - 1)
    v0 = table[ta, tb, idx]
    v1 = table[ta, tb, nxt]
    return v0 + frac * (v1 - v0)
