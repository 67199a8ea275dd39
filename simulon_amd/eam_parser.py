"""EAM potential files (setfl / eam.alloy / eam.fs) and the uniform tables the
EAM kernels read.

Format handling follows io_utils/eam_parser.py:23-210 of the reference, so the
same file yields the same functions:
  * the first non-empty line after the three comment lines decides the layout:
    "<n> <El> <El> ..." -> alloy/FS, anything else -> single-element setfl
  * alloy/FS data are flattened from numeric lines only: a line holding any
    non-numeric token (the "74 183.84 3.16 bcc" element lines) is skipped whole
  * FS layout (file name *.eam.fs, "eam.fs" in the comments, or enough data)
    stores n_el density functions after each embedding function; the k-th one
    in element beta's block is density_splines_by_pair[k][beta]
  * pair data are r*phi(r); phi = (r phi)/r with phi(0) = 0
  * every function is a not-a-knot cubic spline (scipy CubicSpline, as in the
    reference) with its first derivative.

Uniform tables follow core/force/eam_force_cu.py:334-415: float32 knots and
coefficients, grids torch.linspace(0, rc, n_r) and
torch.linspace(rho_min, rho_max, n_rho) in float32.
"""
from collections import defaultdict

import numpy as np
import torch
from scipy.interpolate import CubicSpline


class EAMParser:
    """Parsed EAM file: elements, grids, cutoff and spline objects."""

    def __init__(self, filepath, device=None):
        self.filepath = str(filepath)
        self.device = device
        with open(self.filepath) as f:
            lines = f.readlines()
        first = 0
        for i, line in enumerate(lines):
            if i > 2 and line.strip():
                first = i
                break
        head = lines[first:]
        tok = head[0].split()
        if len(tok) > 1 and all(t.isalpha() for t in tok[1:]):
            self._alloy(head)
        else:
            self._setfl(lines)
        self.element_map = {e: k for k, e in enumerate(self.elements)}

    @staticmethod
    def _spline(x, y):
        sp = CubicSpline(x, y, bc_type="not-a-knot")
        return sp, sp.derivative(1)

    def _init_maps(self):
        self.embedding_splines, self.embedding_deriv_splines = {}, {}
        self.density_splines, self.density_deriv_splines = {}, {}
        self.density_splines_by_pair = defaultdict(dict)
        self.density_deriv_splines_by_pair = defaultdict(dict)
        self.pair_potential_splines = defaultdict(dict)
        self.pair_potential_deriv_splines = defaultdict(dict)

    def _grid_params(self, tok):
        self.n_rho, self.d_rho = int(tok[0]), float(tok[1])
        self.n_r, self.d_r, self.cutoff = int(tok[2]), float(tok[3]), float(tok[4])

    def _setfl(self, lines):
        start = next(i for i, l in enumerate(lines) if l.strip() and l[0] != "#")
        self.elements = [lines[start + 3].split()[3]]
        self._grid_params(lines[start + 4].split())
        vals = [float(v) for l in lines[start + 5:] for v in l.split()]
        self._init_maps()
        self.is_fs_format = False
        rho = np.arange(self.n_rho) * self.d_rho
        r = np.arange(self.n_r) * self.d_r
        self.embedding_splines[0], self.embedding_deriv_splines[0] = self._spline(
            rho, np.array(vals[:self.n_rho]))
        sp, dsp = self._spline(r, np.array(vals[self.n_rho:self.n_rho + self.n_r]))
        self.density_splines[0] = self.density_splines_by_pair[0][0] = sp
        self.density_deriv_splines[0] = self.density_deriv_splines_by_pair[0][0] = dsp
        rphi = np.array(vals[self.n_rho + self.n_r:self.n_rho + 2 * self.n_r])
        phi = np.divide(rphi, r, out=np.zeros_like(rphi), where=r != 0)
        self.pair_potential_splines[0][0], self.pair_potential_deriv_splines[0][0] = \
            self._spline(r, phi)

    def _alloy(self, lines):
        tok = lines[0].split()
        ne = int(tok[0])
        self.elements = tok[1:]
        self._grid_params(lines[1].split())
        self._init_maps()
        vals = []
        for l in lines[2:]:
            l = l.strip()
            if not l:
                continue
            try:
                row = [float(v) for v in l.split()]
            except ValueError:
                continue            # element line ("74 183.84 3.16 bcc")
            vals.extend(row)
        n_alloy = ne * (self.n_rho + self.n_r) + ne * (ne + 1) // 2 * self.n_r
        n_fs = ne * (self.n_rho + ne * self.n_r) + ne * (ne + 1) // 2 * self.n_r
        self.is_fs_format = (self.filepath.lower().endswith(".eam.fs")
                             or any("eam.fs" in l.lower() for l in lines[:3])
                             or (len(vals) >= n_fs and len(vals) - n_alloy >= ne * (ne - 1) * self.n_r))
        rho = np.arange(self.n_rho) * self.d_rho
        r = np.arange(self.n_r) * self.d_r
        p = 0

        def take(n, what):
            nonlocal p
            if p + n > len(vals):
                raise ValueError(f"{what}: need {n} values, {len(vals) - p} left")
            out = np.array(vals[p:p + n])
            p += n
            return out

        for beta in range(ne):
            self.embedding_splines[beta], self.embedding_deriv_splines[beta] = self._spline(
                rho, take(self.n_rho, "embedding function"))
            if self.is_fs_format:
                for alpha in range(ne):
                    sp, dsp = self._spline(r, take(self.n_r, "FS density function"))
                    self.density_splines_by_pair[alpha][beta] = sp
                    self.density_deriv_splines_by_pair[alpha][beta] = dsp
                self.density_splines[beta] = self.density_splines_by_pair[beta][beta]
                self.density_deriv_splines[beta] = self.density_deriv_splines_by_pair[beta][beta]
            else:
                sp, dsp = self._spline(r, take(self.n_r, "density function"))
                self.density_splines[beta], self.density_deriv_splines[beta] = sp, dsp
                for alpha in range(ne):
                    self.density_splines_by_pair[alpha][beta] = sp
                    self.density_deriv_splines_by_pair[alpha][beta] = dsp
        for i in range(ne):
            for j in range(i, ne):
                rphi = take(self.n_r, "pair function")
                phi = np.divide(rphi, r, out=np.zeros_like(rphi), where=r != 0)
                sp, dsp = self._spline(r, phi)
                self.pair_potential_splines[i][j] = self.pair_potential_splines[j][i] = sp
                self.pair_potential_deriv_splines[i][j] = dsp
                self.pair_potential_deriv_splines[j][i] = dsp


def spline_eval_f32(sp, x):
    """Evaluate a scipy PPoly on float32 x with float32 knots and coefficients
    (eam_force_cu.py:397-415)."""
    knots = torch.from_numpy(np.asarray(sp.x)).to(torch.float32)
    c = torch.from_numpy(np.asarray(sp.c)).to(torch.float32)
    x = torch.clamp(x, knots[0], knots[-1])
    idx = torch.clamp(torch.searchsorted(knots, x, right=True) - 1, 0, len(knots) - 2)
    dx = x - knots[idx]
    k = c.shape[0]
    if k == 4:
        return c[3, idx] + dx * (c[2, idx] + dx * (c[1, idx] + dx * c[0, idx]))
    if k == 3:
        return c[2, idx] + dx * (c[1, idx] + dx * c[0, idx])
    if k == 2:
        return c[1, idx] + dx * c[0, idx]
    raise RuntimeError("unsupported spline order")


class EAMTables:
    """float32 uniform tables of one EAM file (eam_force_cu.py:334-395)."""

    def __init__(self, parser: EAMParser, n_r=8192, n_rho=4096):
        self.parser = parser
        self.ne = len(parser.elements)
        self.cutoff = float(parser.cutoff)
        self.n_r, self.n_rho = int(n_r), int(n_rho)
        self.r_max = self.cutoff
        self.dr = self.r_max / (self.n_r - 1)
        self.inv_dr = 1.0 / self.dr
        rg = torch.linspace(0.0, self.r_max, self.n_r)
        ne = self.ne
        z = lambda: torch.zeros((ne, ne, self.n_r), dtype=torch.float32)
        self.density, self.density_deriv, self.pair, self.pair_deriv = z(), z(), z(), z()
        for a in range(ne):
            for b in range(ne):
                self.density[a, b] = spline_eval_f32(parser.density_splines_by_pair[a][b], rg)
                self.density_deriv[a, b] = spline_eval_f32(
                    parser.density_deriv_splines_by_pair[a][b], rg)
                if b in parser.pair_potential_splines[a]:
                    self.pair[a, b] = spline_eval_f32(parser.pair_potential_splines[a][b], rg)
                    self.pair_deriv[a, b] = spline_eval_f32(
                        parser.pair_potential_deriv_splines[a][b], rg)
        self.rho_min, self.drho, self.inv_drho = [], [], []
        self.embed = torch.zeros((ne, self.n_rho), dtype=torch.float32)
        self.embed_deriv = torch.zeros((ne, self.n_rho), dtype=torch.float32)
        for a in range(ne):
            x = torch.from_numpy(np.asarray(parser.embedding_splines[a].x)).to(torch.float32)
            lo, hi = float(x[0]), float(x[-1])
            d = (hi - lo) / (self.n_rho - 1)
            g = torch.linspace(lo, hi, self.n_rho)
            self.embed[a] = spline_eval_f32(parser.embedding_splines[a], g)
            self.embed_deriv[a] = spline_eval_f32(parser.embedding_deriv_splines[a], g)
            self.rho_min.append(lo)
            self.drho.append(d)
            self.inv_drho.append(1.0 / d)

    @staticmethod
    def _pairs(t):
        """[..., n] -> [..., n, 2] holding (t[k], t[min(k+1, n-1)])."""
        nxt = torch.cat([t[..., 1:], t[..., -1:]], dim=-1)
        return torch.stack([t, nxt], dim=-1)

    def device_arrays(self, device, type_map=None):
        """Interleaved tables for the kernels (see pair_eam.hip), re-indexed so
        that row k belongs to engine atom type k = file element type_map[k]."""
        tm = torch.arange(self.ne) if type_map is None else torch.as_tensor(type_map)
        sel = lambda t: t[tm][:, tm]
        p1 = torch.cat([self._pairs(sel(self.density)), self._pairs(sel(self.pair))], dim=-1)
        p2 = torch.cat([self._pairs(sel(self.density_deriv)), self._pairs(sel(self.pair_deriv))],
                       dim=-1)
        return {
            "pass1": p1.contiguous().to(device),        # [nt, nt, n_r, 4]
            "pass2": p2.contiguous().to(device),
            "emb": self._pairs(self.embed[tm]).contiguous().to(device),   # [nt, n_rho, 2]
            "demb": self._pairs(self.embed_deriv[tm]).contiguous().to(device),
        }
