"""Force fields as engine plug-ins: each writes forces and the energy/virial
partial columns of every replica in one or two kernel launches."""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream_ptr
from .eam_parser import EAMParser, EAMTables

LJ_MODES = {"raw": 0, "shift": 1, "switch": 2}


class LJForceField:
    """Lennard-Jones (core/force/lennard_jones_force.py:15-95).

    eps, sig: per-type-pair tables [nt][nt] (eV, Angstrom); rc: cutoff;
    mode: raw | shift (force shift, :65-75) | switch (:48-63, start at
    switch_ratio * rc).
    """
    name = "lj"

    def __init__(self, eps, sig, rc, mode="raw", switch_ratio=0.9):
        eps = np.atleast_2d(np.asarray(eps, np.float64))
        sig = np.atleast_2d(np.asarray(sig, np.float64))
        nt = eps.shape[0]
        if nt > _lib.SIM_MAX_TYPES or eps.shape != (nt, nt) or sig.shape != (nt, nt):
            raise ValueError("eps/sig must be square per-type tables of at most 4 types")
        p = _lib.LJParams()
        for a in range(nt):
            for b in range(nt):
                p.eps[a * _lib.SIM_MAX_TYPES + b] = float(np.float32(eps[a, b]))
                p.sig[a * _lib.SIM_MAX_TYPES + b] = float(np.float32(sig[a, b]))
        p.rc = float(np.float32(rc))
        p.rs = float(np.float32(switch_ratio * rc))
        p.mode = LJ_MODES[mode] if isinstance(mode, str) else int(mode)
        p.ntypes = nt
        self.params = p
        self.ntypes = nt
        self.rc = float(rc)

    # the global-list pass can do the step's second half (sim_lj_force_second)
    fuses_second = True

    def compute(self, eng, virial=True, second=None):
        """second: the engine's second_args(); the global-list pass then also
        does the step's second half. Returns whether it did."""
        if len(eng.masses_amu) > self.ntypes:
            raise ValueError(f"{len(eng.masses_amu)} atom types, LJ tables for {self.ntypes}")
        if eng.use_tiled and eng.family == "rows":
            call("sim_lj_force_rows", *eng.row_args(), len(eng.masses_amu), self.params,
                 ptr(eng.force), ptr(eng.partial_cell), int(bool(virial)), ptr(eng.overflow),
                 stream_ptr())
            return False
        if eng.use_tiled:
            call("sim_lj_force_cell", *eng.cell_args(), self.params, ptr(eng.force),
                 ptr(eng.partial_cell), int(bool(virial)), ptr(eng.overflow), stream_ptr())
            return False
        eng.quantize()
        args = (eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles, ptr(eng.posq), ptr(eng.nbr),
                ptr(eng.nnbr), eng.N, self.params, ptr(eng.force), ptr(eng.partial),
                int(bool(virial)))
        if second is not None and self.fuses_second:
            call("sim_lj_force_second", *args, *second, stream_ptr())
            return True
        call("sim_lj_force", *args, stream_ptr())
        return False


class EAMForceField:
    """Two-pass EAM (core/force/eam_force_cu.py:461-579) from an EAM file."""
    name = "eam"

    def __init__(self, source, n_r=8192, n_rho=4096, device="cuda", type_map=None):
        """type_map[k]: index in the file's element list of engine atom type k
        (default: identity)."""
        parser = source if isinstance(source, EAMParser) else EAMParser(source)
        self.parser = parser
        self.tables = EAMTables(parser, n_r, n_rho)
        self.cutoff = self.tables.cutoff
        tm = list(range(self.tables.ne)) if type_map is None else [int(k) for k in type_map]
        if len(tm) > _lib.SIM_MAX_TYPES:
            raise ValueError("at most 4 atom types")
        self.type_map = tm
        self.dev = self.tables.device_arrays(device, tm)
        t = self.tables
        p = _lib.EAMParams()
        p.pass1 = self.dev["pass1"].data_ptr()
        p.pass2 = self.dev["pass2"].data_ptr()
        p.emb = self.dev["emb"].data_ptr()
        p.demb = self.dev["demb"].data_ptr()
        p.n_r, p.n_rho, p.ntypes = t.n_r, t.n_rho, len(tm)
        p.inv_dr = float(np.float32(t.inv_dr))
        p.r_clamp = float(np.float32(t.r_max * (1 - 1e-7)))
        p.cutoff = float(np.float32(t.cutoff))
        for k, a in enumerate(tm):
            p.rho_lo[k] = float(np.float32(t.rho_min[a]))
            p.rho_hi[k] = float(np.float32(t.rho_min[a] + t.drho[a] * (t.n_rho - 1) * (1 - 1e-7)))
            p.inv_drho[k] = float(np.float32(t.inv_drho[a]))
        self.params = p

    def embedding_energy(self, eng):
        """[R] float64 sums of F_t(rho_i) over each replica's atoms from the
        densities of the last density pass, with the kernels' float32 clamp
        and interpolation (core/force/eam_force_cu.py:448-459, 577)."""
        rho = eng.scratch("eam_rho", torch.float32)
        t = eng.pos4[:, 3].long()
        p = self.params
        lo = torch.tensor([p.rho_lo[k] for k in range(p.ntypes)], dtype=torch.float32,
                          device=eng.device)[t]
        hi = torch.tensor([p.rho_hi[k] for k in range(p.ntypes)], dtype=torch.float32,
                          device=eng.device)[t]
        inv = torch.tensor([p.inv_drho[k] for k in range(p.ntypes)], dtype=torch.float32,
                           device=eng.device)[t]
        xf = (torch.minimum(torch.maximum(rho, lo), hi) - lo) * inv
        idx = xf.to(torch.int64)
        fr = xf - idx.to(torch.float32)
        e = self.dev["emb"][t, idx]                       # [N, 2]: F(k), F(k+1)
        Fi = (e[:, 0] + fr * (e[:, 1] - e[:, 0])).double()
        return torch.stack([Fi[eng.replica_slice(r)].sum() for r in range(eng.R)])

    fuses_second = True

    def fuses_next_first(self, eng):
        """The fused force pass can also do the next step's first half: the
        global-list pass (not a tiled family) with one element type."""
        return (bool(self.fuses_second) and not eng.use_tiled and self.params.ntypes == 1
                and len(eng.masses_amu) == 1 and eng.types_struct.ntypes == 1)

    def compute(self, eng, virial=True, force=None, partial=None, keep_rho=False, second=None):
        """force/partial: optional output arrays in place of the engine's
        (float [N][3] and double [ntiles][16]; the global-list kernels run).
        keep_rho: also store the densities (embedding_energy reads them); the
        global-list step writes only the records the force pass reads.
        second: the engine's second_args(); the global-list force pass then
        also does the step's second half (sim_eam_force_second). Returns
        whether it did."""
        s = stream_ptr()
        own = force is None and partial is None
        force = eng.force if force is None else force
        partial = eng.partial if partial is None else partial
        if own and eng.use_tiled and self.params.ntypes == 1:
            rho = eng.scratch("eam_rho", torch.float32)
            fp = eng.scratch("eam_fp", torch.float32)
            fam = "rows" if eng.family == "rows" else "cell"
            common = (*(eng.row_args() if fam == "rows" else eng.cell_args()), self.params)
            call(f"sim_eam_density_{fam}", *common, ptr(rho), ptr(fp), ptr(eng.partial_cell),
                 ptr(eng.overflow), s)
            call(f"sim_eam_force_{fam}", *common, ptr(fp), ptr(eng.force), ptr(eng.partial_cell),
                 int(bool(virial)), ptr(eng.overflow), s)
            return False
        rec = eng.scratch("eam_rec", torch.int32, (eng.N, 4))
        eng.quantize()
        rho = eng.scratch("eam_rho", torch.float32) if keep_rho else None
        # LDS staging only for the engine's own force field: its steps rebuild
        # the stage with the lists. A call into work arrays (an EAM judge of
        # another force field) reads the lists directly: a stage it created
        # would not be rebuilt by that engine's captured steps and go stale.
        st = eng.stage_arrays() if own and hasattr(eng, "stage_arrays") else None
        dargs = (eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles, ptr(eng.posq), ptr(eng.nbr),
                 ptr(eng.nnbr), eng.N, self.params, ptr(rho), None, ptr(rec), ptr(partial))
        if st is not None:
            call("sim_eam_density_staged", *dargs, ctypes.byref(st), s)
        else:
            call("sim_eam_density", *dargs, s)
        args = (eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles, ptr(eng.posq), ptr(rec),
                ptr(eng.nbr), ptr(eng.nnbr), eng.N, self.params, ptr(force), ptr(partial),
                int(bool(virial)))
        if second is not None and own:
            call("sim_eam_force_second", *args, *second,
                 ctypes.byref(st) if st is not None else None, *eng.next_first_args(), s)
            return True
        if st is not None:
            call("sim_eam_force_staged", *args, ctypes.byref(st), s)
        else:
            call("sim_eam_force", *args, s)
        return False
