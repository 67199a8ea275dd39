"""Device-resident MD engine over packed replicas (one GPU).

One `ReplicaEngine` owns the state of R >= 1 independent systems in HBM and
advances all of them with one launch sequence per step:

    first half (B+A, float64 positions, displacement max)
    -> skin trigger (per-replica rebuild flag, no host sync)
    -> cell-list rebuild, gated per replica by the flag
    -> force field (LJ, or EAM density + force)
    -> second half (O+B, Philox Langevin kicks, KE / kinetic tensor)
    -> float64 compensated reductions -> Berendsen (NPT replicas)

The sequence contains no host synchronisation and no allocation, so it is
captured once into a HIP graph (torch.cuda.CUDAGraph) and replayed. A single
system is R = 1; the batched replica scan is the same code with R = 64.

Reference behaviour mirrored (JinZongxiao/Simulon):
  step order                core/md_model.py:88-166
  skin/2 rebuild trigger    io_utils/reader.py:136-149 (displacement since the
                            last build; with Berendsen the reference resets its
                            reference positions every step, core/barostat.py:67,
                            so it never rebuilds — here the reference positions
                            are scaled with the box instead and the threshold
                            accounts for the accumulated scaling)
  integrator constants      core/integrator/integrator.py:41-87
  Berendsen                 core/barostat.py:21-68
"""
import ctypes
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream_ptr

KB = 8.617333262e-5
N_A = 6.02214076e23
J_PER_EV = 1.60218e-19
MASS_CONV = 10.0 / (N_A * J_PER_EV)
MB_MASS_CONV = 1.036427e-4


def internal_mass_f32(amu):
    """integrator.py:61-64: float32(amu) * float32(10 / (N_A J_per_eV))."""
    return np.float32(np.float32(amu) * np.float32(MASS_CONV))


def langevin_f32(mass_int, dt, gamma, T):
    """integrator.py:84-87,110 in float32: (damp, kick std)."""
    g = np.float32(gamma)
    dtf = np.float32(dt)
    exp_term = np.exp(np.float32(-2.0) * g * dtf)
    rff = np.sqrt((np.float32(KB) * np.float32(T)) / np.float32(mass_int) * (np.float32(1.0) - exp_term))
    damp = np.exp(-g * dtf)
    return np.float32(damp), np.float32(rff)


def list_radius_sq(cutoff, skin):
    """reader.py:50,77-78 + gpu_kdtree.py:99: (fp32 cutoff + fp32 skin)^2 in float32."""
    rl = np.float32(cutoff) + np.float32(skin)
    return float(np.float32(rl * rl)), float(rl)


def morton_order(pos, H, bits=10):
    """Permutation sorting atoms along a Z-order curve of their wrapped
    fractional coordinates (bits per dimension), so that atoms close in
    space are close in memory."""
    s = np.asarray(pos, np.float64) @ np.linalg.inv(np.asarray(H, np.float64))
    s = s - np.floor(s)
    q = np.minimum((s * (1 << bits)).astype(np.uint64), (1 << bits) - 1)
    key = np.zeros(len(q), np.uint64)
    for b in range(bits):
        for d in range(3):
            key |= ((q[:, d] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + d)
    return np.argsort(key, kind="stable")


def noise_key(seed):
    """64-bit Philox key of a replica's noise stream (splitmix64 finaliser of
    its seed, so that nearby seeds give unrelated keys)."""
    z = (int(seed) + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def perpendicular_widths(H):
    H = np.asarray(H, np.float64)
    V = abs(np.linalg.det(H))
    a, b, c = H
    return np.array([V / np.linalg.norm(np.cross(b, c)), V / np.linalg.norm(np.cross(c, a)),
                     V / np.linalg.norm(np.cross(a, b))])


@dataclass
class ReplicaSpec:
    """One independent system: atoms, box and thermodynamic controls."""
    positions: np.ndarray                 # [N,3] Angstrom (float64 or float32)
    box: np.ndarray                       # [3,3] lattice vectors as rows
    types: np.ndarray = None              # [N] int, default all 0
    velocities: np.ndarray = None         # [N,3] A/ps; None -> Maxwell-Boltzmann at t_init
    dt: float = 0.001                     # ps
    ensemble: str = "NVE"                 # NVE | NVT | NPT
    temperature: float = 300.0            # Langevin target, K
    t_init: float = None                  # velocity initialisation, K (default temperature)
    gamma: float = 10.0                   # Langevin friction, 1/ps
    target_pressure: float = 1.0          # bar
    tau_p: float = 1.0                    # ps
    kappa: float = 4.57e-5                # 1/bar (barostat.py:27)
    mu_max: float = 0.05                  # barostat.py:28
    seed: int = None                      # Langevin noise stream; None: engine seed + replica index
    barostat: str = "berendsen"           # NPT controller: berendsen | anisotropic


FAMILIES = ("cell", "rows", "global")
ROW_CAPS = (192, 256, 320, 384, 512)          # atoms per stencil column (row kernels)
CELL_CAPS = (512, 1024, 1280, 1536, 2048)     # atoms per 27-cell stencil (cell kernels)


class ReplicaEngine:
    """Packed multi-replica MD state and step on one GPU."""

    def __init__(self, specs, masses_amu, cutoff, skin, device="cuda", max_nbr=None,
                 precise_list=True, seed=12345, velocity_seed=None, sort=True, tiled=None,
                 family=None, stage=True):
        """tiled / family choose the pair-kernel family: family "cell" (one
        workgroup per cell, 27-cell halo in LDS), "rows" (one workgroup per row
        of cells, sliding LDS window of stencil columns) or "global" (one
        thread per atom over the int32 list); tiled=True means "rows",
        tiled=False "global"; both None = measured choice at setup().
        stage: let the global-list EAM passes stage each group's neighbour
        records in LDS (stage_arrays); the same results bit for bit."""
        self.device = torch.device(device)
        if isinstance(specs, ReplicaSpec):
            specs = [specs]
        self.specs = list(specs)
        self.R = len(self.specs)
        self.masses_amu = [float(m) for m in masses_amu]
        if len(self.masses_amu) > _lib.SIM_MAX_TYPES:
            raise ValueError(f"at most {_lib.SIM_MAX_TYPES} atom types")
        self.cutoff = float(cutoff)
        self.skin = float(skin)
        self.precise_list = bool(precise_list)
        self._seed = int(seed)
        self.sort = bool(sort)
        self.stage_on = bool(stage)
        self.stage_cap = _lib.SIM_STAGE_CAP_MAX   # members per group (groups beyond: unstaged)
        self._stage = None
        _lib.load()
        self._layout()
        self._alloc(max_nbr, velocity_seed)
        # the LDS-staged kernels need >= 3 cells per dimension in every replica
        self.can_tile = all(min(self._grid_of(r)) >= 3 for r in range(self.R))
        if family is None and tiled is not None:
            family = "rows" if tiled else "global"
        if family not in (None,) + FAMILIES:
            raise ValueError(f"family must be one of {FAMILIES}")
        self._family_auto = family is None
        self.family = family or "rows"
        if not self.can_tile:
            self.family = "global"
        self.cap_rows = 0     # atoms per stencil column (row kernels)
        self.cap_cell = 0     # atoms per 27-cell stencil (cell kernels)
        # stencil populations only change at rebuilds; overflow is checked
        self.cap_margin = 1.05
        self.force_field = None
        self._scratch = {}
        self._graph = None
        self._graph_virial = None
        self._fuse_next = False
        self.chain = True          # graph runs of >= 2 steps may chain (run())
        self.steps_done = 0
        self.loading = False        # a tensile loader deforms the box at the start of each step
        # True while no box has changed since construction (cleared by every
        # write of a replica record, a deformation and a box assignment): the
        # cell grids are then the ones can_tile was decided on
        self.box_fixed = True
        self.profile = {"graph_replays": 0, "eager_steps": 0}

    # ---- layout --------------------------------------------------------------
    def _layout(self):
        R = self.R
        self.natoms = [int(len(s.positions)) for s in self.specs]
        self.atom_base = np.concatenate([[0], np.cumsum(self.natoms)]).astype(np.int64)
        self.N = int(self.atom_base[-1])
        if self.N >= 2 ** 28:
            # the pair kernels address 16-byte records by 32-bit byte offsets
            raise ValueError("at most 2^28 - 1 atoms per engine")
        tiles, rep_tiles = [], [0]
        for r in range(R):
            for k in range(0, self.natoms[r], _lib.SIM_TILE):
                tiles.append((r, int(self.atom_base[r]) + k))
            rep_tiles.append(len(tiles))
        self.ntiles = len(tiles)
        self.tiles_np = np.array(tiles, np.int32).reshape(-1, 2)
        self.rep_tiles_np = np.array(rep_tiles, np.int32)
        rl2, rl = list_radius_sq(self.cutoff, self.skin)
        self.rl2, self.rl = rl2, rl
        caps, self.cell_base = [], [0]
        for s in self.specs:
            w = perpendicular_widths(s.box)
            if s.ensemble.upper() == "NPT":
                w = w * 1.25
            nc = np.maximum(1, np.floor(w / (rl * (1 + 1e-6)))).astype(np.int64)
            nc = np.minimum(nc, 1024)
            caps.append(int(nc.prod()))
            self.cell_base.append(self.cell_base[-1] + caps[-1])
        self.cell_cap = caps
        self.ncells = int(self.cell_base[-1])
        self.cell_rep_np = np.repeat(np.arange(R, dtype=np.int32), caps)

    @property
    def family(self):
        """Pair-kernel family: "cell", "rows" or "global"."""
        return self._family

    @family.setter
    def family(self, fam):
        # builds under the global list skip the column slots and the row
        # prefix; a switch to an LDS family rebuilds them at the next build
        if fam != "global" and getattr(self, "_family", fam) == "global":
            self._stale_slots = True
        self._family = fam

    @property
    def tiled(self):
        return self.family != "global"

    @property
    def cap(self):
        """LDS capacity of the current family (0 for the global list)."""
        return {"rows": self.cap_rows, "cell": self.cap_cell}.get(self.family, 0)

    @cap.setter
    def cap(self, v):
        if self.family == "rows":
            self.cap_rows = int(v)
        elif self.family == "cell":
            self.cap_cell = int(v)

    @property
    def use_tiled(self):
        return self.family != "global" and self.cap > 0

    def pick_cap(self):
        """LDS capacities from the populations of the last build (host read,
        setup only): per stencil column for the row kernels, per 27-cell
        stencil for the cell kernels. A column or stencil that outgrows its
        capacity later is read from global memory (slower, still exact); a
        family without a capacity falls back to the global list."""
        if not self.can_tile or self.cap_rows or self.cap_cell:
            return
        cmax, smax = (int(v) for v in self.stencil_max.cpu())
        self.cap_rows = next((c for c in ROW_CAPS if c >= int(cmax * self.cap_margin) + 8), 0)
        cell_caps = CELL_CAPS[:2] if self._family_auto else CELL_CAPS
        self.cap_cell = next((c for c in cell_caps if c >= int(smax * self.cap_margin) + 16), 0)
        if self.cap == 0:
            self.family = "global"

    def autotune(self, reps=5):
        """Measured choice of the pair-kernel family (automatic mode only):
        every available family (cell, rows, global list) evaluates the forces
        of the current configuration with a small random jitter (restored
        afterwards), each evaluation timed with HIP events after the caches were
        flushed by a 512 MB write (as the other kernels of a step flush them),
        and the fastest is kept. Cached per (force
        field, replica count, atom count, list radius, LDS capacities)."""
        if not (self._family_auto and self.can_tile and self.force_field is not None):
            return
        ff = self.force_field
        key = (getattr(ff, "name", type(ff).__name__), self.R, self.N, self.rl, self.cap_rows,
               self.cap_cell)
        choice = _TUNE_CACHE.get(key)
        if choice is None:
            virial = self.has_barostat
            cands = [f for f, c in (("cell", self.cap_cell), ("rows", self.cap_rows),
                                    ("global", 1)) if c]
            flush = torch.empty(128 << 20, dtype=torch.float32, device=self.device)
            # time on thermally disordered positions (a perfect lattice repeats
            # the same pair distances and table entries): +-0.1 A, within the
            # skin, restored afterwards
            saved = self.pos4[:, :3].clone()
            g = torch.Generator(device=self.device)
            g.manual_seed(7)
            self.pos4[:, :3] += (torch.rand(saved.shape, generator=g, device=self.device,
                                            dtype=torch.float64) - 0.5) * 0.2
            times = {}
            for fam in cands:
                self.family = fam
                self.compute_forces(virial=virial)
                total = 0.0
                for _ in range(reps):
                    flush.fill_(1.0)
                    a = torch.cuda.Event(enable_timing=True)
                    b = torch.cuda.Event(enable_timing=True)
                    a.record()
                    self.compute_forces(virial=virial)
                    b.record()
                    b.synchronize()
                    total += a.elapsed_time(b)
                times[fam] = total / reps
            del flush
            self.pos4[:, :3] = saved
            choice = min(times, key=times.get)
            _TUNE_CACHE[key] = choice
            self.profile["autotune_ms"] = times
        self.family = choice
        self._stale_slots = False   # no build since setup's, which wrote every slot

    def cell_args(self):
        """Leading arguments of the sim_*_cell entry points."""
        return (self.ncells, ptr(self.reps), ptr(self.cell_rep), ptr(self.cell_start),
                ptr(self.cell_count), ptr(self.pos4), ptr(self.nbr16), ptr(self.nnbr),
                self.max_nbr, self.cap)

    def row_args(self):
        """Leading arguments of the sim_*_rows entry points."""
        return (self.ncells, ptr(self.reps), ptr(self.cell_rep), ptr(self.row_base), self.R,
                ptr(self.cell_start), ptr(self.cell_count), ptr(self.pos4), ptr(self.nbr16),
                ptr(self.nnbr), self.max_nbr, self.cap)

    def _grid_of(self, r):
        w = perpendicular_widths(self.specs[r].box)
        nc = np.maximum(1, np.floor(w / (self.rl * (1 + 1e-6)))).astype(np.int64)
        return [int(min(x, 1024)) for x in nc]

    def _replica_struct(self, r):
        s = self.specs[r]
        P = _lib.Replica()
        H = np.asarray(s.box, np.float64)
        Hinv = torch.linalg.inv(torch.as_tensor(H, dtype=torch.float64)).numpy()
        for k in range(9):
            P.h[k] = float(H.flat[k])
            P.hinv[k] = float(Hinv.flat[k])
        off = H - np.diag(np.diag(H))
        P.ortho = int(bool(np.all(np.abs(off) <= 1e-8)))
        P.volume = float(abs(np.linalg.det(H)))
        P.rl2, P.rl = self.rl2, self.rl
        half = np.float32(self.skin) / np.float32(2.0)
        P.trig2 = float(np.float64(half) ** 2)
        P.skin = float(np.float32(self.skin))
        P.mu_acc = 1.0
        P.dt = float(s.dt)
        P.dtf = float(np.float32(s.dt))
        P.mu = 1.0
        P.target_p, P.tau_p, P.kappa, P.mu_max = s.target_pressure, s.tau_p, s.kappa, s.mu_max
        ens = s.ensemble.upper()
        P.langevin = int(ens in ("NVT", "NPT"))
        P.barostat = (2 if s.barostat == "anisotropic" else 1) if ens == "NPT" else 0
        damp = np.float32(1.0)
        for t, m in enumerate(self.masses_amu):
            mi = internal_mass_f32(m)
            if P.langevin:
                damp, rff = langevin_f32(mi, s.dt, s.gamma, s.temperature)
                P.rff[t] = float(rff)
        P.damp = float(damp)
        P.rcf = float(np.float32(self.cutoff))
        P.noise_key = self.replica_noise_key(r)
        P.atom_base = int(self.atom_base[r])
        P.natoms = self.natoms[r]
        P.cell_base = int(self.cell_base[r])
        P.cell_cap = int(self.cell_cap[r])
        return P

    def replica_seed(self, r):
        """Seed of replica r's Langevin noise: its spec's seed, else the
        engine seed + r (a replica packed with others draws the same kicks as
        a single-replica engine of that seed)."""
        s = self.specs[r].seed
        return self._seed + r if s is None else int(s)

    def replica_noise_key(self, r):
        return noise_key(self.replica_seed(r))

    @property
    def seed(self):
        return self._seed

    @seed.setter
    def seed(self, value):
        """A new engine seed re-keys the replicas that have no seed of their own."""
        self._seed = int(value)
        if getattr(self, "reps", None) is None:
            return
        o = _lib.Replica.noise_key.offset
        sz = ctypes.sizeof(_lib.Replica)
        keys = np.array([self.replica_noise_key(r) for r in range(self.R)], np.uint64)
        raw = torch.from_numpy(keys.view(np.uint8).reshape(self.R, 8).copy())
        self.reps.view(self.R, sz)[:, o:o + 8].copy_(raw.to(self.device))
        self._graph = None

    def noise_keys(self):
        """[R] int64 view of the replicas' Philox keys (device)."""
        o = _lib.Replica.noise_key.offset
        sz = ctypes.sizeof(_lib.Replica)
        return self.reps.view(self.R, sz)[:, o:o + 8].contiguous().view(torch.int64).view(self.R)

    def _alloc(self, max_nbr, velocity_seed):
        dev, N, R = self.device, self.N, self.R
        i32 = dict(dtype=torch.int32, device=dev)
        pos = np.zeros((N, 4), np.float64)
        vel = np.zeros((N, 3), np.float32)
        order = np.arange(N, dtype=np.int64)
        rng = np.random.default_rng(self._seed if velocity_seed is None else velocity_seed)
        for r, s in enumerate(self.specs):
            b, e = self.atom_base[r], self.atom_base[r + 1]
            x = np.asarray(s.positions, np.float64)
            t = np.zeros(self.natoms[r], np.int64) if s.types is None else np.asarray(s.types)
            if t.size and (t.min() < 0 or t.max() >= len(self.masses_amu)):
                raise ValueError("atom type out of range")
            if s.velocities is not None:
                v = np.asarray(s.velocities, np.float32)
            else:
                T0 = s.temperature if s.t_init is None else s.t_init
                m = np.array([self.masses_amu[k] for k in t], np.float64) * MB_MASS_CONV
                v = rng.standard_normal((len(t), 3)) * np.sqrt(KB * T0 / m)[:, None]
                v -= (v * m[:, None]).sum(0) / m.sum()
                v = v.astype(np.float32)
            perm = morton_order(x, s.box) if self.sort else np.arange(len(x))
            pos[b:e, :3] = x[perm]
            pos[b:e, 3] = t[perm]
            vel[b:e] = v[perm]
            order[b:e] = b + perm
        # order[k]: caller's index of the atom stored in slot k
        self.order = torch.from_numpy(order).to(dev)
        inv = np.empty_like(order)
        inv[order] = np.arange(N)
        self.slot_of = torch.from_numpy(inv.astype(np.int32)).to(dev)
        self.pos4 = torch.from_numpy(pos).to(dev)
        self.pos_last = self.pos4.clone()
        self.vel = torch.from_numpy(vel).to(dev)
        self.vel_half = torch.zeros_like(self.vel)
        self.force = torch.zeros((N, 3), dtype=torch.float32, device=dev)
        self.posq = torch.zeros((N, 4), dtype=torch.int32, device=dev)
        self.tiles = torch.from_numpy(self.tiles_np).to(dev)
        self.rep_tiles = torch.from_numpy(self.rep_tiles_np).to(dev)
        self.reps = _lib.struct_tensor([self._replica_struct(r) for r in range(R)], dev)
        self.cell_rep = torch.from_numpy(self.cell_rep_np).to(dev)
        self.cell_count = torch.zeros(self.ncells + 1, **i32)
        self.cell_start = torch.zeros(self.ncells + 1, **i32)
        self.cell_atoms = torch.zeros(N, **i32)
        self.atom_cell = torch.zeros(N, **i32)
        self.atom_slot = torch.zeros(N, **i32)
        if max_nbr is None:
            est = 0.0
            for r, s in enumerate(self.specs):
                rho = self.natoms[r] / abs(np.linalg.det(np.asarray(s.box, np.float64)))
                est = max(est, 4.0 / 3.0 * math.pi * self.rl ** 3 * rho)
            max_nbr = int(math.ceil((1.3 * est + 16) / 8.0) * 8)
        self.max_nbr = int(math.ceil(int(max_nbr) / 8.0) * 8)   # rows 16-byte aligned
        self.nbr = torch.zeros((self.max_nbr, N), **i32)
        self.nnbr = torch.zeros(N, **i32)
        self.overflow = torch.zeros(1, **i32)
        self.flag = torch.ones(R, **i32)
        self.nbuild = torch.zeros(R, **i32)
        self.act = torch.zeros(R + 1, **i32)      # cell prefix of the staged list build
        self.maxd2 = torch.zeros(R, dtype=torch.float64, device=dev)
        self.partial = torch.zeros((self.ntiles, _lib.SIM_NPART), dtype=torch.float64, device=dev)
        self.thermo = torch.zeros((R, _lib.SIM_NPART), dtype=torch.float64, device=dev)
        self.red_chunks = torch.zeros((R, 64, _lib.SIM_NPART), dtype=torch.float64, device=dev)
        self.step_ctr = torch.zeros(1, dtype=torch.int64, device=dev)
        # physical cell-order reorder at every rebuild (sim_nl_build with a reorder)
        self.image = torch.zeros((N, 4), dtype=torch.int32, device=dev)
        self._ro_tmp = {
            "pos": torch.zeros((N, 4), dtype=torch.float64, device=dev),
            "vel": torch.zeros((N, 3), dtype=torch.float32, device=dev),
            "force": torch.zeros((N, 3), dtype=torch.float32, device=dev),
            "order": torch.zeros(N, dtype=torch.int64, device=dev),
            "image": torch.zeros((N, 4), dtype=torch.int32, device=dev),
            "cell": torch.zeros(N, dtype=torch.int32, device=dev),
            "vel_half": torch.zeros((N, 3), dtype=torch.float32, device=dev),
        }
        self.nbr16 = torch.zeros((N, self.max_nbr), dtype=torch.int16, device=dev)
        self.stencil_max = torch.zeros(2, dtype=torch.int32, device=dev)
        self.partial_cell = torch.zeros((self.ncells, _lib.SIM_NPART), dtype=torch.float64,
                                        device=dev)
        self.rep_cells = torch.as_tensor(np.array(self.cell_base, np.int32), device=dev)
        ro = _lib.NLReorder()
        ro.vel, ro.force = self.vel.data_ptr(), self.force.data_ptr()
        ro.order, ro.slot_of = self.order.data_ptr(), self.slot_of.data_ptr()
        ro.image = self.image.data_ptr()
        ro.pos_tmp, ro.vel_tmp = self._ro_tmp["pos"].data_ptr(), self._ro_tmp["vel"].data_ptr()
        ro.force_tmp, ro.order_tmp = (self._ro_tmp["force"].data_ptr(),
                                      self._ro_tmp["order"].data_ptr())
        ro.image_tmp, ro.cell_tmp = (self._ro_tmp["image"].data_ptr(),
                                     self._ro_tmp["cell"].data_ptr())
        ro.nbr16 = self.nbr16.data_ptr()
        ro.stencil_max = self.stencil_max.data_ptr()
        self.row_base = torch.zeros(R + 1, dtype=torch.int32, device=dev)
        ro.row_base = self.row_base.data_ptr()
        ro.vel_half, ro.vel_half_tmp = (self.vel_half.data_ptr(),
                                        self._ro_tmp["vel_half"].data_ptr())
        ro.posq = self.posq.data_ptr()     # a reorder rewrites the records it moves
        self._ro = ro
        self._posq_fresh = False           # records already current (inside a step)
        self.types_struct = _lib.Types()
        for t, m in enumerate(self.masses_amu):
            self.types_struct.mass[t] = float(internal_mass_f32(m))
        self.types_struct.ntypes = len(self.masses_amu)

    def scratch(self, name, dtype, shape=None):
        """Per-engine work array owned by the engine (allocated on first use,
        outside any graph capture, and kept for the engine's lifetime)."""
        shape = (self.N,) if shape is None else tuple(shape)
        t = self._scratch.get(name)
        if t is None or t.dtype != dtype or tuple(t.shape) != shape:
            if torch.cuda.is_current_stream_capturing():
                raise RuntimeError(f"scratch array {name!r} first requested during graph capture")
            t = torch.zeros(shape, dtype=dtype, device=self.device)
            self._scratch[name] = t
        return t

    # ---- primitives (each a stream-ordered launch) -----------------------------
    def build_list(self, force=False):
        """Skin trigger (force=True: every replica) and the gated rebuild."""
        s = stream_ptr()
        if getattr(self, "_stale_slots", False) and self.use_tiled:
            force = True
            self._stale_slots = False
        # the cell-row prefix and the 16-bit column slots are read by the LDS
        # kernel families only (and are built at the forced build of setup,
        # before the autotuner has chosen)
        lds = force or self.use_tiled
        self._ro.row_base = (self.row_base.data_ptr()
                             if force or (self.use_tiled and self.family == "rows") else None)
        self._ro.nbr16 = self.nbr16.data_ptr() if lds else None
        # fixed boxes whose grids have >= 3 cells per dimension never need the
        # thread-per-atom fallback list kernel (the device grid is the host's
        # _grid_of while no box changes)
        self._ro.stencil3_only = int(self.can_tile and self.box_fixed and not self.has_barostat
                                     and not self.loading)
        # the skin trigger and the gated rebuild in one call (sim_nl_trigger +
        # sim_nl_build in fewer launches)
        call("sim_nl_update", self.R, ptr(self.reps), ptr(self.tiles), self.ntiles, self.N,
             ptr(self.pos4), ptr(self.pos_last), ptr(self.maxd2), ptr(self.flag),
             ptr(self.nbuild), int(bool(force)), ptr(self.cell_rep), self.ncells,
             ptr(self.atom_cell), ptr(self.atom_slot), ptr(self.cell_count), ptr(self.cell_start),
             ptr(self.cell_atoms), ptr(self.nbr), ptr(self.nnbr), self.max_nbr,
             ptr(self.overflow), int(self.precise_list), ptr(self.act), ctypes.byref(self._ro), s)
        if self._stage is not None:
            self._stage_build(self.flag)

    # the persistent EAM kernels, which stage, run at >= 256 tiles
    STAGE_MIN_TILES = 256

    def stage_arrays(self):
        """The sim_stage_t of this engine's lists (created, and built for every
        group, on first use; from then on rebuilt with the lists, for the
        groups of rebuilt replicas), or None when staging is off or the engine
        is too small for the staged kernels."""
        if not self.stage_on or self.ntiles < self.STAGE_MIN_TILES:
            return None
        if self._stage is None:
            ng = -(-self.ntiles // _lib.SIM_STAGE_TILES)
            cap = int(self.stage_cap)
            if not 1 <= cap <= _lib.SIM_STAGE_CAP_MAX:
                raise ValueError(f"stage_cap must be in 1..{_lib.SIM_STAGE_CAP_MAX}")
            dev = self.device
            self.stage_members = torch.zeros(ng * cap, dtype=torch.int32, device=dev)
            self.stage_count = torch.full((ng,), -1, dtype=torch.int32, device=dev)
            self.stage_slot = torch.zeros((self.max_nbr, self.N), dtype=torch.int16, device=dev)
            self.stage_table = torch.zeros(ng * 3 * _lib.SIM_STAGE_TABLE, dtype=torch.int32,
                                           device=dev)
            st = _lib.Stage()
            st.members, st.count = self.stage_members.data_ptr(), self.stage_count.data_ptr()
            st.table = self.stage_table.data_ptr()
            st.slot, st.cap = self.stage_slot.data_ptr(), cap
            self._stage = st
            self._stage_build(None)
        return self._stage

    def _stage_build(self, flag):
        call("sim_nl_stage", self.R, ptr(self.reps), ptr(self.tiles), self.ntiles, ptr(flag),
             ptr(self.nbr), ptr(self.nnbr), self.N, self.max_nbr, ptr(self.cell_rep),
             ptr(self.atom_cell), ptr(self.cell_start), ptr(self.cell_count),
             ctypes.byref(self._stage), stream_ptr())

    def quantize(self):
        """Fractional fixed-point records of the current positions (read by
        the full-list force kernels). Inside a step the first half and the
        list build have written them already (build_and_forces)."""
        if self._posq_fresh:
            return
        call("sim_pos_quantize", ptr(self.reps), ptr(self.tiles), self.ntiles, ptr(self.pos4),
             ptr(self.posq), stream_ptr())

    def build_and_forces(self, virial, noise=None, second=False):
        """Gated rebuild and forces right after first_half(records=True): the
        records of the new positions are current (written by the first half,
        rewritten by a reorder), so the force pass does not quantize again.
        second=True: the step's second half too (noise as in second_half), in
        the force pass where the force field fuses it, else after it."""
        self.build_list(force=False)
        self._posq_fresh = True
        try:
            fused = False
            if second and getattr(self.force_field, "fuses_second", False):
                fused = self.force_field.compute(self, virial, second=self.second_args(noise))
            else:
                self.compute_forces(virial=virial)
        finally:
            self._posq_fresh = False
        if self._fuse_next and not fused:
            raise RuntimeError("the next step's first half was to be fused into a force pass "
                               "that does not fuse the second half")
        if second and not fused:
            self.second_half(noise, advance=False)

    def second_args(self, noise=None):
        """Arguments of the second half after the force arrays: types, v_half,
        v, noise, atom keys, seed, device step counter."""
        return (self.types_struct, ptr(self.vel_half), ptr(self.vel), ptr(noise), ptr(self.order),
                0, ptr(self.step_ctr))

    def next_first_args(self):
        """Arguments of sim_eam_force_second for the next step's first half
        (positions, build positions, displacement maxima, records), or four
        NULLs when the step being recorded does not fuse it."""
        if not self._fuse_next:
            return (None, None, None, None)
        return (ptr(self.pos4), ptr(self.pos_last), ptr(self.maxd2), ptr(self.posq))

    def fuses_next_first(self):
        """Whether graph replays may let the force pass do the next step's first
        half (the EAM global-list pass with one element type)."""
        ff = self.force_field
        return ff is not None and bool(getattr(ff, "fuses_next_first", lambda e: False)(self))

    def compute_forces(self, virial=True):
        if self.force_field is None:
            raise RuntimeError("no force field attached")
        self.force_field.compute(self, virial)

    def first_half(self, records=False, unfused_only=False, scale=False):
        """B+A; records=True also writes the position records (posq).
        unfused_only: only the replicas whose first half the previous step's
        force pass did not do (those with a barostat or loading). scale: the
        Berendsen replicas first get the scaling of x, v and the build
        positions that the previous step's barostat left undone."""
        if records:
            flags = ((_lib.SIM_FIRST_UNFUSED_ONLY if unfused_only else 0)
                     | (_lib.SIM_FIRST_BERENDSEN_SCALE if scale else 0))
            call("sim_integrate_first_q", ptr(self.reps), ptr(self.tiles), self.ntiles,
                 self.types_struct, ptr(self.pos4), ptr(self.vel), ptr(self.force),
                 ptr(self.vel_half), ptr(self.pos_last), ptr(self.maxd2), ptr(self.posq),
                 flags, stream_ptr())
            return
        if unfused_only or scale:
            raise ValueError("unfused_only and scale need records=True")
        call("sim_integrate_first", ptr(self.reps), ptr(self.tiles), self.ntiles, self.types_struct,
             ptr(self.pos4), ptr(self.vel), ptr(self.force), ptr(self.vel_half),
             ptr(self.pos_last), ptr(self.maxd2), stream_ptr())

    def second_half(self, noise=None, advance=True):
        """O+B; advance=False leaves the noise step counter to the step's last
        reduction (reduce(..., advance=True))."""
        s = stream_ptr()
        call("sim_integrate_second", ptr(self.reps), ptr(self.tiles), self.ntiles,
             self.types_struct, ptr(self.pos4), ptr(self.vel_half), ptr(self.force), ptr(self.vel),
             ptr(noise), ptr(self.order), 0, 0, ptr(self.step_ctr), ptr(self.partial), s)
        if advance:
            call("sim_step_advance", ptr(self.step_ctr), s)

    def reduce(self, col0=0, col1=15, ff=None, advance=False):
        """float64 per-replica sums of the partial columns [col0, col1): force
        columns (< T_KE) from the partials of force field ff (default: the
        attached one), kinetic columns from the integrator's. advance=True
        also advances the noise step counter (the step's last reduction)."""
        s = stream_ptr()
        fa, fb = col0, min(col1, _lib.T_KE)
        ka, kb = max(col0, _lib.T_KE), col1
        part, rows = self.force_partials(ff) if fa < fb else (None, None)

        def red(rows_, part_, a, b, last):
            if advance and last:
                call("sim_reduce_step", self.R, ptr(rows_), ptr(part_), a, b,
                     ptr(self.red_chunks), ptr(self.thermo), ptr(self.step_ctr), s)
            else:
                call("sim_reduce", self.R, ptr(rows_), ptr(part_), a, b, ptr(self.red_chunks),
                     ptr(self.thermo), s)

        if part is self.partial and rows is self.rep_tiles and ka < kb:
            # force and kinetic columns share the per-tile rows: one pass
            red(rows, part, fa, kb, True)
            return
        if fa < fb:
            red(rows, part, fa, fb, not ka < kb)
        if ka < kb:
            red(self.rep_tiles, self.partial, ka, kb, True)
        elif advance and not fa < fb:
            call("sim_step_advance", ptr(self.step_ctr), s)

    def force_partials(self, ff=None):
        """(partial rows, per-replica row offsets) the force field writes."""
        ff = self.force_field if ff is None else ff
        if ff is not None and hasattr(ff, "partials"):
            return ff.partials(self)
        if self.use_tiled:
            return self.partial_cell, self.rep_cells
        return self.partial, self.rep_tiles

    def barostat(self, defer_scale=False):
        """The barostat update of the NPT replicas. defer_scale: the Berendsen
        replicas' box only; the next step's first half scales their atoms."""
        kinds = {s.barostat for s in self.specs if s.ensemble.upper() == "NPT"}
        if "berendsen" in kinds and defer_scale:
            call("sim_berendsen_box", self.R, ptr(self.reps), ptr(self.thermo), stream_ptr())
        elif "berendsen" in kinds:
            call("sim_berendsen", self.R, ptr(self.reps), ptr(self.tiles), self.ntiles,
                 ptr(self.thermo), ptr(self.pos4), ptr(self.vel), ptr(self.pos_last), stream_ptr())
        if "anisotropic" in kinds:
            call("sim_aniso_barostat", self.R, ptr(self.reps), ptr(self.tiles), self.ntiles,
                 ptr(self.thermo), ptr(self.pos4), ptr(self.vel), ptr(self.pos_last),
                 0, ptr(self.step_ctr), stream_ptr())

    def deform(self):
        """Scale the lattice rows of the replicas with loading = 1 by their
        deform factors, carrying the atoms with the box."""
        self.box_fixed = False
        call("sim_box_deform", self.R, ptr(self.reps), ptr(self.tiles), self.ntiles,
             ptr(self.pos4), ptr(self.pos_last), stream_ptr())

    @property
    def has_barostat(self):
        return any(s.ensemble.upper() == "NPT" for s in self.specs)

    # ---- step ----------------------------------------------------------------
    def setup(self):
        """Initial list and forces at the current positions (md_model.py:91-102)."""
        self.build_list(force=True)
        self.pick_cap()
        self.autotune()
        self.compute_forces(virial=True)
        self.reduce(0, 8)
        self._ke_now()

    def _ke_now(self):
        """Kinetic columns of thermo from the current velocities (setup only)."""
        m = torch.tensor([float(x) for x in self.types_struct.mass[:len(self.masses_amu)]],
                         dtype=torch.float64, device=self.device)
        mi = m[self.pos4[:, 3].long()]
        v = self.vel.double()
        prods = torch.stack([v[:, 0] * v[:, 0], v[:, 1] * v[:, 1], v[:, 2] * v[:, 2],
                             v[:, 0] * v[:, 1], v[:, 0] * v[:, 2], v[:, 1] * v[:, 2]], 1) * mi[:, None]
        K = torch.stack([prods[self.replica_slice(r)].sum(0) for r in range(self.R)])
        self.thermo[:, _lib.T_K:_lib.T_K + 6] = K
        self.thermo[:, _lib.T_KE] = 0.5 * (K[:, 0] + K[:, 1] + K[:, 2])

    def step(self, virial=None, noise=None):
        """One MD step for every replica (md_model.py:103-153)."""
        self._step(virial, noise)

    def _step(self, virial=None, noise=None, first_unfused=False, fuse_next=False,
              scale_pending=False, defer_scale=False):
        """step() as recorded in the graph replays of run(): first_unfused: the
        previous step's force pass did the first half of the replicas without
        barostat or loading (no first-half launch at all when every replica is
        such); fuse_next: this step's force pass does the next step's first
        half for them; scale_pending: the previous step left its Berendsen
        scaling of the atoms to this step's first half; defer_scale: this
        step leaves it to the next."""
        if virial is None:
            virial = True
        if self.loading:
            self.deform()
        if not (first_unfused and self._all_fusable()):
            self.first_half(records=True, unfused_only=first_unfused, scale=scale_pending)
        self._fuse_next = fuse_next
        try:
            self.build_and_forces(virial or self.has_barostat, noise, second=True)
        finally:
            self._fuse_next = False
        self.reduce(0, 15, advance=True)
        if self.has_barostat:
            self.barostat(defer_scale=defer_scale)
        self.steps_done += 1

    # graph replays: _graph (an unfused step) and, when the force field
    # allows, three steps chained through the fused first half
    # (first step, middle steps, last step), cleared together
    @property
    def _graph(self):
        return self.__dict__.get("_graph_main")

    @_graph.setter
    def _graph(self, g):
        self.__dict__["_graph_main"] = g
        if g is None:
            self.__dict__["_graphs_chain"] = {}
            self.__dict__["_fusable_all"] = None

    def _record(self, virial, first_unfused=False, fuse_next=False, scale_pending=False,
                defer_scale=False):
        if getattr(self, "_stale_slots", False) and self.use_tiled:
            self.build_list(force=True)     # refresh the LDS families' slots first
        g = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            self._step(virial, first_unfused=first_unfused, fuse_next=fuse_next,
                       scale_pending=scale_pending, defer_scale=defer_scale)
        self.steps_done -= 1
        torch.cuda.synchronize()
        return g

    def capture(self, virial=True):
        """Record one step into a HIP graph. Recording executes nothing, so the
        state is unchanged until the first replay."""
        g = self._record(virial)
        self._graph, self._graph_virial = g, virial
        return g

    def _all_fusable(self):
        """Whether no replica record asks for a barostat or loading step (the
        records as the chain graphs were recorded; write_replica_state clears
        those graphs when a record asks for either)."""
        v = self.__dict__.get("_fusable_all")
        if v is None:
            raw = self.reps.cpu().numpy().tobytes()
            sz = ctypes.sizeof(_lib.Replica)
            recs = [_lib.Replica.from_buffer_copy(raw[r * sz:(r + 1) * sz]) for r in range(self.R)]
            v = all(P.barostat == 0 and P.loading == 0 for P in recs)
            self.__dict__["_fusable_all"] = v
        return v

    def _defers_scale(self):
        """Whether a run may leave each Berendsen step's scaling of the atoms
        to the next step's first half (Berendsen replicas, no loading: a
        deformation between the two would see unscaled atoms)."""
        return (not self.loading and any(s.ensemble.upper() == "NPT" and s.barostat == "berendsen"
                                         for s in self.specs))

    def _chain(self, virial):
        """The three graphs of a run of >= 2 steps whose force passes do the
        next step's first half and / or whose Berendsen steps leave the
        scaling of the atoms to the next step's first half: (first, middle,
        last); the last step does both of its own."""
        chains = self.__dict__.setdefault("_graphs_chain", {})
        if virial not in chains:
            self._all_fusable()        # host read of the records, before recording
            f, d = self.fuses_next_first(), self._defers_scale()
            chains[virial] = (self._record(virial, False, f, False, d),
                              self._record(virial, f, f, d, d),
                              self._record(virial, f, False, d, False))
        return chains[virial]

    def run(self, nsteps, graph=True, virial=True, check=True):
        """nsteps steps. graph=True replays captured steps; a run of two or
        more steps chains them where it can: with the EAM global-list pass each
        force pass but the last does the next step's first half, and each
        Berendsen step but the last leaves its scaling of the atoms to the next
        step's first half (the same bits as separate steps; the state between
        runs is that of whole steps; self.chain = False turns chaining off)."""
        if graph:
            if self._graph is None or self._graph_virial != virial:
                self.capture(virial)
            if nsteps >= 2 and self.chain and (self.fuses_next_first() or self._defers_scale()):
                g_first, g_mid, g_last = self._chain(virial)
                g_first.replay()
                for _ in range(nsteps - 2):
                    g_mid.replay()
                g_last.replay()
            else:
                for _ in range(nsteps):
                    self._graph.replay()
            self.steps_done += nsteps
            self.profile["graph_replays"] += nsteps
        else:
            for _ in range(nsteps):
                self.step(virial=virial)
            self.profile["eager_steps"] += nsteps
        if check:
            self.check()

    def check(self):
        """Raise if a neighbor list overflowed or a cell grid became too small
        for the cell-tiled kernels (host read; call between chunks). Stencils
        larger than the LDS tile are not errors (those cells read global
        memory); they are counted in profile["lds_overflow"]."""
        ov = int(self.overflow.item())
        if ov & 2:
            self.profile["lds_overflow"] = self.profile.get("lds_overflow", 0) + 1
            self.overflow.fill_(ov & ~2)
        if ov & 1:
            raise RuntimeError(f"neighbor list overflow: an atom has more than {self.max_nbr} "
                               "neighbors; construct the engine with a larger max_nbr")
        if ov & 8:
            raise RuntimeError("a stencil column holds more than 16383 atoms; construct the "
                               "engine with tiled=False")
        if ov & 4:
            raise RuntimeError("a replica's cell grid fell below 3 cells per dimension; "
                               "construct the engine with tiled=False")

    # ---- host views ------------------------------------------------------------
    def to_caller(self, x):
        """Per-atom array in storage order -> the caller's atom order."""
        return x[self.slot_of]

    def from_caller(self, x):
        """Per-atom array in the caller's order -> storage order."""
        return x[self.order]

    def positions(self, dtype=torch.float64, unwrap=False):
        """Positions in storage order. The rebuilds wrap atoms into the primary
        cell (the forces are computed on these); unwrap=True adds back the
        lattice translations, giving continuous trajectories."""
        x = self.pos4[:, :3]
        if unwrap:
            H = torch.as_tensor(self.box_rows(), device=self.device)      # [R,3,3]
            rep = self.atom_replica()
            x = x + torch.einsum("nk,nkj->nj", self.image[:, :3].double(), H[rep])
        return x.to(dtype)

    def atom_replica(self):
        if getattr(self, "_atom_rep", None) is None:
            self._atom_rep = torch.repeat_interleave(
                torch.arange(self.R, device=self.device),
                torch.as_tensor(self.natoms, device=self.device))
        return self._atom_rep

    def box_rows(self):
        """[R,3,3] current lattice vectors (rows) of every replica (device)."""
        sz = ctypes.sizeof(_lib.Replica)
        rec = self.reps.view(self.R, sz)[:, :72]
        return rec.contiguous().view(torch.float64).view(self.R, 3, 3)

    def set_positions(self, xyz):
        """Replace the positions (storage order, continuous coordinates as
        positions(unwrap=True) returns them). They are stored in the frame of
        the last rebuild (minus the recorded lattice translations), so the
        skin trigger keeps measuring true displacements."""
        xyz = torch.as_tensor(xyz, device=self.device).to(torch.float64)
        H = self.box_rows()[self.atom_replica()]
        self.pos4[:, :3] = xyz - torch.einsum("nk,nkj->nj", self.image[:, :3].double(), H)

    def replica_slice(self, r):
        return slice(int(self.atom_base[r]), int(self.atom_base[r + 1]))

    def replica_state(self, r=0):
        return _lib.struct_from_tensor(self.reps, _lib.Replica, r)

    def write_replica_state(self, r, P):
        """Store a (modified) replica record; a captured graph keeps reading
        the same device memory, so the change applies to later replays."""
        import ctypes
        sz = ctypes.sizeof(_lib.Replica)
        self.box_fixed = False
        raw = torch.frombuffer(bytearray(bytes(P)), dtype=torch.uint8)
        self.reps[r * sz:(r + 1) * sz].copy_(raw.to(self.device))
        self.specs[r].dt = P.dt
        if (P.barostat or P.loading) and self.__dict__.get("_fusable_all"):
            # the chain graphs were recorded without a first-half launch
            self.__dict__["_graphs_chain"] = {}
            self.__dict__["_fusable_all"] = None

    def box(self, r=0):
        return np.array(self.replica_state(r).h[:], np.float64).reshape(3, 3)

    def thermo_host(self):
        return self.thermo.cpu().numpy()

    def energies(self):
        th = self.thermo_host()
        return th[:, _lib.T_EPOT], th[:, _lib.T_KE]

    def virial_tensor(self, r=0):
        th = self.thermo[r]
        w = th[_lib.T_W:_lib.T_W + 6]
        return torch.stack([torch.stack([w[0], w[3], w[4]]), torch.stack([w[3], w[1], w[5]]),
                            torch.stack([w[4], w[5], w[2]])])

    def kinetic_tensor(self, r=0):
        th = self.thermo[r]
        k = th[_lib.T_K:_lib.T_K + 6]
        return torch.stack([torch.stack([k[0], k[3], k[4]]), torch.stack([k[3], k[1], k[5]]),
                            torch.stack([k[4], k[5], k[2]])])

    def half_list(self):
        """Current list as the reference's half list: edge_index int64 [2,E] (i<j,
        lexicographic) and float64 distances, per replica indices are global."""
        return half_list_from_full(self.nbr, self.nnbr, self.max_nbr, self.reps[:ctypes_size()],
                                   self.positions(torch.float64))


# autotune choices: key -> kernel family
_TUNE_CACHE = {}


def ctypes_size():
    import ctypes
    return ctypes.sizeof(_lib.Replica)


def half_list_from_full(nbr, nnbr, max_nbr, rep_bytes, xyz):
    """Full column-major list -> (edge_index [2,E] int64 sorted, distances)."""
    n = xyz.shape[0]
    dev = xyz.device
    hc = torch.zeros(n + 1, dtype=torch.int32, device=dev)
    hs = torch.zeros(n + 1, dtype=torch.int32, device=dev)
    s = stream_ptr()
    call("sim_nl_half_count", ptr(nbr), ptr(nnbr), n, max_nbr, ptr(hc), ptr(hs), s)
    E = int(hs[n].item())
    ei = torch.empty((2, E), dtype=torch.int64, device=dev)
    dist = torch.empty(E, dtype=xyz.dtype, device=dev)
    dtype = 0 if xyz.dtype == torch.float32 else 1
    call("sim_nl_half_fill", ptr(rep_bytes), ptr(xyz.contiguous()), dtype, n, ptr(nbr), ptr(nnbr),
         max_nbr, ptr(hs), ptr(ei) if E else None, E, ptr(dist) if E else None, s)
    return ei, dist
