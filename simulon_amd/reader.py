"""The "molecular" state object with the interface of the reference's
io_utils/reader.py (AtomFileReader), backed by a one-replica device engine.

What stays as in the reference:
  * constructor AtomFileReader(filename, box_length, cutoff, device, parameter,
    skin_thickness=5.0, is_mlp, is_fs, is_switch, switch_ratio, box_vectors)
    (reader.py:22-87), xyz input, element ids by first appearance (:69-75)
  * update_coordinates(x): rebuild the Verlet list when the largest
    displacement since the last build exceeds skin/2 (:136-163)
  * update_neighbor_list(), get_cutoff_mask() (edge_attr <= cutoff, :245-248),
    set_maxwell_boltzmann_velocity(T) (:319-339), update_velocities(v)
  * graph_data.pos / edge_index (i < j, lexicographic) / edge_attr (distances),
    index_pairs, pair_params (per-edge epsilon, sigma in the parameter dict's
    key order, :165-210), profile counters

What differs: coordinates and velocities live in HBM inside the engine
(float64 positions, Morton storage order); the attributes above are views in
the caller's atom order, produced on access. graph_data is derived from the
device's full list when read.
"""
import time
import types

import numpy as np
import torch

from .box import Box
from .engine import ReplicaEngine, ReplicaSpec, half_list_from_full, ctypes_size

ELEMENT_MASS = {
    "H": 1.008, "He": 4.0026, "Li": 6.94, "C": 12.011, "N": 14.007, "O": 15.999, "Ne": 20.180,
    "Na": 22.990, "Mg": 24.305, "Al": 26.982, "Si": 28.085, "Ar": 39.948, "K": 39.098,
    "Ti": 47.867, "Cr": 51.996, "Fe": 55.845, "Co": 58.933, "Ni": 58.693, "Cu": 63.546,
    "Zn": 65.38, "Kr": 83.798, "Zr": 91.224, "Nb": 92.906, "Mo": 95.95, "Ag": 107.87,
    "Xe": 131.29, "Ta": 180.95, "W": 183.84, "Re": 186.21, "Pt": 195.08, "Au": 196.97,
}


def read_xyz(filename):
    with open(filename) as f:
        lines = [l for l in f.read().splitlines()]
    if not lines:
        raise ValueError(f"Error: The file {filename} is empty.")
    n = int(lines[0].split()[0])
    types_, xyz = [], []
    for l in lines[2:2 + n]:
        t = l.split()
        types_.append(t[0])
        xyz.append([float(t[1]), float(t[2]), float(t[3])])
    return types_, np.array(xyz, np.float64)


class AtomFileReader:
    def __init__(self, filename=None, box_length=None, cutoff=None, device="cuda", parameter=None,
                 skin_thickness=5.0, is_mlp=False, is_fs=False, is_switch=False,
                 switch_ratio: float = 0.9, box_vectors=None, *, atom_types=None,
                 positions=None, masses=None, sort=True, max_nbr=None):
        self.device = torch.device(device)
        if filename is not None:
            atom_types, positions = read_xyz(filename)
        if atom_types is None or positions is None:
            raise ValueError("give an xyz filename, or atom_types and positions")
        self.atom_types = list(atom_types)
        self.atom_count = len(self.atom_types)
        self.box = Box(box_vectors if box_vectors is not None else box_length, device=self.device)
        self.box_length_cpu = box_length
        self.box_length = (self.box.diag[0] if self.box.is_orthogonal
                           else torch.tensor(box_length, device=self.device, dtype=torch.float64))
        self.cutoff = torch.tensor(float(cutoff), device=self.device)
        self.parameter = parameter
        self.is_mlp, self.is_fs, self.is_switch = is_mlp, is_fs, is_switch
        self.switch_ratio = float(switch_ratio)
        ids = {}
        self.element_to_id = [ids.setdefault(e, len(ids)) for e in self.atom_types]
        self.elements = list(ids)
        self.element_ids = torch.tensor(self.element_to_id, device=self.device)
        if masses is None:
            masses = [ELEMENT_MASS[e] for e in self.elements]
        self.type_masses = [float(m) for m in masses]
        self.atom_mass = torch.tensor([self.type_masses[k] for k in self.element_to_id],
                                      dtype=torch.float32, device=self.device)
        self.skin_thickness = torch.tensor(float(skin_thickness), device=self.device)
        self.verlet_cutoff = self.cutoff + self.skin_thickness
        spec = ReplicaSpec(positions=np.asarray(positions, np.float64),
                           box=self.box.H.detach().cpu().numpy(),
                           types=np.array(self.element_to_id), velocities=np.zeros((self.atom_count, 3)))
        self.engine = ReplicaEngine([spec], self.type_masses, float(cutoff), float(skin_thickness),
                                    device=self.device, sort=sort, max_nbr=max_nbr)
        self.box.attach(self.engine.reps[:ctypes_size()])
        self.box._listeners.append(self._box_changed)
        self.needs_update = True
        self.profile = {"neighbor_time": 0.0, "neighbor_builds": 0, "effective_edges_last": 0}
        self._graph_cache = None
        self.update_neighbor_list()

    # ---- state views (caller's atom order) -------------------------------------
    @property
    def coordinates(self):
        return self.engine.to_caller(self.engine.positions(torch.float32, unwrap=True))

    @coordinates.setter
    def coordinates(self, x):
        self._import_positions(x)

    @property
    def atom_velocities(self):
        return self.engine.to_caller(self.engine.vel)

    @atom_velocities.setter
    def atom_velocities(self, v):
        self.update_velocities(v)

    @property
    def last_positions(self):
        eng = self.engine
        H = eng.box_rows()[0]
        x = eng.pos_last[:, :3] + eng.image[:, :3].double() @ H
        return eng.to_caller(x.float())

    def _import_positions(self, x):
        x = torch.as_tensor(x, device=self.device)
        self.engine.set_positions(self.engine.from_caller(x.to(torch.float64)))
        self._graph_cache = None

    def update_velocities(self, velocities):
        v = torch.as_tensor(velocities, device=self.device, dtype=torch.float32)
        self.engine.vel.copy_(self.engine.from_caller(v.detach()))

    def _box_changed(self, box):
        self.needs_update = True
        self.engine.box_fixed = False
        self._graph_cache = None

    # ---- neighbor list ------------------------------------------------------------
    def update_coordinates(self, coordinates):
        """reader.py:136-163: move atoms, rebuild when max |dx| > skin/2."""
        self._import_positions(coordinates.detach() if torch.is_tensor(coordinates) else coordinates)
        self._trigger_and_build(force=self.needs_update)

    def _trigger_and_build(self, force=False):
        t0 = time.perf_counter()
        before = int(self.engine.nbuild[0].item())
        # displacement max since the last build, then the gated rebuild
        d = self.engine.pos4[:, :3] - self.engine.pos_last[:, :3]
        self.engine.maxd2.copy_((d * d).sum(1).max().reshape(1))
        self.engine.build_list(force=force)
        after = int(self.engine.nbuild[0].item())
        self.needs_update = False
        if after != before:
            self.profile["neighbor_builds"] += 1
            self.profile["neighbor_time"] += time.perf_counter() - t0
        self._graph_cache = None
        self.engine.check()

    def update_neighbor_list(self):
        self._trigger_and_build(force=True)
        self.engine.pick_cap()
        self.profile["effective_edges_last"] = int(self.get_cutoff_mask().sum().item())

    def _edges(self):
        if self._graph_cache is None:
            eng = self.engine
            ei, dist = half_list_from_full(eng.nbr, eng.nnbr, eng.max_nbr, eng.reps[:ctypes_size()],
                                           eng.positions(torch.float32))
            o = eng.order[ei]                                   # caller indices
            a, b = torch.minimum(o[0], o[1]), torch.maximum(o[0], o[1])
            key = a * self.atom_count + b
            key, p = torch.sort(key)
            self._graph_cache = (torch.stack([a[p], b[p]]), dist[p])
        return self._graph_cache

    @property
    def graph_data(self):
        ei, dist = self._edges()
        gd = types.SimpleNamespace(pos=self.coordinates, edge_index=ei, edge_attr=dist)
        gd.element_edge_0 = self.element_ids[ei[0]]
        gd.element_edge_1 = self.element_ids[ei[1]]
        gd.index_pairs = torch.stack([gd.element_edge_0, gd.element_edge_1], dim=1)
        return gd

    def get_cutoff_mask(self):
        return self._edges()[1] <= self.cutoff

    def effective_edge_count(self):
        return int(self.get_cutoff_mask().sum().item())

    @property
    def pair_params(self):
        """Per-edge parameters in the parameter dict's key order (reader.py:165-210)."""
        if self.parameter is None:
            return []
        ei, _ = self._edges()
        e0, e1 = self.element_ids[ei[0]], self.element_ids[ei[1]]
        names = list(next(iter(self.parameter.values())).keys())
        out = []
        for name in names:
            tab = torch.zeros((len(self.elements), len(self.elements)), dtype=torch.float32,
                              device=self.device)
            for key, vals in self.parameter.items():
                a, b = parse_pair_key(key)
                if a < len(self.elements) and b < len(self.elements):
                    tab[a, b] = float(vals[name])
            out.append(tab[e0, e1])
        return out

    # ---- velocities ------------------------------------------------------------------
    def set_maxwell_boltzmann_velocity(self, temperature):
        """reader.py:319-339 (torch's global generator, COM momentum removed)."""
        kB, conv = 8.617333262e-5, 1.036427e-4
        T = float(temperature)
        m = (self.atom_mass * conv).unsqueeze(-1).cpu()
        v = torch.randn((self.atom_count, 3), dtype=torch.float32) * torch.sqrt(kB * T / m)
        v = v - (v * m).sum(0) / m.sum()
        self.update_velocities(v.to(self.device))


def parse_pair_key(key):
    """'[0 1]' (str(np.array([0, 1])), reader.py:116) -> (0, 1)."""
    t = str(key).strip().strip("[]()").replace(",", " ").split()
    return int(t[0]), int(t[1])
