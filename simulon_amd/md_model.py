"""Step driver with the interface of the reference's core/md_model.py.

BackboneInterface / IntegratorInterface (:9-17), SumBackboneInterface (:19-52)
and BaseModel (:55-166) keep their names, call signatures and output dicts.
BaseModel.forward() runs one step of the device engine: first half (B+A),
skin trigger and gated list rebuild, forces, second half (O+B), float64
reductions and, with a barostat, its device update (Berendsen or
anisotropic) — all stream-ordered,
without the reference's per-step host synchronisations. BaseModel.run(n)
replays a captured HIP graph of that step n times.

A SumBackboneInterface may hold, next to its one device backbone (LJ, EAM or
MLP), any number of host backbones (plain BackboneInterface modules, e.g. an
indenter, run_scripts/w_indent.py:377): their forward() runs every step after
the device forces, and their forces, energy and virial are added before the
second half and the barostat, as the reference sums all backbones.
"""
from abc import ABCMeta, abstractmethod
from typing import Dict

import torch
import torch.nn as nn

from . import _lib


class BackboneInterface(nn.Module, metaclass=ABCMeta):
    @abstractmethod
    def forward(self) -> Dict[str, torch.Tensor]:
        pass


class IntegratorInterface(nn.Module, metaclass=ABCMeta):
    @abstractmethod
    def forward(self, force: torch.Tensor) -> Dict[str, torch.Tensor]:
        pass


class SumBackboneInterface(nn.Module):
    """Sum of backbone outputs (md_model.py:19-52)."""

    def __init__(self, backbones, molecular):
        super().__init__()
        self.backbones = nn.ModuleList(backbones)
        self.molecular = molecular
        self.device = molecular.device
        self.atom_num = molecular.atom_count

    def forward(self):
        out = None
        for b in self.backbones:
            o = b()
            if out is None:
                out = {k: o[k].clone() for k in ("forces", "energy", "virial", "virial_tensor")}
            else:
                for k in ("forces", "energy", "virial", "virial_tensor"):
                    out[k] = out[k] + o[k]
        return out

    def device_force_field(self):
        """The force field of the first device backbone (the one the fused
        step runs in its kernels), or None."""
        for b in self.backbones:
            if hasattr(b, "ff"):
                return b.ff
        return None

    def host_backbones(self):
        """Every backbone other than the device one: evaluated each step by
        its own forward() and added to the device force field's output, as
        the reference sums all backbones (md_model.py:38-46)."""
        dev = None
        out = []
        for b in self.backbones:
            if dev is None and hasattr(b, "ff"):
                dev = b
            else:
                out.append(b)
        return out


class BaseModel(nn.Module):
    def __init__(self, sum_bone, integrator, molecular, barostat=None):
        super().__init__()
        self.sum_bone = sum_bone
        self.Integrator = integrator
        self.molecular = molecular
        self.engine = molecular.engine
        self.barostat = barostat
        self._first_call = True
        self.profile = {"pair_time": 0.0, "integrate_time": 0.0, "pair_calls": 0,
                        "integrate_calls": 0}
        ff = sum_bone.device_force_field()
        if ff is None:
            raise NotImplementedError("the device step needs one LJ, EAM or MLP backbone")
        self.engine.force_field = ff
        self.extra = sum_bone.host_backbones()

    @property
    def barostat(self):
        return self.__dict__.get("_barostat")

    @barostat.setter
    def barostat(self, b):
        """Assigning a barostat switches the device controller to it (None:
        no pressure control), as swapping model.barostat does in the
        reference's workflows (run_scripts/w_tensile.py:232)."""
        old = self.__dict__.get("_barostat")
        if old is not None and old is not b and hasattr(old, "_deactivate"):
            old._deactivate()
        self.__dict__["_barostat"] = b
        if b is not None and hasattr(b, "_activate"):
            b._activate()
        self.engine._graph = None

    def _kinetic_tensor(self):
        return self.engine.kinetic_tensor(0)

    def _extra_forces(self):
        """Forces of the host backbones, added to the device forces; returns
        their summed energy, virial and virial tensor (device tensors)."""
        eng = self.engine
        self.molecular._graph_cache = None
        e = vir = W = None
        for b in self.extra:
            o = b()
            eng.force.add_(eng.from_caller(o["forces"].to(eng.device, torch.float32)))
            oe = o["energy"].to(eng.device, torch.float64)
            ov = o.get("virial", torch.zeros((), device=eng.device)).to(eng.device, torch.float64)
            ow = o.get("virial_tensor", torch.zeros(3, 3, device=eng.device)).to(eng.device,
                                                                                  torch.float64)
            e, vir, W = (oe, ov, ow) if e is None else (e + oe, vir + ov, W + ow)
        return e, vir, W

    def _add_extra_thermo(self, extra):
        """Host backbones' energy and virial into replica 0's reduced row."""
        e, vir, W = extra
        th = self.engine.thermo[0]
        th[_lib.T_EPOT] += e
        th[_lib.T_VIRIAL] += vir
        # the record keeps the symmetric part (all the barostats read); a
        # host backbone's tensor need not be symmetric (sum f (x) x of an
        # indenter), so its antisymmetric part is kept for the output
        S = 0.5 * (W + W.T)
        th[_lib.T_W:_lib.T_W + 6] += torch.stack([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2],
                                                  S[1, 2]])
        self._extra_asym = W - S

    def _initial_forces(self):
        eng = self.engine
        eng.compute_forces(virial=True)
        extra = self._extra_forces() if self.extra else None
        eng.reduce(0, 8)
        if extra is not None:
            self._add_extra_thermo(extra)

    def _step(self):
        eng = self.engine
        mol = self.molecular
        if mol.needs_update:
            mol._trigger_and_build(force=True)
        if eng.loading:
            eng.deform()
        eng.first_half(records=True)
        eng.build_and_forces(virial=True)
        extra = self._extra_forces() if self.extra else None
        eng.second_half(advance=False)
        eng.reduce(0, 15, advance=True)
        if extra is not None:
            self._add_extra_thermo(extra)
        if self.barostat is not None:
            eng.barostat()
        mol._graph_cache = None

    def forward(self):
        eng = self.engine
        if self._first_call:
            self._initial_forces()
            self._first_call = False
        self._step()
        eng.steps_done += 1
        self.profile["pair_calls"] += 1
        self.profile["integrate_calls"] += 1
        return self.result()

    def result(self):
        eng = self.engine
        th = eng.thermo[0]
        ke = th[_lib.T_KE]
        W = eng.virial_tensor(0)
        if self.extra and getattr(self, "_extra_asym", None) is not None:
            W = W + self._extra_asym
        out = {
            "forces": eng.to_caller(eng.force),
            "energy": th[_lib.T_EPOT],
            "virial": th[_lib.T_VIRIAL],
            "virial_tensor": W,
            "updated_coordinates": self.molecular.coordinates,
            "kinetic_energy": ke,
            "temperature": (2.0 / 3.0) * ke / (self.molecular.atom_count * self.Integrator.BOLTZMAN),
        }
        if self.barostat is not None:
            if hasattr(self.barostat, "pressure_axis_bar"):
                out["pressure"] = self.barostat.pressure_axis_bar.to(torch.float32)
            else:
                o = _lib.Replica.press.offset
                out["pressure"] = eng.reps[o:o + 8].view(torch.float64)[0]
        return out

    def run(self, nsteps, graph=True):
        """nsteps steps; with graph=True one captured step is replayed (with
        host backbones, whose forward() runs Python each step, the steps are
        eager)."""
        eng = self.engine
        if self.extra:
            out = None
            for _ in range(nsteps):
                out = self()
            return out if out is not None else self.result()
        if self._first_call:
            self._initial_forces()
            self._first_call = False
        if self.molecular.needs_update:
            self.molecular._trigger_and_build(force=True)
        eng.run(nsteps, graph=graph, virial=True)
        self.molecular._graph_cache = None
        return self.result()
