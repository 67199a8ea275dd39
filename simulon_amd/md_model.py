"""Step driver with the interface of the reference's core/md_model.py.

BackboneInterface / IntegratorInterface (:9-17), SumBackboneInterface (:19-52)
and BaseModel (:55-166) keep their names, call signatures and output dicts.
BaseModel.forward() runs one step of the device engine: first half (B+A),
skin trigger and gated list rebuild, forces, second half (O+B), float64
reductions and, with a barostat, its device update (Berendsen or
anisotropic) — all stream-ordered,
without the reference's per-step host synchronisations. BaseModel.run(n)
replays a captured HIP graph of that step n times.
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
        """The single device force field a fused step can run, or None."""
        if len(self.backbones) == 1 and hasattr(self.backbones[0], "ff"):
            return self.backbones[0].ff
        return None


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
            raise NotImplementedError("the device step runs one LJ or EAM backbone")
        self.engine.force_field = ff

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

    def _step(self):
        eng = self.engine
        mol = self.molecular
        if mol.needs_update:
            mol._trigger_and_build(force=True)
        if eng.loading:
            eng.deform()
        eng.first_half(records=True)
        eng.build_and_forces(virial=True)
        eng.second_half(advance=False)
        eng.reduce(0, 15, advance=True)
        if self.barostat is not None:
            eng.barostat()
        mol._graph_cache = None

    def forward(self):
        eng = self.engine
        if self._first_call:
            eng.compute_forces(virial=True)
            eng.reduce(0, 8)
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
        out = {
            "forces": eng.to_caller(eng.force),
            "energy": th[_lib.T_EPOT],
            "virial": th[_lib.T_VIRIAL],
            "virial_tensor": eng.virial_tensor(0),
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
        """nsteps steps; with graph=True one captured step is replayed."""
        eng = self.engine
        if self._first_call:
            eng.compute_forces(virial=True)
            eng.reduce(0, 8)
            self._first_call = False
        if self.molecular.needs_update:
            self.molecular._trigger_and_build(force=True)
        eng.run(nsteps, graph=graph, virial=True)
        self.molecular._graph_cache = None
        return self.result()
