"""EAM backbone with the interface of the reference's
core/force/eam_force_cu.py:316-579 (EAMForceCUDAExt, alias EAMForceCUDA) and
core/force/eam_force.py (EAMForce).

Tables are built from an EAMParser exactly as the reference builds them
(float32, n_r/n_rho uniform grids); element names of the molecular are mapped
to the file's elements. use_extension is accepted for compatibility: this
backbone always runs the HIP kernels (there is no CPU fallback)."""
import numpy as np
import torch
import torch.nn as nn

from ..eam_parser import EAMParser
from ..forcefield import EAMForceField
from ..md_model import BackboneInterface
from ._common import backbone_outputs


class EAMForceCUDAExt(BackboneInterface, nn.Module):
    def __init__(self, eam_parser, molecular, n_r: int = 8192, n_rho: int = 4096,
                 use_extension: bool = True):
        super().__init__()
        if not isinstance(eam_parser, EAMParser):
            eam_parser = EAMParser(eam_parser)
        self.parser = eam_parser
        self.molecular = molecular
        self.cutoff = float(eam_parser.cutoff)
        self.n_r, self.n_rho = n_r, n_rho
        # molecular element id -> file element index
        try:
            emap = [eam_parser.element_map[e] for e in molecular.elements]
        except KeyError as e:
            raise ValueError(f"element {e} is not in the EAM file {eam_parser.elements}")
        self.ff = EAMForceField(eam_parser, n_r, n_rho, device=molecular.device, type_map=emap)

    def forward(self):
        out = backbone_outputs(self.molecular, self.ff, keep_rho=True)
        # eam_force_cu.py:577-578: the embedding and pair parts of the energy
        emb = self.ff.embedding_energy(self.molecular.engine)[0]
        out["embedding_energy"] = emb
        out["pair_potential"] = out["energy"] - emb
        return out


EAMForceCUDA = EAMForceCUDAExt


class EAMForce(EAMForceCUDAExt):
    """core/force/eam_force.py:EAMForce(eam_parser, molecular): same physics
    (the diagonal FS density equals the host-neighbor table for one element)."""

    def __init__(self, eam_parser, molecular, n_r: int = 8192, n_rho: int = 4096, **kw):
        super().__init__(eam_parser, molecular, n_r=n_r, n_rho=n_rho)
