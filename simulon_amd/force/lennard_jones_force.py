"""Lennard-Jones backbone with the interface of the reference's
core/force/lennard_jones_force.py:7-95 (and lennard_jones_force_cu.py).

Parameters come from molecular.parameter, keyed like the reference
('[0 0]' -> {epsilon, sigma}); as in the reference the first key of each entry
is used as epsilon and the second as sigma (reader.py:190-195 builds
pair_params positionally). is_switch selects the polynomial switch on
[switch_ratio*rc, rc], is_fs the force shift (:48-75)."""
import numpy as np
import torch.nn as nn

from ..forcefield import LJForceField
from ..md_model import BackboneInterface
from ..reader import parse_pair_key
from ._common import backbone_outputs


def lj_tables(parameter, ntypes):
    eps = np.zeros((ntypes, ntypes))
    sig = np.ones((ntypes, ntypes))
    for key, vals in parameter.items():
        a, b = parse_pair_key(key)
        if a < ntypes and b < ntypes:
            v = list(vals.values())
            eps[a, b], sig[a, b] = float(v[0]), float(v[1])
    return eps, sig


class LennardJonesForce(BackboneInterface, nn.Module):
    def __init__(self, molecular):
        super().__init__()
        self.molecular = molecular
        if molecular.parameter is None:
            raise ValueError("LennardJonesForce needs molecular.parameter (per type-pair epsilon, sigma)")
        eps, sig = lj_tables(molecular.parameter, len(molecular.elements))
        mode = "switch" if molecular.is_switch else ("shift" if molecular.is_fs else "raw")
        self.ff = LJForceField(eps, sig, float(molecular.cutoff), mode=mode,
                               switch_ratio=molecular.switch_ratio)

    def forward(self):
        return backbone_outputs(self.molecular, self.ff)


LennardJonesForceCUDA = LennardJonesForce
