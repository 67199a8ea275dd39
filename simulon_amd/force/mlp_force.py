"""Learned-potential backbone with the return contract of the reference's
machine_learning_potentials/machine_learning_force.py:85-142
(MachineLearningForce.forward -> {'energy', 'forces', 'virial'}), plus the
'virial_tensor' the LJ/EAM backbones return. The network is this library's
radial-descriptor MLP (simulon_amd.mlp), evaluated by the HIP kernels on the
molecular's engine; the molecular's cutoff must reach the descriptor cutoff.
"""
import torch.nn as nn

from ..md_model import BackboneInterface
from ..mlp import MLPForceField, MLPNet
from ._common import backbone_outputs


class MLPForce(BackboneInterface, nn.Module):
    def __init__(self, molecular, net):
        """net: an MLPNet, or the path of one saved with MLPNet.save."""
        super().__init__()
        if not isinstance(net, MLPNet):
            net = MLPNet.load(net, device=molecular.device)
        if float(molecular.cutoff) < net.rc:
            raise ValueError(f"molecular cutoff {float(molecular.cutoff)} is below the "
                             f"descriptor cutoff {net.rc}")
        self.molecular = molecular
        self.net = net
        self.ff = MLPForceField(net)

    def forward(self):
        return backbone_outputs(self.molecular, self.ff)
