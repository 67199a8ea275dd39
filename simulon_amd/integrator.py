"""Velocity Verlet with the interface of the reference's
core/integrator/integrator.py (VerletIntegrator).

Constants and units as in the reference (:41-64): kB = 8.617333262e-5 eV/K,
internal mass m = amu * 10 / (N_A * J_per_eV) in eV ps^2 / A^2. NVT and NPT use
the Langevin O step with damping exp(-gamma dt) and kick width
sqrt(kB T / m (1 - exp(-2 gamma dt))) (:84-87, :109-112), gamma in 1/ps;
velocities are initialised from Maxwell-Boltzmann at T_init (:82).
The kicks come from Philox4x32-10 on the device, keyed by (seed, step, atom).
"""
import numpy as np
import torch
import torch.nn as nn

from . import _lib
from .engine import internal_mass_f32, langevin_f32
from .md_model import IntegratorInterface


class VerletIntegrator(IntegratorInterface, nn.Module):
    BOLTZMAN = 8.617333262e-5
    N_a = 6.02214076e23
    J_per_ev = 1.60218e-19

    def __init__(self, molecular, dt: float, ensemble: str = None, temperature=None,
                 gamma: float = None, force_field=None, seed: int = None):
        super().__init__()
        self.molecular = molecular
        self.dt = dt
        self.ensemble = (ensemble or "NVE").upper()
        eng = molecular.engine
        self.engine = eng
        if seed is not None:
            eng.seed = int(seed)
        self.atom_mass = (molecular.atom_mass.unsqueeze(-1).expand(-1, 3)
                          * (10.0 / (self.N_a * self.J_per_ev)))
        self.is_langevin_thermostat = self.ensemble in ("NVT", "NPT")
        P = eng.replica_state(0)
        P.dt = float(dt)
        P.dtf = float(np.float32(dt))
        P.langevin = int(self.is_langevin_thermostat)
        if self.is_langevin_thermostat:
            if temperature is None or gamma is None:
                raise ValueError(f"ensemble='{self.ensemble}' needs temperature and gamma")
            T_init, T_target = float(temperature[0]), float(temperature[1])
            self.temperature = torch.tensor(T_target, device=molecular.device)
            self.gamma = torch.tensor(float(gamma), device=molecular.device)
            molecular.set_maxwell_boltzmann_velocity(T_init)
            for t, m in enumerate(molecular.type_masses):
                damp, rff = langevin_f32(internal_mass_f32(m), dt, gamma, T_target)
                P.rff[t] = float(rff)
            P.damp = float(damp)
        eng.write_replica_state(0, P)

    @property
    def new_coords(self):
        return self.molecular.coordinates

    @property
    def vel_half(self):
        return self.engine.to_caller(self.engine.vel_half)

    def _load_forces(self, forces):
        if forces is not None:
            self.engine.force.copy_(self.engine.from_caller(torch.as_tensor(forces).float()))

    def first_half(self, forces_old):
        """integrator.py:91-99: v_half, x(t+dt), then the skin trigger/rebuild."""
        self._load_forces(forces_old)
        self.engine.first_half()
        self.engine.build_list(force=False)
        self.molecular._graph_cache = None

    def second_half(self, forces_new):
        """integrator.py:102-124."""
        self._load_forces(forces_new)
        self.engine.second_half()
        self.engine.reduce(_lib.T_KE, _lib.T_KE + 7)
        ke = self.engine.thermo[0, _lib.T_KE]
        return {"update_coordinates": self.new_coords, "kinetic_energy": ke,
                "temperature": (2.0 / 3.0) * ke / (self.molecular.atom_count * self.BOLTZMAN)}

    def apply_pbc(self, coordinates):
        return self.molecular.box.wrap_positions(coordinates)

    def forward(self, force):
        self.first_half(force)
        return self.second_half(force)
