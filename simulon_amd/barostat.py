"""Berendsen barostat with the interface of the reference's core/barostat.py:21-75.

P = tr(K + W) / (3V) in bar (1 eV/A^3 = 1,602,176.6208 bar), mu^3 = 1 +
kappa (dt / tau_p)(P - P0) clamped to [(1 - mu_max)^3, (1 + mu_max)^3]; H, x
and v are scaled by mu. Inside BaseModel the update runs on the device after
the second half with K and W from the step's float64 reductions; step() is
the standalone call with explicit tensors and returns P as a float, as the
reference does.
"""
import torch

from . import _lib

EV_A3_TO_BAR = 1_602_176.6208


class BerendsenBarostat:
    def __init__(self, molecular, target_pressure: float, tau_p: float,
                 compressibility: float = 4.57e-5, mu_max: float = 0.05):
        self.molecular = molecular
        self.target_P = float(target_pressure)
        self.tau_p = float(tau_p)
        self.kappa = float(compressibility)
        self.mu_max = float(mu_max)
        eng = molecular.engine
        P = eng.replica_state(0)
        P.barostat = 1
        P.target_p, P.tau_p, P.kappa, P.mu_max = self.target_P, self.tau_p, self.kappa, self.mu_max
        eng.write_replica_state(0, P)
        eng.specs[0].ensemble = "NPT"

    def step(self, dt, kinetic_energy, virial, kinetic_tensor=None, virial_tensor=None) -> float:
        eng = self.molecular.engine
        P = eng.replica_state(0)
        if abs(P.dt - float(dt)) > 0:
            P.dt = float(dt)
            eng.write_replica_state(0, P)
        th = eng.thermo[0]
        if kinetic_tensor is not None and virial_tensor is not None:
            K = torch.as_tensor(kinetic_tensor, dtype=torch.float64, device=eng.device)
            W = torch.as_tensor(virial_tensor, dtype=torch.float64, device=eng.device)
            idx = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
            for k, (a, b) in enumerate(idx):
                th[_lib.T_K + k] = K[a, b]
                th[_lib.T_W + k] = W[a, b]
        else:
            # (2 KE + virial) / (3V): put 2KE and the scalar virial on the diagonal
            ke = float(kinetic_energy)
            vir = float(virial)
            for k in range(3):
                th[_lib.T_K + k] = 2.0 * ke / 3.0
                th[_lib.T_W + k] = vir / 3.0
            for k in range(3, 6):
                th[_lib.T_K + k] = 0.0
                th[_lib.T_W + k] = 0.0
        eng.barostat()
        self.molecular._graph_cache = None
        return float(eng.replica_state(0).press)
