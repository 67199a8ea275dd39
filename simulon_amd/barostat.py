"""Barostats with the interface of the reference's core/barostat.py.

BerendsenBarostat (barostat.py:21-75): P = tr(K + W) / (3V) in bar
(1 eV/A^3 = 1,602,176.6208 bar), mu^3 = 1 + kappa (dt / tau_p)(P - P0)
clamped to [(1 - mu_max)^3, (1 + mu_max)^3]; H, x and v are scaled by mu.

AnisotropicNPTBarostat (barostat.py:78-208): the pressure tensor (K + W)/V
projected on the unit lattice rows, d eta/dt = kappa (P_a - P0_a) / tau_p^2 -
gamma_p eta with a deadband of pressure_tolerance_bar (inside it the rate
decays with friction max(gamma_p, 5) and is zeroed below 1e-7), eta clamped
to +-max_eta_dot and zero on uncontrolled axes; lattice rows scaled by
exp(dt eta), atoms carried with the box, velocities scaled by exp(-dt eta)
per Cartesian axis.

Inside BaseModel both run on the device after the second half, with K and W
from the step's float64 reductions (no host synchronisation; the controller
state eta_dot lives in the replica record). step() is the standalone call
with explicit tensors. A barostat is active on the device from its
construction and whenever it is assigned to model.barostat; assigning another
barostat (or None) switches the device controller.
"""
import torch

from . import _lib

EV_A3_TO_BAR = 1_602_176.6208
KB_EV_K = 8.617333262e-5


def _load_tensors(eng, kinetic_energy, virial, kinetic_tensor, virial_tensor):
    """Put K and W (or 2KE and the scalar virial on the diagonal) into the
    thermo row the device barostats read."""
    th = eng.thermo[0]
    if kinetic_tensor is not None and virial_tensor is not None:
        K = torch.as_tensor(kinetic_tensor, dtype=torch.float64, device=eng.device)
        W = torch.as_tensor(virial_tensor, dtype=torch.float64, device=eng.device)
        idx = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
        for k, (a, b) in enumerate(idx):
            th[_lib.T_K + k] = K[a, b]
            th[_lib.T_W + k] = W[a, b]
    else:
        ke = float(kinetic_energy)
        vir = float(virial)
        for k in range(3):
            th[_lib.T_K + k] = 2.0 * ke / 3.0
            th[_lib.T_W + k] = vir / 3.0
        for k in range(3, 6):
            th[_lib.T_K + k] = 0.0
            th[_lib.T_W + k] = 0.0


def _baro_dt(eng, dt):
    """Time step of the next barostat update of replica 0 (0: the
    integrator's). Only the barostat reads this field, so a standalone
    step(dt) never changes the integrator's time step."""
    o = _lib.Replica.baro_dt.offset
    eng.reps[o:o + 8].view(torch.float64).fill_(float(dt))


class _DeviceBarostat:
    """Activation of a barostat's controller in replica 0's record."""
    kind = None
    mode = 0

    def _write(self, P):
        raise NotImplementedError

    def _activate(self):
        """Write this controller's parameters and state into the record; the
        controller active until now keeps its state on the host."""
        eng = self.molecular.engine
        P = eng.replica_state(0)
        prev = getattr(eng, "_active_barostat", None)
        if prev is not None and P.barostat == prev.mode:
            prev._save(P)
        P.barostat = self.mode
        self._write(P)
        eng.write_replica_state(0, P)
        eng.specs[0].ensemble = "NPT"
        eng.specs[0].barostat = self.kind
        eng._active_barostat = self
        eng._graph = None

    def _deactivate(self):
        eng = self.molecular.engine
        if getattr(eng, "_active_barostat", None) is not self:
            return
        P = eng.replica_state(0)
        self._save(P)
        P.barostat = 0
        eng.write_replica_state(0, P)
        eng.specs[0].ensemble = "NVT" if P.langevin else "NVE"
        eng._active_barostat = None
        eng._graph = None

    def _save(self, P):
        pass

    def _ensure_active(self):
        if getattr(self.molecular.engine, "_active_barostat", None) is not self:
            self._activate()


class BerendsenBarostat(_DeviceBarostat):
    kind, mode = "berendsen", 1

    def __init__(self, molecular, target_pressure: float, tau_p: float,
                 compressibility: float = 4.57e-5, mu_max: float = 0.05):
        self.molecular = molecular
        self.target_P = float(target_pressure)
        self.tau_p = float(tau_p)
        self.kappa = float(compressibility)
        self.mu_max = float(mu_max)
        self._activate()

    def _write(self, P):
        P.target_p, P.tau_p, P.kappa, P.mu_max = self.target_P, self.tau_p, self.kappa, self.mu_max

    def step(self, dt, kinetic_energy, virial, kinetic_tensor=None, virial_tensor=None) -> float:
        eng = self.molecular.engine
        self._ensure_active()
        _baro_dt(eng, dt)
        _load_tensors(eng, kinetic_energy, virial, kinetic_tensor, virial_tensor)
        eng.barostat()
        _baro_dt(eng, 0.0)
        self.molecular._graph_cache = None
        return float(eng.replica_state(0).press)


class AnisotropicNPTBarostat(_DeviceBarostat):
    kind, mode = "anisotropic", 2

    def __init__(self, molecular, target_pressure_bar, temperature_k: float, tau_p: float = 0.2,
                 gamma_p: float = 2.0, control_axes=(True, True, True), max_eta_dot: float = 0.05,
                 compressibility_bar_inv: float = 3.2e-7, pressure_tolerance_bar: float = 25.0,
                 stochastic: bool = False):
        self.molecular = molecular
        self.device = molecular.device
        t = torch.as_tensor(target_pressure_bar, dtype=torch.float64).flatten()
        if t.numel() == 1:
            t = t.repeat(3)
        if t.numel() != 3:
            raise ValueError("target_pressure_bar must be scalar or length-3")
        self.target_pressure_bar = t.to(self.device)
        ca = torch.as_tensor(control_axes, dtype=torch.bool).flatten()
        if ca.numel() != 3:
            raise ValueError("control_axes must have length 3")
        self.control_axes = ca.to(self.device)
        self.temperature_k = float(temperature_k)
        self.tau_p = float(tau_p)
        self.gamma_p = float(gamma_p)
        self.max_eta_dot = float(max_eta_dot)
        self.compressibility_bar_inv = float(compressibility_bar_inv)
        self.compressibility_ev_ang3_inv = self.compressibility_bar_inv * EV_A3_TO_BAR
        self.pressure_tolerance_bar = float(pressure_tolerance_bar)
        self.stochastic = bool(stochastic)
        self._eta = [0.0, 0.0, 0.0]
        self._activate()

    def _write(self, P):
        tgt = self.target_pressure_bar.tolist()
        for a in range(3):
            P.p_target3[a] = tgt[a]
            P.eta_dot[a] = self._eta[a]
        P.tau_p = self.tau_p
        P.kappa = self.compressibility_bar_inv
        P.gamma_p = self.gamma_p
        P.max_eta_dot = self.max_eta_dot
        P.p_tol = self.pressure_tolerance_bar
        P.temp_k = self.temperature_k
        P.axes = sum(1 << a for a, c in enumerate(self.control_axes.tolist()) if c)
        P.stochastic = int(self.stochastic)

    def _save(self, P):
        self._eta = [float(P.eta_dot[a]) for a in range(3)]

    def _field(self, name):
        o = getattr(_lib.Replica, name).offset
        return self.molecular.engine.reps[o:o + 24].view(torch.float64)

    @property
    def eta_dot(self) -> torch.Tensor:
        """Current strain rates (a view of the device record while active)."""
        if getattr(self.molecular.engine, "_active_barostat", None) is self:
            return self._field("eta_dot")
        return torch.tensor(self._eta, dtype=torch.float64, device=self.device)

    @property
    def pressure_axis_bar(self) -> torch.Tensor:
        """The last pressure projected on the lattice rows, bar (device view)."""
        return self._field("p_axis")

    def step(self, dt, kinetic_energy, virial, kinetic_tensor=None, virial_tensor=None):
        if kinetic_tensor is None or virial_tensor is None:
            raise ValueError("AnisotropicNPTBarostat requires kinetic_tensor and virial_tensor")
        eng = self.molecular.engine
        self._ensure_active()
        _baro_dt(eng, dt)
        _load_tensors(eng, kinetic_energy, virial, kinetic_tensor, virial_tensor)
        eng.barostat()
        _baro_dt(eng, 0.0)
        self.molecular._graph_cache = None
        return self.pressure_axis_bar.to(torch.float32)
