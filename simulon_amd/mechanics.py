"""Uniaxial tensile loading with the interface of the reference's
core/mechanics/loading.py (UniaxialTensileLoader, :6-81).

Each loading step scales the lattice row `axis` by 1 + strain_rate dt and
the two lateral rows by 1 ("fixed") or by max(1 - poisson_ratio strain_rate
dt, 1 - max_lateral_scale_step) ("poisson"), carrying the atoms with the box
(fractional coordinates unchanged). Velocities are not scaled.

step(dt) is the reference's call, made before every model() call: it runs
the row scaling on the device at once and returns the engineering strain of
the loading axis (one host read). attach(model) instead makes the scaling the
first launch of every model step, so model.run(n) replays n loading steps
inside the captured graph with no host work between them; strain() reads the
strain reached.

The reference wraps the carried coordinates into the box and resets the
rebuild reference to them (loading.py:66, :74), so its list is never rebuilt
for the deformation. Here positions stay continuous and the rebuild trigger
bounds the deformation since the last build: separations can have shrunk by
at most the smallest singular value of Hb^-1 H, which enters the skin test.
"""
from __future__ import annotations

import numpy as np
import torch


class UniaxialTensileLoader:
    def __init__(self, molecular, axis: int = 0, strain_rate: float = 1.0e-3,
                 lateral_mode: str = "fixed", poisson_ratio: float = 0.28,
                 max_lateral_scale_step: float = 2.0e-3):
        self.molecular = molecular
        self.axis = int(axis)
        self.strain_rate = float(strain_rate)
        self.lateral_mode = str(lateral_mode).lower()
        self.poisson_ratio = float(poisson_ratio)
        self.max_lateral_scale_step = float(max_lateral_scale_step)
        if self.axis not in (0, 1, 2):
            raise ValueError(f"axis must be 0/1/2, got {axis}")
        if not hasattr(molecular, "box"):
            raise NotImplementedError("UniaxialTensileLoader requires molecular.box")
        if self.lateral_mode not in ("fixed", "poisson"):
            raise ValueError(f"unsupported lateral_mode={lateral_mode}")
        self.initial_lengths = molecular.box.lengths.detach().clone().to(torch.float64)
        self.total_engineering_strain = 0.0
        self._dt = None
        self._model = None

    def factors(self, dt: float) -> np.ndarray:
        """Row scale factors of one loading step (loading.py:47-55)."""
        d = self.strain_rate * float(dt)
        s = np.ones(3)
        s[self.axis] = 1.0 + d
        if self.lateral_mode == "poisson":
            lat = max(1.0 - self.poisson_ratio * d, 1.0 - self.max_lateral_scale_step)
            for a in range(3):
                if a != self.axis:
                    s[a] = lat
        return s

    def _arm(self, dt):
        """Store the factors for dt in the replica record (once per dt)."""
        if self._dt == float(dt):
            return
        eng = self.molecular.engine
        P = eng.replica_state(0)
        for a, v in enumerate(self.factors(dt)):
            P.deform[a] = float(v)
        P.loading = 1
        eng.write_replica_state(0, P)
        self._dt = float(dt)
        eng._graph = None

    def _sync_host(self):
        mol = self.molecular
        lengths = self.current_lengths()
        mol.box_length_cpu = float(lengths[0])
        if torch.is_tensor(getattr(mol, "box_length", None)):
            mol.box_length = lengths[0].to(device=mol.device, dtype=mol.box_length.dtype)
        self.total_engineering_strain = (float(lengths[self.axis])
                                         / float(self.initial_lengths[self.axis]) - 1.0)
        return self.total_engineering_strain

    def step(self, dt: float) -> float:
        if self._model is not None:
            raise RuntimeError("this loader is attached: every model step already loads")
        self._arm(dt)
        self.molecular.engine.deform()
        self.molecular._graph_cache = None
        return self._sync_host()

    def current_lengths(self) -> torch.Tensor:
        return self.molecular.box.lengths.detach().clone()

    # ---- loading inside the device step ------------------------------------------------
    def attach(self, model, dt: float | None = None):
        """Load at the start of every step of model (forward and run)."""
        self._arm(model.Integrator.dt if dt is None else dt)
        model.engine.loading = True
        model.engine._graph = None
        self._model = model

    def detach(self):
        if self._model is not None:
            self._model.engine.loading = False
            self._model.engine._graph = None
            self._model = None

    def strain(self) -> float:
        """Engineering strain of the loading axis now (one host read)."""
        return self._sync_host()
