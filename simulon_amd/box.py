"""Periodic box with the interface of the reference's core/box.py.

H holds the lattice vectors as rows (float64). minimum_image/wrap_positions
follow core/box.py:109-142 (orthogonal: r - d round(r/d); triclinic:
fractional rounding through H^-1), scale/scale_anisotropic :145-153.

A box attached to a device engine (AtomFileReader does this) keeps H inside
the engine's per-replica record, so a Berendsen step on the GPU is visible
through `box.H` without any copy, and assigning `box.H` updates the record.
"""
from __future__ import annotations

from typing import Union

import torch

_OFF_H, _OFF_HINV, _OFF_VOL = 0, 72, 144


class Box:
    def __init__(self, h: Union[float, list, torch.Tensor], device: torch.device = None):
        if device is None:
            device = torch.device("cuda" if torch.cuda.is_available() else "cpu")
        self.device = torch.device(device)
        t = torch.as_tensor(h, dtype=torch.float64)
        if t.dim() == 0 or t.numel() == 1:
            H = torch.eye(3, dtype=torch.float64) * float(t)
        elif t.dim() == 1 and t.numel() == 3:
            H = torch.diag(t)
        elif tuple(t.shape) == (3, 3):
            H = t.clone()
        else:
            raise ValueError(f"h must be a scalar, [3] or [3,3]; got shape {tuple(t.shape)}")
        self._rec = None              # engine record (uint8 tensor) when attached
        self._listeners = []          # called after H changes (e.g. to force a list rebuild)
        self._H = H.to(self.device)
        self._refresh()

    # ---- derived quantities -----------------------------------------------------
    def _refresh(self):
        self._H_inv = torch.linalg.inv(self._H)
        off = self._H - torch.diag(torch.diag(self._H))
        self._is_orthogonal = bool(torch.all(off.abs() <= 1e-8))
        if self._rec is not None:
            from ._lib import Replica
            self._rec[_OFF_HINV:_OFF_HINV + 72].view(torch.float64).copy_(self._H_inv.reshape(9))
            self._rec[_OFF_VOL:_OFF_VOL + 8].view(torch.float64).fill_(self.volume)
            o = Replica.ortho.offset
            self._rec[o:o + 4].view(torch.int32).fill_(int(self._is_orthogonal))
        for fn in self._listeners:
            fn(self)

    def attach(self, record):
        """Share H with an engine record (sim_replica_t bytes on the device)."""
        record[_OFF_H:_OFF_H + 72].view(torch.float64).copy_(self._H.reshape(9).to(record.device))
        self._rec = record
        self.device = record.device
        self._H = record[_OFF_H:_OFF_H + 72].view(torch.float64).view(3, 3)
        self._refresh()

    def _pull(self):
        # the engine may have scaled H (Berendsen); H^-1 is kept in the record
        if self._rec is not None:
            self._H_inv = self._rec[_OFF_HINV:_OFF_HINV + 72].view(torch.float64).view(3, 3)

    @property
    def H(self) -> torch.Tensor:
        return self._H

    @H.setter
    def H(self, value: torch.Tensor):
        v = torch.as_tensor(value, dtype=torch.float64).to(self.device)
        if self._rec is not None:
            self._H.copy_(v)
        else:
            self._H = v.clone()
        self._refresh()

    @property
    def H_inv(self) -> torch.Tensor:
        self._pull()
        return self._H_inv

    @property
    def is_orthogonal(self) -> bool:
        return self._is_orthogonal

    @property
    def volume(self) -> float:
        return float(torch.det(self._H).abs())

    @property
    def lengths(self) -> torch.Tensor:
        return torch.norm(self._H, dim=1)

    @property
    def diag(self) -> torch.Tensor:
        return torch.diag(self._H)

    @property
    def box_length(self) -> torch.Tensor:
        d = torch.diag(self._H)
        if not self._is_orthogonal or not torch.allclose(d, d[0].expand_as(d), rtol=1e-5):
            raise ValueError("box_length is defined for cubic boxes only; use box.diag or box.H")
        return d[0].to(torch.float32)

    @property
    def box_length_cpu(self) -> float:
        return float(self.box_length)

    # ---- operations ------------------------------------------------------------------
    def minimum_image(self, rij: torch.Tensor) -> torch.Tensor:
        self._pull()
        if self._is_orthogonal:
            d = torch.diag(self._H).to(rij)
            return rij - d * torch.round(rij / d)
        s = rij @ self._H_inv.to(rij)
        s = s - torch.round(s)
        return s @ self._H.to(rij)

    def wrap_positions(self, pos: torch.Tensor) -> torch.Tensor:
        self._pull()
        if self._is_orthogonal:
            d = torch.diag(self._H).to(pos)
            return pos - torch.floor(pos / d) * d
        s = pos @ self._H_inv.to(pos)
        s = s - torch.floor(s)
        return s @ self._H.to(pos)

    def scale(self, factor: float):
        self.H = self._H * factor

    def scale_anisotropic(self, mu: torch.Tensor):
        self.H = mu.double().to(self.device) @ self._H

    def to(self, device) -> "Box":
        if self._rec is None:
            self.device = torch.device(device)
            self._H = self._H.to(self.device)
            self._H_inv = self._H_inv.to(self.device)
        return self

    def state_dict(self) -> dict:
        return {"H": self._H.detach().cpu().clone()}

    @classmethod
    def from_state_dict(cls, d: dict, device=None) -> "Box":
        return cls(d["H"], device=device or torch.device("cpu"))

    def __repr__(self) -> str:
        d = torch.diag(self._H)
        kind = ("cubic" if self._is_orthogonal and torch.allclose(d, d[0].expand(3), rtol=1e-4)
                else ("orthogonal" if self._is_orthogonal else "triclinic"))
        return f"Box({kind}, H=\n{self._H.cpu().numpy()})"
