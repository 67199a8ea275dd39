"""Checkpoint / restart with the interface of the reference's io_utils/restart.py.

save_checkpoint(model, step, path) writes the reference's keys
(restart.py:32-74): step, coordinates, velocities, box_H, rng_cpu, rng_cuda,
vel_half, force_cache, energy_cache, virial_cache, all in the caller's atom
order. Under the extra key "simulon_amd" it adds what the device step needs to
continue the same trajectory: the float64 coordinates, the Langevin noise seed,
the replica's Philox key and the device step counter (the noise of an atom is keyed by its caller index and
that counter, so it continues unchanged), the virial tensor, the last
pressure, the anisotropic barostat's strain rates and the reduced thermo row.

load_checkpoint(model, path) restores them, forces a list rebuild at the next
step as the reference does (restart.py:111-118), skips the first force
evaluation (the forces come from the file, restart.py:128-136) and returns
step + 1. Files are read with torch.load(weights_only=True): a checkpoint
holds tensors and numbers only.
"""
from __future__ import annotations

from pathlib import Path

import torch

from . import _lib

_KEY = "simulon_amd"


def save_checkpoint(model, step: int, path: str | Path) -> None:
    mol = model.molecular
    eng = model.engine
    th = eng.thermo[0]
    o = _lib.Replica.press.offset
    oe = _lib.Replica.eta_dot.offset
    ckpt = {
        "step": int(step),
        "coordinates": mol.coordinates.detach().cpu(),
        "velocities": mol.atom_velocities.detach().cpu(),
        "box_H": mol.box.H.detach().cpu().clone(),
        "rng_cpu": torch.get_rng_state(),
        "vel_half": eng.to_caller(eng.vel_half).cpu(),
        "force_cache": eng.to_caller(eng.force).cpu(),
        "energy_cache": th[_lib.T_EPOT].detach().cpu().clone(),
        "virial_cache": th[_lib.T_VIRIAL].detach().cpu().clone(),
        _KEY: {
            "coordinates_f64": eng.to_caller(eng.positions(torch.float64, unwrap=True)).cpu(),
            "seed": int(eng.seed),
            "noise_key": eng.noise_keys().cpu().clone(),
            "step_counter": eng.step_ctr.detach().cpu().clone(),
            "thermo": th.detach().cpu().clone(),
            "pressure": eng.reps[o:o + 8].view(torch.float64).detach().cpu().clone(),
            "eta_dot": eng.reps[oe:oe + 24].view(torch.float64).detach().cpu().clone(),
        },
    }
    if torch.cuda.is_available():
        ckpt["rng_cuda"] = torch.cuda.get_rng_state(eng.device)
    torch.save(ckpt, path)


def load_checkpoint(model, path: str | Path) -> int:
    ckpt = torch.load(path, map_location="cpu", weights_only=True)
    mol = model.molecular
    eng = model.engine
    dev = eng.device
    if ckpt["coordinates"].shape[0] != eng.N:
        raise ValueError(f"checkpoint has {ckpt['coordinates'].shape[0]} atoms, the model {eng.N}")
    ext = ckpt.get(_KEY, {})
    if "box_H" in ckpt:
        mol.box.H = ckpt["box_H"].to(device=dev, dtype=torch.float64)
    elif "box_length" in ckpt:
        # scalar-box checkpoints (io_utils/restart.py:59-60): a cubic box
        bl = ckpt["box_length"]
        L = float(bl.reshape(-1)[0].double()) if torch.is_tensor(bl) else float(bl)
        mol.box.H = torch.eye(3, dtype=torch.float64, device=dev) * L
    else:
        raise KeyError("checkpoint holds neither 'box_H' nor 'box_length'")
    x = ext.get("coordinates_f64", ckpt["coordinates"])
    mol.coordinates = x.to(dev)
    mol.atom_velocities = ckpt["velocities"].to(dev)
    if "vel_half" in ckpt:
        eng.vel_half.copy_(eng.from_caller(ckpt["vel_half"].to(dev)))
    if "force_cache" in ckpt:
        eng.force.copy_(eng.from_caller(ckpt["force_cache"].to(dev)))
    if "thermo" in ext:
        eng.thermo[0].copy_(ext["thermo"].to(dev))
    else:
        if "energy_cache" in ckpt:
            eng.thermo[0, _lib.T_EPOT] = float(ckpt["energy_cache"])
        if "virial_cache" in ckpt:
            eng.thermo[0, _lib.T_VIRIAL] = float(ckpt["virial_cache"])
    if "seed" in ext:
        eng.seed = int(ext["seed"])
    if "noise_key" in ext:
        o = _lib.Replica.noise_key.offset
        eng.reps[o:o + 8].view(torch.int64).copy_(ext["noise_key"][:1].to(dev))
    if "step_counter" in ext:
        eng.step_ctr.copy_(ext["step_counter"].to(dev))
    if "pressure" in ext:
        o = _lib.Replica.press.offset
        eng.reps[o:o + 8].view(torch.float64).copy_(ext["pressure"].to(dev))
    if "eta_dot" in ext:                  # anisotropic barostat strain rates
        oe = _lib.Replica.eta_dot.offset
        eng.reps[oe:oe + 24].view(torch.float64).copy_(ext["eta_dot"].to(dev))
        b = getattr(eng, "_active_barostat", None)
        if b is not None and hasattr(b, "_eta"):
            b._eta = ext["eta_dot"].tolist()
    if "rng_cpu" in ckpt:
        torch.set_rng_state(ckpt["rng_cpu"])
    if "rng_cuda" in ckpt and torch.cuda.is_available():
        torch.cuda.set_rng_state(ckpt["rng_cuda"], dev)
    # rebuild the list at the next step; the forces of the file stand for this one
    mol.needs_update = True
    mol._graph_cache = None
    eng._graph = None              # the seed is baked into a captured step
    model._first_call = False
    return int(ckpt["step"]) + 1
