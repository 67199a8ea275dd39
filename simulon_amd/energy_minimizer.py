"""Energy minimisation with the interface of the reference's core/energy_minimizer.py.

Both minimisers keep every coordinate, force and trial position on the device
and evaluate the force field with the engine's kernels (one launch sequence
per evaluation, the neighbour list kept valid by the skin trigger). Only the
energy of a trial and the largest atomic force there cross to the host, in
one read, where the line-search decisions of the reference are taken:

- minimize_energy_steepest_descent (energy_minimizer.py:52-133): per-atom unit
  force directions, trial steps of step_size halved up to max_backtracks times,
  each tried along +F then -F, accepted when the energy drops; stops when the
  largest force is below force_threshold or no descending step is found. The
  accepted trial's forces start the next step, so a step whose first trial
  descends costs one evaluation and one host read.
- minimize_energy_bfgs_scipy (energy_minimizer.py:4-49): torch.optim.LBFGS on
  the coordinates with the gradient -F from the kernels (the reference
  differentiates its torch energy).

Positions move in float64 (the reference moves float32 coordinates).
"""
import torch

from . import _lib


def _energy(model):
    """Force-field energy (device scalar) and forces at the current positions."""
    eng = model.engine
    eng.compute_forces(virial=False)
    eng.reduce(_lib.T_EPOT, _lib.T_EPOT + 1)
    return eng.thermo[0, _lib.T_EPOT]


def _coords(eng):
    """Continuous float64 coordinates in the caller's atom order (stable across
    the reorders a rebuild makes)."""
    return eng.to_caller(eng.positions(torch.float64, unwrap=True))


def _place(model, x):
    """Move the atoms to caller-order coordinates x, then the skin trigger and
    the gated rebuild, all stream-ordered on the device (no host read; list
    overflow is checked once at the end of a minimisation)."""
    eng = model.engine
    eng.set_positions(eng.from_caller(x))
    d = eng.pos4[:, :3] - eng.pos_last[:, :3]
    eng.maxd2.copy_((d * d).sum(1).max().reshape(1))     # one replica
    eng.build_list(force=False)
    model.molecular._graph_cache = None


def _stats(model, e_dev):
    """(energy, largest atomic force) as host floats: one device-to-host read."""
    fmax = model.engine.force.norm(dim=1).max().double()
    e, f = torch.stack([e_dev.double(), fmax]).cpu().tolist()
    model.minimize_reads = getattr(model, "minimize_reads", 0) + 1
    return e, f


def minimize_energy_steepest_descent(model, max_steps=10000, step_size=0.002,
                                     force_threshold=1e-2, print_interval=10, max_backtracks=12,
                                     min_step=1e-6, verbose=False):
    """The reference's backtracking steepest descent (energy_minimizer.py:52-133)
    with the trial sequence +s, -s, +s/2, -s/2, ... and the acceptance test
    E_trial < E_base. The forces of an accepted trial are those of the next
    step's base point, so an accepted step costs one force evaluation and one
    host read (its energy and largest force together); only backtracking
    trials add evaluations and reads."""
    eng = model.engine
    mol = model.molecular
    mol._trigger_and_build(force=mol.needs_update)
    history = []
    model.minimize_reads = 0
    e_base, max_force = _stats(model, _energy(model))
    for step in range(max_steps):
        history.append((e_base, max_force))
        if verbose and ((step + 1) % print_interval == 0 or step == 0):
            print(f"Step {step + 1}/{max_steps} -> Energy: {e_base:.4f} eV, "
                  f"Max Force: {max_force:.6f} eV/Ang")
        if max_force < force_threshold:
            break
        fn = eng.force.norm(dim=1).clamp_min(1e-12)
        f_dir = eng.to_caller((eng.force / fn[:, None]).double())
        x0 = _coords(eng)
        step_try = float(step_size)
        accepted = False
        for _ in range(max_backtracks):
            for sign in (1.0, -1.0):
                _place(model, x0 + (sign * step_try) * f_dir)
                e_t, f_t = _stats(model, _energy(model))
                if e_t < e_base:
                    accepted = True
                    e_base, max_force = e_t, f_t
                    break
            if accepted:
                break
            step_try *= 0.5
            if step_try < min_step:
                break
        if not accepted:                  # no descending step: back to the last point
            _place(model, x0)
            _energy(model)
            break
    final_energy = float(_energy(model))
    eng.check()
    model.minimize_history = history
    if verbose:
        print(f"Final Energy: {final_energy:.4f} eV, Final Max Force: "
              f"{float(eng.force.norm(dim=1).max()):.6f} eV/Ang")
    return final_energy


def minimize_energy_bfgs_scipy(model, max_iter=20, lr=0.5, history_size=10,
                               line_search_fn="strong_wolfe"):
    eng = model.engine
    mol = model.molecular
    mol._trigger_and_build(force=mol.needs_update)
    x = torch.nn.Parameter(_coords(eng).clone())
    opt = torch.optim.LBFGS([x], lr=lr, max_iter=max_iter, history_size=history_size,
                            line_search_fn=line_search_fn)

    def closure():
        opt.zero_grad()
        _place(model, x.detach())
        e = _energy(model)
        x.grad = -eng.to_caller(eng.force).double()       # dE/dx
        return e.detach().double().clone()

    opt.step(closure)
    _place(model, x.detach())
    e = float(_energy(model))
    eng.check()
    return e
