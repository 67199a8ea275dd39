"""Simulation driver with the interface of the reference's core/md_simulation.py.

MDSimulator keeps the reference's constructor (md_simulation.py:30-97), its
output files and their formats (traj.xyz, energies.csv with the header
step,potential,kinetic,total,temperature, forces/forces_NNNNNN.pt;
:125-167), the per-step store buffers of runs up to 2000 steps (:57-68), the
progress line (:196-207), run()'s return dict (:240-246) and
summarize_profile() (:252-293).

The loop differs in how it reaches the device. The reference calls
model() once per step and reads every scalar back with a host
synchronisation. Here the steps between two output events are replays of one
captured HIP graph of the device step, and an event only enqueues device
copies of what it needs (positions, forces, the thermo scalars, the overflow
flags) into pinned host buffers behind an event marker. The host formats and
writes a snapshot while the next segment runs, so output overlaps compute.
Work that has to see every step on the host (an RDF accumulator, the
graph-dataset list, energies written every step) makes each step its own
segment; the store buffers are filled on the device without a host wait.
"""
from __future__ import annotations

import math
import time
import types
from collections import deque
from pathlib import Path

import torch

from . import _lib
from .energy_minimizer import minimize_energy_steepest_descent

AVOGADRO = 6.02214076e23
MAX_STORE_STEPS = 2000          # md_simulation.py:57: store buffers up to this many steps
MAX_SEGMENT = 500               # steps between two overflow checks without output


class _Snapshot:
    """Device copies of one step's output, in flight to pinned host memory."""

    def __init__(self, step_i, scalars, coords, forces, flags):
        self.step_i = step_i
        self.scalars = scalars      # [epot, ke, T, volume] float64
        self.coords = coords        # [N,3] float32, caller order, or None
        self.forces = forces        # [N,3] float32, caller order, or None
        self.flags = flags          # what to do with it
        self.done = torch.cuda.Event()
        self.done.record()


def _to_pinned(x):
    h = torch.empty(x.shape, dtype=x.dtype, pin_memory=True)
    h.copy_(x, non_blocking=True)
    return h


class MDSimulator:

    def __init__(self, model, num_steps: int, print_interval: int = 10,
                 save_to_graph_dataset=False, spread_mode: str | None = None,
                 output_dir: str | None = None, dump_interval: int = 0, dump_format: str = "xyz",
                 write_forces: bool = False, write_energies: bool = False,
                 traj_interval: int | None = None, forces_interval: int | None = None,
                 energies_interval: int | None = None, rdf_accumulator=None):
        self.model = model
        self.num_steps = int(num_steps)
        self.print_interval = int(print_interval or 0)
        self.device = model.engine.device
        self.save_to_graph_dataset = save_to_graph_dataset
        if save_to_graph_dataset:
            self.dataset = []
        self.trajectory, self.energy_list, self.force_list = [], [], []
        self.spread_mode = spread_mode
        self.rdf_accumulator = rdf_accumulator

        n = model.molecular.atom_count
        self._store_steps = min(self.num_steps, MAX_STORE_STEPS)
        self._store_buffers = self.num_steps <= MAX_STORE_STEPS
        if self._store_buffers:
            dt = torch.float16 if self.device.type == "cuda" else torch.float32
            self.gpu_trajectory = torch.empty((self._store_steps, n, 3), device=self.device, dtype=dt)
            self.gpu_forces = torch.empty((self._store_steps, n, 3), device=self.device, dtype=dt)
            self.gpu_energies = torch.empty(self._store_steps, device=self.device, dtype=dt)
        else:
            self.gpu_trajectory = self.gpu_forces = self.gpu_energies = None

        self.output_dir = Path(output_dir) if output_dir else None
        self.dump_format = (dump_format or "xyz").lower()
        self.write_forces = bool(write_forces)
        self.write_energies = bool(write_energies)
        self.dump_interval = int(dump_interval or 0)
        self.traj_interval = int(traj_interval) if traj_interval is not None else self.dump_interval
        self.forces_interval = (int(forces_interval) if forces_interval is not None
                                else self.dump_interval)
        # None: an energies row every step (md_simulation.py:80)
        self.energies_interval = int(energies_interval) if energies_interval is not None else None

        any_traj = self.traj_interval > 0
        any_forces = self.write_forces and self.forces_interval > 0
        self.output_enabled = bool(self.output_dir and (any_traj or any_forces or self.write_energies))
        self._paths = {}
        if self.output_enabled:
            self.output_dir.mkdir(parents=True, exist_ok=True)
            if any_traj and self.dump_format == "xyz":
                p = self.output_dir / "traj.xyz"
                p.write_text("", encoding="utf-8")
                self._paths["traj"] = str(p)
            if self.write_energies:
                p = self.output_dir / "energies.csv"
                p.write_text("step,potential,kinetic,total,temperature\n", encoding="utf-8")
                self._paths["energies"] = str(p)
            if any_forces:
                d = self.output_dir / "forces"
                d.mkdir(parents=True, exist_ok=True)
                self._paths["forces_dir"] = str(d)
        self.sim_profile = {"output_time": 0.0}

    # ---- set-up --------------------------------------------------------------------
    def _apply_spread(self):
        """md_simulation.py:99-123: rescale the coordinates to fill the box
        ('scale') or draw them uniformly in it ('random')."""
        if self.spread_mode not in ("scale", "random"):
            return
        mol = self.model.molecular
        L = float(mol.box_length_cpu)
        with torch.no_grad():
            x = mol.coordinates.clone()
            if self.spread_mode == "scale":
                lo = x.min(dim=0).values
                ext = (x.max(dim=0).values - lo).max().clamp(min=1e-6)
                x = (x - lo) * ((L - 1e-6) / ext)
            else:
                x = torch.rand_like(x) * L
            mol.update_coordinates(x)
        rho_n = mol.atom_count / L ** 3
        half = (4.0 / 3.0) * math.pi * float(mol.cutoff) ** 3 * rho_n
        print(f"[Spread] Mode={self.spread_mode}, edges={mol.graph_data.edge_index.shape[1]}, "
              f"expected ave neigh (half, cutoff)~{half:.4f}")

    # ---- event schedule ------------------------------------------------------------
    def _every_step(self):
        """Host work that has to see every step."""
        return (self.rdf_accumulator is not None or self.save_to_graph_dataset
                or (self.output_enabled and self.write_energies and self.energies_interval is None))

    def _flags(self, step_i):
        """What happens after step step_i (1-based)."""
        f = set()
        if self.print_interval > 0 and step_i % self.print_interval == 0:
            f.add("print")
        if self.output_enabled:
            if (self.traj_interval > 0 and self.dump_format == "xyz"
                    and step_i % self.traj_interval == 0 and "traj" in self._paths):
                f.add("traj")
            if self.write_forces and self.forces_interval > 0 and step_i % self.forces_interval == 0:
                f.add("forces")
            if self.write_energies and (self.energies_interval is None or (
                    self.energies_interval > 0 and step_i % self.energies_interval == 0)):
                f.add("energies")
        return f

    def _next_stop(self, step):
        """The step after which the loop next leaves the device: the next step
        with an output event, every step when host or store work needs it, and
        at least every MAX_SEGMENT steps (to check the list overflow flags)."""
        last = min(self.num_steps, step + MAX_SEGMENT)
        if self._store_buffers or self._every_step():
            return step + 1
        for s in range(step + 1, last + 1):
            if self._flags(s):
                return s
        return last

    # ---- device state ------------------------------------------------------------
    def _prime(self):
        """BaseModel.run's start: the first force evaluation and a pending rebuild."""
        model, eng, mol = self.model, self.model.engine, self.model.molecular
        if model._first_call:
            model._initial_forces()
            model._first_call = False
        if mol.needs_update:
            mol._trigger_and_build(force=True)

    def _advance(self, k):
        if getattr(self.model, "extra", None):
            self.model.run(k)          # host backbones: eager steps
        else:
            self.model.engine.run(k, graph=True, virial=True, check=False)
        self.model.molecular._graph_cache = None

    def _scalars(self):
        """[epot, ke, T, volume, overflow flags] of replica 0 on the device."""
        eng = self.model.engine
        th = eng.thermo[0]
        ke = th[_lib.T_KE]
        T = (2.0 / 3.0) * ke / (self.model.molecular.atom_count * self.model.Integrator.BOLTZMAN)
        H = eng.box_rows()[0]
        vol = torch.dot(H[0], torch.linalg.cross(H[1], H[2]))     # no host wait, unlike det
        return torch.stack([th[_lib.T_EPOT], ke, T, vol,
                            eng.overflow.reshape(-1)[0].to(torch.float64)])

    def _store(self, step):
        idx = step % self._store_steps
        eng = self.model.engine
        self.gpu_forces[idx] = eng.to_caller(eng.force).to(self.gpu_forces.dtype)
        self.gpu_energies[idx] = eng.thermo[0, _lib.T_EPOT].to(self.gpu_energies.dtype)
        self.gpu_trajectory[idx] = self.model.molecular.coordinates.to(self.gpu_trajectory.dtype)

    def _snapshot(self, step_i, flags):
        eng = self.model.engine
        coords = _to_pinned(self.model.molecular.coordinates) if "traj" in flags else None
        forces = _to_pinned(eng.to_caller(eng.force)) if "forces" in flags else None
        return _Snapshot(step_i, _to_pinned(self._scalars()), coords, forces, flags)

    # ---- host side -------------------------------------------------------------------
    def _finish(self, snap):
        snap.done.synchronize()
        t0 = time.perf_counter()
        try:
            epot, ke, T, vol, ov = snap.scalars.tolist()
            if int(ov) & ~2:                         # an overflow: raise with its message
                self.model.engine.check()
            s = snap.step_i
            if "print" in snap.flags:
                mass = float(self.model.molecular.atom_mass.sum())
                rho = mass / AVOGADRO / (vol * 1e-24)
                print(f"Step {s}/{self.num_steps}:Tot_E={epot + ke:.4f} ev, Pot_E = {epot:.4f} ev, "
                      f"Kin_E = {ke:.4f} ev, T = {T:.4f} K, Density = {rho: .4f} g/cm3")
            if "traj" in snap.flags:
                self._write_xyz(s, epot, snap.coords)
            if "forces" in snap.flags:
                torch.save(snap.forces.clone(), Path(self._paths["forces_dir"]) / f"forces_{s:06d}.pt")
            if "energies" in snap.flags:
                with open(self._paths["energies"], "a", encoding="utf-8") as f:
                    f.write(f"{s},{epot:.8f},{ke:.8f},{epot + ke:.8f},{T:.6f}\n")
        finally:
            self.sim_profile["output_time"] += time.perf_counter() - t0

    def _write_xyz(self, step_i, epot, coords):
        mol = self.model.molecular
        n = coords.shape[0]
        names = mol.atom_types if getattr(mol, "atom_types", None) and len(mol.atom_types) == n \
            else ["X"] * n
        lines = [f"{n}\n", f"Step {step_i}, Energy = {epot:.6f}\n"]
        lines += [f"{t} {x:.6f} {y:.6f} {z:.6f}\n" for t, (x, y, z) in zip(names, coords.tolist())]
        with open(self._paths["traj"], "a", encoding="utf-8") as f:
            f.writelines(lines)

    def _host_every_step(self, step):
        """The per-step host work of md_simulation.py:187-189 and :211-221."""
        mol = self.model.molecular
        if self.rdf_accumulator is not None:
            self.rdf_accumulator.update(step, mol)
        if self.save_to_graph_dataset:
            gd = mol.graph_data
            eng = self.model.engine
            self.dataset.append(types.SimpleNamespace(
                x=mol.element_ids.clone(), pos=gd.pos, edge_index=gd.edge_index,
                edge_attr=gd.edge_attr, forces=eng.to_caller(eng.force).clone(),
                energy=eng.thermo[0, _lib.T_EPOT].clone()))

    # ---- the loop ------------------------------------------------------------------
    def run(self, enable_minimize_energy: bool = True):
        if self.device.type == "cuda":
            print(f"Simulation will run on GPU: {torch.cuda.get_device_name(self.device)}")
        self._apply_spread()
        if enable_minimize_energy:
            print("=== Minimizing energy after spreading ===")
            minimize_energy_steepest_descent(self.model)
            print("=== Energy Minimization Completed ===\n")

        self._prime()
        pending = deque()
        torch.cuda.synchronize(self.device)
        loop_t0 = time.perf_counter()
        step = 0
        while step < self.num_steps:
            stop = self._next_stop(step)
            self._advance(stop - step)
            step = stop
            if self._store_buffers:
                self._store(step - 1)
            if self._every_step():
                self._host_every_step(step - 1)
            pending.append(self._snapshot(step, self._flags(step)))
            # the previous snapshot's copies are done once this segment is queued
            while len(pending) > 1 or (pending and pending[0].done.query()):
                self._finish(pending.popleft())
        while pending:
            self._finish(pending.popleft())
        torch.cuda.synchronize(self.device)
        loop_time = time.perf_counter() - loop_t0

        if self.rdf_accumulator is not None:
            self.rdf_accumulator.finalize()
        if self.gpu_trajectory is not None:
            self.trajectory = self.gpu_trajectory.cpu().numpy()
            self.force_list = self.gpu_forces.cpu().numpy()
            self.energy_list = self.gpu_energies.cpu().numpy()
        print(f"{self.num_steps} steps simulation finished.")
        print(f"Total simulation time: {loop_time:.2f} seconds.")
        self.sim_profile["loop_time"] = loop_time
        return {
            "trajectory": self.trajectory,
            "energies": self.energy_list,
            "forces": self.force_list,
            "output_paths": self._paths if self.output_enabled else {},
            "rdf_file": getattr(self.rdf_accumulator, "outfile", None) if self.rdf_accumulator else None,
        }

    # ---- reporting ---------------------------------------------------------------------
    def _force_time(self, reps=5):
        """Seconds of one force evaluation at the current positions (device
        events; the evaluation is repeatable, so the state is unchanged)."""
        eng = self.model.engine
        eng.compute_forces(virial=True)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            eng.compute_forces(virial=True)
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / reps * 1e-3

    def summarize_profile(self):
        """md_simulation.py:252-293. A replayed graph has no host-side phase
        boundaries, so Pair is the steps times one event-timed force
        evaluation; Neigh is the host time of rebuilds requested outside the
        graph; Output is the host time spent formatting and writing; Integr
        (integration, in-graph rebuilds and reductions) is the rest of the loop."""
        mol = self.model.molecular
        loop_time = self.sim_profile.get("loop_time", 0.0)
        steps = self.num_steps
        pair_t = min(steps * self._force_time(), loop_time)
        neigh_t = mol.profile["neighbor_time"] if loop_time else 0.0
        out_t = self.sim_profile["output_time"]
        integ_t = max(loop_time - pair_t - out_t, 0.0)
        atoms = mol.atom_count
        dt_ps = getattr(self.model.Integrator, "dt", 0.0)
        sim_ns = steps * dt_ps / 1000.0
        steps_per_s = steps / loop_time if loop_time > 0 else 0.0
        matom = steps * atoms / loop_time / 1e6 if loop_time > 0 else 0.0
        ns_day = sim_ns / (loop_time / 86400.0) if loop_time > 0 else 0.0
        total = max(loop_time, 1e-12)
        print("\nMPI task timing breakdown (internal):")
        print("Section |  time (s) | %total")
        for name, t in (("Pair", pair_t), ("Neigh", neigh_t), ("Integr", integ_t),
                        ("Output", out_t)):
            print(f"{name:<7}| {t:9.4f} | {100.0 * t / total:6.2f}")
        print(f"Loop   | {loop_time:9.4f} | 100.00")
        builds = int(self.model.engine.nbuild[0])
        print(f"Neighbor list builds: {builds}")
        eff = mol.effective_edge_count()
        H = self.model.engine.box_rows()[0]
        vol = float(torch.dot(H[0], torch.linalg.cross(H[1], H[2])))
        full = (4.0 / 3.0) * math.pi * float(mol.cutoff) ** 3 * atoms / vol
        print(f"Effective neighbors (<=cutoff): {eff}  Half-ave/atom={eff / atoms:.4f}  "
              f"Full-ave/atom={2 * eff / atoms:.4f}")
        print(f"Theoretical (uniform): full-ave/atom~{full:.4f}  half-ave/atom~{full / 2:.4f}  "
              f"expected half edges total~{int(full / 2 * atoms):d}")
        print(f"Performance: {ns_day:.3f} ns/day, {steps_per_s:.3f} steps/s, "
              f"{matom:.3f} Matom-step/s")
        self.sim_profile.update(pair_time=pair_t, integrate_time=integ_t, neighbor_time=neigh_t,
                                ns_per_day=ns_day, steps_per_s=steps_per_s,
                                matom_steps_per_s=matom, neighbor_builds=builds)
        return dict(self.sim_profile)

    # ---- post-run helpers -------------------------------------------------------------
    def save_energy_curve(self, path: str):
        """md_simulation.py:296-345: potential energy against step, from
        energies.csv when it was written, else from the store buffer."""
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
        import numpy as np

        steps, pot = [], []
        csv = self._paths.get("energies", "")
        if csv and Path(csv).exists():
            rows = Path(csv).read_text().splitlines()[1:]
            for r in rows:
                p = r.split(",")
                if len(p) >= 2:
                    steps.append(int(p[0]))
                    pot.append(float(p[1]))
        if not pot and len(self.energy_list) > 0:
            e = np.asarray(self.energy_list, dtype=np.float64).ravel()
            steps, pot = list(range(1, len(e) + 1)), e.tolist()
        if not pot:
            print("[save_energy_curve] No energy data available; skipping plot.")
            return
        fig, ax = plt.subplots(figsize=(8, 4))
        ax.plot(steps, pot, linewidth=1.0, color="steelblue")
        ax.set_xlabel("Step")
        ax.set_ylabel("Potential Energy (eV)")
        ax.set_title("MD Potential Energy")
        ax.grid(True, linestyle="--", alpha=0.5)
        fig.tight_layout()
        Path(path).resolve().parent.mkdir(parents=True, exist_ok=True)
        fig.savefig(path, dpi=150)
        plt.close(fig)
