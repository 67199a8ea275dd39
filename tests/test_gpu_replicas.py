"""The batched temperature scan at its real size: 64 W EAM replicas of
2,000-118,638 atoms (1,783,658 in all) with Langevin temperatures over
100-1500 K, packed in one engine (the reference runs them one by one,
run_scripts/w_dbtt_scan.py:97-102).

- every replica's forces and energy at setup against the float64 C oracle;
- after 20 graph-replayed steps, the forces of 8 replicas spanning the sizes
  against the oracle at their new positions;
- the smallest, a middle and the largest replica advance bit for bit as in
  single-replica engines with the same per-replica noise seed and the same
  pair-kernel family (the LDS families agree with the global list only to
  float32 rounding, so the autotuner's per-engine choice is pinned here).
"""
import os

import numpy as np
import pytest
import torch

from helpers import force_error
from oracle import cell as ocell
from oracle import eam as oeam
from simulon_amd import _lib
from simulon_amd import lattice as lat
from simulon_amd.engine import ReplicaEngine, ReplicaSpec
from simulon_amd.forcefield import EAMForceField

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5


@pytest.fixture(scope="module")
def setfl(tmp_path_factory):
    return lat.write_w_synthetic_setfl(str(tmp_path_factory.mktemp("w") / "W_synthetic.eam.alloy"))


def _spec(pos, H, T, **kw):
    return ReplicaSpec(positions=pos, box=H, dt=0.001, ensemble="NVT", temperature=T, gamma=10.0,
                       **kw)


def test_replicas64_parity_and_batched_equals_single(setfl):
    # thermal disorder: +-0.05 A jitter (a perfect lattice has zero net forces)
    rs = [(lat.jitter(pos, 0.05, 100 + k), H, T) for k, (pos, H, T) in enumerate(lat.replica_set(64))]
    eng = ReplicaEngine([_spec(*r) for r in rs], [lat.W_MASS], cutoff=5.5, skin=1.0, seed=2024)
    ff = EAMForceField(setfl, device=eng.device)
    eng.force_field = ff
    eng.setup()
    torch.cuda.synchronize()
    natoms = np.array(eng.natoms)
    print(f"[case] 64 replicas, {eng.N} atoms ({natoms.min()}..{natoms.max()}), family "
          f"{eng.family}", flush=True)
    tab = oeam.EAMTables(oeam.EAMFile(setfl), dtype=np.float64)

    def check(reps, tag):
        x_all = eng.positions().cpu().numpy()
        f_all = eng.force.cpu().numpy()
        th = eng.thermo_host()
        worst = 0.0
        for r in reps:
            sl = eng.replica_slice(r)
            ref = ocell.eam(x_all[sl], eng.box(r), None, tab, 1.0)
            e = force_error(f_all[sl], ref["forces"])
            worst = max(worst, e)
            assert e <= FORCE_TOL, f"replica {r} ({natoms[r]} atoms): force err {e:.3e}"
            if tag == "setup":
                assert th[r, _lib.T_EPOT] == pytest.approx(ref["energy"], rel=1e-6)
                assert int(eng.nnbr[sl].sum().item()) // 2 == ref["npairs"]
        print(f"{tag}: worst force err over {len(reps)} replicas {worst:.3e}", flush=True)

    check(range(eng.R), "setup")
    order = np.argsort(natoms)
    subset = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    v_caller = eng.to_caller(eng.vel).cpu().numpy()
    singles = []
    for r in subset:
        pos, H, T = rs[r]
        e1 = ReplicaEngine([_spec(pos, H, T, velocities=v_caller[eng.replica_slice(r)],
                                  seed=eng.replica_seed(r))],
                           [lat.W_MASS], cutoff=5.5, skin=1.0, family=eng.family)
        e1.force_field = ff
        e1.setup()
        singles.append(e1)
    eng.run(20, graph=True, virial=False)
    for e1 in singles:
        e1.run(20, graph=True, virial=False)
    torch.cuda.synchronize()
    eng.check()
    # forces at the positions after 20 steps (the last step's force pass)
    spread = [int(order[k]) for k in np.linspace(0, len(order) - 1, 8).astype(int)]
    eng.compute_forces(virial=True)
    check(spread, "after 20 steps")
    x_all = eng.to_caller(eng.positions(unwrap=True)).cpu().numpy()
    v_all = eng.to_caller(eng.vel).cpu().numpy()
    for r, e1 in zip(subset, singles):
        sl = eng.replica_slice(r)
        np.testing.assert_array_equal(e1.to_caller(e1.positions(unwrap=True)).cpu().numpy(),
                                      x_all[sl])
        np.testing.assert_array_equal(e1.to_caller(e1.vel).cpu().numpy(), v_all[sl])
    print(f"batched == single bit for bit on replicas {subset} ({natoms[subset].tolist()} atoms)",
          flush=True)
