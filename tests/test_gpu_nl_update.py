"""sim_nl_update (the skin trigger and the gated, reordering list build in
fewer launches) against sim_nl_trigger followed by sim_nl_build: two engines
advanced alike, the same displacement maxima (some replicas past the skin
test, some not, then a forced build of all), and every array the build writes
compared bit for bit: flags, build counts, maxima, replica records, the
reordered atoms with their velocities, forces, caller order and images, the
cells and the lists.
"""
import ctypes

import pytest
import torch

from simulon_amd import lattice as lat
from simulon_amd._lib import call, ptr, stream_ptr
from simulon_amd.engine import ReplicaEngine, ReplicaSpec
from simulon_amd.forcefield import EAMForceField

pytestmark = pytest.mark.gpu

SIDES = (9, 12, 7, 10)
COMPARED = ("flag", "nbuild", "maxd2", "reps", "pos4", "pos_last", "vel", "vel_half", "force",
            "order", "image", "slot_of", "posq", "atom_cell", "cell_count", "cell_start",
            "cell_atoms", "nbr", "nnbr", "overflow")


@pytest.fixture(scope="module")
def setfl(tmp_path_factory):
    return lat.write_w_synthetic_setfl(str(tmp_path_factory.mktemp("w") / "W_synthetic.eam.alloy"))


def _engine(setfl):
    specs = []
    for k, n in enumerate(SIDES):
        pos, H = lat.w_bcc(n, seed=70 + k, amp=0.05)
        specs.append(ReplicaSpec(positions=pos, box=H, dt=0.002, ensemble="NVT",
                                 temperature=300.0 + 400 * k, gamma=5.0))
    eng = ReplicaEngine(specs, [lat.W_MASS], cutoff=5.5, skin=1.0, seed=11, family="global",
                        stage=False)
    eng.force_field = EAMForceField(setfl, device=eng.device)
    eng.setup()
    eng.run(12, graph=False)
    return eng


def _old_build(eng, force_all):
    s = stream_ptr()
    call("sim_nl_trigger", eng.R, ptr(eng.reps), ptr(eng.maxd2), ptr(eng.flag), ptr(eng.nbuild),
         int(force_all), s)
    call("sim_nl_build", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles, eng.N, ptr(eng.pos4),
         ptr(eng.pos_last), ptr(eng.flag), ptr(eng.cell_rep), eng.ncells, ptr(eng.atom_cell),
         ptr(eng.atom_slot), ptr(eng.cell_count), ptr(eng.cell_start), ptr(eng.cell_atoms),
         ptr(eng.nbr), ptr(eng.nnbr), eng.max_nbr, ptr(eng.overflow), int(eng.precise_list),
         ctypes.byref(eng._ro), s)


def _new_build(eng, force_all):
    call("sim_nl_update", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles, eng.N, ptr(eng.pos4),
         ptr(eng.pos_last), ptr(eng.maxd2), ptr(eng.flag), ptr(eng.nbuild), int(force_all),
         ptr(eng.cell_rep), eng.ncells, ptr(eng.atom_cell), ptr(eng.atom_slot),
         ptr(eng.cell_count), ptr(eng.cell_start), ptr(eng.cell_atoms), ptr(eng.nbr),
         ptr(eng.nnbr), eng.max_nbr, ptr(eng.overflow), int(eng.precise_list), ptr(eng.act),
         ctypes.byref(eng._ro), stream_ptr())


def _compare(a, b, what):
    for name in COMPARED:
        x, y = getattr(a, name, None), getattr(b, name, None)
        if x is None:
            continue
        assert torch.equal(x, y), f"{what}: {name} differs"


def test_update_matches_trigger_then_build(setfl):
    a, b = _engine(setfl), _engine(setfl)
    print(f"[case] {a.R} replicas, {a.N} atoms", flush=True)
    _compare(a, b, "before")
    # displacement maxima: replicas 0 and 2 past the skin test, 1 and 3 not
    trip = (0.6 ** 2, 0.01, 0.7 ** 2, 0.2 ** 2)
    for eng in (a, b):
        eng.maxd2.copy_(torch.tensor(trip, dtype=torch.float64, device=eng.device))
    _old_build(a, 0)
    _new_build(b, 0)
    torch.cuda.synchronize()
    assert a.flag.cpu().tolist() == [1, 0, 1, 0]
    _compare(a, b, "gated build")
    # the reorder moved the atoms of replicas 0 and 2: a force pass and more
    # steps on both, then a forced build of every replica
    for eng in (a, b):
        eng.compute_forces(virial=True)
        eng.run(5, graph=False)
    _old_build(a, 1)
    _new_build(b, 1)
    torch.cuda.synchronize()
    assert a.flag.cpu().tolist() == [1, 1, 1, 1]
    _compare(a, b, "forced build")
    assert int(a.overflow.item()) == 0
