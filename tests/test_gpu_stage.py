"""LDS staging of the EAM passes' neighbour records (sim_nl_stage,
sim_eam_density_staged, sim_eam_force_second with a stage).

- the stage arrays describe the lists: every group's members are whole cells
  and include every atom its atoms' lists name, each list entry's slot
  names its neighbour, and rows past a list's end up to its wave's longest
  list (rounded up to 8) hold the atom's own slot;
- engines with staging on (every group staged), off, and with a member cap
  that leaves the large groups unstaged advance bit for bit alike over graph
  replays, gated rebuilds of some replicas (with the reorder that moves their
  atoms) and a forced rebuild of all.
"""
import numpy as np
import pytest
import torch

from simulon_amd import _lib
from simulon_amd import lattice as lat
from simulon_amd.engine import ReplicaEngine, ReplicaSpec
from simulon_amd.forcefield import EAMForceField

pytestmark = pytest.mark.gpu

SIDES = (40, 9, 30, 12, 25)


@pytest.fixture(scope="module")
def setfl(tmp_path_factory):
    return lat.write_w_synthetic_setfl(str(tmp_path_factory.mktemp("w") / "W_synthetic.eam.alloy"))


def _specs():
    out = []
    for k, n in enumerate(SIDES):
        pos, H = lat.w_bcc(n, seed=30 + k, amp=0.05)
        if k == 1:    # small and hot: its lists rebuild within the run
            kw = dict(ensemble="NVT", temperature=5000.0, gamma=1.0)
        elif k == 2:
            kw = dict(ensemble="NPT", temperature=300.0, target_pressure=1.0, tau_p=0.5)
        else:
            kw = dict(ensemble="NVT", temperature=100.0 + 50 * k, gamma=10.0)
        out.append(ReplicaSpec(positions=pos, box=H, dt=0.002, **kw))
    return out


def _engine(setfl, stage, cap=None):
    eng = ReplicaEngine(_specs(), [lat.W_MASS], cutoff=5.5, skin=1.0, seed=5, family="global",
                        stage=stage)
    if cap is not None:
        eng.stage_cap = cap
    eng.force_field = EAMForceField(setfl, device=eng.device)
    eng.setup()
    return eng


def _check_stage(eng):
    """Host check of the stage arrays against the lists."""
    cap = int(eng._stage.cap)
    cnt = eng.stage_count.cpu().numpy()
    mem = eng.stage_members.cpu().numpy().reshape(-1, cap)
    slot = eng.stage_slot.cpu().numpy().view(np.uint16).astype(np.int64)
    nbr = eng.nbr.cpu().numpy()
    acell = eng.atom_cell.cpu().numpy()
    ccount = eng.cell_count.cpu().numpy()
    crep = eng.cell_rep.cpu().numpy()
    reps = [_lib.struct_from_tensor(eng.reps, _lib.Replica, r) for r in range(eng.R)]
    nn = eng.nnbr.cpu().numpy()
    tiles = eng.tiles_np
    end = np.array([eng.atom_base[r] + eng.natoms[r] for r in range(eng.R)])
    staged = 0
    for G in range(len(cnt)):
        atoms = []
        for tb in range(G * 4, min(G * 4 + 4, eng.ntiles)):
            r, s0 = tiles[tb]
            atoms.extend(range(s0, min(s0 + 256, end[r])))
        atoms = np.array(atoms)
        want = set(atoms.tolist())
        for i in atoms:
            want.update(nbr[:nn[i], i].tolist())
        # the cells of the 27-cell stencils of the group's atoms' cells
        ucells = set()
        for h in np.unique(acell[atoms]):
            P = reps[crep[h]]
            nx, ny, nz = P.ncell[0], P.ncell[1], P.ncell[2]
            loc = h - P.cell_base
            cx, cy, cz = loc % nx, (loc // nx) % ny, loc // (nx * ny)
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        ucells.add(P.cell_base + (((cz + dz) % nz) * ny + (cy + dy) % ny) * nx
                                   + (cx + dx) % nx)
        ucells = np.array(sorted(ucells))
        total = int(ccount[ucells].sum())
        if cnt[G] < 0:
            assert total > cap, f"group {G} unstaged with {total} members"
            continue
        staged += 1
        m = mem[G, :cnt[G]]
        ms = set(m.tolist())
        assert cnt[G] == total and len(ms) == total, f"group {G}"
        assert set(np.unique(acell[m]).tolist()) == set(ucells.tolist()) and want <= ms
        for i in atoms:
            k = nn[i]
            assert np.array_equal(m[slot[:k, i]], nbr[:k, i]), f"atom {i}"
        # padding rows: the longest list of the atom's wave (64 lanes of its
        # tile) rounded up to 8
        for tb in range(G * 4, min(G * 4 + 4, eng.ntiles)):
            r, s0 = tiles[tb]
            for w0 in range(s0, min(s0 + 256, end[r]), 64):
                wa = np.arange(w0, min(w0 + 64, end[r]))
                pad = min((nn[wa].max() + 7) // 8 * 8, eng.max_nbr)
                for i in wa:
                    assert np.all(m[slot[nn[i]:pad, i]] == i), f"padding of atom {i}"
    return staged, len(cnt)


def test_stage_arrays_and_bitwise_equal_steps(setfl):
    a = _engine(setfl, True)
    b = _engine(setfl, False)
    c = _engine(setfl, True, cap=3000)
    print(f"[case] sides {SIDES}: {a.N} atoms, {a.ntiles} tiles", flush=True)
    assert a._stage is not None and b._stage is None and c._stage is not None
    sa, ng = _check_stage(a)
    sc, _ = _check_stage(c)
    print(f"staged groups: all-cap {sa}/{ng}, cap 3000 {sc}/{ng}", flush=True)
    assert sa == ng and 0 < sc < ng
    for name in ("force", "vel", "pos4"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name
        assert torch.equal(getattr(c, name), getattr(b, name)), name
    for eng in (a, b, c):
        eng.run(40, graph=True)
        eng.run(5, graph=False)
    nb = a.nbuild.cpu().numpy()
    print(f"rebuilds per replica: {nb.tolist()}", flush=True)
    # the hot replica rebuilt within the run, some others kept their lists
    assert nb[1] > 1 and nb.max() > nb.min()
    _check_stage(a)
    for eng in (a, b, c):
        eng.build_list(force=True)
        eng.run(5, graph=False)
    torch.cuda.synchronize()
    _check_stage(a)
    _check_stage(c)
    for name in ("pos4", "vel", "force"):
        x = getattr(b, name)
        for e in (a, c):
            y = getattr(e, name)
            assert torch.equal(x, y), f"{name}: max diff {(x - y).abs().max().item():.3e}"
    np.testing.assert_array_equal(a.thermo_host(), b.thermo_host())
    np.testing.assert_array_equal(c.thermo_host(), b.thermo_host())
