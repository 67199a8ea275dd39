"""The EAM force pass with the integrator's second half fused in
(sim_eam_force_second) against the separate second-half kernel
(sim_integrate_second after sim_eam_force): positions, velocities, forces and
every thermo column agree bit for bit over graph-replayed and eager steps, for
Langevin, NVE and Berendsen NPT replicas packed together, on the per-tile
global tables (small launch) and on the persistent LDS-table kernel (more
than 256 tiles). Graph runs of two or more steps also let the force pass do
the next step's first half (the NPT replica's stays in its own kernel): the
same bits as runs that do not, across list rebuilds and runs of 1, 2 and more
steps, and the Berendsen replicas' scaling left to the next step's first half
equals the barostat's own. The LJ global-list pass with the second half fused in
(sim_lj_force_second) equals k_lj followed by the separate kernel, bit for
bit, on triclinic argon cells packed the same way.
"""
import numpy as np
import pytest
import torch

from simulon_amd import lattice as lat
from simulon_amd.engine import ReplicaEngine, ReplicaSpec
from simulon_amd.forcefield import EAMForceField, LJForceField

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setfl(tmp_path_factory):
    return lat.write_w_synthetic_setfl(str(tmp_path_factory.mktemp("w") / "W_synthetic.eam.alloy"))


def _specs(sides):
    kinds = [dict(ensemble="NVT", temperature=600.0, gamma=10.0),
             dict(ensemble="NVE"),
             dict(ensemble="NPT", temperature=300.0, target_pressure=1.0, barostat="berendsen",
                  tau_p=0.5)]
    out = []
    for k, n in enumerate(sides):
        pos, H = lat.w_bcc(n, seed=7 + k, amp=0.05)
        out.append(ReplicaSpec(positions=pos, box=H, dt=0.001, **kinds[k % 3]))
    return out


def _engine(setfl, sides, fused, skin=1.0, chain=True):
    eng = ReplicaEngine(_specs(sides), [lat.W_MASS], cutoff=5.5, skin=skin, seed=99,
                        family="global")
    ff = EAMForceField(setfl, device=eng.device)
    if not fused:
        ff.fuses_second = False            # instance attribute: the separate kernel
    if not chain:
        ff.fuses_next_first = lambda e: False
    eng.force_field = ff
    eng.chain = chain
    eng.setup()
    return eng


@pytest.mark.parametrize("sides", [(6, 5, 7), (40, 12, 30)], ids=["per-tile", "persistent"])
def test_fused_second_half_bitwise(setfl, sides):
    a = _engine(setfl, sides, True)
    b = _engine(setfl, sides, False)
    print(f"[case] sides {sides}: {a.N} atoms, {a.ntiles} tiles", flush=True)
    assert torch.equal(a.vel, b.vel) and torch.equal(a.pos4, b.pos4)
    for eng in (a, b):
        eng.run(15, graph=True)
        eng.run(3, graph=False)
    torch.cuda.synchronize()
    for name in ("pos4", "vel", "force"):
        x, y = getattr(a, name), getattr(b, name)
        assert torch.equal(x, y), f"{name}: max diff {(x - y).abs().max().item():.3e}"
    ta, tb = a.thermo_host(), b.thermo_host()
    np.testing.assert_array_equal(ta, tb)
    assert np.all(np.isfinite(ta))


@pytest.mark.parametrize("sides", [(6, 5, 7), (40, 12, 30)], ids=["per-tile", "persistent"])
def test_next_first_half_chain_bitwise(setfl, sides):
    a = _engine(setfl, sides, True, skin=0.3)
    b = _engine(setfl, sides, True, skin=0.3, chain=False)
    print(f"[case] chain, sides {sides}: {a.N} atoms, {a.ntiles} tiles", flush=True)
    assert a.fuses_next_first() and not b.fuses_next_first()
    nb0 = a.nbuild.clone()
    for n in (2, 1, 25, 3):
        a.run(n, graph=True)
        b.run(n, graph=True)
    a.run(2, graph=False)
    b.run(2, graph=False)
    torch.cuda.synchronize()
    assert int((a.nbuild - nb0).min()) > 0, "no list rebuild in the compared steps"
    for name in ("pos4", "vel", "vel_half", "force", "posq", "pos_last", "maxd2", "nbuild",
                 "reps"):
        x, y = getattr(a, name), getattr(b, name)
        assert torch.equal(x, y), f"{name} differs"
    np.testing.assert_array_equal(a.thermo_host(), b.thermo_host())


def _lj_engine(fused, sides=(6, 5, 7)):
    kinds = [dict(ensemble="NVT", temperature=120.0, gamma=5.0),
             dict(ensemble="NVE"),
             dict(ensemble="NPT", temperature=90.0, target_pressure=1.0, barostat="berendsen",
                  tau_p=0.5)]
    specs = []
    for k, n in enumerate(sides):
        pos, H = lat.lj_argon_npt_tri(ncell=n, seed=21 + k)
        specs.append(ReplicaSpec(positions=pos, box=H, dt=0.002, **kinds[k % 3]))
    eng = ReplicaEngine(specs, [lat.AR_MASS], cutoff=8.5, skin=1.0, seed=17, family="global")
    ff = LJForceField([[lat.AR_EPS]], [[lat.AR_SIG]], 8.5)
    if not fused:
        ff.fuses_second = False            # instance attribute: the separate kernel
    eng.force_field = ff
    eng.setup()
    return eng


def test_lj_fused_second_half_bitwise():
    a = _lj_engine(True)
    b = _lj_engine(False)
    print(f"[case] LJ fused second half: {a.N} atoms, {a.ntiles} tiles", flush=True)
    for eng in (a, b):
        eng.run(25, graph=True)
        eng.run(3, graph=False)
    torch.cuda.synchronize()
    for name in ("pos4", "vel", "vel_half", "force"):
        x, y = getattr(a, name), getattr(b, name)
        assert torch.equal(x, y), f"{name}: max diff {(x - y).abs().max().item():.3e}"
    ta, tb = a.thermo_host(), b.thermo_host()
    np.testing.assert_array_equal(ta, tb)
    assert np.all(np.isfinite(ta))


def test_lj_npt_chain_bitwise():
    """Graph runs that leave each Berendsen step's scaling of the atoms to the
    next step's first half equal unchained runs, bit for bit, LJ global list."""
    a = _lj_engine(True)
    b = _lj_engine(True)
    b.chain = False
    print(f"[case] LJ Berendsen chain: {a.N} atoms", flush=True)
    assert a._defers_scale()
    for n in (2, 1, 25, 3):
        a.run(n, graph=True)
        b.run(n, graph=True)
    a.run(2, graph=False)
    b.run(2, graph=False)
    torch.cuda.synchronize()
    for name in ("pos4", "vel", "vel_half", "force", "posq", "pos_last", "maxd2", "nbuild",
                 "reps"):
        x, y = getattr(a, name), getattr(b, name)
        assert torch.equal(x, y), f"{name} differs"
    np.testing.assert_array_equal(a.thermo_host(), b.thermo_host())
