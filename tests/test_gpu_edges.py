"""Edge cases of the device path: boxes with one or two cells per dimension
(minimum image only, as the reference's brute force), a two-atom system
against the closed form, ragged batches mixing a two-atom replica with large
ones, and the neighbor-list overflow error."""
import numpy as np
import pytest
import torch

from helpers import force_error, oracle_half_list, oracle_lj
from oracle import lj as olj
from simulon_amd import lattice as lat
from simulon_amd.engine import ReplicaEngine, ReplicaSpec
from simulon_amd.forcefield import LJForceField

pytestmark = pytest.mark.gpu


def _engine(specs, **kw):
    eng = ReplicaEngine(specs, [lat.AR_MASS], cutoff=8.5, skin=1.0, **kw)
    eng.force_field = LJForceField([[lat.AR_EPS]], [[lat.AR_SIG]], 8.5)
    eng.setup()
    return eng


@pytest.mark.parametrize("ncell,tri", [(4, False), (5, False), (6, False), (5, True)])
def test_small_box_pairs_and_forces(ncell, tri):
    """Boxes narrower than two list radii (1-2 cells per dimension): the pair
    set equals the minimum-image brute force and the forces the FP64 oracle."""
    L = 3.64 * ncell
    H = np.diag([L, L, L])
    if tri:
        H = np.array([[L, 0, 0], [0.15 * L, L, 0], [-0.1 * L, 0.05 * L, L]])
    pos = lat.jitter(lat.lattice(lat.SC_BASIS, ncell, H), 0.05, 21)
    print(f"[case] small box ncell={ncell} tri={tri} n={len(pos)}", flush=True)
    eng = _engine([ReplicaSpec(positions=pos, box=H, dt=0.002)])
    ei, _ = eng.half_list()
    x = eng.positions().cpu().numpy()
    ref, _ = oracle_half_list(x, H, 8.5, 1.0)
    assert np.array_equal(ei.cpu().numpy(), ref)
    r = oracle_lj(x, H, np.zeros(len(x), int), [[lat.AR_EPS]], [[lat.AR_SIG]], 8.5, 1.0)
    err = force_error(eng.force.cpu().numpy(), r["forces"])
    th = eng.thermo_host()[0]
    print(f"cells {eng._grid_of(0)}: force err {err:.2e}, E {th[0]:.8f} vs {r['energy']:.8f}",
          flush=True)
    assert err <= 1e-5
    assert abs(th[0] - r["energy"]) <= 1e-6 * abs(r["energy"])
    eng.run(20)                           # steps and rebuilds in the small box stay valid
    x = eng.positions().cpu().numpy()
    eng.compute_forces(virial=True)
    r = oracle_lj(x, H, np.zeros(len(x), int), [[lat.AR_EPS]], [[lat.AR_SIG]], 8.5, 1.0)
    assert force_error(eng.force.cpu().numpy(), r["forces"]) <= 1e-5


def test_two_atoms_closed_form():
    H = np.eye(3) * 40.0
    for r in (3.2, 3.82, 6.0, 8.49):
        pos = np.array([[10.0, 10.0, 10.0], [10.0 + r, 10.0, 10.0]])
        eng = _engine([ReplicaSpec(positions=pos, box=H, dt=0.002)])
        U, F = olj.lj_two_atom(r, lat.AR_EPS, lat.AR_SIG)
        f = eng.to_caller(eng.force).cpu().numpy()
        th = eng.thermo_host()[0]
        print(f"[case] two atoms r={r}: U {th[0]:.9e} vs {U:.9e}, F {f[1, 0]:.6e} vs {F:.6e}",
              flush=True)
        # one float32 pair in the repulsive wall: s^12 - s^6 cancels, so the
        # bar is 1e-5 (system energies average such errors down to 1e-6)
        assert th[0] == pytest.approx(U, rel=1e-5)
        # near the minimum (r = 3.82) F itself cancels to ~1e-4: the absolute
        # bar is 1e-5 of the force scale 24 eps / sigma = 0.073 eV/A
        assert f[1, 0] == pytest.approx(F, rel=1e-5, abs=1e-6)
        assert f[0, 0] == pytest.approx(-F, rel=1e-5, abs=1e-6)
        assert np.abs(f[:, 1:]).max() <= 1e-9


def test_ragged_batch_with_tiny_replica():
    """A two-atom replica packed between large ones: every replica's forces and
    energy equal those of its own single-replica engine, bit for bit."""
    big, Hb = lat.lj_argon_nvt(seed=3)
    mid, Hm = lat.lj_argon_npt_tri(ncell=4, seed=4)
    two = np.array([[5.0, 5.0, 5.0], [8.6, 5.2, 5.1]])
    Ht = np.eye(3) * 30.0
    specs = [ReplicaSpec(positions=p, box=h, dt=0.002)
             for p, h in ((big, Hb), (two, Ht), (mid, Hm))]
    print(f"[case] ragged batch sizes {[len(s.positions) for s in specs]}", flush=True)
    eng = _engine(specs, family="global")     # one kernel family on both sides
    f_all = eng.force.cpu().numpy()
    e_all = eng.thermo_host()[:, 0]
    for k, s in enumerate(specs):
        e1 = _engine([ReplicaSpec(positions=s.positions, box=s.box, dt=0.002)], family="global")
        np.testing.assert_array_equal(e1.force.cpu().numpy(), f_all[eng.replica_slice(k)])
        assert e1.thermo_host()[0, 0] == e_all[k]
    U, _ = olj.lj_two_atom(float(np.linalg.norm(two[1] - two[0])), lat.AR_EPS, lat.AR_SIG)
    assert e_all[1] == pytest.approx(U, rel=1e-5)


def test_neighbor_overflow_raises():
    """A list capacity below the real neighbor count is reported, not
    truncated silently (the list kernels bound every write by max_nbr)."""
    from simulon_amd.reader import AtomFileReader
    pos, H = lat.lj_argon_nvt(seed=5)
    print("[case] neighbor overflow max_nbr=16", flush=True)
    with pytest.raises(RuntimeError, match="neighbor list overflow"):
        AtomFileReader(atom_types=["Ar"] * len(pos), positions=pos, box_length=36.4, cutoff=8.5,
                       device="cuda", parameter={"[0 0]": {"epsilon": lat.AR_EPS,
                                                          "sigma": lat.AR_SIG}},
                       skin_thickness=1.0, max_nbr=16)
    torch.cuda.synchronize()
