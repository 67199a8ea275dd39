"""GPU parity of the NPT step (core/barostat.py:36-68, core/md_model.py:83-86)
and of the 1,048,576-atom triclinic LJ NPT configuration at full size.

- the device Berendsen step (k_baro_box + k_baro_scale) against the
  reference's own outputs (tests/golden/golden_barostat.npz, written by the
  reference's BerendsenBarostat.step) and the oracle's closed form;
- the kinetic-tensor columns of the second half against the float64 sum;
- the full 64^3 fcc triclinic argon cell: pair count, every atom's force,
  energy and virial against the float64 C oracle, then 5 NPT steps with
  injected noise, each checked against the oracle's BAOB + Berendsen step.
"""
import os

import numpy as np
import pytest
import torch

from helpers import lj_force_error, sym3
from oracle import box as obox
from oracle import cell as ocell
from oracle import integrate as oint
from simulon_amd import _lib
from simulon_amd import lattice as lat
from simulon_amd.engine import ReplicaEngine, ReplicaSpec, internal_mass_f32
from simulon_amd.forcefield import LJForceField

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FORCE_TOL = 1e-5


def _golden_mol():
    from simulon_amd.reader import AtomFileReader
    g = np.load(os.path.join(G, "golden_barostat.npz"))
    x0 = g["baro_x0"].astype(np.float64)
    mol = AtomFileReader(atom_types=["Ar"] * len(x0), positions=x0, box_length=float(g["baro_H0"][0, 0]),
                         cutoff=8.5, device="cuda", box_vectors=torch.tensor(g["baro_H0"]),
                         parameter={"[0 0]": {"epsilon": 0.0104, "sigma": 3.405}},
                         skin_thickness=1.0, sort=True, max_nbr=256)
    mol.atom_velocities = torch.tensor(g["baro_v0"])
    return g, mol


def test_berendsen_step_matches_reference():
    """One BerendsenBarostat.step(dt, KE, W, K, W_t) on the inputs of the
    reference's golden call: pressure, box, coordinates and velocities equal
    the reference's outputs and the oracle's closed form."""
    from simulon_amd.barostat import BerendsenBarostat
    g, mol = _golden_mol()
    eng = mol.engine
    baro = BerendsenBarostat(mol, target_pressure=1.0, tau_p=1.0)
    last0 = mol.last_positions.cpu().numpy().astype(np.float64)
    P = baro.step(0.002, torch.tensor(0.0), torch.tensor(0.0),
                  kinetic_tensor=torch.tensor(g["baro_K"]), virial_tensor=torch.tensor(g["baro_W"]))
    mu, P_or = oint.berendsen_mu(g["baro_K"], g["baro_W"], obox.volume(g["baro_H0"]), 0.002, 1.0, 1.0)
    print(f"Berendsen P {P:.6f} reference {float(g['baro_P']):.6f} oracle {P_or:.6f} mu {mu!r}",
          flush=True)
    assert P == pytest.approx(float(g["baro_P"]), rel=1e-12)
    assert P == pytest.approx(P_or, rel=1e-12)
    assert eng.replica_state(0).mu == pytest.approx(mu, rel=1e-15)
    np.testing.assert_allclose(mol.box.H.cpu().numpy(), g["baro_H1"], rtol=1e-14, atol=0)
    np.testing.assert_allclose(mol.box.H.cpu().numpy(), g["baro_H0"] * mu, rtol=1e-15, atol=0)
    # coordinates: float64 x0 mu here, float32 x0 * mu in the reference
    x1 = mol.coordinates.cpu().numpy()
    np.testing.assert_allclose(x1, g["baro_x1"], rtol=2.4e-7, atol=1e-6)
    xd = eng.to_caller(eng.positions(torch.float64, unwrap=True)).cpu().numpy()
    np.testing.assert_allclose(xd, g["baro_x0"].astype(np.float64) * mu, rtol=1e-14, atol=1e-12)
    # velocities: float32 v * float32(mu), as the reference's in-place mul_
    np.testing.assert_array_equal(mol.atom_velocities.cpu().numpy(), g["baro_v1"])
    # the rebuild reference moves with the box (the reference resets it to x,
    # core/barostat.py:67; here it is scaled, so the skin trigger keeps
    # measuring displacements since the last build)
    np.testing.assert_allclose(mol.last_positions.cpu().numpy(), last0 * mu, rtol=2e-7, atol=1e-5)


def test_berendsen_standalone_dt_leaves_integrator_dt():
    """step(dt) with a dt of its own uses it for the barostat only: the
    integrator's time step in the replica record is unchanged."""
    from simulon_amd.barostat import BerendsenBarostat
    from simulon_amd.integrator import VerletIntegrator
    g, mol = _golden_mol()
    integ = VerletIntegrator(mol, dt=0.002, ensemble="NVE")
    baro = BerendsenBarostat(mol, target_pressure=1.0, tau_p=1.0)
    P = baro.step(0.005, torch.tensor(0.0), torch.tensor(0.0),
                  kinetic_tensor=torch.tensor(g["baro_K"]), virial_tensor=torch.tensor(g["baro_W"]))
    mu, _ = oint.berendsen_mu(g["baro_K"], g["baro_W"], obox.volume(g["baro_H0"]), 0.005, 1.0, 1.0)
    st = mol.engine.replica_state(0)
    assert st.mu == pytest.approx(mu, rel=1e-15)
    assert st.dt == 0.002 and st.dtf == np.float32(0.002) and st.baro_dt == 0.0
    assert integ.dt == 0.002 and np.isfinite(P)


def test_kinetic_tensor_columns():
    """thermo[K_xx..K_yz] and KE after a step equal the float64 sums
    sum m v_a v_b of the velocities the step wrote (md_model.py:83-86)."""
    pos, H = lat.lj_argon_npt_tri(ncell=6, seed=9)
    spec = ReplicaSpec(positions=pos, box=H, dt=0.002, ensemble="NVT", temperature=120.0,
                       gamma=10.0)
    eng = ReplicaEngine([spec], [lat.AR_MASS], cutoff=8.5, skin=1.0)
    eng.force_field = LJForceField([[lat.AR_EPS]], [[lat.AR_SIG]], 8.5)
    eng.setup()
    eng.run(5, graph=True)
    v = eng.vel.cpu().numpy()
    m = np.full(eng.N, np.float32(eng.types_struct.mass[0]), np.float32)
    K = oint.kinetic_tensor(v, m)
    th = eng.thermo_host()[0]
    Kd = sym3(th[_lib.T_K:_lib.T_K + 6])
    print(f"kinetic tensor rel err {np.max(np.abs(Kd - K)) / np.max(np.abs(K)):.2e}", flush=True)
    np.testing.assert_allclose(Kd, K, rtol=1e-12, atol=1e-12 * np.max(np.abs(K)))
    assert th[_lib.T_KE] == pytest.approx(0.5 * np.trace(K), rel=1e-12)


# ---- the 1,048,576-atom triclinic LJ NPT cell -----------------------------------

def _lj_oracle(x, H):
    return ocell.lj(x, H, None, [[lat.AR_EPS]], [[lat.AR_SIG]], 8.5, 1.0)


def test_lj_npt_triclinic_full_size():
    """64^3 fcc argon (1,048,576 atoms), box rows (L,0,0), (0.1L,L,0),
    (0.05L,0,L), 90 K Langevin + Berendsen at 1 bar (tau_P 1 ps), dt 2 fs.

    At setup: the half-list pair count equals the float64 oracle's; every
    atom's force is within 1e-5 (relative max-atom) of the oracle's, except
    atoms with a pair within 1e-5 A of rc, which float32 may count on the
    other side of the cutoff (held to |Fmag(rc)| per such pair); energy and
    virial tensor within 1e-6. Then 5 steps with injected Gaussian kicks,
    each checked against the oracle's step from the same state: first half
    (B+A), forces at the new positions (float64 oracle, all atoms), O+B with
    the same kicks, kinetic tensor, Berendsen mu, box, positions and
    velocities."""
    pos, H = lat.lj_argon_npt_tri(ncell=64, seed=0)
    n = len(pos)
    assert n == 1_048_576
    spec = ReplicaSpec(positions=pos, box=H, dt=0.002, ensemble="NPT", temperature=90.0,
                       gamma=10.0, target_pressure=1.0, tau_p=1.0)
    eng = ReplicaEngine([spec], [lat.AR_MASS], cutoff=8.5, skin=1.0)
    eng.force_field = LJForceField([[lat.AR_EPS]], [[lat.AR_SIG]], 8.5)
    eng.setup()
    torch.cuda.synchronize()
    print(f"cfg2 engine: family {eng.family}, max_nbr {eng.max_nbr}", flush=True)

    def state():
        x = eng.to_caller(eng.positions(torch.float64, unwrap=True)).cpu().numpy()
        v = eng.to_caller(eng.vel).cpu().numpy()
        f = eng.to_caller(eng.force).cpu().numpy()
        return x, v, f, eng.box(0), eng.thermo_host()[0]

    x, v, f, Hk, th = state()
    ref = _lj_oracle(x, Hk)
    npairs = int(eng.nnbr.sum().item()) // 2
    err, nties = lj_force_error(f, ref, lat.AR_EPS, lat.AR_SIG, 8.5)
    W = sym3(th[_lib.T_W:_lib.T_W + 6])
    wer = np.max(np.abs(W - ref["virial_tensor"])) / np.max(np.abs(ref["virial_tensor"]))
    print(f"cfg2 setup: pairs {npairs} oracle {ref['npairs']}, force err {err:.3e} "
          f"({nties} atoms with a pair within 1e-5 A of rc held to |Fmag(rc)|), "
          f"E {th[0]:.6f} oracle {ref['energy']:.6f}, W err {wer:.2e}", flush=True)
    assert npairs == ref["npairs"]
    assert err <= FORCE_TOL
    assert th[_lib.T_EPOT] == pytest.approx(ref["energy"], rel=1e-6)
    assert wer <= 1e-6

    m = np.full(n, internal_mass_f32(lat.AR_MASS), np.float32)
    dt, gamma, T = 0.002, 10.0, 90.0
    rng = np.random.default_rng(2026)
    builds0 = int(eng.nbuild.cpu()[0])
    for k in range(5):
        noise = rng.standard_normal((n, 3)).astype(np.float32)
        eng.step(noise=eng.from_caller(torch.from_numpy(noise).to(eng.device)))
        torch.cuda.synchronize()
        x1, v1, f1, H1, th1 = state()
        # oracle step from the state before
        vh, xa = oint.first_half(x, v, f, m, dt)
        refa = _lj_oracle(xa, Hk)
        ferr, nties = lj_force_error(f1, refa, lat.AR_EPS, lat.AR_SIG, 8.5)
        vo, ke, _ = oint.second_half(vh, f1, m, dt, gamma=gamma, T=T, noise=noise)
        K = oint.kinetic_tensor(vo, m)
        Wd = sym3(th1[_lib.T_W:_lib.T_W + 6])
        wer = np.max(np.abs(Wd - refa["virial_tensor"])) / np.max(np.abs(refa["virial_tensor"]))
        mu, P = oint.berendsen_mu(K, refa["virial_tensor"], obox.volume(Hk), dt, 1.0, 1.0)
        st = eng.replica_state(0)
        print(f"cfg2 step {k}: force err {ferr:.3e} ({nties} tie atoms), W err {wer:.2e}, "
              f"P {st.press:.4f} "
              f"oracle {P:.4f} bar, mu-1 {st.mu - 1:.3e} oracle {mu - 1:.3e}", flush=True)
        assert ferr <= FORCE_TOL
        assert wer <= 1e-6
        assert np.max(np.abs(sym3(th1[_lib.T_K:_lib.T_K + 6]) - K)) <= 1e-5 * np.max(np.abs(K))
        assert st.press == pytest.approx(P, rel=1e-5, abs=0.05)
        assert st.mu == pytest.approx(mu, abs=1e-10)
        np.testing.assert_allclose(H1, Hk * st.mu, rtol=1e-15, atol=0)
        np.testing.assert_allclose(x1, xa * st.mu, rtol=0, atol=1e-9)
        np.testing.assert_allclose(v1, vo * np.float32(st.mu), rtol=2e-6, atol=2e-6)
        x, v, f, Hk, th = x1, v1, f1, H1, th1
    assert int(eng.nbuild.cpu()[0]) == builds0     # the check assumed no reorder mid-step
