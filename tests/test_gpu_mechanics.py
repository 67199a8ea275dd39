"""Anisotropic barostat and uniaxial tensile loading on the device
(simulon_amd.barostat.AnisotropicNPTBarostat, simulon_amd.mechanics;
the reference's core/barostat.py:78-208 and core/mechanics/loading.py:39-78):
standalone calls against the float64 oracle, the loading fused into the
captured step against the reference's loop (loader.step(dt) before every
model() call), and the rebuild trigger under deformation."""
import numpy as np
import pytest
import torch

from helpers import force_error, oracle_eam, sym3
from oracle import mechanics as om
from simulon_amd import lattice as lat

pytestmark = pytest.mark.gpu


def _lj_tri():
    from simulon_amd.force import LennardJonesForce
    from simulon_amd.integrator import VerletIntegrator
    from simulon_amd.md_model import BaseModel, SumBackboneInterface
    from simulon_amd.reader import AtomFileReader
    torch.manual_seed(0)
    pos, H = lat.lj_argon_npt_tri(ncell=4, seed=2)
    mol = AtomFileReader(atom_types=["Ar"] * len(pos), positions=pos, box_length=float(H[0, 0]),
                         cutoff=8.5, device="cuda",
                         parameter={"[0 0]": {"epsilon": lat.AR_EPS, "sigma": lat.AR_SIG}},
                         skin_thickness=1.0, box_vectors=torch.tensor(H))
    ff = LennardJonesForce(mol)
    integ = VerletIntegrator(mol, dt=0.002, ensemble="NVT", temperature=(94.4, 94.4), gamma=10.0)
    return mol, BaseModel(SumBackboneInterface([ff], mol), integ, mol)


def _state(mol):
    eng = mol.engine
    x = eng.to_caller(eng.positions(torch.float64, unwrap=True)).cpu().numpy()
    return mol.box.H.cpu().numpy().copy(), x, mol.atom_velocities.cpu().numpy().copy()


def _tensors(rng, H, p_bar, axis0_at=None, target=0.0):
    """K, W (eV) whose pressure tensor is a random symmetric one of scale
    p_bar; with axis0_at, shifted so lattice row 0 reads target + axis0_at."""
    A = rng.normal(size=(3, 3))
    S = (A + A.T) * p_bar
    if axis0_at is not None:
        u = H[0] / np.linalg.norm(H[0])
        S = S + np.eye(3) * (target + axis0_at - u @ S @ u)
    V = abs(np.linalg.det(H))
    K = np.diag(rng.uniform(0.5, 1.0, 3))
    W = S / om.EV_A3_TO_BAR * V - K
    return K, W


def test_aniso_barostat_step_matches_oracle():
    from simulon_amd.barostat import AnisotropicNPTBarostat
    print("[case] aniso barostat standalone, triclinic n=256", flush=True)
    mol, model = _lj_tri()
    ctl = (True, False, True)
    kw = dict(tau_p=0.2, gamma_p=2.0, max_eta_dot=0.05, kappa=3.2e-4, tol=25.0)
    target = np.array([0.0, 100.0, -50.0])
    baro = AnisotropicNPTBarostat(mol, target_pressure_bar=target.tolist(), temperature_k=94.4,
                                  tau_p=0.2, gamma_p=2.0, control_axes=ctl,
                                  compressibility_bar_inv=3.2e-4, pressure_tolerance_bar=25.0)
    H, x, v = _state(mol)
    eta = np.zeros(3)
    rng = np.random.default_rng(9)
    cases = [(800.0, None), (3000.0, None), (500.0, 10.0), (500.0, -20.0), (50.0, None)]
    for k, (scale, a0) in enumerate(cases):
        K, W = _tensors(rng, H, scale, a0, target[0])
        p = baro.step(0.002, 0.0, 0.0, kinetic_tensor=torch.tensor(K), virial_tensor=torch.tensor(W))
        H, x, v, eta, p_ref = om.aniso_step(H, x, v, eta, K, W, 0.002, target, control=ctl, **kw)
        Hd, xd, vd = _state(mol)
        ed = baro.eta_dot.cpu().numpy()
        print(f"call {k}: p_axis {p.cpu().numpy()} eta {ed}", flush=True)
        np.testing.assert_allclose(p.cpu().numpy(), p_ref.astype(np.float32), rtol=1e-6)
        np.testing.assert_allclose(baro.pressure_axis_bar.cpu().numpy(), p_ref, rtol=1e-10)
        np.testing.assert_allclose(ed, eta, rtol=1e-12, atol=1e-18)
        np.testing.assert_allclose(Hd, H, rtol=1e-13, atol=1e-13)
        assert np.abs(xd - x).max() <= 1e-9
        np.testing.assert_allclose(vd, v, rtol=1e-7, atol=0)
    assert eta[1] == 0.0 and np.abs(eta).max() > 0.0
    assert np.abs(H - _lj_tri_H()).max() > 1e-4          # the box did change

    # the stochastic term: one more call from the same state, with and without it
    K, W = _tensors(rng, H, 1500.0)
    e_det = om.aniso_update(eta, om.axis_pressure(K, W, H), 0.002, target, control=ctl, **kw)
    baro.stochastic = True
    baro._activate()
    baro.step(0.002, 0.0, 0.0, kinetic_tensor=torch.tensor(K), virial_tensor=torch.tensor(W))
    e_sto = baro.eta_dot.cpu().numpy()
    g = (e_sto - e_det) / (0.002 * 3.2e-4 * 8.617333262e-5 * 94.4 / 0.2)
    print(f"stochastic draws {g}", flush=True)
    assert g[1] == 0.0 and np.all(np.abs(g) < 6.0) and np.abs(g).max() > 1e-3


def _lj_tri_H():
    return lat.lj_argon_npt_tri(ncell=4, seed=2)[1]


def test_tensile_loader_step_matches_oracle():
    from simulon_amd.mechanics import UniaxialTensileLoader
    print("[case] tensile loader standalone, triclinic n=256", flush=True)
    mol, model = _lj_tri()
    loader = UniaxialTensileLoader(mol, axis=1, strain_rate=2.0, lateral_mode="poisson",
                                   poisson_ratio=0.3)
    H, x, v = _state(mol)
    L0 = np.linalg.norm(H[1])
    for n in range(10):
        strain = loader.step(0.002)
        H, x, strain_ref = om.loader_step(H, x, 1, 2.0, 0.002, L0, lateral_mode="poisson",
                                          poisson_ratio=0.3)
        Hd, xd, vd = _state(mol)
        np.testing.assert_allclose(Hd, H, rtol=1e-14, atol=1e-13)
        assert np.abs(xd - x).max() <= 1e-9
        np.testing.assert_array_equal(vd, v)                 # loading leaves v alone
        assert strain == pytest.approx(strain_ref, rel=1e-12)
    assert strain == pytest.approx(1.004 ** 10 - 1, rel=1e-9)
    assert mol.box_length_cpu == pytest.approx(np.linalg.norm(H[0]), rel=1e-14)


@pytest.fixture(scope="module")
def w_setfl(tmp_path_factory):
    p = tmp_path_factory.mktemp("eam") / "W_synthetic.eam.alloy"
    return str(lat.write_w_synthetic_setfl(p))


def _w_model(setfl, pos, H, ensemble, barostat=None):
    from simulon_amd.eam_parser import EAMParser
    from simulon_amd.force import EAMForceCUDAExt
    from simulon_amd.integrator import VerletIntegrator
    from simulon_amd.md_model import BaseModel, SumBackboneInterface
    from simulon_amd.reader import AtomFileReader
    torch.manual_seed(0)
    parser = EAMParser(setfl)
    mol = AtomFileReader(atom_types=["W"] * len(pos), positions=pos, box_length=float(H[0, 0]),
                         cutoff=parser.cutoff, device="cuda", skin_thickness=1.0, is_mlp=True,
                         box_vectors=torch.tensor(H))
    ff = EAMForceCUDAExt(parser, mol)
    kw = dict(temperature=(300.0, 300.0), gamma=1.0, seed=3) if ensemble != "NVE" else {}
    integ = VerletIntegrator(mol, dt=0.001, ensemble=ensemble, **kw)
    b = barostat(mol) if barostat else None
    return mol, BaseModel(SumBackboneInterface([ff], mol), integ, mol, barostat=b)


def _lateral_barostat(mol):
    from simulon_amd.barostat import AnisotropicNPTBarostat
    return AnisotropicNPTBarostat(mol, target_pressure_bar=0.0, temperature_k=300.0, tau_p=0.2,
                                  gamma_p=2.0, control_axes=(False, True, True),
                                  compressibility_bar_inv=3.2e-6, pressure_tolerance_bar=25.0)


def test_fused_loading_matches_reference_loop(w_setfl):
    """W tensile step (run_scripts/w_tensile.py:337-340): loader.step(dt) then
    model() with a lateral stress-free barostat, 300 steps, against the same
    loading fused into the captured graph."""
    from simulon_amd.mechanics import UniaxialTensileLoader
    print("[case] fused tensile loading, W bcc n=1024, 300 steps", flush=True)
    pos, H = lat.w_bcc(8, seed=3, amp=0.03)
    nsteps, dt, rate = 300, 0.001, 0.5

    mol_a, model_a = _w_model(w_setfl, pos, H, "NPT", _lateral_barostat)
    model_a()
    load_a = UniaxialTensileLoader(mol_a, axis=0, strain_rate=rate)
    for _ in range(nsteps):
        strain_a = load_a.step(dt)
        out_a = model_a()
    Ha, xa, va = _state(mol_a)
    eta_a = model_a.barostat.eta_dot.cpu().numpy()

    mol_b, model_b = _w_model(w_setfl, pos, H, "NPT", _lateral_barostat)
    model_b()
    load_b = UniaxialTensileLoader(mol_b, axis=0, strain_rate=rate)
    load_b.attach(model_b)
    out_b = model_b.run(nsteps)
    strain_b = load_b.strain()
    Hb, xb, vb = _state(mol_b)
    eta_b = model_b.barostat.eta_dot.cpu().numpy()
    print(f"strain {strain_a:.6f} / {strain_b:.6f}; lateral lengths "
          f"{np.linalg.norm(Hb, axis=1)[1:]} (from {H[1, 1]:.4f}); eta {eta_b}; "
          f"p_axis {out_b['pressure'].cpu().numpy()} bar", flush=True)
    assert strain_b == pytest.approx((1 + rate * dt) ** nsteps - 1, rel=1e-9)
    assert strain_a == pytest.approx(strain_b, rel=1e-12)
    np.testing.assert_allclose(Hb, Ha, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(eta_b, eta_a, rtol=1e-9, atol=1e-15)
    assert np.abs(xb - xa).max() <= 1e-6
    assert abs(float(out_a["energy"]) - float(out_b["energy"])) <= 1e-6 * abs(float(out_a["energy"]))
    assert np.abs(np.linalg.norm(Hb, axis=1)[1:] - H[1, 1]).max() > 1e-4   # lateral relaxation

    # the list after loading holds every pair: forces, energy and virial vs FP64
    eng = model_b.engine
    eng.compute_forces(virial=True)
    eng.reduce(0, 8)
    x = eng.positions().cpu().numpy()
    ref = oracle_eam(x, Hb, np.zeros(len(x), int), w_setfl, 1.0)
    th = eng.thermo_host()[0]
    err = force_error(eng.force.cpu().numpy(), ref["forces"])
    print(f"after loading: force err {err:.2e}, E {th[0]:.6f} vs {ref['energy']:.6f}", flush=True)
    assert err <= 1e-5
    assert abs(th[0] - ref["energy"]) <= 1e-6 * abs(ref["energy"])
    W = sym3(th[2:8])
    assert np.abs(W - ref["virial_tensor"]).max() <= 1e-5 * np.abs(ref["virial_tensor"]).max()


def test_compression_triggers_rebuild(w_setfl):
    """A perfect bcc crystal at 0 K compressed 18 % along z by the fused
    loader: atoms only move with the box (zero co-moving displacement), so
    only the deformation bound can rebuild the list; the energy and virial
    afterwards equal the FP64 oracle's (no pair missing) and the atoms keep
    their fractional coordinates."""
    from simulon_amd.mechanics import UniaxialTensileLoader
    print("[case] compression rebuild, W bcc n=1024, 200 steps", flush=True)
    pos, H = lat.w_bcc(8)
    mol, model = _w_model(w_setfl, pos, H, "NVE")
    model()
    eng = model.engine
    builds0 = int(eng.nbuild[0])
    frac0 = _state(mol)[1] @ np.linalg.inv(mol.box.H.cpu().numpy())
    loader = UniaxialTensileLoader(mol, axis=2, strain_rate=-1.0)
    loader.attach(model)
    model.run(200)
    strain = loader.strain()
    Hn, x, _ = _state(mol)
    builds = int(eng.nbuild[0]) - builds0
    frac = x @ np.linalg.inv(Hn)
    d = frac - frac0
    d -= np.rint(d)
    print(f"strain {strain:.4f}, rebuilds {builds}, max frac drift {np.abs(d).max():.2e}", flush=True)
    assert strain == pytest.approx(0.999 ** 200 - 1, rel=1e-9)
    assert builds >= 1
    assert np.abs(d).max() <= 1e-6
    eng.compute_forces(virial=True)
    eng.reduce(0, 8)
    xs = eng.positions().cpu().numpy()
    ref = oracle_eam(xs, Hn, np.zeros(len(xs), int), w_setfl, 1.0)
    th = eng.thermo_host()[0]
    print(f"E {th[0]:.6f} vs {ref['energy']:.6f}", flush=True)
    assert abs(th[0] - ref["energy"]) <= 1e-6 * abs(ref["energy"])
    W = sym3(th[2:8])
    assert np.abs(W - ref["virial_tensor"]).max() <= 1e-5 * np.abs(ref["virial_tensor"]).max()
