"""Checkpoint / restart round trip (simulon_amd.restart; the reference's
io_utils/restart.py:32-149): a run interrupted by save + load into a fresh
model continues the uninterrupted trajectory."""
import numpy as np
import pytest
import torch

from simulon_amd import lattice as lat

pytestmark = pytest.mark.gpu


def _model(ensemble, seed):
    from simulon_amd.force import LennardJonesForce
    from simulon_amd.integrator import VerletIntegrator
    from simulon_amd.md_model import BaseModel, SumBackboneInterface
    from simulon_amd.reader import AtomFileReader
    pos, H = lat.lj_argon_nvt(seed=4)
    mol = AtomFileReader(atom_types=["Ar"] * len(pos), positions=pos, box_length=36.4, cutoff=8.5,
                         device="cuda", parameter={"[0 0]": {"epsilon": 0.0104, "sigma": 3.405}},
                         skin_thickness=1.0)
    ff = LennardJonesForce(mol)
    kw = dict(temperature=(94.4, 94.4), gamma=10.0) if ensemble != "NVE" else {}
    integ = VerletIntegrator(mol, dt=0.002, ensemble=ensemble, seed=seed, **kw)
    baro = None
    if ensemble == "NPT":           # anisotropic: its strain rates travel in the file
        from simulon_amd.barostat import AnisotropicNPTBarostat
        baro = AnisotropicNPTBarostat(mol, target_pressure_bar=[0.0, 50.0, -50.0],
                                      temperature_k=94.4, compressibility_bar_inv=3.2e-6)
    return mol, BaseModel(SumBackboneInterface([ff], mol), integ, mol, barostat=baro)


@pytest.mark.parametrize("ensemble", ["NVE", "NVT", "NPT"])
def test_restart_continues_trajectory(ensemble, tmp_path):
    from simulon_amd.restart import load_checkpoint, save_checkpoint
    print(f"[case] restart {ensemble}", flush=True)
    torch.manual_seed(0)
    mol_a, model_a = _model(ensemble, seed=11)
    model_a()
    model_a.run(40)
    path = tmp_path / "restart.pt"
    save_checkpoint(model_a, 41, path)
    x_saved = model_a.engine.to_caller(model_a.engine.positions(torch.float64, unwrap=True)).cpu()
    v_saved = mol_a.atom_velocities.cpu().clone()
    out_a = model_a.run(40)
    x_a = model_a.engine.to_caller(model_a.engine.positions(torch.float64, unwrap=True)).cpu()

    # a fresh model with other velocities and seed: everything comes from the file
    mol_b, model_b = _model(ensemble, seed=99)
    nxt = load_checkpoint(model_b, path)
    assert nxt == 42
    x_b0 = model_b.engine.to_caller(model_b.engine.positions(torch.float64, unwrap=True)).cpu()
    assert torch.equal(x_b0, x_saved)
    assert torch.equal(mol_b.atom_velocities.cpu(), v_saved)
    out_b = model_b.run(40)
    x_b = model_b.engine.to_caller(model_b.engine.positions(torch.float64, unwrap=True)).cpu()
    if ensemble == "NPT":
        ea, eb = model_a.barostat.eta_dot.cpu(), model_b.barostat.eta_dot.cpu()
        print(f"NPT eta_dot {ea.tolist()} vs {eb.tolist()}", flush=True)
        assert float(ea.abs().max()) > 0 and torch.allclose(ea, eb, rtol=1e-5, atol=1e-9)
        assert torch.allclose(mol_a.box.H.cpu(), mol_b.box.H.cpu(), rtol=1e-9, atol=0)
    d = float((x_a - x_b).norm(dim=1).max())
    dE = abs(float(out_a["energy"]) - float(out_b["energy"]))
    print(f"{ensemble}: max |dx| after 40 steps {d:.2e} A, |dE| {dE:.2e} eV", flush=True)
    # the continued runs differ only by float32 summation order (the restart
    # rebuilds the list), not by noise or state
    assert d < 1e-4
    assert dE < 1e-5 * abs(float(out_a["energy"]))
    assert np.isfinite(float(out_b["temperature"]))


@pytest.mark.parametrize("name", ["boxH", "boxlen"])
def test_load_reference_checkpoint(name):
    """A checkpoint written by the reference's own save_checkpoint
    (tests/golden/restart_ref_*.pt; io_utils/restart.py:32-78) loads into a
    model built on other coordinates and another box, and the run continues
    along the float64 oracle's NVE trajectory from the file's state: its
    coordinates and velocities, its box ('box_H', or the scalar
    'box_length'), and its force_cache as the forces of the first step
    (restart.py:128-136)."""
    import os
    from oracle import cell as ocell
    from oracle import integrate as oint
    from simulon_amd.force import LennardJonesForce
    from simulon_amd.integrator import VerletIntegrator
    from simulon_amd.md_model import BaseModel, SumBackboneInterface
    from simulon_amd.reader import AtomFileReader
    from simulon_amd.restart import load_checkpoint
    print(f"[case] reference checkpoint {name}", flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        f"restart_ref_{name}.pt")
    ck = torch.load(path, map_location="cpu", weights_only=True)
    pos, _ = lat.lj_argon_nvt(seed=4)
    mol = AtomFileReader(atom_types=["Ar"] * len(pos), positions=pos * (37.0 / 36.4),
                         box_length=37.0, cutoff=8.5, device="cuda",
                         parameter={"[0 0]": {"epsilon": 0.0104, "sigma": 3.405}},
                         skin_thickness=1.0)
    integ = VerletIntegrator(mol, dt=0.002, ensemble="NVE")
    model = BaseModel(SumBackboneInterface([LennardJonesForce(mol)], mol), integ, mol)
    assert load_checkpoint(model, path) == 42
    H = ck["box_H"].numpy() if "box_H" in ck else np.eye(3) * float(ck["box_length"])
    np.testing.assert_array_equal(mol.box.H.cpu().numpy(), H)
    x0 = ck["coordinates"].numpy().astype(np.float64)
    v0 = ck["velocities"].numpy()
    f0 = ck["force_cache"].numpy()
    m32 = np.full(len(x0), np.float32(model.engine.types_struct.mass[0]), np.float32)
    # step 1 from the file's forces, exactly as the oracle's first half
    out = model()
    vh, x1 = oint.first_half(x0, v0, f0, m32, 0.002)
    np.testing.assert_allclose(out["updated_coordinates"].cpu().numpy().astype(np.float64), x1,
                               rtol=0, atol=5e-6)
    xe = model.engine.to_caller(model.engine.positions(torch.float64, unwrap=True)).cpu().numpy()
    np.testing.assert_allclose(xe, x1, rtol=0, atol=1e-11)
    # 60 more steps against the float64 oracle trajectory
    m = float(m32[0])

    def lj(x):
        return ocell.lj(x, H, None, [[lat.AR_EPS]], [[lat.AR_SIG]], 8.5, 1.0)

    o = lj(x1)
    f = o["forces"]
    v = vh.astype(np.float64) + 0.5 * f / m * 0.002
    x = x1
    worst = abs(float(out["energy"]) - o["energy"])
    for _ in range(60):
        out = model()
        vh = v + 0.5 * f / m * 0.002
        x = x + vh * 0.002
        o = lj(x)
        f = o["forces"]
        v = vh + 0.5 * f / m * 0.002
        worst = max(worst, abs(float(out["energy"]) - o["energy"]),
                    abs(float(out["kinetic_energy"]) - 0.5 * m * float((v * v).sum())))
    xe = model.engine.to_caller(model.engine.positions(torch.float64, unwrap=True)).cpu().numpy()
    dx = float(np.abs(xe - x).max())
    print(f"reference checkpoint {name}: 61 steps, max |dE| {worst:.2e} eV, |dx| {dx:.2e} A",
          flush=True)
    assert worst <= 1e-3
    assert dx <= 1e-4
