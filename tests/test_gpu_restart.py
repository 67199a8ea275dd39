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
