"""GPU tests of the reference-compatible API (simulon_amd.reader / force /
integrator / barostat / md_model / neighbor), written like the reference's
own script tests (cuda_test/test_eam_fs_parser_regression.py,
cuda_test/test_w_elastic_static.py) plus parity with the reference's outputs
stored in tests/golden/."""
import gzip
import os
import shutil

import numpy as np
import pytest
import torch

from simulon_amd import lattice as lat

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EV_ANG3_TO_BAR = 1_602_176.6208
EV_ANG3_TO_GPA = 160.21766208


def load(name):
    return np.load(os.path.join(G, f"golden_{name}.npz"))


@pytest.fixture(scope="module")
def wre_file(tmp_path_factory):
    p = tmp_path_factory.mktemp("w") / "WRe_YC2.eam.fs"
    with gzip.open(os.path.join(G, "data", "WRe_YC2.eam.fs.gz"), "rb") as fi, open(p, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    return str(p)


@pytest.fixture(scope="module")
def syn_file(tmp_path_factory):
    p = tmp_path_factory.mktemp("g") / "W_synthetic.eam.alloy"
    p.write_bytes(load("eam")["eam_syn_file"].tobytes())
    return str(p)


@pytest.mark.parametrize("case", ["sc", "tri", "liquid"])
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_find_neighbors_equals_reference(case, dt):
    """Pair sets identical to the reference's find_neighbors_gpu_pbc_pytorch +
    reader.py:477-484 on the same inputs, FP32 and FP64."""
    from simulon_amd.box import Box
    from simulon_amd.neighbor import find_neighbors_gpu_pbc
    g = load("pairs")
    pos = torch.tensor(g[f"nl_{case}_{dt}_pos"]).cuda()
    box = Box(torch.tensor(g[f"nl_{case}_H"]), device="cuda")
    rl = torch.tensor(8.5) + torch.tensor(1.0)
    print(f"[case] find_neighbors {case} {dt} n={len(pos)}", flush=True)
    ei, d = find_neighbors_gpu_pbc(pos, rl, box)
    np.testing.assert_array_equal(ei.cpu().numpy(), g[f"nl_{case}_{dt}_ei"].astype(np.int64))
    np.testing.assert_allclose(d.cpu().numpy(), g[f"nl_{case}_{dt}_dist"],
                               rtol=4 * np.finfo(pos.cpu().numpy().dtype).eps)


def _w_reader(pos, H, cutoff, **kw):
    from simulon_amd.reader import AtomFileReader
    return AtomFileReader(atom_types=["W"] * len(pos), positions=pos, box_length=float(H[0, 0]),
                          cutoff=cutoff, device="cuda", skin_thickness=1.0, is_mlp=True,
                          box_vectors=torch.tensor(H), **kw)


def test_eam_fs_parser_regression(wre_file):
    """cuda_test/test_eam_fs_parser_regression.py on the device path."""
    from simulon_amd.eam_parser import EAMParser
    from simulon_amd.force import EAMForceCUDAExt
    parser = EAMParser(wre_file, device=torch.device("cuda"))
    assert parser.is_fs_format
    assert parser.density_splines_by_pair[0][0] is not parser.density_splines_by_pair[0][1]
    pos, H = lat.w_bcc(4, a=3.1652)
    mol = _w_reader(pos, H, parser.cutoff)
    out = EAMForceCUDAExt(parser, mol, use_extension=False)()
    stress = out["virial_tensor"].diag().to(torch.float64) / float(mol.box.volume) * EV_ANG3_TO_BAR
    assert float(torch.max(torch.abs(stress))) < 20.0, stress.tolist()


def test_w_elastic_static(wre_file):
    """cuda_test/test_w_elastic_static.py on the device path."""
    from simulon_amd.eam_parser import EAMParser
    from simulon_amd.force import EAMForceCUDAExt
    parser = EAMParser(wre_file)
    pos, H = lat.w_bcc(4, a=3.1652)
    mol = _w_reader(pos, H, parser.cutoff)
    force = EAMForceCUDAExt(parser, mol, n_r=10_000, n_rho=10_000, use_extension=False)
    h0 = mol.box.H.detach().clone()
    frac0 = mol.coordinates.detach().clone().to(h0.dtype) @ torch.linalg.inv(h0)
    volume0 = float(mol.box.volume)
    eps = 1.0e-3

    def measure():
        out = force()
        s = out["virial_tensor"].diag().to(torch.float64) / float(mol.box.volume) * EV_ANG3_TO_BAR
        return float(out["energy"]), s

    def strain(e11):
        S = torch.eye(3, dtype=h0.dtype, device=h0.device)
        S[0, 0] = 1.0 + e11
        mol.box.H = S @ h0
        mol.update_coordinates((frac0 @ mol.box.H).to(torch.float32))

    e0, _ = measure()
    strain(eps)
    ep, sp = measure()
    strain(-eps)
    em, sm = measure()
    c11_e = ((ep + em - 2.0 * e0) / (volume0 * eps * eps)) * EV_ANG3_TO_GPA
    c11_s = -float(sp[0] - sm[0]) / (2.0 * eps) / 10_000.0
    c12_s = -float(sp[1] - sm[1]) / (2.0 * eps) / 10_000.0
    print(f"C11_energy={c11_e:.2f} C11_stress={c11_s:.2f} C12_stress={c12_s:.2f} GPa", flush=True)
    assert 350.0 < c11_e < 700.0 and 350.0 < c11_s < 700.0 and 100.0 < c12_s < 350.0
    assert abs(c11_e - c11_s) / max(abs(c11_e), 1.0) < 0.08


@pytest.mark.parametrize("mode", ["raw", "shift", "switch"])
def test_lj_backbone_matches_reference(mode):
    from simulon_amd.force import LennardJonesForce
    from simulon_amd.reader import AtomFileReader
    g = load("lj")
    pos = g["lj_pos"]
    mol = AtomFileReader(atom_types=["Ar"] * len(pos), positions=pos, box_length=36.4, cutoff=8.5,
                         device="cuda", parameter={"[0 0]": {"epsilon": 0.0104, "sigma": 3.405}},
                         skin_thickness=1.0, is_fs=(mode == "shift"), is_switch=(mode == "switch"))
    np.testing.assert_array_equal(mol.graph_data.edge_index.cpu().numpy(), g["lj_ei"])
    out = LennardJonesForce(mol)()
    ref = g[f"lj_{mode}_forces"].astype(np.float64)
    f = out["forces"].cpu().numpy()
    # float64 oracle on the same (float32-valued) configuration
    from oracle import lj as olj
    o64 = olj.lj_forward(pos.astype(np.float64), g["lj_ei"], np.zeros(len(pos), int), [[0.0104]],
                         [[3.405]], 8.5, g["lj_H"], mode=mode)["forces"]

    def err(a, b):
        return np.max(np.linalg.norm(a - b, axis=1)) / np.max(np.linalg.norm(b, axis=1))

    e_ours, e_ref = err(f, o64), err(ref, o64)
    print(f"LJ {mode}: ours vs fp64 {e_ours:.2e}, reference fp32 vs fp64 {e_ref:.2e}", flush=True)
    assert e_ours <= 1e-5
    assert err(f, ref) <= e_ref + 1e-5
    assert float(out["energy"]) == pytest.approx(float(g[f"lj_{mode}_energy"]), rel=1e-5)
    np.testing.assert_allclose(out["virial_tensor"].cpu().numpy(), g[f"lj_{mode}_virial_tensor"],
                               atol=1e-4 * np.abs(g[f"lj_{mode}_virial_tensor"]).max())


@pytest.mark.parametrize("case", ["syn", "wre_jit"])
def test_eam_backbone_matches_reference(case, syn_file, wre_file):
    from simulon_amd.eam_parser import EAMParser
    from simulon_amd.force import EAMForceCUDAExt
    g = load("eam")
    pos = g[f"eam_{case}_pos"]
    H = g[f"eam_{case}_H"]
    parser = EAMParser(syn_file if case == "syn" else wre_file)
    mol = _w_reader(pos, H, parser.cutoff)
    out = EAMForceCUDAExt(parser, mol)()
    ref = g[f"eam_{case}_forces"].astype(np.float64)
    f = out["forces"].cpu().numpy()
    err = np.max(np.linalg.norm(f - ref, axis=1)) / np.max(np.linalg.norm(ref, axis=1))
    print(f"eam {case} err vs reference fp32 {err:.2e}", flush=True)
    assert err < 2e-5
    assert float(out["energy"]) == pytest.approx(float(g[f"eam_{case}_energy"]), rel=1e-6)
    # the energy split of eam_force_cu.py:577-578
    for k in ("embedding_energy", "pair_potential"):
        assert float(out[k]) == pytest.approx(float(g[f"eam_{case}_{k}"]), rel=1e-5, abs=1e-4)
    assert float(out["embedding_energy"] + out["pair_potential"]) == \
        pytest.approx(float(out["energy"]), rel=1e-12)


def test_base_model_nvt_thermalises():
    from simulon_amd.force import LennardJonesForce
    from simulon_amd.integrator import VerletIntegrator
    from simulon_amd.md_model import BaseModel, SumBackboneInterface
    from simulon_amd.reader import AtomFileReader
    torch.manual_seed(0)
    pos, H = lat.lj_argon_nvt(seed=1)
    mol = AtomFileReader(atom_types=["Ar"] * len(pos), positions=pos, box_length=36.4, cutoff=8.5,
                         device="cuda", parameter={"[0 0]": {"epsilon": 0.0104, "sigma": 3.405}},
                         skin_thickness=1.0)
    ff = LennardJonesForce(mol)
    integ = VerletIntegrator(mol, dt=0.002, ensemble="NVT", temperature=(94.4, 94.4), gamma=10.0)
    model = BaseModel(SumBackboneInterface([ff], mol), integ, mol)
    out = model()
    for k in ["forces", "energy", "virial", "virial_tensor", "updated_coordinates",
              "kinetic_energy", "temperature"]:
        assert k in out
    temps = []
    for _ in range(10):
        out = model.run(100)
        temps.append(float(out["temperature"]))
    T = np.mean(temps[3:])
    print(f"NVT mean T {T:.1f} K, list builds {mol.profile['neighbor_builds']}", flush=True)
    assert abs(T - 94.4) < 8.0
    assert torch.isfinite(out["forces"]).all()


def test_base_model_npt_barostat_moves_box():
    from simulon_amd.barostat import BerendsenBarostat
    from simulon_amd.force import LennardJonesForce
    from simulon_amd.integrator import VerletIntegrator
    from simulon_amd.md_model import BaseModel, SumBackboneInterface
    from simulon_amd.reader import AtomFileReader
    pos, H = lat.lj_argon_npt_tri(ncell=6, seed=3)
    mol = AtomFileReader(atom_types=["Ar"] * len(pos), positions=pos, box_length=float(H[0, 0]),
                         cutoff=8.5, device="cuda", box_vectors=torch.tensor(H),
                         parameter={"[0 0]": {"epsilon": 0.0104, "sigma": 3.405}},
                         skin_thickness=1.0)
    ff = LennardJonesForce(mol)
    integ = VerletIntegrator(mol, dt=0.002, ensemble="NPT", temperature=(90.0, 90.0), gamma=10.0)
    baro = BerendsenBarostat(mol, target_pressure=1.0, tau_p=1.0)
    model = BaseModel(SumBackboneInterface([ff], mol), integ, mol, barostat=baro)
    V0 = mol.box.volume
    out = model()
    assert "pressure" in out
    out = model.run(200)
    V1 = mol.box.volume
    P = float(out["pressure"])
    print(f"NPT V0 {V0:.1f} V1 {V1:.1f} P {P:.1f} bar", flush=True)
    assert V1 != V0 and np.isfinite(P)
    # box, coordinates and the device record stay consistent
    H1 = mol.box.H.cpu().numpy()
    np.testing.assert_allclose(H1 / H1[0, 0], H / H[0, 0], rtol=1e-12)


def test_maxwell_boltzmann_initialisation():
    """reader.py:319-339: velocities ~ N(0, kB T / (m 1.036427e-4)) per
    component with the centre-of-mass velocity removed, both for the
    engine's own initialisation (ReplicaSpec without velocities, at t_init)
    and for AtomFileReader.set_maxwell_boltzmann_velocity (called by the NVT
    integrator with T_init, integrator.py:82)."""
    from simulon_amd.engine import KB, MB_MASS_CONV, ReplicaEngine, ReplicaSpec
    from simulon_amd.integrator import VerletIntegrator
    from simulon_amd.reader import AtomFileReader
    pos, H = lat.w_bcc(20)
    n = len(pos)

    def check(v, mass_amu, T, tag):
        v = v.astype(np.float64)
        m = mass_amu * MB_MASS_CONV
        sig = np.sqrt(KB * T / m)
        com = np.abs(v.sum(0)) / (n * sig)
        var = (v * v).mean(0) / sig ** 2
        print(f"{tag}: |COM| / (N sigma) {com.max():.2e}, variance ratio {var}", flush=True)
        assert com.max() < 1e-6
        # 3 sigma of the sample variance of n draws is 3 sqrt(2/n) = 3.4% at n = 16000
        assert np.all(np.abs(var - 1.0) < 0.035)
        # and the normal shape: fourth moment 3 sigma^4
        k4 = (v ** 4).mean(0) / sig ** 4 / var ** 2
        assert np.all(np.abs(k4 - 3.0) < 0.2)

    eng = ReplicaEngine([ReplicaSpec(positions=pos, box=H, ensemble="NVT", temperature=900.0,
                                     t_init=300.0)], [lat.W_MASS], cutoff=5.5, skin=1.0, seed=5)
    check(eng.to_caller(eng.vel).cpu().numpy(), lat.W_MASS, 300.0, "engine init")
    torch.manual_seed(3)
    mol = AtomFileReader(atom_types=["W"] * n, positions=pos, box_length=float(H[0, 0]), cutoff=5.5,
                         device="cuda", skin_thickness=1.0)
    VerletIntegrator(mol, dt=0.001, ensemble="NVT", temperature=(450.0, 300.0), gamma=10.0)
    check(mol.atom_velocities.cpu().numpy(), lat.W_MASS, 450.0, "reader init")


class _Wall(torch.nn.Module):
    """A host backbone: a repulsive sphere U = K/3 (R - d)^3 for d < R around
    c, F_i = K (R - d)^2 (x_i - c)/d (the form of the reference's
    core/mechanics/indenter.py), evaluated in torch on the caller's
    coordinates."""

    def __init__(self, mol, center, radius=4.0, k=0.05):
        super().__init__()
        self.mol, self.c, self.R, self.K = mol, torch.tensor(center, dtype=torch.float64), radius, k

    @staticmethod
    def terms(x, c, R, K):
        d_vec = x - c
        d = np.linalg.norm(d_vec, axis=1)
        p = np.maximum(R - d, 0.0)
        f = (K * p ** 2 / np.maximum(d, 1e-12))[:, None] * d_vec
        return K / 3.0 * float((p ** 3).sum()), f, np.einsum("ni,nj->ij", f, x)

    def forward(self):
        x = self.mol.coordinates.double()
        c = self.c.to(x.device)
        d_vec = x - c
        d = d_vec.norm(dim=1).clamp_min(1e-12)
        p = (self.R - d).clamp_min(0.0)
        f = (self.K * p ** 2 / d)[:, None] * d_vec
        W = torch.einsum("ni,nj->ij", f, x)
        return {"energy": self.K / 3.0 * (p ** 3).sum(), "forces": f.float(),
                "virial": torch.trace(W), "virial_tensor": W}


def test_base_model_sums_device_and_host_backbones():
    """SumBackboneInterface([device LJ, host wall]) under BaseModel
    (md_model.py:19-52): the step integrates the summed forces, and the
    outputs carry the summed energy, forces and virial. 100 NVE steps
    against the float64 oracle trajectory of LJ + wall."""
    from oracle import cell as ocell
    from simulon_amd.force import LennardJonesForce
    from simulon_amd.integrator import VerletIntegrator
    from simulon_amd.md_model import BaseModel, SumBackboneInterface
    from simulon_amd.reader import AtomFileReader
    torch.manual_seed(1)
    pos, H = lat.lj_argon_nvt(seed=5)
    mol = AtomFileReader(atom_types=["Ar"] * len(pos), positions=pos, box_length=36.4, cutoff=8.5,
                         device="cuda", parameter={"[0 0]": {"epsilon": 0.0104, "sigma": 3.405}},
                         skin_thickness=1.0)
    integ = VerletIntegrator(mol, dt=0.002, ensemble="NVE")
    mol.atom_velocities = torch.zeros((len(pos), 3))
    c = [20.02, 20.02, 20.02]        # the centre of a lattice cube: 8 atoms at 3.15 A
    wall = _Wall(mol, c)
    model = BaseModel(SumBackboneInterface([LennardJonesForce(mol), wall], mol), integ, mol)
    assert model.extra == [wall]
    eng = model.engine
    x = eng.to_caller(eng.positions(torch.float64, unwrap=True)).cpu().numpy()
    v = np.zeros_like(x)
    m = float(eng.types_struct.mass[0])

    def total(x):
        o = ocell.lj(x, H, None, [[lat.AR_EPS]], [[lat.AR_SIG]], 8.5, 1.0)
        ew, fw, ww = _Wall.terms(x, np.array(c), wall.R, wall.K)
        return o["energy"] + ew, o["forces"] + fw, o["virial_tensor"] + ww

    e, f, W = total(x)
    worst_f = worst_e = 0.0
    for s in range(100):
        out = model()
        vh = v + 0.5 * f / m * 0.002
        x = x + vh * 0.002
        e, f, W = total(x)
        v = vh + 0.5 * f / m * 0.002
        if s in (0, 99):
            fe = np.max(np.linalg.norm(out["forces"].cpu().numpy() - f, axis=1)) / \
                np.max(np.linalg.norm(f, axis=1))
            worst_f = max(worst_f, fe)
            # float32 pair terms of 1000 atoms, and a float32-coordinate wall
            assert np.max(np.abs(out["virial_tensor"].cpu().numpy() - W)) <= 2e-5 * np.max(np.abs(W))
        worst_e = max(worst_e, abs(float(out["energy"]) - e))
    ew0 = _Wall.terms(x, np.array(c), wall.R, wall.K)[0]
    print(f"LJ + wall: force err {worst_f:.2e}, max |dE| {worst_e:.2e} eV, wall energy {ew0:.3e} eV",
          flush=True)
    assert ew0 > 1e-3                       # the wall acts on the trajectory
    assert worst_f <= 1e-5
    assert worst_e <= 1e-3
