"""The oracle against the reference's own outputs (tests/golden/*.npz, written
by tests/golden/make_golden.py from the reference's pure-Python modules)."""
import os

import numpy as np
import pytest

from oracle import box as obox
from oracle import eam as oeam
from oracle import integrate as oint
from oracle import lj as olj
from oracle import neighbor as onb

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, f"golden_{name}.npz"))


@pytest.mark.parametrize("box", ["ortho", "tri"])
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_min_image_bitwise(box, dt):
    g = load("minimage")
    r = g[f"mi_{box}_{dt}_in"]
    out = obox.minimum_image(r, g[f"mi_{box}_H"], g[f"mi_{box}_Hinv"])
    ref = g[f"mi_{box}_{dt}_out"]
    if box == "ortho":
        np.testing.assert_array_equal(out, ref)
    else:
        # triclinic: the reference multiplies by H^-1 and H with a GEMM whose
        # summation order (and FMA use) is the BLAS's; the oracle's fixed order
        # agrees to within a few roundings of the box length
        L = np.abs(g[f"mi_{box}_H"]).max()
        eps = np.finfo(r.dtype).eps
        assert np.abs(out - ref).max() <= 4 * eps * L
        assert np.mean(np.all(out == ref, axis=1)) > 0.3


@pytest.mark.parametrize("case", ["sc", "tri", "liquid"])
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_pair_sets_equal_reference(case, dt):
    g = load("pairs")
    pos = g[f"nl_{case}_{dt}_pos"]
    H = g[f"nl_{case}_H"]
    Hinv = np.linalg.inv(H)
    ref_ei = g[f"nl_{case}_{dt}_ei"].astype(np.int64)
    ei, d = onb.brute_force_half_list(pos, H, onb.list_radius_sq(8.5, 1.0), Hinv)
    np.testing.assert_array_equal(ei, ref_ei)
    np.testing.assert_allclose(d, g[f"nl_{case}_{dt}_dist"],
                               rtol=4 * np.finfo(pos.dtype).eps, atol=0)


@pytest.mark.parametrize("mode", ["raw", "shift", "switch"])
def test_lj_matches_reference(mode):
    g = load("lj")
    pos = g["lj_pos"]
    H = g["lj_H"]
    ei = g["lj_ei"]
    ref_f = g[f"lj_{mode}_forces"].astype(np.float64)
    # float32 restatement vs the reference's float32 torch result
    o32 = olj.lj_forward(pos, ei, np.zeros(len(pos), int), [[0.0104]], [[3.405]], 8.5, H, mode=mode)
    scale = np.abs(ref_f).max()
    # same formulas in float32; per-atom sums of ~56 pair terms in another order
    assert np.abs(o32["forces"] - ref_f).max() <= 1e-5 * scale
    # float64 oracle on the same configuration: the reference's float32 is within 1e-5
    o64 = olj.lj_forward(pos.astype(np.float64), ei, np.zeros(len(pos), int), [[0.0104]],
                         [[3.405]], 8.5, H, mode=mode)
    err = np.max(np.linalg.norm(ref_f - o64["forces"], axis=1)) / np.max(
        np.linalg.norm(o64["forces"], axis=1))
    assert err < 1e-4
    assert o64["energy"] == pytest.approx(float(g[f"lj_{mode}_energy"]), rel=1e-5)
    assert o64["virial"] == pytest.approx(float(g[f"lj_{mode}_virial"]), rel=1e-4)
    np.testing.assert_allclose(o64["virial_tensor"], g[f"lj_{mode}_virial_tensor"],
                               atol=1e-4 * np.abs(o64["virial_tensor"]).max())


@pytest.fixture(scope="module")
def syn_file(tmp_path_factory):
    g = load("eam")
    p = tmp_path_factory.mktemp("g") / "W_synthetic.eam.alloy"
    p.write_bytes(g["eam_syn_file"].tobytes())
    return str(p)


@pytest.fixture(scope="module")
def wre_file(tmp_path_factory):
    import gzip
    import shutil
    p = tmp_path_factory.mktemp("w") / "WRe_YC2.eam.fs"
    with gzip.open(os.path.join(G, "data", "WRe_YC2.eam.fs.gz"), "rb") as fi, open(p, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    return str(p)


def test_eam_tables_bitwise(syn_file):
    """float32 tables equal the reference's (eam_force_cu.py:334-395) bit for bit."""
    g = load("eam")
    t = oeam.EAMTables(oeam.EAMFile(syn_file), dtype=np.float32)
    np.testing.assert_array_equal(t.dens, g["eam_syn_density_table"])
    np.testing.assert_array_equal(t.ddens, g["eam_syn_density_deriv_table"])
    np.testing.assert_array_equal(t.phi, g["eam_syn_pair_table"])
    np.testing.assert_array_equal(t.dphi, g["eam_syn_pair_deriv_table"])
    np.testing.assert_array_equal(t.F[0], g["eam_syn_embed_table"])
    np.testing.assert_array_equal(t.dF[0], g["eam_syn_embed_deriv_table"])


@pytest.mark.parametrize("case", ["syn", "wre", "wre_jit"])
def test_eam_matches_reference(case, syn_file, wre_file):
    g = load("eam")
    path = syn_file if case == "syn" else wre_file
    f = oeam.EAMFile(path)
    pos = g[f"eam_{case}_pos"]
    H = g[f"eam_{case}_H"]
    ei, _ = onb.brute_force_half_list(pos, H, onb.list_radius_sq(f.cutoff, 1.0))
    ref_f = g[f"eam_{case}_forces"].astype(np.float64)
    t32 = oeam.EAMTables(f, dtype=np.float32)
    o32 = oeam.eam_forward(pos, ei, np.zeros(len(pos), int), t32, H)
    # perfect lattice ("wre"): net forces are round-off, so the bar is absolute
    scale = max(np.abs(ref_f).max(), 0.1)
    assert np.abs(o32["forces"] - ref_f).max() <= 1e-4 * scale
    assert o32["energy"] == pytest.approx(float(g[f"eam_{case}_energy"]), rel=1e-6)
    # float32 sums of cancelling pair terms: relative bar, floored at 2e-4 eV
    # (0.16 bar of stress on the 128-atom cell)
    Wref = g[f"eam_{case}_virial_tensor"]
    np.testing.assert_allclose(o32["virial_tensor"], Wref, rtol=0,
                               atol=max(1e-5 * np.abs(Wref).max(), 2e-4))
    assert o32["virial"] == pytest.approx(float(g[f"eam_{case}_virial"]), rel=1e-5, abs=2e-4)
    assert o32["embedding_energy"] == pytest.approx(float(g[f"eam_{case}_embedding_energy"]),
                                                    rel=1e-6)


@pytest.mark.parametrize("ens", ["NVE", "NVT"])
def test_integrator_matches_reference(ens):
    g = load("integrator")
    x0, v0, f0, f1 = g["int_x0"], g["int_v0"], g["int_f0"], g["int_f1"]
    m = oint.internal_mass(np.full(len(x0), 39.948))
    np.testing.assert_array_equal(m, g[f"int_{ens}_mass_internal"][:, 0])
    vh, x1 = oint.first_half(x0, v0, f0, m, 0.002)
    np.testing.assert_array_equal(vh, g[f"int_{ens}_vel_half"])
    np.testing.assert_array_equal(x1, g[f"int_{ens}_x1"])
    if ens == "NVE":
        v1, ke, T = oint.second_half(vh, f1, m, 0.002)
    else:
        v1, ke, T = oint.second_half(vh, f1, m, 0.002, gamma=10.0, T=94.4, noise=g["int_NVT_noise"])
        _, sig = oint.langevin_factors(m.astype(np.float64), 0.002, 10.0, 94.4)
        np.testing.assert_allclose(sig, g["int_NVT_rff"][:, 0], rtol=1e-6)
    np.testing.assert_allclose(v1, g[f"int_{ens}_v1"], rtol=1e-6, atol=1e-6)
    assert ke == pytest.approx(float(g[f"int_{ens}_ke"]), rel=1e-6)
    assert T == pytest.approx(float(g[f"int_{ens}_T"]), rel=1e-6)


def test_berendsen_matches_reference():
    g = load("barostat")
    H0 = g["baro_H0"]
    mu, P = oint.berendsen_mu(g["baro_K"], g["baro_W"], obox.volume(H0), 0.002, 1.0, 1.0)
    assert P == pytest.approx(float(g["baro_P"]), rel=1e-6)
    np.testing.assert_allclose(H0 * mu, g["baro_H1"], rtol=1e-14)
    np.testing.assert_allclose(g["baro_x0"] * np.float32(mu), g["baro_x1"], rtol=1e-6)
    np.testing.assert_allclose(g["baro_v0"] * np.float32(mu), g["baro_v1"], rtol=1e-6)
