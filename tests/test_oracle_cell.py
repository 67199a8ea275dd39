"""The C cell-list oracle (oracle/cell_md.c) against the numpy oracle it
restates: pair counts exactly, forces, energies, virials and densities to
1e-12 relative, on orthogonal and triclinic systems (CPU only)."""
import os

import numpy as np
import pytest

from oracle import cell as ocell
from oracle import eam as oeam
from oracle import lj as olj
from oracle import neighbor as onb
from simulon_amd import lattice as lat

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def syn_file(tmp_path_factory):
    p = tmp_path_factory.mktemp("g") / "W_synthetic.eam.alloy"
    p.write_bytes(np.load(os.path.join(G, "golden_eam.npz"))["eam_syn_file"].tobytes())
    return str(p)


@pytest.mark.parametrize("box", ["sc", "tri"])
@pytest.mark.parametrize("mode", ["raw", "shift", "switch"])
def test_cell_lj_equals_numpy_oracle(box, mode):
    if box == "sc":
        pos, H = lat.lj_argon_nvt(seed=11)
    else:
        pos, H = lat.lj_argon_npt_tri(ncell=5, seed=12)
    types = np.zeros(len(pos), int)
    ei, _ = onb.brute_force_half_list(pos, H, onb.list_radius_sq(8.5, 1.0))
    ref = olj.lj_forward(pos, ei, types, [[0.0104]], [[3.405]], 8.5, H, mode=mode)
    got = ocell.lj(pos, H, types, [[0.0104]], [[3.405]], 8.5, 1.0, mode=mode)
    assert got["npairs"] == ei.shape[1]
    assert rel(got["forces"], ref["forces"]) < 1e-12
    assert got["energy"] == pytest.approx(ref["energy"], rel=1e-12)
    assert got["virial"] == pytest.approx(ref["virial"], rel=1e-12)
    assert rel(got["virial_tensor"], ref["virial_tensor"]) < 1e-12


def test_cell_lj_two_types_small_box():
    """Two atom types and a 2-cell grid along one axis (offsets {0, 1})."""
    rng = np.random.default_rng(5)
    H = np.array([[20.0, 0, 0], [3.0, 30.0, 0], [1.0, 2.0, 31.0]])
    pos = rng.uniform(0, 1, (300, 3)) @ H
    types = rng.integers(0, 2, 300)
    eps = [[0.0104, 0.012], [0.012, 0.02]]
    sig = [[3.405, 3.2], [3.2, 3.0]]
    ei, _ = onb.brute_force_half_list(pos, H, onb.list_radius_sq(8.5, 1.0))
    ref = olj.lj_forward(pos, ei, types, eps, sig, 8.5, H)
    got = ocell.lj(pos, H, types, eps, sig, 8.5, 1.0)
    assert got["npairs"] == ei.shape[1]
    assert rel(got["forces"], ref["forces"]) < 1e-12
    assert got["energy"] == pytest.approx(ref["energy"], rel=1e-12)


@pytest.mark.parametrize("box", ["bcc", "tri"])
def test_cell_eam_equals_numpy_oracle(box, syn_file):
    pos, H = lat.w_bcc(6, seed=3, amp=0.1)
    if box == "tri":
        S = np.array([[1.0, 0, 0], [0.08, 1.0, 0], [0.04, -0.05, 1.0]])
        pos, H = pos @ S, H @ S
    tab = oeam.EAMTables(oeam.EAMFile(syn_file), dtype=np.float64)
    types = np.zeros(len(pos), int)
    ei, _ = onb.brute_force_half_list(pos, H, onb.list_radius_sq(tab.cutoff, 1.0))
    ref = oeam.eam_forward(pos, ei, types, tab, H)
    got = ocell.eam(pos, H, types, tab, 1.0)
    assert got["npairs"] == ei.shape[1]
    assert rel(got["forces"], ref["forces"]) < 1e-12
    assert rel(got["rho"], ref["rho"]) < 1e-12
    assert rel(got["dF"], ref["dF"]) < 1e-12
    for k in ("energy", "virial", "embedding_energy", "pair_potential"):
        assert got[k] == pytest.approx(ref[k], rel=1e-12)
    assert rel(got["virial_tensor"], ref["virial_tensor"]) < 1e-12


def test_cell_rows_equal_brute_force():
    pos, H = lat.lj_argon_npt_tri(ncell=5, seed=13)
    rl2 = onb.list_radius_sq(8.5, 1.0)
    ei, _ = onb.brute_force_half_list(pos, H, rl2)
    counts, nb = onb.full_list_from_half(ei, len(pos))
    start = np.concatenate([[0], np.cumsum(counts)])
    idx = np.array([0, 7, 100, 499, len(pos) - 1])
    k, j = ocell.rows(pos, H, rl2, idx)
    for q, i in enumerate(idx):
        np.testing.assert_array_equal(j[k == q], nb[start[i]:start[i + 1]])


def test_cell_lj_cutoff_ties():
    """Pairs within tie_tol of rc are counted on both atoms, others not."""
    H = np.diag([40.0, 40.0, 40.0])
    for r, want in [(8.5 - 3e-6, 1), (8.5 + 3e-6, 1), (8.4, 0), (8.5 + 2e-5, 0)]:
        pos = np.array([[10.0, 10.0, 10.0], [10.0 + r, 10.0, 10.0]])
        out = ocell.lj(pos, H, None, [[0.0104]], [[3.405]], 8.5, 1.0)
        assert out["tie"].tolist() == [want, want]
        assert out["ncut"] == (1 if r < 8.5 else 0)
