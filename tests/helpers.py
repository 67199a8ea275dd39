"""Shared helpers for the parity tests: oracle evaluations and error metrics."""
import numpy as np

from oracle import box as obox
from oracle import eam as oeam
from oracle import lj as olj
from oracle import neighbor as onb


def force_error(f_test, f_ref):
    """max_i |F_test_i - F_ref_i| / max_i |F_ref_i| (the relative max-atom error)."""
    f_test = np.asarray(f_test, np.float64)
    f_ref = np.asarray(f_ref, np.float64)
    den = np.max(np.linalg.norm(f_ref, axis=1))
    return float(np.max(np.linalg.norm(f_test - f_ref, axis=1)) / den)


def pair_set(edge_index):
    ei = np.asarray(edge_index)
    return set(zip(ei[0].tolist(), ei[1].tolist()))


def oracle_half_list(pos, H, cutoff, skin, dtype=np.float64):
    rl2 = onb.list_radius_sq(cutoff, skin)
    Hinv = np.linalg.inv(np.asarray(H, np.float64))
    pos = np.asarray(pos, dtype)
    if len(pos) <= 6000:
        return onb.brute_force_half_list(pos, H, rl2, Hinv)
    return onb.cell_half_list(pos, H, rl2, Hinv)


def oracle_lj(pos, H, types, eps, sig, rc, skin, mode="raw"):
    ei, _ = oracle_half_list(pos, H, rc, skin)
    return olj.lj_forward(np.asarray(pos, np.float64), ei, types, eps, sig, rc, H, mode=mode)


def oracle_eam(pos, H, types, eam_path, skin, n_r=8192, n_rho=4096):
    f = oeam.EAMFile(eam_path)
    tab = oeam.EAMTables(f, n_r, n_rho, dtype=np.float64)
    ei, _ = oracle_half_list(pos, H, tab.cutoff, skin)
    return oeam.eam_forward(np.asarray(pos, np.float64), ei, types, tab, H)


def sym3(w6):
    xx, yy, zz, xy, xz, yz = w6
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])


def lj_force_error(f_test, ref, eps, sig, rc):
    """force_error against the C oracle's LJ output `ref`, where atoms with a
    pair within the oracle's tie tolerance of rc (ref["tie"]) are held to
    their own bound: float32 may put such a pair on the other side of the
    r < rc test, and raw LJ jumps by |Fmag(rc)| there. Returns (max-atom
    relative error over the other atoms, number of tie atoms)."""
    from oracle import lj as olj
    f_test = np.asarray(f_test, np.float64)
    fr = np.asarray(ref["forces"], np.float64)
    den = np.max(np.linalg.norm(fr, axis=1))
    d = np.linalg.norm(f_test - fr, axis=1)
    tie = np.asarray(ref["tie"]) > 0
    jump = abs(olj.lj_two_atom(rc, eps, sig)[1])
    assert np.all(d[tie] <= np.asarray(ref["tie"])[tie] * jump * (1 + 1e-3) + 1e-5 * den)
    return float(np.max(d[~tie]) / den), int(tie.sum())
