"""Oracle (test infrastructure only): ctypes front end of cell_md.c, the
float64 cell-list restatement of the half-list LJ and EAM for large systems.

Same quantities as oracle.lj.lj_forward and oracle.eam.eam_forward on the
pair set of oracle.neighbor (pairs i < j with |mi(x_i - x_j)|^2 < r_l^2), in
float64, computed in C so that million-atom configurations take seconds.
`rows()` returns the full neighbor rows of chosen atoms.

pinned by: equality with the numpy oracle (pair counts exactly, forces,
energies, virials and densities to 1e-12 relative) on small orthogonal and
triclinic systems (tests/test_oracle_cell.py).

The library is compiled from cell_md.c by build() (gcc); __graft_entry__.build()
calls it.
"""
import ctypes
import os
import subprocess

import numpy as np

from . import box as obox
from . import neighbor as onb

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cell_md.c")
LIB = os.path.join(HERE, "liboracle_cell.so")
_lib = None


def build():
    tmp = LIB + ".tmp"
    subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared", "-o", tmp, SRC, "-lm"],
                   check=True)
    os.replace(tmp, LIB)
    return LIB


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        lib.orc_lj.argtypes = [ctypes.c_int, P, P, ctypes.c_int, P, P, ctypes.c_double,
                               ctypes.c_int, ctypes.c_double, P, P, ctypes.c_int,
                               ctypes.c_double, P, P, P, ctypes.c_double]
        lib.orc_lj.restype = ctypes.c_int
        lib.orc_eam.argtypes = [ctypes.c_int, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double, P, P, P, P, ctypes.c_int, P,
                                P, P, P, P, P, P, ctypes.c_int, ctypes.c_double, P, P, P, P]
        lib.orc_eam.restype = ctypes.c_int
        lib.orc_rows.argtypes = [ctypes.c_int, P, P, P, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, P, ctypes.c_longlong, P, P]
        lib.orc_rows.restype = ctypes.c_longlong
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _box(H):
    H = np.ascontiguousarray(H, np.float64)
    Hi = np.ascontiguousarray(obox.inverse(H))
    return H, Hi, int(obox.is_orthogonal(H))


def _sym(w):
    return np.asarray(w, np.float64).reshape(3, 3)


def lj(pos, H, types, eps_table, sig_table, rc, skin, mode="raw", switch_ratio=0.9,
       tie_tol=1e-5):
    """dict(energy, forces, virial, virial_tensor, npairs, ncut, tie) in
    float64. tie[i]: pairs of atom i with |r - rc| < tie_tol, whose r < rc
    decision float32 arithmetic may take the other way (raw LJ jumps by
    Fmag(rc) at the cutoff)."""
    lib = _load()
    pos = np.ascontiguousarray(pos, np.float64)
    n = len(pos)
    H, Hi, ortho = _box(H)
    eps = np.ascontiguousarray(eps_table, np.float64)
    sig = np.ascontiguousarray(sig_table, np.float64)
    nt = eps.shape[0]
    t = None if types is None else np.ascontiguousarray(types, np.int32)
    forces = np.zeros((n, 3), np.float64)
    out = np.zeros(16, np.float64)
    tie = np.zeros(n, np.int32)
    rl2 = float(onb.list_radius_sq(rc, skin))
    m = {"raw": 0, "shift": 1, "switch": 2}[mode]
    rv = lib.orc_lj(n, _p(pos), None if t is None else _p(t), nt, _p(eps), _p(sig), float(rc), m,
                    switch_ratio, _p(H), _p(Hi), ortho, rl2, _p(forces), _p(out), _p(tie),
                    float(tie_tol))
    if rv != 0:
        raise MemoryError("orc_lj")
    return {"energy": out[0], "virial": out[1], "virial_tensor": _sym(out[2:11]),
            "forces": forces, "npairs": int(out[11]), "ncut": int(out[12]), "tie": tie}


def eam(pos, H, types, tab, skin):
    """dict(energy, forces, virial, virial_tensor, rho, dF, embedding_energy,
    pair_potential, npairs, ncut) for float64 tables `tab` (oracle.eam.EAMTables)."""
    lib = _load()
    if tab.dtype != np.float64:
        raise ValueError("the C oracle takes float64 tables")
    pos = np.ascontiguousarray(pos, np.float64)
    n = len(pos)
    H, Hi, ortho = _box(H)
    t = None if types is None else np.ascontiguousarray(types, np.int32)
    c = lambda a: np.ascontiguousarray(a, np.float64)
    dens, ddens, phi, dphi = c(tab.dens), c(tab.ddens), c(tab.phi), c(tab.dphi)
    F, dF = c(tab.F), c(tab.dF)
    lo, d, inv = c(tab.rho_min), c(tab.drho), c(tab.inv_drho)
    forces = np.zeros((n, 3), np.float64)
    rho = np.zeros(n, np.float64)
    dFi = np.zeros(n, np.float64)
    out = np.zeros(16, np.float64)
    rl2 = float(onb.list_radius_sq(tab.cutoff, skin))
    rv = lib.orc_eam(n, _p(pos), None if t is None else _p(t), tab.ne, tab.n_r, tab.cutoff,
                     tab.r_max, tab.inv_dr, _p(dens), _p(ddens), _p(phi), _p(dphi),
                     int(tab.n_rho[0]), _p(F), _p(dF), _p(lo), _p(d), _p(inv), _p(H), _p(Hi),
                     ortho, rl2, _p(forces), _p(rho), _p(dFi), _p(out))
    if rv != 0:
        raise MemoryError("orc_eam")
    return {"energy": out[0], "virial": out[1], "virial_tensor": _sym(out[2:11]),
            "forces": forces, "rho": rho, "dF": dFi, "embedding_energy": out[13],
            "pair_potential": out[14], "npairs": int(out[11]), "ncut": int(out[12])}


def rows(pos, H, rl2, idx):
    """Full neighbor rows of atoms idx: (row position k into idx, j) int arrays,
    j ascending within a row."""
    lib = _load()
    pos = np.ascontiguousarray(pos, np.float64)
    H, Hi, ortho = _box(H)
    idx = np.ascontiguousarray(idx, np.int32)
    cap = max(1, len(idx) * 256)
    while True:
        prow = np.zeros(cap, np.int32)
        pcol = np.zeros(cap, np.int32)
        m = lib.orc_rows(len(pos), _p(pos), _p(H), _p(Hi), ortho, float(np.float32(rl2)),
                         len(idx), _p(idx), cap, _p(prow), _p(pcol))
        if m < 0:
            raise MemoryError("orc_rows")
        if m <= cap:
            return prow[:m], pcol[:m]
        cap = int(m)
