"""Oracle: edge-list Lennard-Jones energy, forces and virial (test infrastructure only).

Restates core/force/lennard_jones_force.py:15-95 on a half list, in the dtype
of `pos` (float32 or float64):

  rij = mi(pos_i - pos_j); r2 = max(|rij|^2, 1e-24); r = sqrt(r2)    (:32-37)
  U = 4 eps (s^12 - s^6), Fmag = 24 eps / r (2 s^12 - s^6), s = sigma/r  (:42-46)
  switch  S(x) = 1 - 10x^3 + 15x^4 - 6x^5 on [rs, rc], Fmag = Fmag S - U dS (:48-63)
  shift   U - U(rc) + (r - rc) Fmag(rc), Fmag - Fmag(rc)             (:65-75)
  only r < rc counts                                                  (:82-84)
  F_i += Fmag rij/r, F_j -= Fmag rij/r                                (:87-90)
  E = sum U, virial = sum Fmag r, virial_tensor[a,b] = sum f_a rij_b  (:92-94)

Per-edge epsilon and sigma come from a per-type-pair table indexed by the
two atom types (reader.py:165-210 builds them from the parameter dict).

pinned by: the two-atom closed form, Newton's third law, F = -dE/dx by FP64
central differences, tr(W) = virial, lattice sums (tests/test_oracle_physics.py)
and the reference's own outputs (tests/golden/).
"""
import numpy as np

from . import box as obox


def lj_pair_terms(r, eps, sig, rc, mode="raw", switch_ratio=0.9):
    """Energy and force magnitude per edge, lennard_jones_force.py:36-84."""
    dt = r.dtype
    one = dt.type(1.0)
    inv_r = one / r
    sr = sig * inv_r
    sr6 = sr ** 6
    sr12 = sr6 ** 2
    U = dt.type(4.0) * eps * (sr12 - sr6)
    F = dt.type(24.0) * eps * inv_r * (dt.type(2.0) * sr12 - sr6)
    rc_ = dt.type(rc)
    if mode == "switch":
        rs = dt.type(switch_ratio * rc)
        x = np.clip((r - rs) / (rc_ - rs), 0.0, 1.0).astype(dt)
        S_poly = one - dt.type(10.0) * x ** 3 + dt.type(15.0) * x ** 4 - dt.type(6.0) * x ** 5
        dS_poly = (dt.type(-30.0) * x ** 2 + dt.type(60.0) * x ** 3 - dt.type(30.0) * x ** 4) / (rc_ - rs)
        mid = (r > rs) & (r < rc_)
        S = np.where(r <= rs, one, np.where(mid, S_poly, dt.type(0.0))).astype(dt)
        dS = np.where(mid, dS_poly, dt.type(0.0)).astype(dt)
        Ueff = U * S
        Feff = F * S - U * dS
    elif mode == "shift":
        src = sig / rc_
        s6 = src ** 6
        s12 = s6 ** 2
        Urc = dt.type(4.0) * eps * (s12 - s6)
        Frc = dt.type(24.0) * eps * (one / rc_) * (dt.type(2.0) * s12 - s6)
        Ueff = U - Urc + (r - rc_) * Frc
        Feff = F - Frc
    else:
        Ueff, Feff = U, F
    mask = (r < rc_).astype(dt)
    return Ueff * mask, Feff * mask


def lj_forward(pos, edge_index, types, eps_table, sig_table, rc, H, Hinv=None,
               ortho=None, mode="raw", switch_ratio=0.9):
    """lennard_jones_force.py:15-95. Returns dict(energy, forces, virial, virial_tensor)."""
    pos = np.asarray(pos)
    dt = pos.dtype
    i, j = np.asarray(edge_index[0]), np.asarray(edge_index[1])
    ti, tj = np.asarray(types)[i], np.asarray(types)[j]
    eps = np.asarray(eps_table, np.float64)[ti, tj].astype(dt)
    sig = np.asarray(sig_table, np.float64)[ti, tj].astype(dt)
    rij = obox.minimum_image(pos[i] - pos[j], H, Hinv, ortho)
    r2 = np.maximum((rij[:, 0] * rij[:, 0] + rij[:, 1] * rij[:, 1]) + rij[:, 2] * rij[:, 2],
                    dt.type(1e-24))
    r = np.sqrt(r2)
    U, F = lj_pair_terms(r, eps, sig, rc, mode, switch_ratio)
    e_ij = rij * (dt.type(1.0) / r)[:, None]
    fij = F[:, None] * e_ij
    forces = np.zeros_like(pos)
    np.add.at(forces, i, fij)
    np.add.at(forces, j, -fij)
    return {
        "energy": U.astype(np.float64).sum(),
        "forces": forces,
        "virial": (F * r).astype(np.float64).sum(),
        "virial_tensor": np.einsum("ei,ej->ij", fij.astype(np.float64), rij.astype(np.float64)),
    }


def lj_two_atom(r, eps, sig):
    """Closed form for one pair: (U, Fmag) with Fmag = -dU/dr."""
    s6 = (sig / r) ** 6
    return 4 * eps * (s6 * s6 - s6), 24 * eps / r * (2 * s6 * s6 - s6)
