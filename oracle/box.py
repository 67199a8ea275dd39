"""Oracle: periodic box operations (test infrastructure only).

Restates core/box.py:109-148. H is the 3x3 lattice matrix with the lattice
vectors as ROWS, held in float64; operations cast H (and H^-1, which the
caller supplies exactly as the box computed it, core/box.py:55) to the
operand dtype first, as the reference's `.to(rij)` does (core/box.py:123,126).

Floating-point order is fixed so that the FP32 results are reproducible
bit for bit:
  orthogonal  r - d * rint(r / d)                      (core/box.py:122-124)
  triclinic   s_k = (r0*Hi[0,k] + r1*Hi[1,k]) + r2*Hi[2,k]
              s  -= rint(s)
              r_m = (s0*H[0,m] + s1*H[1,m]) + s2*H[2,m]  (core/box.py:125-130)
numpy performs every elementwise op with one IEEE rounding and no
contraction, and np.rint rounds half to even like torch.round.

pinned by: closed-form cases and invariants (|s|<=1/2, idempotence,
agreement with a 27-image brute force) in tests/test_oracle_geometry.py, and the
reference's own FP32/FP64 minimum-image outputs in tests/golden/.
"""
import numpy as np


def is_orthogonal(H):
    """core/box.py:56-57 — off-diagonal entries all within 1e-8 of zero."""
    H = np.asarray(H, dtype=np.float64)
    off = H - np.diag(np.diag(H))
    return bool(np.all(np.abs(off) <= 1e-8))


def inverse(H):
    """H^-1 in float64 (the reference uses torch.linalg.inv, core/box.py:55)."""
    return np.linalg.inv(np.asarray(H, dtype=np.float64))


def _rowvec_mat(r, M):
    """Row vectors r[...,3] times M[3,3] with the fixed summation order."""
    return (r[..., 0:1] * M[0] + r[..., 1:2] * M[1]) + r[..., 2:3] * M[2]


def minimum_image(rij, H, Hinv=None, ortho=None):
    """core/box.py:109-130 in the dtype of `rij`."""
    rij = np.asarray(rij)
    dt = rij.dtype
    H = np.asarray(H, dtype=np.float64)
    if ortho is None:
        ortho = is_orthogonal(H)
    if ortho:
        d = np.diag(H).astype(dt)
        return rij - d * np.rint(rij / d)
    if Hinv is None:
        Hinv = inverse(H)
    Hi = np.asarray(Hinv, dtype=np.float64).astype(dt)
    Hd = H.astype(dt)
    s = _rowvec_mat(rij, Hi)
    s = s - np.rint(s)
    return _rowvec_mat(s, Hd)


def wrap_positions(pos, H, Hinv=None, ortho=None):
    """core/box.py:132-142: fold positions into the primary cell."""
    pos = np.asarray(pos)
    dt = pos.dtype
    H = np.asarray(H, dtype=np.float64)
    if ortho is None:
        ortho = is_orthogonal(H)
    if ortho:
        d = np.diag(H).astype(dt)
        return pos - np.floor(pos / d) * d
    if Hinv is None:
        Hinv = inverse(H)
    s = _rowvec_mat(pos, np.asarray(Hinv, dtype=np.float64).astype(dt))
    s = s - np.floor(s)
    return _rowvec_mat(s, H.astype(dt))


def volume(H):
    """core/box.py:79-80."""
    return float(abs(np.linalg.det(np.asarray(H, dtype=np.float64))))


def perpendicular_widths(H):
    """Distance between opposite faces of the cell, one per lattice vector."""
    H = np.asarray(H, dtype=np.float64)
    V = volume(H)
    a, b, c = H
    return np.array([V / np.linalg.norm(np.cross(b, c)),
                     V / np.linalg.norm(np.cross(c, a)),
                     V / np.linalg.norm(np.cross(a, b))])
