"""Seeded structure generators and the synthetic W EAM file used by the
benchmark workloads (all synthetic: no dataset or potential is downloaded).

Workloads (sizes from the benchmark definition):
  lj_argon_nvt      1000 Ar on a 10^3 simple-cubic lattice, spacing 3.64 A
  w_eam_nve         54,000 W as 30^3 bcc cells, a = 3.165 A
  lj_argon_npt_tri  1,048,576 Ar on a 64^3 fcc lattice, sheared xy = 0.1 L, xz = 0.05 L
  w_eam_replicas    64 W bcc cells of 2k..128k atoms (log-uniform), 100..1500 K
  w_mlp_cell        128,000 W (40^3 bcc)
Masses: Ar 39.948, W 183.84 (core/element_info.py:23,25 of the reference).
"""
import math

import numpy as np

AR_MASS = 39.948
W_MASS = 183.84
AR_EPS, AR_SIG = 0.0104, 3.405

SC_BASIS = np.array([[0.0, 0.0, 0.0]])
BCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]])
FCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])


def lattice(basis, ncell, H):
    """Fractional lattice (ncell^3 cells of `basis`) mapped through box rows H.
    Atoms are ordered cell by cell (x fastest), a spatially compact order."""
    n = int(ncell)
    g = np.arange(n)
    cz, cy, cx = np.meshgrid(g, g, g, indexing="ij")
    cells = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], 1).astype(np.float64)
    frac = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3) / n
    return frac @ np.asarray(H, np.float64)


def jitter(pos, amp, seed):
    rng = np.random.default_rng(seed)
    return pos + rng.uniform(-amp, amp, size=pos.shape)


def lj_argon_nvt(seed=0):
    """1000 Ar, simple cubic 10^3, box 36.4 A, jitter U(-0.05, 0.05)."""
    H = np.diag([36.4, 36.4, 36.4])
    return jitter(lattice(SC_BASIS, 10, H), 0.05, seed), H


def w_bcc(ncell, a=3.165, seed=None, amp=0.0):
    H = np.diag([ncell * a] * 3)
    pos = lattice(BCC_BASIS, ncell, H)
    if seed is not None and amp > 0:
        pos = jitter(pos, amp, seed)
    return pos, H


def lj_argon_npt_tri(ncell=64, density=0.0207, seed=0):
    """fcc ncell^3 (4-atom basis) in a box with rows (L,0,0), (0.1L,L,0),
    (0.05L,0,L), L from the number density; jitter 0.05 A."""
    n = 4 * ncell ** 3
    L = (n / density) ** (1.0 / 3.0)
    H = np.array([[L, 0.0, 0.0], [0.1 * L, L, 0.0], [0.05 * L, 0.0, L]])
    return jitter(lattice(FCC_BASIS, ncell, H), 0.05, seed), H


def replica_sizes(nrep=64, lo=1000, hi=64000, seed=3):
    """Whole cubic bcc blocks with cell counts drawn log-uniform in [lo, hi]:
    side n = round(cbrt(cells)) -> 2 n^3 atoms (2,000..128,000 atoms)."""
    rng = np.random.default_rng(seed)
    cells = np.exp(rng.uniform(math.log(lo), math.log(hi), nrep))
    side = np.clip(np.rint(np.cbrt(cells)), round(lo ** (1 / 3)), round(hi ** (1 / 3)))
    return side.astype(int)


def replica_set(nrep=64, t_lo=100.0, t_hi=1500.0):
    """The temperature-scan workload: nrep W bcc cells (a = 3.165 A) of the
    sizes replica_sizes(nrep) draws, with Langevin temperatures evenly spaced
    over [t_lo, t_hi]. Returns [(positions, box, temperature)]."""
    temps = np.linspace(t_lo, t_hi, nrep)
    return [(*w_bcc(int(n)), float(T)) for n, T in zip(replica_sizes(nrep), temps)]


def taper(r, r0=5.0, r1=5.5):
    """Cubic taper 1 - 3t^2 + 2t^3 from 1 at r0 to 0 at r1 (C1 at both ends)."""
    t = np.clip((np.asarray(r) - r0) / (r1 - r0), 0.0, 1.0)
    return 1.0 - 3.0 * t * t + 2.0 * t * t * t


def w_synthetic_functions(r, rho):
    """f(r) = exp(-1.5(r-2.74)), phi(r) = 0.6[exp(-4(r-2.74)) - 2 exp(-2(r-2.74))],
    both tapered to zero over 5.0-5.5 A; F(rho) = -2.5 sqrt(rho)."""
    S = taper(r)
    f = np.exp(-1.5 * (r - 2.74)) * S
    phi = 0.6 * (np.exp(-4.0 * (r - 2.74)) - 2.0 * np.exp(-2.0 * (r - 2.74))) * S
    F = -2.5 * np.sqrt(rho)
    return f, phi, F


def write_w_synthetic_setfl(path, n=5000, rc=5.5, rho_max=40.0):
    """Single-element eam/alloy (setfl) file with n points per function.
    The element line carries the text token "bcc", so readers that skip
    non-numeric lines (io_utils/eam_parser.py:135-144) see only table data."""
    dr = rc / (n - 1)
    drho = rho_max / (n - 1)
    r = np.arange(n) * dr
    rho = np.arange(n) * drho
    f, phi, F = w_synthetic_functions(r, rho)
    with open(path, "w") as fh:
        fh.write("synthetic W EAM: f=exp(-1.5(r-2.74)), phi=0.6[exp(-4(r-2.74))-2exp(-2(r-2.74))]\n")
        fh.write("F(rho)=-2.5 sqrt(rho); cubic taper to zero over 5.0-5.5 A\n")
        fh.write("written by simulon_amd.lattice.write_w_synthetic_setfl\n")
        fh.write("1 W\n")
        fh.write(f"{n} {drho:.16e} {n} {dr:.16e} {rc:.16e}\n")
        fh.write(f"74 {W_MASS} 3.165 bcc\n")
        for arr in (F, f, r * phi):
            for k in range(0, n, 5):
                fh.write(" ".join(f"{v:.16e}" for v in arr[k:k + 5]) + "\n")
    return path
