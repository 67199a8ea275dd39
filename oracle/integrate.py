"""Oracle: velocity Verlet with Langevin (BAOB) and Berendsen (test infrastructure only).

Restates
  core/integrator/integrator.py:41-64   constants and the internal mass unit
                                        m = amu * 10 / (N_A * J_per_eV)  [eV ps^2 / A^2]
  core/integrator/integrator.py:84-87   sigma_v = sqrt(kB T / m (1 - exp(-2 gamma dt)))
  core/integrator/integrator.py:91-99   first half (B + A)
  core/integrator/integrator.py:102-124 second half (O + B), KE, T
  core/md_model.py:83-86                kinetic tensor K_ab = sum m v_a v_b
  core/barostat.py:36-68                Berendsen: P = tr(K + W) / (3V), mu^3 clamp,
                                        H, x and v scaled by mu
  io_utils/reader.py:319-339            Maxwell-Boltzmann velocities minus COM

The Gaussian kicks are an input here (`noise`, one N(0,1) draw per atom and
component), so one step can be compared bit for bit; statistics of the
generator are tested separately.

pinned by: the ballistic closed form under constant force, harmonic-oscillator
energy conservation, the injected-noise formula, equipartition and the
Berendsen closed form (tests/test_oracle_physics.py).
"""
import numpy as np

KB = 8.617333262e-5          # eV/K, integrator.py:41
N_A = 6.02214076e23          # integrator.py:42
J_PER_EV = 1.60218e-19       # integrator.py:43
MASS_CONV = 10.0 / (N_A * J_PER_EV)   # integrator.py:63
MB_MASS_CONV = 1.036427e-4   # reader.py:327
EV_A3_TO_BAR = 1_602_176.6208  # barostat.py:16


def internal_mass(mass_amu, dtype=np.float32):
    """integrator.py:61-64 (a float32 tensor times a Python float)."""
    return (np.asarray(mass_amu, dtype) * dtype(MASS_CONV)).astype(dtype)


def first_half(x, v, f, m, dt):
    """integrator.py:91-99: returns (v_half, x_new).

    Velocity arithmetic is done in the dtype of v (float32 in the reference),
    the position update in the dtype of x."""
    t = v.dtype.type
    a = f / m[:, None]
    vh = v + t(0.5) * a * t(dt)
    tx = x.dtype.type
    return vh, x + vh.astype(x.dtype) * tx(dt)


def langevin_factors(m, dt, gamma, T):
    """integrator.py:84-87, 110: (damp, sigma_v[N])."""
    damp = np.exp(-gamma * dt)
    sig = np.sqrt(KB * T / m * (1.0 - np.exp(-2.0 * gamma * dt)))
    return damp, sig


def second_half(vh, f, m, dt, gamma=None, T=None, noise=None):
    """integrator.py:102-124: returns (v_new, KE, T_inst)."""
    t = vh.dtype.type
    v = vh
    f = f.astype(vh.dtype)
    if gamma is not None:
        damp, sig = langevin_factors(m.astype(np.float64), dt, gamma, T)
        v = v * t(damp) + noise.astype(vh.dtype) * sig.astype(vh.dtype)[:, None]
    v = v + t(0.5) * (f / m[:, None]) * t(dt)
    ke = (0.5 * m.astype(np.float64)[:, None] * v.astype(np.float64) ** 2).sum()
    return v, ke, (2.0 / 3.0) * ke / (len(m) * KB)


def kinetic_tensor(v, m):
    """md_model.py:83-86."""
    v = v.astype(np.float64)
    return np.einsum("ni,nj->ij", m.astype(np.float64)[:, None] * v, v)


def berendsen_mu(K, W, V, dt, tau_p, P0, kappa=4.57e-5, mu_max=0.05):
    """barostat.py:46-58: returns (mu, P_bar)."""
    P = np.trace(np.asarray(K, np.float64) + np.asarray(W, np.float64)) / V / 3.0 * EV_A3_TO_BAR
    mu3 = 1.0 + kappa * (dt / tau_p) * (P - P0)
    mu3 = max(mu3, (1.0 - mu_max) ** 3)
    mu3 = min(mu3, (1.0 + mu_max) ** 3)
    return mu3 ** (1.0 / 3.0), P


def maxwell_boltzmann(mass_amu, T, rng, dtype=np.float32):
    """reader.py:319-339 with an explicit generator."""
    m = np.asarray(mass_amu, np.float64) * MB_MASS_CONV
    v = rng.standard_normal((len(m), 3)) * np.sqrt(KB * T / m)[:, None]
    p = (v * m[:, None]).sum(0)
    return (v - p / m.sum()).astype(dtype)
