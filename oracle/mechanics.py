"""Oracle: anisotropic barostat and uniaxial tensile loading (test infrastructure only).

Restates, in float64 numpy,
  core/barostat.py:78-208          AnisotropicNPTBarostat.step: the pressure
                                   tensor (K + W) / V projected on the unit
                                   lattice rows, the strain-rate update
                                   d eta/dt = kappa dP / tau_p^2 - gamma_p eta
                                   with the deadband hold, the clamp and the
                                   control mask; lattice rows scaled by
                                   exp(dt eta), coordinates carried with the
                                   box, velocities scaled by exp(-dt eta) per
                                   Cartesian axis.
  core/mechanics/loading.py:39-78  UniaxialTensileLoader.step: the loading row
                                   scaled by 1 + rate dt, the lateral rows
                                   fixed or Poisson-scaled (clamped), the
                                   engineering strain.

Lattice vectors are the rows of H (x = s H for fractional s); a row scaling
is H' = diag(s) H and carries the atoms as x' = x H^-1 H'. The reference maps
float32 coordinates and then wraps them into the box; here coordinates are
mapped in float64 and not wrapped (compare them modulo the lattice).

pinned by: closed forms of one and two updates from rest (outside and inside
the deadband, clamped, uncontrolled axes), invariance of the projected
pressure under a rotation of box and tensor together, conservation of
fractional coordinates and volume = det H under row scaling, and the
loading recurrence (1 + rate dt)^n (tests/test_oracle_mechanics.py).
"""
import numpy as np

EV_A3_TO_BAR = 1_602_176.6208   # barostat.py:16
KB = 8.617333262e-5             # barostat.py:18


def scale_rows(H, x, s):
    """H' = diag(s) H and x' = x H^-1 H' (fractional coordinates kept)."""
    H = np.asarray(H, np.float64)
    Hn = np.asarray(s, np.float64)[:, None] * H
    M = np.linalg.inv(H) @ Hn
    return Hn, np.asarray(x, np.float64) @ M


def axis_pressure(K, W, H):
    """barostat.py:146-155: (K + W)/V projected on the unit rows of H, in bar."""
    H = np.asarray(H, np.float64)
    V = abs(np.linalg.det(H))
    S = (np.asarray(K, np.float64) + np.asarray(W, np.float64)) / V
    u = H / np.maximum(np.linalg.norm(H, axis=1, keepdims=True), 1e-12)
    return np.einsum("ai,ij,aj->a", u, S, u) * EV_A3_TO_BAR


def aniso_update(eta, p_axis, dt, target, tau_p=0.2, gamma_p=2.0, control=(True, True, True),
                 max_eta_dot=0.05, kappa=3.2e-7, tol=25.0, noise=None, temperature=1.0):
    """barostat.py:157-184: the new eta_dot from the projected pressure.
    noise: three N(0,1) draws when the controller is stochastic."""
    eta = np.asarray(eta, np.float64)
    control = np.asarray(control, bool)
    dp = np.asarray(p_axis, np.float64) - np.broadcast_to(np.asarray(target, np.float64), (3,))
    within = np.abs(dp) <= tol
    accel = kappa * np.where(within, 0.0, dp) / max(tau_p ** 2, 1e-12)
    if gamma_p > 0.0:
        accel = accel - gamma_p * eta
    if noise is not None:
        accel = accel + np.asarray(noise, np.float64) * (kappa * KB * max(temperature, 1.0)
                                                         / max(tau_p, 1e-12))
    accel = np.where(within, -max(gamma_p, 5.0) * eta, accel)
    new = np.clip(eta + dt * accel, -max_eta_dot, max_eta_dot)
    new = np.where(control, new, 0.0)
    new = np.where(within & (np.abs(new) < 1e-7), 0.0, new)
    return new


def aniso_step(H, x, v, eta, K, W, dt, target, control=(True, True, True), **kw):
    """One AnisotropicNPTBarostat.step (barostat.py:131-208).
    Returns (H', x', v' float32, eta', p_axis bar)."""
    p_axis = axis_pressure(K, W, H)
    eta_new = aniso_update(eta, p_axis, dt, target, control=control, **kw)
    control = np.asarray(control, bool)
    s = np.where(control, np.exp(dt * eta_new), 1.0)
    Hn, xn = scale_rows(H, x, s)
    vs = np.where(control, np.exp(-dt * eta_new), 1.0).astype(np.float32)
    vn = (np.asarray(v, np.float32) * vs[None, :]).astype(np.float32)
    return Hn, xn, vn, eta_new, p_axis


def loader_factors(axis, rate, dt, lateral_mode="fixed", poisson_ratio=0.28,
                   max_lateral_scale_step=2.0e-3):
    """loading.py:47-55: the row scale factors of one loading step."""
    d = rate * float(dt)
    s = np.ones(3)
    s[axis] = 1.0 + d
    if lateral_mode == "poisson":
        lat = max(1.0 - poisson_ratio * d, 1.0 - max_lateral_scale_step)
        for a in range(3):
            if a != axis:
                s[a] = lat
    return s


def loader_step(H, x, axis, rate, dt, L0, **kw):
    """One UniaxialTensileLoader.step (loading.py:47-78): (H', x', strain)."""
    Hn, xn = scale_rows(H, x, loader_factors(axis, rate, dt, **kw))
    strain = np.linalg.norm(Hn[axis]) / L0 - 1.0
    return Hn, xn, strain
