"""Pins of the anisotropic-barostat and tensile-loading oracle
(oracle/mechanics.py) by closed forms and invariants."""
import numpy as np
import pytest

from oracle import mechanics as om

H_TRI = np.array([[20.0, 0.0, 0.0], [3.0, 19.0, 0.0], [1.5, -2.0, 21.0]])


def _rot(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    Rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    Ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    Rx = np.array([[1, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return Rz @ Ry @ Rx


def test_axis_pressure_diagonal_box_reads_the_diagonal():
    H = np.diag([10.0, 12.0, 14.0])
    K = np.diag([1.0, 2.0, 3.0])
    W = np.array([[0.5, 7.0, -3.0], [7.0, 0.25, 9.0], [-3.0, 9.0, -1.0]])
    p = om.axis_pressure(K, W, H)
    V = 10.0 * 12.0 * 14.0
    np.testing.assert_allclose(p, np.array([1.5, 2.25, 2.0]) / V * om.EV_A3_TO_BAR, rtol=1e-14)


def test_axis_pressure_is_rotation_invariant():
    rng = np.random.default_rng(3)
    A = rng.normal(size=(3, 3))
    K = A @ A.T
    W = rng.normal(size=(3, 3))
    W = W + W.T
    R = _rot(0.3, -0.7, 1.1)
    # rotate the lattice rows and the tensors together: x -> x R^T
    p0 = om.axis_pressure(K, W, H_TRI)
    p1 = om.axis_pressure(R @ K @ R.T, R @ W @ R.T, H_TRI @ R.T)
    np.testing.assert_allclose(p1, p0, rtol=1e-12)
    # a transposed projection (columns instead of rows) would differ here
    u = H_TRI.T / np.linalg.norm(H_TRI.T, axis=1, keepdims=True)
    wrong = np.einsum("ai,ij,aj->a", u, (K + W) / abs(np.linalg.det(H_TRI)), u) * om.EV_A3_TO_BAR
    assert np.abs(wrong - p0).max() > 1e-3 * np.abs(p0).max()


def test_update_from_rest_closed_forms():
    dt, tau, kap, gam = 0.001, 0.2, 3.2e-7, 2.0
    tgt = np.array([0.0, 0.0, 0.0])
    p = np.array([1000.0, -400.0, 10.0])       # outside, outside, inside the 25 bar band
    e1 = om.aniso_update(np.zeros(3), p, dt, tgt, tau_p=tau, gamma_p=gam, kappa=kap)
    want = dt * kap * p / tau ** 2
    want[2] = 0.0
    np.testing.assert_allclose(e1, want, rtol=1e-14, atol=0)
    assert e1[0] > 0 and e1[1] < 0              # compressive stress expands the axis
    # second update: friction acts on the first rate
    e2 = om.aniso_update(e1, p, dt, tgt, tau_p=tau, gamma_p=gam, kappa=kap)
    np.testing.assert_allclose(e2[:2], e1[:2] + dt * (kap * p[:2] / tau ** 2 - gam * e1[:2]),
                               rtol=1e-14)
    # inside the band a nonzero rate decays with the hold friction max(gamma, 5)
    e3 = om.aniso_update(np.array([0.0, 0.0, 0.01]), p, dt, tgt, gamma_p=gam)
    assert e3[2] == pytest.approx(0.01 * (1 - 5.0 * dt), rel=1e-14)
    # ... and a tiny one is zeroed
    e4 = om.aniso_update(np.array([0.0, 0.0, 5e-8]), p, dt, tgt, gamma_p=gam)
    assert e4[2] == 0.0
    # clamp and control mask
    e5 = om.aniso_update(np.zeros(3), np.array([1e9, -1e9, 1e9]), 1.0, tgt,
                         control=(True, True, False))
    np.testing.assert_array_equal(e5, [0.05, -0.05, 0.0])


def test_aniso_step_scales_rows_and_keeps_fractions():
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 20, (50, 3))
    v = rng.normal(0, 2, (50, 3)).astype(np.float32)
    K = np.eye(3) * 5.0
    W = np.diag([400.0, -300.0, 2.0])
    eta0 = np.array([0.01, -0.02, 0.03])
    H1, x1, v1, eta1, p = om.aniso_step(H_TRI, x, v, eta0, K, W, 0.002, 0.0,
                                        control=(True, False, True))
    s = np.where([True, False, True], np.exp(0.002 * eta1), 1.0)
    np.testing.assert_allclose(H1, s[:, None] * H_TRI, rtol=1e-15)
    np.testing.assert_allclose(x1 @ np.linalg.inv(H1), x @ np.linalg.inv(H_TRI), atol=1e-12)
    assert abs(np.linalg.det(H1)) == pytest.approx(abs(np.linalg.det(H_TRI)) * s.prod(), rel=1e-13)
    assert eta1[1] == 0.0
    vs = np.where([True, False, True], np.exp(-0.002 * eta1), 1.0).astype(np.float32)
    np.testing.assert_array_equal(v1, v * vs)


def test_loader_recurrence_and_poisson():
    H = H_TRI.copy()
    x = np.random.default_rng(1).uniform(0, 20, (20, 3))
    L0 = np.linalg.norm(H, axis=1)
    rate, dt = 1e-3, 0.002
    for n in range(1, 51):
        H, x, strain = om.loader_step(H, x, 1, rate, dt, L0[1])
        assert strain == pytest.approx((1 + rate * dt) ** n - 1, rel=1e-9)
    np.testing.assert_allclose(np.linalg.norm(H, axis=1)[[0, 2]], L0[[0, 2]], rtol=1e-15)
    s = om.loader_factors(0, 1e-3, 0.002, "poisson", poisson_ratio=0.3)
    np.testing.assert_allclose(s, [1 + 2e-6, 1 - 0.3 * 2e-6, 1 - 0.3 * 2e-6], rtol=1e-15)
    s = om.loader_factors(2, 10.0, 1.0, "poisson", poisson_ratio=0.3)
    np.testing.assert_allclose(s, [1 - 2e-3, 1 - 2e-3, 11.0])      # lateral clamp
