"""Pins of the float64 MLP-potential oracle (oracle/mlp.py) on the CPU.

The network has no reference counterpart, so the oracle is pinned by what the
mathematics fixes: the descriptor of a perfect bcc lattice from its shell
coordination numbers, forces as minus the gradient (central differences),
invariance under rigid motions, and the loss gradient against central
differences of the loss.
"""
import math

import numpy as np
import pytest
import torch

from oracle import mlp as om


def _params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, scale / math.sqrt(32), (64, 32))
    b1 = rng.normal(0, 0.1, 64)
    W2 = rng.normal(0, scale / math.sqrt(64), (64, 64))
    b2 = rng.normal(0, 0.1, 64)
    w3 = rng.normal(0, 1 / 8, 64)
    b3 = -8.0
    return [om.bf16_round(W1), b1, om.bf16_round(W2), b2, w3, b3]


def _bcc(ncell=5, a=3.165, amp=0.0, seed=0):
    g = np.stack(np.meshgrid(*[np.arange(ncell)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = np.concatenate([g, g + 0.5]) * a
    pos = pos + np.random.default_rng(seed).uniform(-amp, amp, pos.shape)
    return pos, np.eye(3) * a * ncell


def _phi(r):
    fc = 0.5 * (math.cos(math.pi * r / om.RC) + 1.0) if r < om.RC else 0.0
    return np.exp(-((r - om.MU) ** 2) / (2 * om.WIDTH ** 2)) * fc


def test_descriptor_of_perfect_bcc_is_the_shell_sum():
    """Every atom of a perfect bcc lattice sees the same shells:
    8 at a*sqrt(3)/2, 6 at a, 12 at a*sqrt(2), 24 at a*sqrt(11)/2, 8 at a*sqrt(3)
    (all inside 5.5 A for a = 3.165; the next shell, 6 at 2a = 6.33 A, is not)."""
    a = 3.165
    pos, H = _bcc(5, a)
    i, j = om.pairs(pos, H)
    g = om.descriptors(torch.as_tensor(pos), H, i, j).numpy()
    shells = [(8, a * math.sqrt(3) / 2), (6, a), (12, a * math.sqrt(2)),
              (24, a * math.sqrt(11) / 2), (8, a * math.sqrt(3))]
    want = sum(n * _phi(r) for n, r in shells)
    np.testing.assert_allclose(g, np.broadcast_to(want, g.shape), rtol=1e-12, atol=1e-14)


def test_forces_are_minus_the_energy_gradient():
    pos, H = _bcc(5, amp=0.15, seed=1)
    P = om.to_torch(_params(1))
    mean, std = np.zeros(32), np.full(32, 2.0)
    E, F, _, _ = om.energy_forces(pos, H, P, mean, std)
    h = 1e-5
    for a in (0, 17, 123):
        for d in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[a, d] += h
            p2[a, d] -= h
            e1 = float(om.energy_forces(p1, H, P, mean, std)[0])
            e2 = float(om.energy_forces(p2, H, P, mean, std)[0])
            assert abs(-(e1 - e2) / (2 * h) - float(F[a, d])) < 1e-6 * max(1.0, abs(float(F[a, d])))


def test_energy_is_invariant_under_rigid_motion():
    pos, H = _bcc(5, amp=0.1, seed=2)
    P = om.to_torch(_params(2))
    mean, std = np.zeros(32), np.full(32, 2.0)
    E0, F0, _, _ = om.energy_forces(pos, H, P, mean, std)
    E1, F1, _, _ = om.energy_forces(pos + np.array([1.3, -0.7, 2.9]), H, P, mean, std)
    assert abs(float(E0) - float(E1)) < 1e-9 * abs(float(E0))
    np.testing.assert_allclose(F0.numpy(), F1.numpy(), atol=1e-9)
    # net force of a periodic system vanishes (Newton's third law pairwise)
    assert np.abs(F0.numpy().sum(0)).max() < 1e-9


def test_zero_first_layer_gives_constant_energy_and_no_force():
    pos, H = _bcc(5, amp=0.1, seed=3)
    P = _params(3)
    P[0] = np.zeros((64, 32))
    E, F, e, _ = om.energy_forces(pos, H, om.to_torch(P), np.zeros(32), np.ones(32))
    assert np.abs(F.numpy()).max() == 0.0
    assert float(e.std()) < 1e-12


def test_loss_gradient_against_central_differences():
    frames = []
    for s in range(2):
        pos, H = _bcc(5, amp=0.12, seed=10 + s)
        rng = np.random.default_rng(s)
        frames.append((pos, H, -8.0 * len(pos) + rng.normal(), rng.normal(0, 0.5, pos.shape)))
    P = _params(4)
    mean, std = np.zeros(32), np.full(32, 2.0)
    L, grads = om.loss_and_grad(frames, P, mean, std, wE=0.3, wF=1.0)
    rng = np.random.default_rng(5)
    h = 1e-6
    for k, shape in enumerate([(64, 32), (64,), (64, 64), (64,), (64,), ()]):
        for _ in range(2):
            idx = tuple(rng.integers(0, n) for n in shape)
            Pp = [np.array(p, np.float64, copy=True) for p in P]
            Pm = [np.array(p, np.float64, copy=True) for p in P]
            Pp[k][idx] += h
            Pm[k][idx] -= h
            lp = float(om.loss(frames, om.to_torch(Pp), mean, std, 0.3, 1.0))
            lm = float(om.loss(frames, om.to_torch(Pm), mean, std, 0.3, 1.0))
            fd = (lp - lm) / (2 * h)
            g = float(np.asarray(grads[k])[idx])
            assert abs(fd - g) < 1e-5 * max(1.0, abs(g)), (k, idx, fd, g)


def test_virial_is_minus_the_strain_derivative():
    """W_ab = -dE/d(eps_ab) for x -> x (1 + eps), H -> H (1 + eps) (row vectors)."""
    pos, H = _bcc(5, amp=0.12, seed=6)
    H = H + np.array([[0, 0, 0], [0.8, 0, 0], [0.3, 0.0, 0]])     # triclinic
    P = om.to_torch(_params(6))
    mean, std = np.zeros(32), np.full(32, 2.0)
    E, W, vir = om.energy_virial(pos, H, P, mean, std)
    E2 = float(om.energy_forces(pos, H, P, mean, std)[0])
    assert abs(E - E2) < 1e-9 * abs(E)
    h = 1e-6
    for a in range(3):
        for b in range(3):
            eps = np.zeros((3, 3))
            eps[a, b] = h
            M = np.eye(3) + eps
            ep = om.energy_virial(pos @ M, H @ M, P, mean, std)[0]
            em = om.energy_virial(pos @ (2 * np.eye(3) - M), H @ (2 * np.eye(3) - M), P, mean,
                                  std)[0]
            assert abs(-(ep - em) / (2 * h) - W[a, b]) < 1e-5 * max(1.0, np.abs(W).max())
    assert abs(vir + np.trace(W)) < 1e-9 * max(1.0, abs(vir))
