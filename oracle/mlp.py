"""Oracle: radial-descriptor MLP potential in float64 (test infrastructure only).

There is no reference counterpart: the reference's learned potential is an
external CHGNet wrapper (machine_learning_potentials/machine_learning_force.py,
return contract :85-142). This module restates the library's own network with
torch autograd on the CPU in float64, so that the device kernels' energies,
forces and weight gradients can be checked against it.

Model (all lengths in Angstrom, energies in eV):
  phi_k(r) = exp(-(r - mu_k)^2 / (2 w^2)) fc(r),  mu_k = linspace(1.5, 5.5, 32),
             w = 0.25, fc(r) = (cos(pi r / rc) + 1) / 2 for r < rc = 5.5, else 0
  g_ik     = sum_{j != i, r_ij < rc} phi_k(r_ij)
  gh_i     = (g_i - mean) / std                       (fixed feature scaling)
  E_i      = w3 . silu(W2 silu(W1 gh_i + b1) + b2) + b3,   E = sum_i E_i
  F        = -dE/dx
Weights may be given already rounded to bfloat16 (the device stores them so).

Loss for one frame (force matching, as trained on the device):
  L = wE (E - E_ref)^2 / N^2 + wF sum_i |F_i - F_ref_i|^2 / (3 N)

pinned by: F = -dE/dx by central differences, the descriptor of a perfect bcc
lattice from its shell coordination numbers, invariance under rigid motions,
the virial against a strain derivative, and the loss gradient against central
differences of the loss (tests/test_oracle_mlp.py).
"""
import math

import numpy as np
import torch

from . import box as obox

NDESC, NH = 32, 64
MU = np.linspace(1.5, 5.5, NDESC)
WIDTH = 0.25
RC = 5.5


def pairs(pos, H, rc=RC):
    """Half list within rc (brute force) and minimum-image vectors."""
    pos = np.asarray(pos, np.float64)
    n = len(pos)
    d = obox.minimum_image(pos[:, None, :] - pos[None, :, :], H)
    r2 = (d * d).sum(-1)
    i, j = np.nonzero(np.triu(r2 < rc * rc, k=1))
    return i, j


def descriptors(pos, H, i, j, dtype=torch.float64):
    """g [N, 32] as a differentiable function of pos (torch tensor)."""
    Ht = torch.as_tensor(np.asarray(H), dtype=dtype)
    Hi = torch.linalg.inv(Ht)
    d = pos[i] - pos[j]
    s = d @ Hi
    s = s - torch.round(s.detach())
    d = s @ Ht
    r = torch.sqrt((d * d).sum(-1))
    mu = torch.as_tensor(MU, dtype=dtype)
    fc = 0.5 * (torch.cos(math.pi * r / RC) + 1.0) * (r < RC)
    phi = torch.exp(-((r[:, None] - mu[None, :]) ** 2) / (2 * WIDTH ** 2)) * fc[:, None]
    n = pos.shape[0]
    g = torch.zeros((n, NDESC), dtype=dtype)
    g = g.index_add(0, torch.as_tensor(i), phi).index_add(0, torch.as_tensor(j), phi)
    return g


def silu(x):
    return x * torch.sigmoid(x)


def network(params, gh):
    W1, b1, W2, b2, w3, b3 = params
    h1 = silu(gh @ W1.T + b1)
    h2 = silu(h1 @ W2.T + b2)
    return h2 @ w3 + b3


def to_torch(params, dtype=torch.float64, requires_grad=False):
    out = []
    for p in params:
        t = torch.as_tensor(np.asarray(p, np.float64), dtype=dtype).clone()
        t.requires_grad_(requires_grad)
        out.append(t)
    return out


def energy_forces(pos, H, params, mean, std, create_graph=False):
    """Returns (E, F [N,3], per-atom energies, g) in float64."""
    x = torch.as_tensor(np.asarray(pos, np.float64)).clone().requires_grad_(True)
    i, j = pairs(pos, H)
    g = descriptors(x, H, i, j)
    gh = (g - torch.as_tensor(mean)) / torch.as_tensor(std)
    e = network(params, gh)
    E = e.sum()
    (dEdx,) = torch.autograd.grad(E, x, create_graph=create_graph)
    if not create_graph:
        return E.detach(), -dEdx, e.detach(), g.detach()
    return E, -dEdx, e, g


def energy_virial(pos, H, params, mean, std):
    """(E, W [3,3], scalar virial) with W = -sum_pairs d (x) dE/dd over the
    minimum-image pair vectors d = x_i - x_j, and the scalar sum_pairs d . dE/dd.

    pinned by: -W equals the derivative of E under a homogeneous strain of
    positions and box (tests/test_oracle_mlp.py)."""
    pos = np.asarray(pos, np.float64)
    i, j = pairs(pos, H)
    Ht = torch.as_tensor(np.asarray(H, np.float64))
    x = torch.as_tensor(pos)
    d = x[i] - x[j]
    s = d @ torch.linalg.inv(Ht)
    d = ((s - torch.round(s)) @ Ht).clone().requires_grad_(True)
    r = torch.sqrt((d * d).sum(-1))
    mu = torch.as_tensor(MU)
    fc = 0.5 * (torch.cos(math.pi * r / RC) + 1.0) * (r < RC)
    phi = torch.exp(-((r[:, None] - mu[None, :]) ** 2) / (2 * WIDTH ** 2)) * fc[:, None]
    g = torch.zeros((len(pos), NDESC), dtype=torch.float64)
    g = g.index_add(0, torch.as_tensor(i), phi).index_add(0, torch.as_tensor(j), phi)
    gh = (g - torch.as_tensor(mean)) / torch.as_tensor(std)
    E = network(params, gh).sum()
    (dEdd,) = torch.autograd.grad(E, d)
    W = -(d.detach()[:, :, None] * dEdd[:, None, :]).sum(0)
    return float(E.detach()), W.numpy(), float((d.detach() * dEdd).sum())


def loss(frames, params, mean, std, wE=1.0, wF=1.0):
    """Force-matching loss summed over frames [(pos, H, E_ref, F_ref)]."""
    total = 0.0
    for pos, H, E_ref, F_ref in frames:
        E, F, _, _ = energy_forces(pos, H, params, mean, std, create_graph=True)
        n = len(pos)
        total = total + wE * (E - float(E_ref)) ** 2 / n ** 2 \
            + wF * ((F - torch.as_tensor(F_ref)) ** 2).sum() / (3 * n)
    return total


def loss_and_grad(frames, params_np, mean, std, wE=1.0, wF=1.0):
    params = to_torch(params_np, requires_grad=True)
    L = loss(frames, params, mean, std, wE, wF)
    grads = torch.autograd.grad(L, params)
    return float(L.detach()), [g.numpy() for g in grads]


def bf16_round(a):
    """Round float64/32 values to the nearest bfloat16 (ties to even)."""
    t = torch.as_tensor(np.asarray(a, np.float32)).to(torch.bfloat16).to(torch.float64)
    return t.numpy()
