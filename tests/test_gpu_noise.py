"""The device Langevin noise (Philox4x32-10 + Box-Muller on the hardware
log2, square root and sine/cosine, sim_common.h gauss3) drawn through
sim_integrate_second with zero force and zero half-step velocity, so that
v = xi * rff exactly (core/integrator/integrator.py:102-124 with F = 0,
v_half = 0): over 4 x 131,072 atom draws the components have mean 0,
variance 1, third moment 0 and fourth moment 3 of N(0, 1) within sampling
bounds; the three components of an atom, and the draws of two steps, are
uncorrelated; the tail beyond 4 sigma has the normal's mass."""
import math

import numpy as np
import pytest
import torch

from simulon_amd import lattice as lat
from simulon_amd.engine import ReplicaEngine, ReplicaSpec

pytestmark = pytest.mark.gpu


def _draws(eng, rff):
    eng.vel_half.zero_()
    eng.force.zero_()
    eng.second_half(None, advance=True)
    torch.cuda.synchronize()
    return (eng.vel.double() / rff).cpu().numpy()


def test_langevin_noise_is_standard_normal():
    pos, H = lat.lj_argon_npt_tri(ncell=32)           # 131,072 atoms
    spec = ReplicaSpec(positions=pos, box=H, dt=0.002, ensemble="NVT", temperature=300.0,
                       gamma=5.0)
    eng = ReplicaEngine([spec], [lat.AR_MASS], cutoff=8.5, skin=1.0, seed=41, family="global")
    P = eng.replica_state(0)
    rff = float(P.rff[0])
    assert rff > 0.0
    print(f"[case] noise moments over {eng.N} atoms x 4 steps", flush=True)
    g = np.concatenate([_draws(eng, rff) for _ in range(4)], axis=0)   # [4N, 3]
    x = g.ravel()
    n = x.size
    se = 1.0 / math.sqrt(n)
    assert abs(x.mean()) < 5 * se
    assert abs(x.var() - 1.0) < 5 * math.sqrt(2.0) * se
    assert abs(np.mean(x ** 3)) < 5 * math.sqrt(15.0) * se
    assert abs(np.mean(x ** 4) - 3.0) < 5 * math.sqrt(96.0) * se
    # components of one atom, and the same atom on consecutive steps
    N = eng.N
    for a, b in ((g[:, 0], g[:, 1]), (g[:, 0], g[:, 2]), (g[:, 1], g[:, 2]),
                 (g[:N, 0], g[N:2 * N, 0])):
        r = float(np.corrcoef(a, b)[0, 1])
        assert abs(r) < 5 / math.sqrt(a.size)
    # tail mass beyond 4 sigma: 2 (1 - Phi(4)) = 6.334e-5
    tail = float(np.mean(np.abs(x) > 4.0))
    expect = 6.334e-5
    assert abs(tail - expect) < 5 * math.sqrt(expect / n)
