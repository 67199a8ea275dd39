"""The allocation-free EAM entry point (sim_eam_energy_force_noalloc) against
the entry points it combines with caller buffers (sim_eam_density +
sim_eam_force + sim_reduce): forces and the energy/virial columns bit for
bit, on two calls in a row, the second on a smaller system (the work array
keeps the first call's contents), and its size limit."""
import ctypes

import numpy as np
import pytest
import torch

from simulon_amd import _lib
from simulon_amd import lattice as lat
from simulon_amd._lib import ptr, stream_ptr
from simulon_amd.engine import ReplicaEngine, ReplicaSpec
from simulon_amd.forcefield import EAMForceField

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setfl(tmp_path_factory):
    return lat.write_w_synthetic_setfl(str(tmp_path_factory.mktemp("w") / "W_synthetic.eam.alloy"))


def _engine(sides, setfl, seed):
    specs = [ReplicaSpec(positions=lat.jitter(lat.w_bcc(n)[0], 0.08, seed + k), box=lat.w_bcc(n)[1],
                         dt=0.001) for k, n in enumerate(sides)]
    eng = ReplicaEngine(specs, [lat.W_MASS], cutoff=5.5, skin=1.0, family="global")
    eng.force_field = EAMForceField(setfl, device=eng.device)
    eng.setup()
    eng.quantize()
    return eng


def _with_buffers(eng, virial):
    ff = eng.force_field
    rec = torch.zeros((eng.N, 4), dtype=torch.int32, device=eng.device)
    part = torch.zeros((eng.ntiles, _lib.SIM_NPART), dtype=torch.float64, device=eng.device)
    chunks = torch.zeros((eng.R, 64, _lib.SIM_NPART), dtype=torch.float64, device=eng.device)
    force = torch.zeros((eng.N, 3), dtype=torch.float32, device=eng.device)
    thermo = torch.zeros((eng.R, _lib.SIM_NPART), dtype=torch.float64, device=eng.device)
    s = stream_ptr()
    _lib.call("sim_eam_density", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles, ptr(eng.posq),
              ptr(eng.nbr), ptr(eng.nnbr), eng.N, ff.params, None, None, ptr(rec), ptr(part), s)
    _lib.call("sim_eam_force", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles, ptr(eng.posq),
              ptr(rec), ptr(eng.nbr), ptr(eng.nnbr), eng.N, ff.params, ptr(force), ptr(part),
              int(virial), s)
    _lib.call("sim_reduce", eng.R, ptr(eng.rep_tiles), ptr(part), 0, _lib.T_KE, ptr(chunks),
              ptr(thermo), s)
    return force, thermo


def _noalloc(eng, virial):
    force = torch.full((eng.N, 3), float("nan"), dtype=torch.float32, device=eng.device)
    thermo = torch.full((eng.R, _lib.SIM_NPART), -7.0, dtype=torch.float64, device=eng.device)
    _lib.call("sim_eam_energy_force_noalloc", eng.R, ptr(eng.reps), ptr(eng.tiles), eng.ntiles,
              ptr(eng.rep_tiles), ptr(eng.posq), ptr(eng.nbr), ptr(eng.nnbr), eng.N,
              eng.force_field.params, ptr(force), ptr(thermo), int(virial), stream_ptr())
    return force, thermo


def test_noalloc_equals_buffered_entry_points(setfl):
    # first call: 3 replicas, 2,000-16,000 atoms (>= 256 tiles: the
    # persistent kernels); second: one 1,024-atom replica (4 tiles)
    for sides, virial, seed in (([10, 20, 16], 1, 3), ([8], 0, 9)):
        eng = _engine(sides, setfl, seed)
        print(f"[case] noalloc EAM: replicas {sides}, {eng.N} atoms, {eng.ntiles} tiles, "
              f"virial {virial}", flush=True)
        f0, t0 = _with_buffers(eng, virial)
        f1, t1 = _noalloc(eng, virial)
        torch.cuda.synchronize()
        assert torch.equal(f0, f1)
        assert torch.equal(t0[:, :_lib.T_KE], t1[:, :_lib.T_KE])
        assert torch.all(t1[:, _lib.T_KE:] == -7.0)      # columns it does not own are untouched
        assert float(t1[:, 0].abs().min()) > 0


def test_noalloc_rejects_sizes_beyond_its_work_space(setfl):
    eng = _engine([8], setfl, 1)
    lib = _lib.load()
    args = lambda n, ntiles, R: (R, ptr(eng.reps), ptr(eng.tiles), ntiles, ptr(eng.rep_tiles),
                                 ptr(eng.posq), ptr(eng.nbr), ptr(eng.nnbr), n,
                                 eng.force_field.params, ptr(eng.force), ptr(eng.thermo), 1,
                                 stream_ptr())
    assert lib.sim_eam_energy_force_noalloc(*args((1 << 21) + 1, eng.ntiles, 1)) == -4
    assert lib.sim_eam_energy_force_noalloc(*args(eng.N, 16385, 1)) == -4
    assert lib.sim_eam_energy_force_noalloc(*args(eng.N, eng.ntiles, 257)) == -4
