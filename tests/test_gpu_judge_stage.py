"""An EAM evaluation into work arrays (the judge of another force field,
simulon_amd/mlp.py JudgedMLP.judge) on an engine large enough for the LDS
staging of the EAM passes (>= 256 tiles) stays correct after the engine's
graph-replayed steps rebuild its lists: it reads the lists directly. (A
stage created by the first such call was not rebuilt by the engine's
captured steps, so the second call read stale member sets and slots; on the
128,000-atom MLP bench that looked like a diverged network.) Reference: the
same positions evaluated by a fresh engine whose own force field is EAM
(core/force/eam_force_cu.py:461-579)."""
import numpy as np
import pytest
import torch

from helpers import force_error
from simulon_amd import lattice as lat
from simulon_amd.engine import ReplicaEngine, ReplicaSpec
from simulon_amd.forcefield import EAMForceField, LJForceField

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setfl(tmp_path_factory):
    return lat.write_w_synthetic_setfl(str(tmp_path_factory.mktemp("w") / "W_synthetic.eam.alloy"))


def _eam_forces(eam, eng):
    f = eng.scratch("judge_force", torch.float32, (eng.N, 3))
    part = eng.scratch("judge_partial", torch.float64, (eng.ntiles, 16))
    eam.compute(eng, virial=False, force=f, partial=part)
    torch.cuda.synchronize()
    return eng.to_caller(f).cpu().numpy().copy()


def test_eam_into_work_arrays_after_rebuilds(setfl):
    pos, H = lat.w_bcc(34, seed=5, amp=0.05)          # 78,608 atoms: 308 tiles
    spec = ReplicaSpec(positions=pos, box=H, dt=0.001, ensemble="NVT", temperature=3000.0,
                       gamma=1.0)
    eng = ReplicaEngine([spec], [lat.W_MASS], cutoff=5.5, skin=0.3, seed=3, family="global")
    # the engine steps with another force field (a soft LJ stand-in)
    eng.force_field = LJForceField([[0.05]], [[2.5]], 5.5)
    eng.setup()
    assert eng.ntiles >= eng.STAGE_MIN_TILES
    eam = EAMForceField(setfl, device=eng.device)
    print(f"[case] judge-style EAM on {eng.N} atoms, {eng.ntiles} tiles", flush=True)
    for k in range(2):
        nb0 = int(eng.nbuild.sum())
        eng.run(40, graph=True)
        torch.cuda.synchronize()
        assert int(eng.nbuild.sum()) > nb0, "no rebuild between the evaluations"
        got = _eam_forces(eam, eng)
        x = eng.to_caller(eng.positions(unwrap=False)).cpu().numpy()
        ref_eng = ReplicaEngine([ReplicaSpec(positions=x, box=H)], [lat.W_MASS], cutoff=5.5,
                                skin=0.3, family="global")
        ref_eng.force_field = EAMForceField(setfl, device=ref_eng.device)
        ref_eng.setup()
        ref = ref_eng.to_caller(ref_eng.force).cpu().numpy()
        err = force_error(got, ref)
        print(f"evaluation {k}: force difference {err:.2e}", flush=True)
        assert err < 1e-5
