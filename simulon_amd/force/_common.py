"""Shared backbone plumbing: run one force field on the molecular's engine and
return the reference's output dict in the caller's atom order."""
import torch

from .. import _lib


def backbone_outputs(mol, ff, virial=True, **kw):
    eng = mol.engine
    ff.compute(eng, virial, **kw)
    eng.reduce(0, 8, ff=ff)
    th = eng.thermo[0]
    return {
        "energy": th[_lib.T_EPOT],
        "forces": eng.to_caller(eng.force),
        "virial": th[_lib.T_VIRIAL],
        "virial_tensor": eng.virial_tensor(0),
    }
