"""ctypes binding of libsimulon_hip.so (the C ABI declared in include/simulon_hip.h).

The library is loaded from this package directory only. A missing library is
an error, never a reason to fall back to CPU code: every force, neighbor and
integration call in this package goes through the HIP kernels.
"""
import ctypes
import os

import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsimulon_hip.so")

SIM_TILE = 256
SIM_MAX_TYPES = 4
SIM_NPART = 16
SIM_STAGE_TILES = 4
SIM_STAGE_CAP_MAX = 8192
SIM_STAGE_TABLE = 2048
MLP_NDESC, MLP_NH, MLP_ROWS_PER_TILE = 32, 64, 4
MLP_NPARAM = 6337
MLP_NGRAD = MLP_NPARAM + 1
T_EPOT, T_VIRIAL, T_W, T_KE, T_K = 0, 1, 2, 8, 9
SIM_FIRST_UNFUSED_ONLY, SIM_FIRST_BERENDSEN_SCALE = 1, 2

_c_int_p = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


class Replica(ctypes.Structure):
    """Mirror of sim_replica_t."""
    _fields_ = [
        ("h", ctypes.c_double * 9), ("hinv", ctypes.c_double * 9), ("volume", ctypes.c_double),
        ("rl2", ctypes.c_double), ("trig2", ctypes.c_double), ("rl", ctypes.c_double),
        ("skin", ctypes.c_double), ("mu_acc", ctypes.c_double), ("dt", ctypes.c_double),
        ("mu", ctypes.c_double), ("press", ctypes.c_double), ("target_p", ctypes.c_double),
        ("tau_p", ctypes.c_double), ("kappa", ctypes.c_double), ("mu_max", ctypes.c_double),
        ("damp", ctypes.c_float), ("rff", ctypes.c_float * SIM_MAX_TYPES), ("rcf", ctypes.c_float),
        ("dtf", ctypes.c_float),
        ("atom_base", ctypes.c_int), ("natoms", ctypes.c_int), ("cell_base", ctypes.c_int),
        ("cell_cap", ctypes.c_int), ("ncell", ctypes.c_int * 3), ("ortho", ctypes.c_int),
        ("langevin", ctypes.c_int), ("barostat", ctypes.c_int), ("pad", ctypes.c_int),
        ("eta_dot", ctypes.c_double * 3), ("p_target3", ctypes.c_double * 3),
        ("p_axis", ctypes.c_double * 3), ("gamma_p", ctypes.c_double),
        ("max_eta_dot", ctypes.c_double), ("p_tol", ctypes.c_double), ("temp_k", ctypes.c_double),
        ("deform", ctypes.c_double * 3), ("amap", ctypes.c_double * 9),
        ("hb_inv", ctypes.c_double * 9), ("vscale", ctypes.c_float * 3), ("axes", ctypes.c_int),
        ("stochastic", ctypes.c_int), ("loading", ctypes.c_int), ("pad2", ctypes.c_int),
        ("noise_key", ctypes.c_uint64), ("baro_dt", ctypes.c_double),
    ]


class Types(ctypes.Structure):
    _fields_ = [("mass", ctypes.c_float * SIM_MAX_TYPES), ("ntypes", ctypes.c_int),
                ("pad", ctypes.c_int)]


class LJParams(ctypes.Structure):
    _fields_ = [("eps", ctypes.c_float * (SIM_MAX_TYPES * SIM_MAX_TYPES)),
                ("sig", ctypes.c_float * (SIM_MAX_TYPES * SIM_MAX_TYPES)),
                ("rc", ctypes.c_float), ("rs", ctypes.c_float), ("mode", ctypes.c_int),
                ("ntypes", ctypes.c_int)]


class EAMParams(ctypes.Structure):
    _fields_ = [("pass1", _vp), ("pass2", _vp), ("emb", _vp), ("demb", _vp),
                ("n_r", ctypes.c_int), ("n_rho", ctypes.c_int), ("ntypes", ctypes.c_int),
                ("pad", ctypes.c_int), ("inv_dr", ctypes.c_float), ("r_clamp", ctypes.c_float),
                ("cutoff", ctypes.c_float), ("rho_lo", ctypes.c_float * SIM_MAX_TYPES),
                ("rho_hi", ctypes.c_float * SIM_MAX_TYPES),
                ("inv_drho", ctypes.c_float * SIM_MAX_TYPES)]


class MLPNet(ctypes.Structure):
    """Mirror of sim_mlp_net_t."""
    _fields_ = [(n, _vp) for n in ("w1", "w2", "b1", "b2", "w3", "b3", "mean", "inv_std")] + [
        ("rc", ctypes.c_float), ("mu0", ctypes.c_float), ("dmu", ctypes.c_float),
        ("inv_w2", ctypes.c_float)]


class NLReorder(ctypes.Structure):
    """Mirror of sim_nl_reorder_t."""
    _fields_ = [(n, _vp) for n in ("vel", "force", "order", "slot_of", "image", "pos_tmp",
                                   "vel_tmp", "force_tmp", "order_tmp", "image_tmp", "cell_tmp",
                                   "nbr16", "stencil_max", "vel_half", "vel_half_tmp",
                                   "row_base", "posq")] + [("stencil3_only", ctypes.c_int),
                                                           ("pad_ro", ctypes.c_int)]


class Stage(ctypes.Structure):
    """Mirror of sim_stage_t."""
    _fields_ = [("members", _vp), ("count", _vp), ("table", _vp), ("slot", _vp),
                ("cap", ctypes.c_int), ("pad", ctypes.c_int)]


# name -> (restype, argtypes)
_I, _U64, _LL = ctypes.c_int, ctypes.c_ulonglong, ctypes.c_longlong
SIGNATURES = {
    "sim_version": (_I, [ctypes.c_char_p, _I]),
    "sim_nl_build": (_I, [_I, _vp, _vp, _I, _I, _vp, _vp, _vp, _vp, _I, _vp, _vp, _vp, _vp, _vp,
                          _vp, _vp, _I, _vp, _I, _vp, _vp]),
    "sim_lj_force_cell": (_I, [_I, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _I, _I, LJParams, _vp, _vp,
                               _I, _vp, _vp]),
    "sim_eam_density_cell": (_I, [_I, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _I, _I, EAMParams, _vp,
                                  _vp, _vp, _vp, _vp]),
    "sim_eam_force_cell": (_I, [_I, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _I, _I, EAMParams, _vp, _vp,
                                _vp, _I, _vp, _vp]),
    "sim_lj_force_rows": (_I, [_I, _vp, _vp, _vp, _I, _vp, _vp, _vp, _vp, _vp, _I, _I, _I,
                               LJParams, _vp, _vp, _I, _vp, _vp]),
    "sim_eam_density_rows": (_I, [_I, _vp, _vp, _vp, _I, _vp, _vp, _vp, _vp, _vp, _I, _I,
                                  EAMParams, _vp, _vp, _vp, _vp, _vp]),
    "sim_eam_force_rows": (_I, [_I, _vp, _vp, _vp, _I, _vp, _vp, _vp, _vp, _vp, _I, _I,
                                EAMParams, _vp, _vp, _vp, _I, _vp, _vp]),
    "sim_nl_build_xyz": (_I, [_vp, _vp, _I, _I, _vp, _vp, _vp, _vp, _vp, _I, _vp, _vp, _I, _vp,
                              _vp]),
    "sim_nl_half_count": (_I, [_vp, _vp, _I, _I, _vp, _vp, _vp]),
    "sim_nl_half_fill": (_I, [_vp, _vp, _I, _I, _vp, _vp, _I, _vp, _vp, _LL, _vp, _vp]),
    "sim_nl_trigger": (_I, [_I, _vp, _vp, _vp, _vp, _I, _vp]),
    "sim_nl_update": (_I, [_I, _vp, _vp, _I, _I, _vp, _vp, _vp, _vp, _vp, _I, _vp, _I, _vp, _vp,
                           _vp, _vp, _vp, _vp, _vp, _I, _vp, _I, _vp, _vp, _vp]),
    "sim_lj_force": (_I, [_I, _vp, _vp, _I, _vp, _vp, _vp, _I, LJParams, _vp, _vp, _I, _vp]),
    "sim_lj_force_second": (_I, [_I, _vp, _vp, _I, _vp, _vp, _vp, _I, LJParams, _vp, _vp, _I,
                                 Types, _vp, _vp, _vp, _vp, _U64, _vp, _vp]),
    "sim_pos_quantize": (_I, [_vp, _vp, _I, _vp, _vp, _vp]),
    "sim_eam_density": (_I, [_I, _vp, _vp, _I, _vp, _vp, _vp, _I, EAMParams, _vp, _vp, _vp, _vp,
                             _vp]),
    "sim_eam_force": (_I, [_I, _vp, _vp, _I, _vp, _vp, _vp, _vp, _I, EAMParams, _vp, _vp, _I,
                           _vp]),
    "sim_eam_energy_force_noalloc": (_I, [_I, _vp, _vp, _I, _vp, _vp, _vp, _vp, _I, EAMParams,
                                          _vp, _vp, _I, _vp]),
    "sim_eam_force_second": (_I, [_I, _vp, _vp, _I, _vp, _vp, _vp, _vp, _I, EAMParams, _vp, _vp,
                                  _I, Types, _vp, _vp, _vp, _vp, _U64, _vp, _vp, _vp, _vp, _vp,
                                  _vp, _vp]),
    "sim_eam_density_staged": (_I, [_I, _vp, _vp, _I, _vp, _vp, _vp, _I, EAMParams, _vp, _vp,
                                    _vp, _vp, _vp, _vp]),
    "sim_eam_force_staged": (_I, [_I, _vp, _vp, _I, _vp, _vp, _vp, _vp, _I, EAMParams, _vp, _vp,
                                  _I, _vp, _vp]),
    "sim_nl_stage": (_I, [_I, _vp, _vp, _I, _vp, _vp, _vp, _I, _I, _vp, _vp, _vp, _vp, _vp,
                          _vp]),
    "sim_integrate_first": (_I, [_vp, _vp, _I, Types, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sim_integrate_first_q": (_I, [_vp, _vp, _I, Types, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _I,
                                   _vp]),
    "sim_integrate_second": (_I, [_vp, _vp, _I, Types, _vp, _vp, _vp, _vp, _vp, _vp, _U64, _U64, _vp,
                                  _vp, _vp]),
    "sim_step_advance": (_I, [_vp, _vp]),
    "sim_reduce": (_I, [_I, _vp, _vp, _I, _I, _vp, _vp, _vp]),
    "sim_reduce_step": (_I, [_I, _vp, _vp, _I, _I, _vp, _vp, _vp, _vp]),
    "sim_berendsen": (_I, [_I, _vp, _vp, _I, _vp, _vp, _vp, _vp, _vp]),
    "sim_berendsen_box": (_I, [_I, _vp, _vp, _vp]),
    "sim_aniso_barostat": (_I, [_I, _vp, _vp, _I, _vp, _vp, _vp, _vp, _U64, _vp, _vp]),
    "sim_box_deform": (_I, [_I, _vp, _vp, _I, _vp, _vp, _vp]),
    "sim_pos_export": (_I, [_vp, _I, _vp, _I, _vp]),
    "sim_pos_import": (_I, [_vp, _I, _vp, _I, _vp]),
    "sim_mlp_energy": (_I, [_I, _vp, _vp, _I, _vp, _vp, _vp, _I, MLPNet, _vp, _vp, _vp, _vp, _vp]),
    "sim_mlp_force": (_I, [_I, _vp, _vp, _I, _vp, _vp, _vp, _I, MLPNet, _vp, _vp, _vp, _I, _vp]),
    "sim_mlp_adjoint": (_I, [_I, _vp, _vp, _I, _vp, _vp, _vp, _I, MLPNet, _vp, _vp, ctypes.c_float,
                             _vp, _vp]),
    "sim_mlp_grad": (_I, [_I, _vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_float, MLPNet, _vp, _vp,
                          _vp, _vp, _vp, _I, _vp, _vp]),
    "sim_mlp_adam": (_I, [_vp, _vp, _vp, _vp, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                          ctypes.c_float, MLPNet, _vp]),
}

_lib = None


class SimError(RuntimeError):
    pass


def load():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SimError(f"{LIB_PATH} is missing: build it with "
                       "`python -c 'import __graft_entry__ as g; g.build()'` "
                       "or `python simulon_amd/build.py`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def call(name, *args):
    """Call an entry point and raise on a non-zero status."""
    rc = getattr(load(), name)(*args)
    if rc != 0:
        raise SimError(f"{name} returned {rc}")


def ptr(t):
    """Device (or host) address of a tensor, or None."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def version():
    buf = ctypes.create_string_buffer(64)
    call("sim_version", buf, 64)
    return buf.value.decode()


def struct_tensor(structs, device):
    """Pack ctypes structs into a uint8 device tensor."""
    raw = b"".join(bytes(s) for s in structs)
    return torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)


def struct_from_tensor(t, cls, index=0):
    sz = ctypes.sizeof(cls)
    raw = t[index * sz:(index + 1) * sz].cpu().numpy().tobytes()
    return cls.from_buffer_copy(raw)
