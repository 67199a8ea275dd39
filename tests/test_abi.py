"""The C-ABI library loads without a GPU and exports every entry point that
include/simulon_hip.h declares; the ctypes struct mirrors match the C layout."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "simulon_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*int\s+(sim_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "sim_nl_build" in names and "sim_eam_force" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    from simulon_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function has a ctypes signature in the binding
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_version_without_gpu():
    from simulon_amd import _lib
    assert _lib.version().startswith("simulon_hip")


def test_argument_errors_without_gpu():
    """Argument checks run before any launch and return SIM_E_ARG (-1)."""
    from simulon_amd import _lib
    lib = _lib.load()
    assert lib.sim_reduce(0, None, None, 0, 1, None, None, None) == -1
    assert lib.sim_nl_trigger(1, None, None, None, None, 0, None) == -1
    assert lib.sim_pos_export(None, 10, None, 0, None) == -1
    net = _lib.MLPNet()
    assert lib.sim_mlp_energy(1, None, None, 1, None, None, None, 1, net, None, None, None, None,
                              None) == -1
    assert lib.sim_mlp_grad(0, None, None, None, None, 1.0, 1.0, net, None, None, None, None,
                            None, 1, None, None) == -1
    assert lib.sim_mlp_adam(None, None, None, None, 1, 1e-3, 0.9, 0.999, 1e-8, net, None) == -1


def test_struct_layouts_match_c(tmp_path):
    from simulon_amd import _lib
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "simulon_hip.h"\n'
                   'int main(){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %d %d\\n", sizeof(sim_replica_t),'
                   'sizeof(sim_types_t), sizeof(sim_lj_params_t), sizeof(sim_eam_params_t),'
                   'offsetof(sim_replica_t, atom_base), offsetof(sim_eam_params_t, inv_dr),'
                   'offsetof(sim_replica_t, damp), sizeof(sim_mlp_net_t),'
                   'offsetof(sim_mlp_net_t, rc), SIM_MLP_NPARAM, SIM_MLP_NGRAD);return 0;}\n')
    exe = tmp_path / "sizes"
    r = subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("no host C compiler: " + r.stderr)
    got = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    want = [ctypes.sizeof(_lib.Replica), ctypes.sizeof(_lib.Types), ctypes.sizeof(_lib.LJParams),
            ctypes.sizeof(_lib.EAMParams), _lib.Replica.atom_base.offset,
            _lib.EAMParams.inv_dr.offset, _lib.Replica.damp.offset, ctypes.sizeof(_lib.MLPNet),
            _lib.MLPNet.rc.offset, _lib.MLP_NPARAM, _lib.MLP_NGRAD]
    assert got == want


@pytest.mark.parametrize("ctype, cname", [("Replica", "sim_replica_t"), ("Stage", "sim_stage_t")])
def test_record_every_field_offset(tmp_path, ctype, cname):
    """Every field of sim_replica_t and sim_stage_t sits where the ctypes
    mirror puts it (the device kernels and the host read the same bytes)."""
    from simulon_amd import _lib
    cls = getattr(_lib, ctype)
    names = [f[0] for f in cls._fields_]
    body = "".join(f'printf("%zu\\n", offsetof({cname}, {n}));' for n in names)
    src = tmp_path / "offs.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "simulon_hip.h"\n'
                   f'int main(){{{body}printf("%zu\\n", sizeof({cname}));return 0;}}\n')
    exe = tmp_path / "offs"
    r = subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("no host C compiler: " + r.stderr)
    got = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    want = [getattr(cls, n).offset for n in names] + [ctypes.sizeof(cls)]
    assert got == want
