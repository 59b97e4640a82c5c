"""CPU checks of the C ABI: the library loads, exports every entry point that
include/node_hip.h declares, and its structs have the C layout (checked with a
gcc-compiled probe of the header).  No compute call is made here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from diffeqflux_hip import _lib
from diffeqflux_hip.neural_ode import save_plan as host_save_plan
from oracle import tsit5 as otsit5
from oracle.mlp import MLP

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "node_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w]+\*?\s+\**(node_\w+)\s*\(", src, re.M)))


def test_header_declares_the_abi():
    fns = declared_functions()
    for f in ("node_create", "node_destroy", "node_set_params", "node_forward",
              "node_backward", "node_last_error", "node_forward_device",
              "node_backward_device", "node_rhs", "node_vjp"):
        assert f in fns


def test_library_exports_every_declared_symbol(hip_lib):
    for f in declared_functions():
        assert hasattr(hip_lib, f), f
    # and the ctypes table covers them all
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == declared_functions()


def test_struct_layouts_match_c(tmp_path):
    probe = tmp_path / "probe.c"
    probe.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "node_hip.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu\\n\", sizeof(node_desc),"
        " offsetof(node_desc, batch), sizeof(node_solve_opts), offsetof(node_solve_opts, maxiters),"
        " sizeof(node_stats), offsetof(node_stats, t_final));"
        " printf(\"%zu %zu\\n\", sizeof(node_ragged_stats), offsetof(node_ragged_stats, n_success));"
        " return 0;}\n")
    exe = tmp_path / "probe"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(probe), "-o", str(exe)])
    got = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    want = [ctypes.sizeof(_lib.NodeDesc), _lib.NodeDesc.batch.offset,
            ctypes.sizeof(_lib.SolveOpts), _lib.SolveOpts.maxiters.offset,
            ctypes.sizeof(_lib.Stats), _lib.Stats.t_final.offset,
            ctypes.sizeof(_lib.RaggedStats), _lib.RaggedStats.n_success.offset]
    assert got == want


def test_abi_version_and_param_count(hip_lib):
    assert hip_lib.node_abi_version() == 1
    for widths, acts in [([2, 50, 2], ["tanh", "identity"]),
                         ([64, 1024, 1024, 1024, 64], ["softplus"] * 3 + ["identity"])]:
        d = _lib.make_desc(widths, acts)
        assert hip_lib.node_num_params(ctypes.byref(d)) == MLP(widths, acts).nparams
    assert MLP([64, 1024, 1024, 1024, 64], ["softplus"] * 3 + ["identity"]).nparams == 2231360


def test_invalid_desc_rejected(hip_lib):
    d = _lib.make_desc([2, 50, 3], ["tanh", "identity"])  # output width != D
    assert hip_lib.node_num_params(ctypes.byref(d)) == -1


@pytest.mark.parametrize("kw", [
    dict(saveat=None, save_everystep=False, save_start=False),
    dict(saveat=np.arange(0, 11) / 10),
    dict(saveat=0.1),
    dict(saveat=np.linspace(0, 1.5, 30)),
    dict(),
])
def test_host_save_plan_matches_oracle(kw):
    t1 = 1.5 if kw.get("saveat") is not None and np.ndim(kw["saveat"]) and len(kw["saveat"]) == 30 else 1.0
    a = host_save_plan(0.0, t1, **kw)
    b = otsit5.save_plan(0.0, t1, **kw)
    assert a[1:] == b[1:]
    assert np.array_equal(np.array(a[0]), np.array(b[0]))
