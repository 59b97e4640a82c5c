"""node_rhs_device_noalloc (the entry point that allocates nothing: model
input and hidden activations in the library's fixed per-GPU array) against
node_rhs, the entry point it is a variant of, through the C ABI: two calls
in a row on one handle, the second with fewer trajectories (the array keeps
the first call's contents), FP32 and split BF16.  Tolerance: 1e-6 of the
largest output (the two entry points run the same GEMM kernels; only the
tile plan for the smaller batch may differ)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("widths,acts,imap", [
    ([32, 256, 256, 32], ["tanh", "tanh", "identity"], 0),
    ([2, 50, 2], ["tanh", "identity"], 1),
])
def test_noalloc_rhs_matches_node_rhs(widths, acts, imap, prec):
    import torch
    from diffeqflux_hip import Handle
    from oracle.mlp import MLP
    B = 512
    m = MLP(widths, acts, input_map=imap)
    p = m.glorot_init(np.random.default_rng(0)).astype(np.float32)
    h = Handle(widths, acts, imap, prec, B)
    h.set_params(p)
    D = widths[0]
    rng = np.random.default_rng(1)
    for n in (B, 77):      # second call smaller than the first
        print(f"\n[noalloc widths={widths} prec={prec} n={n}]", flush=True)
        u = rng.standard_normal((n, D)).astype(np.float32)
        full = np.zeros((B, D), np.float32)
        full[:n] = u
        ref = h.rhs(full)[:n]                         # node_rhs (host pointers)
        du = torch.from_numpy(u).cuda()
        out = torch.full((n, D), float("nan"), device="cuda")
        torch.cuda.synchronize()
        rc = h.lib.node_rhs_device_noalloc(h.h, ctypes.c_void_p(du.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr()), n)
        h._check(rc, "node_rhs_device_noalloc")
        got = out.cpu().numpy()
        e = _rel(got, ref)
        print(f"  rel diff vs node_rhs {e:.2e}", flush=True)
        assert np.all(np.isfinite(got)) and e < 1e-6


def test_noalloc_rhs_rejects_beyond_capacity():
    from diffeqflux_hip import Handle
    widths, acts = [64, 1024, 1024, 1024, 64], ["softplus"] * 3 + ["identity"]
    from oracle.mlp import MLP
    h = Handle(widths, acts, 0, "bf16x3", 256)
    h.set_params(MLP(widths, acts).glorot_init(np.random.default_rng(0)).astype(np.float32))
    import torch
    buf = torch.zeros(64, device="cuda")     # never read: the size check comes first
    rc = h.lib.node_rhs_device_noalloc(h.h, ctypes.c_void_p(buf.data_ptr()),
                                       ctypes.c_void_p(buf.data_ptr()), 1 << 20)
    assert rc == -4                          # NODE_ERR_CAPACITY
    assert b"no-alloc array" in h.lib.node_last_error(h.h)
