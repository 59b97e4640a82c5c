"""GPU parity of the FFJORD forward (node_cnf_forward): log-likelihood, end
state and Δlogp against the oracle (oracle/ffjord.py, FP64) with the same
fixed Rademacher probe and identical FP32 parameters.

Tolerances: z(T) and Δlogp(T) within 20 * (abstol + reltol*|value|) of the
oracle (adaptive step sequences may differ at the FP32 rounding floor); the
linear-flow closed form Δlogp = -eᵀAe T to 1e-5 relative.
"""
import numpy as np
import pytest

from oracle import ffjord as FF
from oracle.mlp import MLP

pytestmark = pytest.mark.gpu

CASES = {
    # name: widths, acts, B, tol
    "time_small": ([3, 16, 16, 2], ["tanh", "tanh", "identity"], 64, 1e-5),
    "z_only": ([2, 24, 2], ["tanh", "identity"], 37, 1e-5),
    "config5_shape": ([17, 512, 512, 16], ["tanh", "tanh", "identity"], 128, 1e-5),
}


def setup(name, prec, seed=0):
    from diffeqflux_hip import Handle, _lib
    widths, acts, B, tol = CASES[name]
    D = widths[-1]
    m = MLP(widths, acts)
    rng = np.random.default_rng(seed)
    p32 = m.glorot_init(rng).astype(np.float32)
    p32[m.offsets[0][1]:m.offsets[0][2]] = rng.uniform(-0.2, 0.2, widths[1])
    x = rng.standard_normal((B, D)).astype(np.float32)
    e = rng.choice(np.array([-1.0, 1.0], np.float32), size=(B, D))
    tc = widths[0] == D + 1
    h = Handle(widths, acts, _lib.MAP_CNF_TIME if tc else _lib.MAP_CNF, prec, B)
    h.set_params(p32)
    print(f"\n[ffjord {name} prec={prec} B={B} widths={widths}]", flush=True)
    return m, p32, x, e, tc, h, tol


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("name", list(CASES))
def test_cnf_forward_matches_oracle(name, prec):
    if prec == "bf16x3" and name != "config5_shape":
        pytest.skip("split BF16 checked on the workload shape only")
    m, p32, x, e, tc, h, tol = setup(name, prec)
    lp, z, dl, st = h.cnf_forward(x, e, 0.0, 1.0, abstol=tol, reltol=tol)
    assert st["retcode"] == 1
    rlp, rz, rdl, sol = FF.forward(m, p32.astype(np.float64), x.astype(np.float64), 0.0, 1.0,
                                   probe=e.astype(np.float64), time_concat=tc,
                                   abstol=tol, reltol=tol)
    ez = np.abs(z - rz) / (tol + tol * np.abs(rz))
    ed = np.abs(dl - rdl) / (tol + tol * np.abs(rdl))
    el = np.abs(lp - rlp).max() / np.abs(rlp).max()
    print(f"  steps gpu {st['naccept']}+{st['nreject']} oracle {sol.naccept}+{sol.nreject}; "
          f"scaled err z {ez.max():.3f} dlogp {ed.max():.3f}; logpx rel {el:.2e}", flush=True)
    assert ez.max() < 20 and ed.max() < 20
    assert el < 1e-3


def test_cnf_linear_flow_closed_form():
    from diffeqflux_hip import Handle, _lib
    D, B, T = 4, 16, 0.75
    m = MLP([D, D, D], ["identity", "identity"])
    rng = np.random.default_rng(5)
    W1 = 0.3 * rng.standard_normal((D, D)); W2 = 0.3 * rng.standard_normal((D, D))
    p = np.zeros(m.nparams, np.float32)
    (w0, b0, _), (w1, b1, _) = m.offsets
    p[w0:b0] = W1.T.ravel(); p[w1:b1] = W2.T.ravel()
    A = W2 @ W1
    x = rng.standard_normal((B, D)).astype(np.float32)
    e = rng.choice(np.array([-1.0, 1.0], np.float32), size=(B, D))
    h = Handle([D, D, D], ["identity", "identity"], _lib.MAP_CNF, "fp32", B)
    h.set_params(p)
    lp, z, dl, st = h.cnf_forward(x, e, 0.0, T, abstol=1e-7, reltol=1e-7)
    want = -np.einsum("bi,ij,bj->b", e.astype(np.float64), A, e.astype(np.float64)) * T
    err = np.abs(dl - want).max() / np.abs(want).max()
    print(f"\n  linear flow: dlogp rel err {err:.2e}", flush=True)
    assert err < 1e-5


def test_cnf_handles_refuse_the_plain_entry_points():
    from diffeqflux_hip import Handle, NodeHipError, _lib
    h = Handle([3, 8, 2], ["tanh", "identity"], _lib.MAP_CNF_TIME, "fp32", 4)
    h.set_params(np.zeros(h.P, np.float32))
    with pytest.raises(NodeHipError, match="node_cnf_forward"):
        h.forward(np.zeros((4, 3), np.float32), Handle.make_opts(0.0, 1.0))
    with pytest.raises(NodeHipError):
        Handle([3, 8, 2], ["tanh", "tanh"], _lib.MAP_CNF_TIME, "fp32", 4)


def test_ffjord_layer_api():
    """FFJORD(model, tspan, input_dims)(x, ps, st) -> ((logpx, λ₁, λ₂), st),
    logpx of shape (1, B) (src/ffjord.jl:136-178)."""
    from diffeqflux_hip import FFJORD, Chain, Dense
    model = Chain(Dense(3, 16, "tanh"), Dense(16, 2))
    ff = FFJORD(model, (0.0, 1.0), (2,), abstol=1e-5, reltol=1e-5)
    ps, st = ff.setup(np.random.default_rng(0))
    x = np.random.default_rng(1).standard_normal((10, 2)).astype(np.float32)
    st["rng"] = np.random.default_rng(2)
    (lp, l1, l2), st2 = ff(x, ps, st)
    assert lp.shape == (1, 10) and np.all(np.isfinite(lp)) and l1 is l2
    e = np.random.default_rng(2).choice(np.array([-1.0, 1.0], np.float32), size=(10, 2))
    m = MLP([3, 16, 2], ["tanh", "identity"])
    rlp, _, _, _ = FF.forward(m, ps.astype(np.float64), x.astype(np.float64), 0.0, 1.0,
                              probe=e.astype(np.float64), time_concat=True, abstol=1e-5, reltol=1e-5)
    assert np.abs(lp[0] - rlp).max() / np.abs(rlp).max() < 1e-3


def test_cnf_handle_refuses_plain_forward_entry_points():
    """A CNF handle holds the state [z; log p] and needs its probe: the plain
    NeuralODE forward entry points (host and device) refuse it before touching
    any buffer."""
    import ctypes
    import torch
    from diffeqflux_hip import Handle, _lib
    *_, h, _ = setup("z_only", "fp32")
    opts = Handle.make_opts(0.0, 1.0)
    st = _lib.Stats()
    u = torch.zeros((h.B, 2), device="cuda")
    out = torch.zeros((2, h.B, 2), device="cuda")
    rc = h.lib.node_forward_device(h.h, ctypes.byref(opts), ctypes.c_void_p(u.data_ptr()), None,
                                   0, ctypes.c_void_p(out.data_ptr()), 2, ctypes.byref(st))
    assert rc == -5, rc
    assert b"node_cnf_forward" in h.lib.node_last_error(h.h)
    uh = np.zeros((h.B, 2), np.float32)
    oh = np.zeros((2, h.B, 2), np.float32)
    rc = h.lib.node_forward(h.h, ctypes.byref(opts), uh.ctypes.data_as(ctypes.c_void_p), None, 0,
                            oh.ctypes.data_as(ctypes.c_void_p), 2, ctypes.byref(st))
    assert rc == -5, rc
