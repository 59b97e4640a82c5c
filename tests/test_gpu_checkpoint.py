"""GPU parity of InterpolatingAdjoint(checkpointing = true): the forward keeps
only the checkpoint states (a ring of four tape entries), the adjoint
re-solves every interval between consecutive checkpoints (t0, the saved
times, t1) before integrating over it.

Checked through the C ABI against the oracle's restatement
(``oracle.adjoint.Checkpoints``) on the same inputs, against the dense-tape
adjoint of the same handle, and with a tape budget far below what the dense
tape needs (device memory bounded by the longest interval, not the solve).
Tolerances: fixed dt FP32 1e-4 (gradients, relative to the largest entry);
adaptive solves 2e-3 against the FP64 oracle at rtol 1e-6 (the FP32 error
estimate may take another valid step sequence), 1e-3 at the headline shape
(rtol 1e-4) between checkpointed and dense split-BF16 gradients.
"""
import numpy as np
import pytest

from oracle import adjoint as OA
from oracle.mlp import MLP, MAP_CUBE

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _spiral(prec="fp32", B=3, seed=0):
    from diffeqflux_hip import Handle
    widths, acts = [2, 50, 2], ["tanh", "identity"]
    m = MLP(widths, acts, input_map=MAP_CUBE)
    rng = np.random.default_rng(seed)
    p = m.glorot_init(rng).astype(np.float32)
    p[m.offsets[0][1]:m.offsets[0][2]] = rng.uniform(-0.1, 0.1, widths[1])
    u0 = np.array([[2.0, 0.0], [1.0, 0.5], [0.0, 1.0]], np.float32)[:B]
    h = Handle(widths, acts, MAP_CUBE, prec, B)
    h.set_params(p)
    return m, p, u0, h


def _solve(h, u0, opts, interior, cot_seed=3, ckpt=False):
    h.set_option("checkpointing", ckpt)
    out, ts, st = h.forward(u0, opts, interior)
    assert st["retcode"] == 1
    c = np.random.default_rng(cot_seed).standard_normal(out.shape).astype(np.float32)
    du0, dp, sb = h.backward(c)
    assert sb["retcode"] == 1
    return out, ts, c, du0, dp, sb


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_fixed_dt_checkpointed_equals_dense_and_oracle(prec):
    from diffeqflux_hip import Handle
    m, p, u0, h = _spiral(prec)
    print(f"\n[checkpoint fixed dt spiral B=3 {prec}]", flush=True)
    opts = Handle.make_opts(0.0, 1.5, adaptive=False, dt=0.05)
    interior = [0.15 * i for i in range(1, 10)]
    out0, ts, c, g0, gp0, _ = _solve(h, u0, opts, interior, ckpt=False)
    out1, ts1, c1, g1, gp1, sb = _solve(h, u0, opts, interior, ckpt=True)
    assert np.array_equal(out0, out1) and np.array_equal(ts, ts1)
    ref = OA.forward(m, p.astype(np.float64), u0.astype(np.float64), 0.0, 1.5,
                     saveat=list(ts), adaptive=False, dt=0.05)
    gu, gp, a = OA.backward(m, p.astype(np.float64), ref, [x.astype(np.float64) for x in c],
                            checkpointing=True, adaptive=False, dt=0.05)
    print(f"  ckpt vs dense du0 {rel(g1, g0):.2e} dp {rel(gp1, gp0):.2e}; vs oracle du0 "
          f"{rel(g1, gu):.2e} dp {rel(gp1, gp):.2e}; resolves {a.ck_resolves}; adjoint steps "
          f"{sb['naccept']}+{sb['nreject']} (oracle {a.naccept})", flush=True)
    assert rel(gp1, gp0) < 1e-5 and rel(g1, g0) < 1e-5
    bar = 1e-4 if prec == "fp32" else 2e-4
    assert rel(gp1, gp) < bar and rel(g1, gu) < bar


@pytest.mark.parametrize("prec,rtol", [("fp32", 1e-6), ("bf16x3", 1e-4)])
def test_adaptive_checkpointed_vs_oracle(prec, rtol):
    # split BF16 keeps ~16 mantissa bits: its step-size control is run at the
    # tolerances it is built for (1e-4), FP32 at 1e-6
    from diffeqflux_hip import Handle
    m, p, u0, h = _spiral(prec)
    print(f"\n[checkpoint adaptive spiral B=3 {prec} rtol {rtol}]", flush=True)
    tol = dict(abstol=1e-6, reltol=rtol)
    opts = Handle.make_opts(0.0, 1.5, **tol)
    interior = list(np.linspace(0.0, 1.5, 30)[1:-1])
    out, ts, c, g1, gp1, sb = _solve(h, u0, opts, interior, ckpt=True)
    ref = OA.forward(m, p.astype(np.float64), u0.astype(np.float64), 0.0, 1.5,
                     saveat=list(ts), **tol)
    gu, gp, a = OA.backward(m, p.astype(np.float64), ref, [x.astype(np.float64) for x in c],
                            checkpointing=True, **tol)
    _, _, _, g0, gp0, sb0 = _solve(h, u0, opts, interior, ckpt=False)
    print(f"  ckpt vs oracle du0 {rel(g1, gu):.2e} dp {rel(gp1, gp):.2e}; ckpt vs dense dp "
          f"{rel(gp1, gp0):.2e}; adjoint steps {sb['naccept']}+{sb['nreject']} (oracle "
          f"{a.naccept}+{a.nreject}, dense {sb0['naccept']})", flush=True)
    bar = 2e-3 if prec == "fp32" else 5e-3
    assert rel(gp1, gp) < bar and rel(g1, gu) < bar
    assert rel(gp1, gp0) < bar


def test_tape_budget_needs_checkpointing():
    # FP32 at rtol 1e-6: the forward takes ~14 steps, each of the 19
    # checkpoint intervals a few; a budget of 12 tape entries holds an
    # interval but not the whole solve
    from diffeqflux_hip import Handle
    m, p, u0, h = _spiral()
    print("\n[checkpoint tape budget spiral B=3 fp32]", flush=True)
    tol = dict(abstol=1e-6, reltol=1e-6)
    opts = Handle.make_opts(0.0, 1.5, **tol)
    interior = list(np.linspace(0.0, 1.5, 20)[1:-1])
    out, _, _, g0, gp0, sb0 = _solve(h, u0, opts, interior, ckpt=False)
    fwd_steps = h.last_stats["naccept"]
    ent = 8 * 128 * 2 * 4                       # bytes per tape entry (Bp = 128, D = 2)
    h2 = _spiral()[3]
    h2.set_option("tape_budget", 12 * ent)
    h2.set_option("checkpointing", False)
    with pytest.raises(RuntimeError, match="tape budget"):
        h2.forward(u0, opts, interior)
    _, _, _, g1, gp1, sb = _solve(h2, u0, opts, interior, ckpt=True)
    print(f"  forward steps {fwd_steps}; budget 12 entries; ckpt vs dense dp "
          f"{rel(gp1, gp0):.2e} du0 {rel(g1, g0):.2e}", flush=True)
    assert rel(gp1, gp0) < 1e-4 and rel(g1, g0) < 1e-4


def test_checkpointing_refusals():
    from diffeqflux_hip import Handle
    m, p, u0, h = _spiral()
    h.set_option("checkpointing", True)
    with pytest.raises(RuntimeError, match="save_everystep"):
        h.forward(u0, Handle.make_opts(0.0, 1.0, save_everystep=True), [])
    h.forward(u0, Handle.make_opts(0.0, 1.0), [0.5])
    h.set_option("sensealg", "backsolve")
    with pytest.raises(RuntimeError, match="checkpoints only"):
        h.backward(np.zeros((3, 3, 2), np.float32))


def test_layer_api_checkpointing_autograd():
    import torch
    from diffeqflux_hip import (Chain, Cube, Dense, NeuralODE, Tsit5, InterpolatingAdjoint,
                                ZygoteVJP)
    print("\n[checkpoint layer API spiral]", flush=True)
    model = Chain(Cube(), Dense(2, 50, "tanh"), Dense(50, 2))
    ts = list(np.linspace(0.0, 1.5, 30))
    grads = []
    for ck in (False, True):
        node = NeuralODE(model, (0.0, 1.5), Tsit5(), saveat=ts, abstol=1e-6, reltol=1e-6,
                         sensealg=InterpolatingAdjoint(autojacvec=ZygoteVJP(), checkpointing=ck))
        p0, _ = node.setup(np.random.default_rng(1), init="glorot")
        p = torch.tensor(p0, requires_grad=True)
        x = torch.tensor([[2.0, 0.0]], requires_grad=True)
        sol, _ = node(x, p)
        loss = (sol.array ** 2).sum()
        loss.backward()
        grads.append(p.grad.numpy().copy())
    print(f"  layer ckpt vs dense dp {rel(grads[1], grads[0]):.2e}", flush=True)
    assert rel(grads[1], grads[0]) < 1e-4


def test_headline_shape_checkpointed_vs_dense():
    # D=64, MLP 64-1024x3-64 softplus, 32 saves, rtol 1e-4, split BF16
    from diffeqflux_hip import Handle
    widths, acts = [64, 1024, 1024, 1024, 64], ["softplus"] * 3 + ["identity"]
    B = 4096
    m = MLP(widths, acts)
    p = m.glorot_init(np.random.default_rng(0)).astype(np.float32)
    u0 = np.random.default_rng(1).standard_normal((B, 64)).astype(np.float32)
    print(f"\n[checkpoint headline shape B={B} bf16x3]", flush=True)
    h = Handle(widths, acts, 0, "bf16x3", B)
    h.set_params(p)
    opts = Handle.make_opts(0.0, 1.0, abstol=1e-6, reltol=1e-4)
    interior = list(np.linspace(0.0, 1.0, 32)[1:-1])
    _, _, c, g0, gp0, sb0 = _solve(h, u0, opts, interior, ckpt=False)
    _, _, c1, g1, gp1, sb1 = _solve(h, u0, opts, interior, ckpt=True)
    e_p = float(np.linalg.norm(gp1 - gp0) / np.linalg.norm(gp0))
    e_u = float(np.linalg.norm(g1 - g0) / np.linalg.norm(g0))
    print(f"  ckpt vs dense: dp {e_p:.2e} du0 {e_u:.2e}; adjoint steps dense "
          f"{sb0['naccept']}+{sb0['nreject']} ckpt {sb1['naccept']}+{sb1['nreject']}", flush=True)
    assert e_p < 1e-3 and e_u < 1e-3


def test_single_trajectory_checkpointed_vs_oracle():
    """The docs spiral as the layer test uses it (B = 1, Glorot weights with
    zero biases, u0 = [2, 0], 30 saves, tol 1e-6, cotangent 2u): the
    checkpointed and dense adjoints each against the oracle's, with and
    without captured graphs."""
    from diffeqflux_hip import Handle
    widths, acts = [2, 50, 2], ["tanh", "identity"]
    m = MLP(widths, acts, input_map=MAP_CUBE)
    p = m.glorot_init(np.random.default_rng(1)).astype(np.float32)
    u0 = np.array([[2.0, 0.0]], np.float32)
    tol = dict(abstol=1e-6, reltol=1e-6)
    ts = list(np.linspace(0.0, 1.5, 30))
    ref = OA.forward(m, p.astype(np.float64), u0.astype(np.float64), 0.0, 1.5, saveat=ts, **tol)
    cot = [2.0 * u for u in ref.u]
    _, gd, ad = OA.backward(m, p.astype(np.float64), ref, cot, **tol)
    _, gc, ac = OA.backward(m, p.astype(np.float64), ref, cot, checkpointing=True, **tol)
    print(f"\n[checkpoint single trajectory] oracle ckpt vs dense {rel(gc, gd):.2e} "
          f"(adjoint steps {ac.naccept} / {ad.naccept})", flush=True)
    res = {}
    for graphs in (True, False):
        for ck in (False, True):
            h = Handle(widths, acts, MAP_CUBE, "fp32", 1)
            h.set_params(p)
            h.set_option("graphs", graphs)
            h.set_option("checkpointing", ck)
            out, t, st = h.forward(u0, Handle.make_opts(0.0, 1.5, **tol), ts[1:-1])
            c = (2.0 * out).astype(np.float32)
            _, dp, sb = h.backward(c)
            res[(graphs, ck)] = dp
            print(f"  graphs={graphs} ckpt={ck}: vs oracle dense {rel(dp, gd):.2e} vs oracle ckpt "
                  f"{rel(dp, gc):.2e}; steps fwd {st['naccept']} adj {sb['naccept']}+{sb['nreject']}",
                  flush=True)
    for (graphs, ck), dp in res.items():
        assert rel(dp, gc if ck else gd) < 1e-4
