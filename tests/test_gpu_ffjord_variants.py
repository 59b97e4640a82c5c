"""GPU parity of the FFJORD variants of the layer state (src/ffjord.jl:60-65,
81-178): monte_carlo = false (the exact trace: one unit-probe pass per state
dimension in every right-hand side, one tangent chain per unit probe in every
adjoint VJP) and regularize = true (state rows r1' = |f|^2, r2' = |e^T df/dz|,
returned as λ₁, λ₂), forward and adjoint, against the oracle's restatement
(oracle.ffjord.forward / backward with the same probe; pinned there by torch
FP64 double-backward, closed forms and finite differences).

Tolerances as the Hutchinson FFJORD tests: fixed dt FP32 1e-4 and split BF16
2e-4 (relative, max norm) on states, λ's and gradients; adaptive FP32 within
50 * tol.  The regularisers carry no gradient on the HIP path (zero
cotangents, as the reference's own losses -mean(logpx) give them)."""
import numpy as np
import pytest

from oracle import ffjord as FF
from oracle.mlp import MLP

from test_gpu_parity import TOL, rel

pytestmark = pytest.mark.gpu

CASES = {
    "z_only": ([3, 24, 24, 3], ["tanh", "softplus", "identity"], 41, False),
    "time": ([5, 32, 32, 4], ["tanh", "sigmoid", "identity"], 64, True),
}


def _setup(name, prec, reg, mc, seed=0):
    from diffeqflux_hip import Handle, _lib
    widths, acts, B, tc = CASES[name]
    D = widths[-1]
    m = MLP(widths, acts)
    rng = np.random.default_rng(seed)
    p32 = m.glorot_init(rng).astype(np.float32)
    for li in range(len(acts) - 1):
        w0, b0, e = m.offsets[li]
        p32[b0:e] = rng.uniform(-0.3, 0.3, e - b0)
    x = rng.standard_normal((B, D)).astype(np.float32)
    e = rng.choice(np.array([-1.0, 1.0], np.float32), size=(B, D))
    imap = (_lib.MAP_CNF_TIME if tc else _lib.MAP_CNF) | (_lib.MAP_CNF_REGULARIZE if reg else 0) \
        | (0 if mc else _lib.MAP_CNF_EXACT_TRACE)
    h = Handle(widths, acts, imap, prec, B)
    h.set_params(p32)
    print(f"\n[ffjord {name} prec={prec} regularize={reg} monte_carlo={mc} B={B}]", flush=True)
    return m, p32, x, e, tc, h


@pytest.mark.parametrize("reg,mc", [(False, False), (True, True), (True, False)])
@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("name", list(CASES))
def test_variant_forward_and_adjoint_fixed_dt(name, prec, reg, mc):
    m, p32, x, e, tc, h = _setup(name, prec, reg, mc)
    B, D = x.shape
    kw = dict(adaptive=False, dt=1.0 / 16)
    probe = e if (mc or reg) else None
    lp, z, dl, st = h.cnf_forward(x, probe, 0.0, 1.0, **kw)
    assert st["retcode"] == 1 and st["naccept"] == 16
    p64 = p32.astype(np.float64)
    rlp, rz, rdl, sol = FF.forward(m, p64, x.astype(np.float64), 0.0, 1.0, probe=e.astype(np.float64),
                                   time_concat=tc, regularize=reg, monte_carlo=mc, **kw)
    errs = {"z": rel(z, rz), "dlogp": rel(dl, rdl), "logpx": rel(lp, rlp)}
    if reg:
        l1, l2 = h.cnf_regularizers()
        errs["lam1"], errs["lam2"] = rel(l1, sol.lam1), rel(l2, sol.lam2)
    rng = np.random.default_rng(3)
    gz = rng.standard_normal((B, D)).astype(np.float32)
    gd = rng.standard_normal(B).astype(np.float32)
    dx, dp, sb = h.cnf_backward(gz, gd)
    assert sb["retcode"] == 1
    gx, gp, _ = FF.backward(m, p64, sol, e.astype(np.float64), gz.astype(np.float64),
                            gd.astype(np.float64), time_concat=tc, regularize=reg, monte_carlo=mc,
                            **kw)
    errs["dx"], errs["dp"] = rel(dx, gx), rel(dp, gp)
    print("  " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    bar = TOL[prec]["grad"]
    assert all(v < bar for v in errs.values()), errs


@pytest.mark.parametrize("reg,mc", [(False, False), (True, False)])
def test_variant_adaptive_fp32(reg, mc):
    m, p32, x, e, tc, h = _setup("time", "fp32", reg, mc, seed=1)
    B, D = x.shape
    tol = dict(abstol=1e-6, reltol=1e-6)
    lp, z, dl, st = h.cnf_forward(x, e, 0.0, 1.0, **tol)
    p64 = p32.astype(np.float64)
    rlp, rz, rdl, sol = FF.forward(m, p64, x.astype(np.float64), 0.0, 1.0, probe=e.astype(np.float64),
                                   time_concat=tc, regularize=reg, monte_carlo=mc, **tol)
    gz, gd = FF.nll_cotangents(z.astype(np.float64), B)
    dx, dp, sb = h.cnf_backward(gz.astype(np.float32), gd.astype(np.float32))
    gx, gp, adj = FF.backward(m, p64, sol, e.astype(np.float64), gz, gd, time_concat=tc,
                              regularize=reg, monte_carlo=mc, **tol)
    errs = {"logpx": rel(lp, rlp), "dx": rel(dx, gx), "dp": rel(dp, gp)}
    print("  " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()) +
          f"; steps fwd {st['naccept']}+{st['nreject']} (oracle {sol.naccept}) adj "
          f"{sb['naccept']}+{sb['nreject']} (oracle {adj.naccept})", flush=True)
    assert all(v < 50 * 1e-6 + 1e-4 for v in errs.values()), errs


def test_exact_trace_linear_flow_closed_form():
    # f = A z (two linear layers): Δlogp(T) = -tr(A) T exactly
    from diffeqflux_hip import Handle, _lib
    rng = np.random.default_rng(5)
    D, B, T = 3, 40, 0.8
    m = MLP([D, 8, D], ["identity", "identity"])
    p = np.zeros(m.nparams)
    (w0, b0, _), (w1, b1, _) = m.offsets
    W1, W2 = 0.4 * rng.standard_normal((8, D)), 0.4 * rng.standard_normal((D, 8))
    p[w0:b0] = W1.T.ravel()
    p[w1:b1] = W2.T.ravel()
    A = W2 @ W1
    h = Handle([D, 8, D], ["identity", "identity"], _lib.MAP_CNF | _lib.MAP_CNF_EXACT_TRACE,
               "fp32", B)
    h.set_params(p.astype(np.float32))
    x = rng.standard_normal((B, D)).astype(np.float32)
    _, _, dl, st = h.cnf_forward(x, None, 0.0, T, abstol=1e-8, reltol=1e-8)
    e = float(np.abs(dl + np.trace(A) * T).max())
    print(f"\n  exact-trace linear flow: |Δlogp + tr(A) T| max {e:.2e}", flush=True)
    assert e < 1e-5


def test_layer_api_variants_autograd():
    """FFJORD(...)(x, ps, st) with st.regularize / st.monte_carlo through
    torch autograd: -mean(logpx) gradient vs the oracle; λ's are the
    regulariser rows; a loss on λ raises."""
    import torch
    from diffeqflux_hip import FFJORD, Chain, Dense, Tsit5
    rng = np.random.default_rng(8)
    model = Chain(Dense(2, 16, "tanh"), Dense(16, 2))
    ffj = FFJORD(model, (0.0, 1.0), (2,), Tsit5(), abstol=1e-7, reltol=1e-7)
    ps, st = ffj.setup(rng, init="glorot")
    x = rng.standard_normal((30, 2)).astype(np.float32)
    e = rng.choice(np.array([-1.0, 1.0], np.float32), size=(30, 2))
    m = MLP([2, 16, 2], ["tanh", "identity"])
    for reg, mc in [(True, True), (False, False), (True, False)]:
        s = dict(st, regularize=reg, monte_carlo=mc, probe=e)
        pt = torch.tensor(ps, requires_grad=True)
        (lp, l1, l2), _ = ffj(torch.tensor(x), pt, s)
        (-lp.mean()).backward()
        _, rz, _, sol = FF.forward(m, ps.astype(np.float64), x.astype(np.float64), 0.0, 1.0,
                                   probe=e.astype(np.float64), regularize=reg, monte_carlo=mc,
                                   abstol=1e-7, reltol=1e-7)
        gz, gd = FF.nll_cotangents(rz, 30)
        _, gp, _ = FF.backward(m, ps.astype(np.float64), sol, e.astype(np.float64), gz, gd,
                               regularize=reg, monte_carlo=mc, abstol=1e-7, reltol=1e-7)
        err = rel(pt.grad.numpy(), gp)
        print(f"\n  layer regularize={reg} monte_carlo={mc}: dL/dp rel err {err:.2e}", flush=True)
        assert err < 1e-3
        if reg:
            assert rel(l1.detach().numpy().ravel(), sol.lam1) < 1e-3
            pt2 = torch.tensor(ps, requires_grad=True)
            (lp2, l12, _), _ = ffj(torch.tensor(x), pt2, s)
            with pytest.raises(NotImplementedError):
                (-lp2.mean() + l12.mean()).backward()
