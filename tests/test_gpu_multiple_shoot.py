"""GPU parity of multiple shooting (``src/multiple_shooting.jl``) on the
reference's documented example (``docs/src/examples/multiple_shooting.md``:
30 points of the cubic spiral on (0, 5), ``Chain(x -> x.^3, Dense(2, 16,
tanh), Dense(16, 2))``, group size 3, continuity term 200, loss
``sum(abs2, data - pred)``).  The reference's own multiple-shooting test set
is marked broken (``test/BasicNeuralDE/multiple_shoot_tests.jl``), so the
oracle is the expected value: every group solved alone by the oracle's Tsit5
from the data point that starts it, its InterpolatingAdjoint for the
gradient, and the loss assembled as the reference does.

Tolerances (rtol = atol = 1e-6 solves, FP32 on the GPU): group predictions
within 20 * (atol + rtol |u|) of the FP64 oracle, loss and dL/dp within
50 * tol (relative, max norm)."""
import numpy as np
import pytest

from oracle import adjoint as OA
from oracle.mlp import MLP, MAP_CUBE

pytestmark = pytest.mark.gpu

TOL = 1e-6
TERM = 200.0


def spiral_data(u0, t):
    from scipy.integrate import solve_ivp
    A = np.array([[-0.1, 2.0], [-2.0, -0.1]])   # true_A of the docs example
    s = solve_ivp(lambda _t, u: (u ** 3) @ A, (t[0], t[-1]), u0, t_eval=t, method="DOP853",
                  rtol=1e-11, atol=1e-12)
    return s.y.T.astype(np.float32)


def oracle_multiple_shoot(m, p64, data, t, gs, tol=TOL, term=TERM):
    from diffeqflux_hip import group_ranges
    rs = group_ranges(len(t), gs)
    fwd, preds = [], []
    for rg in rs:
        s = [float(x) for x in t[rg.start:rg.stop]]
        ref = OA.forward(m, p64, data[rg.start][None].astype(np.float64), s[0], s[-1], saveat=s,
                         abstol=tol, reltol=tol)
        fwd.append(ref)
        preds.append(np.array(ref.u)[:, 0, :])
    cots = [2.0 * (P - data[rg.start:rg.stop]) for P, rg in zip(preds, rs)]
    loss = 0.0
    for i, rg in enumerate(rs):
        loss += float(((data[rg.start:rg.stop] - preds[i]) ** 2).sum())
        if i > 0:
            d = preds[i - 1][-1] - data[rg.start]
            loss += term * float(np.abs(d).sum())
            cots[i - 1][-1] = cots[i - 1][-1] + term * np.sign(d)
    gp = np.zeros_like(p64)
    for ref, c in zip(fwd, cots):
        _, g, _ = OA.backward(m, p64, ref, [c[k][None] for k in range(len(c))], abstol=tol,
                              reltol=tol)
        gp += g
    return loss, preds, gp


def setup(hidden=16, seed=0):
    m = MLP([2, hidden, 2], ["tanh", "identity"], input_map=MAP_CUBE)
    rng = np.random.default_rng(seed)
    p32 = m.glorot_init(rng).astype(np.float32)
    t = np.linspace(0.0, 5.0, 30)
    return m, p32, t


def gpu_multiple_shoot(p32, data, t, gs, hidden=16, ensemble=False, **kw):
    import torch
    from diffeqflux_hip import (Chain, Cube, Dense, NeuralODE, Tsit5, multiple_shoot,
                                EnsembleProblem, EnsembleThreads)
    nn = Chain(Cube(), Dense(2, hidden, "tanh"), Dense(hidden, 2))
    node = NeuralODE(nn, (float(t[0]), float(t[-1])), Tsit5(), saveat=list(t), abstol=TOL,
                     reltol=TOL)
    pt = torch.tensor(p32, dtype=torch.float32, requires_grad=True)
    loss_fn = lambda u, v: ((u - v) ** 2).sum()                   # noqa: E731
    if ensemble:
        loss, preds = multiple_shoot(pt, data, t, EnsembleProblem(node), EnsembleThreads(),
                                     loss_fn, Tsit5(), gs, continuity_term=TERM,
                                     trajectories=data.shape[0], **kw)
    else:
        loss, preds = multiple_shoot(pt, data, t, node, loss_fn, Tsit5(), gs,
                                     continuity_term=TERM, **kw)
    if not np.isfinite(float(loss)):
        return float(loss), preds, None
    loss.backward()
    return float(loss), [p.detach().cpu().numpy() for p in preds], pt.grad.numpy().astype(np.float64)


def relmax(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("gs", [3, 7, 30])
def test_multiple_shoot_docs_example_matches_oracle(gs):
    m, p32, t = setup()
    data = spiral_data(np.array([2.0, 0.0]), t)
    loss, preds, grad = gpu_multiple_shoot(p32, data, t, gs)
    rl, rp, rg = oracle_multiple_shoot(m, p32.astype(np.float64), data, t, gs)
    worst = max(float((np.abs(a - b) / (TOL + TOL * np.abs(b))).max()) for a, b in zip(preds, rp))
    el, eg = abs(loss - rl) / abs(rl), relmax(grad, rg)
    print(f"\n[multiple_shoot gs={gs}] groups {len(rp)} loss {loss:.6e} (oracle {rl:.6e}, rel "
          f"{el:.2e}); worst scaled pred err {worst:.2f}; dL/dp rel err {eg:.2e}", flush=True)
    assert [p.shape for p in preds] == [q.shape for q in rp]
    for a, rg_ in zip(preds, __import__("diffeqflux_hip").group_ranges(30, gs)):
        assert np.array_equal(a[0], data[rg_.start])              # each group starts on data
    assert worst < 20
    assert el < 50 * TOL and eg < 50 * TOL + 1e-5


def test_ensemble_form_is_the_sum_over_trajectories():
    """Ensemble multiple shooting (src/multiple_shooting.jl:131-195) with a
    separable loss equals the sum of the single-trajectory results; each
    trajectory also against the oracle."""
    m, p32, t = setup(seed=1)
    u0s = [np.array([2.0, 0.0]), np.array([1.5, 0.5]), np.array([-1.0, 1.2])]
    data = np.stack([spiral_data(u, t) for u in u0s])              # (3, 30, 2)
    loss, preds, grad = gpu_multiple_shoot(p32, data, t, 5, ensemble=True)
    tl, tg = 0.0, np.zeros_like(grad)
    ol, og = 0.0, np.zeros_like(grad)
    for j in range(3):
        l1, p1, g1 = gpu_multiple_shoot(p32, data[j], t, 5)
        tl, tg = tl + l1, tg + g1
        for a, b in zip(preds, p1):
            assert np.array_equal(a[j], b)
        rl, _, rg = oracle_multiple_shoot(m, p32.astype(np.float64), data[j], t, 5)
        ol, og = ol + rl, og + rg
    print(f"\n[ensemble] loss {loss:.6e} sum {tl:.6e} oracle {ol:.6e}; grad vs sum "
          f"{relmax(grad, tg):.2e} vs oracle {relmax(grad, og):.2e}", flush=True)
    assert abs(loss - tl) <= 1e-5 * abs(tl)
    assert relmax(grad, tg) < 1e-5
    assert abs(loss - ol) / abs(ol) < 50 * TOL and relmax(grad, og) < 50 * TOL + 1e-5


def test_wide_model_uses_batched_group_solves():
    """Hidden width 160 is beyond the per-trajectory kernels (<= 128): every
    group runs as a batched NeuralODE solve; same oracle parity."""
    m, p32, t = setup(hidden=160, seed=2)
    data = spiral_data(np.array([2.0, 0.0]), t)
    loss, preds, grad = gpu_multiple_shoot(p32, data, t, 10, hidden=160)
    rl, rp, rg = oracle_multiple_shoot(m, p32.astype(np.float64), data, t, 10)
    el, eg = abs(loss - rl) / abs(rl), relmax(grad, rg)
    print(f"\n[batched groups] loss rel err {el:.2e}; dL/dp rel err {eg:.2e}", flush=True)
    assert el < 50 * TOL and eg < 50 * TOL + 1e-5


def test_failed_group_solve_gives_infinite_loss():
    """A group that does not reach its end (maxiters = 2) makes the loss
    infinite (src/multiple_shooting.jl:66-67)."""
    m, p32, t = setup()
    data = spiral_data(np.array([2.0, 0.0]), t)
    loss, preds, grad = gpu_multiple_shoot(p32, data, t, 10, maxiters=2)
    assert loss == float("inf") and grad is None and len(preds) == 4


def test_backward_of_a_superseded_forward_raises():
    """A handle keeps the tape of its latest solve only: adjointing an older
    solve of the same layer raises instead of using the wrong tape."""
    import torch
    from diffeqflux_hip import Chain, Cube, Dense, NeuralODE, Tsit5, multiple_shoot
    m, p32, t = setup()
    data = spiral_data(np.array([2.0, 0.0]), t)
    node = NeuralODE(Chain(Cube(), Dense(2, 16, "tanh"), Dense(16, 2)), (0.0, 5.0), Tsit5(),
                     abstol=TOL, reltol=TOL)
    pt = torch.tensor(p32, requires_grad=True)
    loss_fn = lambda u, v: ((u - v) ** 2).sum()                   # noqa: E731
    l1, _ = multiple_shoot(pt, data, t, node, loss_fn, Tsit5(), 5)
    l2, _ = multiple_shoot(pt, data, t, node, loss_fn, Tsit5(), 5)
    l2.backward()                                                 # latest: fine
    with pytest.raises(RuntimeError):
        l1.backward()
