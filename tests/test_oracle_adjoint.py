"""Pins the oracle InterpolatingAdjoint by closed forms (linear ODE), central
finite differences of the oracle's own forward solve, and torch-FP64
autograd through the discrete fixed-dt forward solver."""
import numpy as np
import pytest
import scipy.linalg as sla
import torch

from oracle import adjoint as adj
from oracle import tableau as tb
from oracle.mlp import MLP, MAP_CUBE
from oracle.problems import spiral


def _linear_problem(D=3, B=2, seed=0):
    rng = np.random.default_rng(seed)
    m = MLP([D, D], ["identity"])
    W = rng.standard_normal((D, D)) * 0.5
    p = np.zeros(m.nparams)
    p[:D * D] = W.T.ravel()
    return m, p, W, rng.standard_normal((B, D))


def test_linear_closed_forms():
    m, p, W, u0 = _linear_problem()
    T, D = 1.3, 3
    tol = dict(abstol=1e-12, reltol=1e-12)
    fwd = adj.forward(m, p, u0, 0.0, T, save_everystep=False, save_start=False, **tol)
    gu0, gp, _ = adj.backward(m, p, fwd, [fwd.u[-1]], **tol)
    E = sla.expm(W * T)
    assert np.abs(gu0 - u0 @ (E.T @ E).T).max() < 1e-11
    G = np.zeros((D, D))
    for b in range(u0.shape[0]):
        uT = E @ u0[b]
        M = np.zeros((2 * D, 2 * D))
        M[:D, :D], M[:D, D:], M[D:, D:] = W.T, np.outer(uT, u0[b]), W.T
        G += sla.expm(M * T)[:D, D:]          # Van Loan block exponential
    assert np.abs(gp[:D * D].reshape(D, D).T - G).max() < 1e-10


def test_saveat_jumps_closed_form():
    # L = sum_i <c_i, u(t_i)>: dL/du0 = sum_i e^{W^T t_i} c_i
    m, p, W, u0 = _linear_problem(seed=1)
    ts = [0.0, 0.4, 0.9, 1.2]
    rng = np.random.default_rng(2)
    cs = [rng.standard_normal(u0.shape) for _ in ts]
    tol = dict(abstol=1e-12, reltol=1e-12)
    fwd = adj.forward(m, p, u0, 0.0, 1.2, saveat=ts, **tol)
    assert fwd.t == ts
    gu0, _, _ = adj.backward(m, p, fwd, cs, **tol)
    expect = sum(c @ sla.expm(W * t) for c, t in zip(cs, ts))
    assert np.abs(gu0 - expect).max() < 1e-10


def _fd_grad(m, p, u0, tspan, saveat, target, tol, idx, h=1e-6):
    def loss(pp):
        f = adj.forward(m, pp, u0, *tspan, saveat=saveat, **tol)
        return float(np.sum((np.array(f.u) - target) ** 2))
    out = []
    for j in idx:
        e = np.zeros_like(p); e[j] = h
        out.append((loss(p + e) - loss(p - e)) / (2 * h))
    return np.array(out)


def test_spiral_adjoint_vs_central_fd():
    pr = spiral()
    m, p, u0 = pr["model"], pr["p"], pr["u0"]
    tol = dict(abstol=1e-10, reltol=1e-10)
    fwd = adj.forward(m, p, u0, *pr["tspan"], saveat=pr["saveat"], **tol)
    target = np.array(fwd.u) * 0.9
    cot = [2 * (u - t) for u, t in zip(fwd.u, target)]
    gu0, gp, _ = adj.backward(m, p, fwd, cot, **tol)
    idx = np.random.default_rng(0).choice(m.nparams, 10, replace=False)
    fd = _fd_grad(m, p, u0, pr["tspan"], pr["saveat"], target, tol, idx)
    assert np.abs(fd - gp[idx]).max() < 1e-5 * max(1.0, np.abs(fd).max())


def _torch_tsit5_fixed(f, u0, T, n):
    a, bb = tb.A, tb.B
    dt = T / n
    u = u0
    for _ in range(n):
        k = [f(u)]
        for s in range(1, 7):
            acc = sum(a[s][j] * k[j] for j in range(s))
            k.append(f(u + dt * acc))
        u = u + dt * sum(bb[j] * k[j] for j in range(6))
    return u


def test_fixed_dt_continuous_vs_discrete_adjoint():
    rng = np.random.default_rng(4)
    m = MLP([3, 16, 3], ["tanh", "identity"], input_map=MAP_CUBE)
    p = m.glorot_init(rng)
    u0 = rng.standard_normal((2, 3)) * 0.5
    T, n = 1.0, 64
    fwd = adj.forward(m, p, u0, 0.0, T, save_everystep=False, save_start=False,
                      adaptive=False, dt=T / n)
    gu0, gp, a = adj.backward(m, p, fwd, [fwd.u[-1]], adaptive=False, dt=T / n)
    assert a.naccept == n
    tu = torch.tensor(u0, requires_grad=True)
    tp = torch.tensor(p, requires_grad=True)

    def f(u):
        h = u ** 3
        for li, (i, j) in enumerate(zip(m.widths[:-1], m.widths[1:])):
            w0, b0, e = m.offsets[li]
            h = h @ tp[w0:b0].reshape(i, j) + tp[b0:e]
            if li == 0:
                h = torch.tanh(h)
        return h
    uT = _torch_tsit5_fixed(f, tu, T, n)
    (0.5 * (uT ** 2).sum()).backward()
    assert np.abs(uT.detach().numpy() - fwd.u[-1]).max() < 1e-13
    # continuous adjoint of a 5th-order discretisation: O(dt^5) agreement
    assert np.abs(gp - tp.grad.numpy()).max() < 1e-7
    assert np.abs(gu0 - tu.grad.numpy()).max() < 1e-7


def test_dropped_sign_would_be_caught():
    """A plausible mistake planted in a copy of the oracle -- the mu equation
    with the wrong sign -- fails the closed-form check."""
    m, p, W, u0 = _linear_problem()
    tol = dict(abstol=1e-10, reltol=1e-10)
    fwd = adj.forward(m, p, u0, 0.0, 1.0, save_everystep=False, save_start=False, **tol)
    _, gp, _ = adj.backward(m, p, fwd, [fwd.u[-1]], **tol)
    orig = m.vjp
    m.vjp = lambda u, pp, lam: (orig(u, pp, lam)[0], -orig(u, pp, lam)[1])
    try:
        _, gbad, _ = adj.backward(m, p, fwd, [fwd.u[-1]], **tol)
    finally:
        m.vjp = orig
    assert np.abs(gbad - gp).max() > 1e-2


# ---- BacksolveAdjoint restatement: same closed forms / finite differences
@pytest.mark.parametrize("ckpt", [True, False])
def test_backsolve_linear_closed_forms(ckpt):
    m, p, W, u0 = _linear_problem()
    T, D = 1.3, 3
    tol = dict(abstol=1e-12, reltol=1e-12)
    fwd = adj.forward(m, p, u0, 0.0, T, save_everystep=False, save_start=False, **tol)
    gu0, gp, _ = adj.backward_backsolve(m, p, fwd, [fwd.u[-1]], checkpointing=ckpt, **tol)
    E = sla.expm(W * T)
    assert np.abs(gu0 - u0 @ (E.T @ E).T).max() < 1e-10
    G = np.zeros((D, D))
    for b in range(u0.shape[0]):
        uT = E @ u0[b]
        M = np.zeros((2 * D, 2 * D))
        M[:D, :D], M[:D, D:], M[D:, D:] = W.T, np.outer(uT, u0[b]), W.T
        G += sla.expm(M * T)[:D, D:]
    assert np.abs(gp[:D * D].reshape(D, D).T - G).max() < 1e-9


@pytest.mark.parametrize("ckpt", [True, False])
def test_backsolve_saveat_jumps_closed_form(ckpt):
    m, p, W, u0 = _linear_problem(seed=1)
    ts = [0.0, 0.4, 0.9, 1.2]
    rng = np.random.default_rng(2)
    cs = [rng.standard_normal(u0.shape) for _ in ts]
    tol = dict(abstol=1e-12, reltol=1e-12)
    fwd = adj.forward(m, p, u0, 0.0, 1.2, saveat=ts, **tol)
    gu0, _, _ = adj.backward_backsolve(m, p, fwd, cs, checkpointing=ckpt, **tol)
    expect = sum(c @ sla.expm(W * t) for c, t in zip(cs, ts))
    assert np.abs(gu0 - expect).max() < 1e-9


def test_backsolve_spiral_vs_central_fd_and_interpolating():
    pr = spiral()
    m, p, u0 = pr["model"], pr["p"], pr["u0"]
    tol = dict(abstol=1e-10, reltol=1e-10)
    fwd = adj.forward(m, p, u0, *pr["tspan"], saveat=pr["saveat"], **tol)
    target = np.array(fwd.u) * 0.9
    cot = [2 * (u - t) for u, t in zip(fwd.u, target)]
    gu0, gp, _ = adj.backward_backsolve(m, p, fwd, cot, **tol)
    gi0, gi, _ = adj.backward(m, p, fwd, cot, **tol)
    assert np.abs(gp - gi).max() < 1e-6 * max(1.0, np.abs(gi).max())
    assert np.abs(gu0 - gi0).max() < 1e-6 * max(1.0, np.abs(gi0).max())
    idx = np.random.default_rng(0).choice(m.nparams, 10, replace=False)
    fd = _fd_grad(m, p, u0, pr["tspan"], pr["saveat"], target, tol, idx)
    assert np.abs(fd - gp[idx]).max() < 1e-5 * max(1.0, np.abs(fd).max())


def test_backsolve_without_u_integration_would_be_caught():
    """A restatement that forgot to integrate u (kept u(T) frozen) fails the
    closed form: plant it by zeroing f's u-part."""
    m, p, W, u0 = _linear_problem()
    T = 1.3
    tol = dict(abstol=1e-12, reltol=1e-12)
    fwd = adj.forward(m, p, u0, 0.0, T, save_everystep=False, save_start=False, **tol)
    orig = m.forward
    try:
        m.forward = lambda u, pp: np.zeros_like(u)
        _, gp, _ = adj.backward_backsolve(m, p, fwd, [fwd.u[-1]], checkpointing=False, **tol)
    finally:
        m.forward = orig
    _, gref, _ = adj.backward_backsolve(m, p, fwd, [fwd.u[-1]], checkpointing=False, **tol)
    assert np.abs(gp - gref).max() > 1e-3


# ---------------------------------------------------------------- checkpointing
def test_checkpointed_linear_closed_forms():
    # interior saves with zero cotangent only place checkpoints; L = |u(T)|^2 / 2
    m, p, W, u0 = _linear_problem(seed=3)
    T, D = 1.3, 3
    tol = dict(abstol=1e-12, reltol=1e-12)
    ts = [0.0, 0.3, 0.65, 1.0, T]
    fwd = adj.forward(m, p, u0, 0.0, T, saveat=ts, **tol)
    cot = [np.zeros_like(u0)] * (len(ts) - 1) + [fwd.u[-1]]
    gu0, gp, a = adj.backward(m, p, fwd, cot, checkpointing=True, **tol)
    assert a.ck_resolves == [3, 2, 1, 0]
    E = sla.expm(W * T)
    assert np.abs(gu0 - u0 @ (E.T @ E).T).max() < 1e-10
    G = np.zeros((D, D))
    for b in range(u0.shape[0]):
        uT = E @ u0[b]
        M = np.zeros((2 * D, 2 * D))
        M[:D, :D], M[:D, D:], M[D:, D:] = W.T, np.outer(uT, u0[b]), W.T
        G += sla.expm(M * T)[:D, D:]
    assert np.abs(gp[:D * D].reshape(D, D).T - G).max() < 1e-9


def test_checkpointed_saveat_jumps_closed_form():
    m, p, W, u0 = _linear_problem(seed=1)
    ts = [0.0, 0.4, 0.9, 1.2]
    rng = np.random.default_rng(2)
    cs = [rng.standard_normal(u0.shape) for _ in ts]
    tol = dict(abstol=1e-12, reltol=1e-12)
    fwd = adj.forward(m, p, u0, 0.0, 1.2, saveat=ts, **tol)
    gu0, _, _ = adj.backward(m, p, fwd, cs, checkpointing=True, **tol)
    expect = sum(c @ sla.expm(W * t) for c, t in zip(cs, ts))
    assert np.abs(gu0 - expect).max() < 1e-10


def test_checkpointed_fixed_dt_equals_dense():
    # save times on step boundaries: each interval re-solve repeats the
    # forward's own steps, so u(t) -- and the gradients -- are the dense ones
    pr = spiral()
    m, p, u0 = pr["model"], pr["p"], pr["u0"]
    kw = dict(adaptive=False, dt=0.05)
    ts = [0.15 * i for i in range(11)]
    fwd = adj.forward(m, p, u0, 0.0, 1.5, saveat=ts, **kw)
    cot = [2 * (u - 0.5) for u in fwd.u]
    g0, gp0, _ = adj.backward(m, p, fwd, cot, **kw)
    g1, gp1, a = adj.backward(m, p, fwd, cot, checkpointing=True, **kw)
    assert a.ck_resolves == list(range(9, -1, -1))
    assert np.abs(gp1 - gp0).max() < 1e-11 * np.abs(gp0).max()
    assert np.abs(g1 - g0).max() < 1e-11 * np.abs(g0).max()


def test_checkpointed_adaptive_matches_dense_and_fd():
    pr = spiral()
    m, p, u0 = pr["model"], pr["p"], pr["u0"]
    tol = dict(abstol=1e-9, reltol=1e-9)
    fwd = adj.forward(m, p, u0, *pr["tspan"], saveat=pr["saveat"], **tol)
    target = np.array(fwd.u) * 0.9
    cot = [2 * (u - t) for u, t in zip(fwd.u, target)]
    _, gp0, _ = adj.backward(m, p, fwd, cot, **tol)
    _, gp1, a = adj.backward(m, p, fwd, cot, checkpointing=True, **tol)
    assert a.ck_nf > 0 and len(a.ck_resolves) == len(pr["saveat"]) - 1
    assert np.abs(gp1 - gp0).max() < 1e-6 * np.abs(gp0).max()
    idx = np.random.default_rng(5).choice(m.nparams, 6, replace=False)
    fd = _fd_grad(m, p, u0, pr["tspan"], pr["saveat"], target, tol, idx)
    assert np.abs(fd - gp1[idx]).max() < 1e-5 * max(1.0, np.abs(fd).max())


def test_checkpointed_planted_wrong_checkpoint_fails(monkeypatch):
    # re-solving each interval from the next checkpoint's state is caught
    class Shifted(adj.Checkpoints):
        def _resolve(self, k, dt):
            self.u = self.u[1:] + self.u[-1:] if not hasattr(self, "_sh") else self.u
            self._sh = True
            return super()._resolve(k, dt)
    monkeypatch.setattr(adj, "Checkpoints", Shifted)
    m, p, W, u0 = _linear_problem(seed=1)
    ts = [0.0, 0.4, 0.9, 1.2]
    rng = np.random.default_rng(2)
    cs = [rng.standard_normal(u0.shape) for _ in ts]
    tol = dict(abstol=1e-12, reltol=1e-12)
    fwd = adj.forward(m, p, u0, 0.0, 1.2, saveat=ts, **tol)
    gu0, gp, _ = adj.backward(m, p, fwd, cs, checkpointing=True, **tol)
    # dL/du0 of a linear ODE does not depend on u(t); dL/dW does
    _, gp_ok, _ = adj.backward(m, p, fwd, cs, **tol)
    assert np.abs(gp - gp_ok).max() > 1e-3 * np.abs(gp_ok).max()
