"""Pins of the FFJORD oracle (oracle/ffjord.py, src/ffjord.jl:81-178).

* linear flow f = A z: z(T) = e^{AT} x (scipy expm), Δlogp(T) = -tr(A) T
  exactly with the exact trace and -eᵀAe T with a fixed probe; logpx is the
  change-of-variables density logN(z(T)) + tr(A) T;
* Hutchinson: the mean of eᵀJe over all 2^D Rademacher probes is exactly
  tr(J) (E[e eᵀ] = I), J the brute-force finite-difference Jacobian of a
  nonlinear MLP;
* time concatenation: the model sees [z; t] (a planted drop of t fails);
* sampling (reversed-time solve): z = e^{-AT} x on the linear flow, and
  forward(sample(x)) = x on a nonlinear time-concatenated flow (a sample
  solved forward in time instead, the plausible slip, fails both).
"""
import math

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import ffjord as FF
from oracle.mlp import MLP


def linear_model(D, rng):
    m = MLP([D, D, D], ["identity", "identity"])
    p = np.zeros(m.nparams)
    (w0, b0, e0), (w1, b1, e1) = m.offsets
    W1 = 0.4 * rng.standard_normal((D, D))
    W2 = 0.4 * rng.standard_normal((D, D))
    p[w0:b0] = W1.T.ravel()     # Wt (in, out) storage == Julia W (out, in) column-major
    p[w1:b1] = W2.T.ravel()
    return m, p, W2 @ W1


def test_linear_flow_closed_forms():
    rng = np.random.default_rng(0)
    D, B, T = 3, 5, 0.8
    m, p, A = linear_model(D, rng)
    x = rng.standard_normal((B, D))
    # exact trace
    lp, z, dl, _ = FF.forward(m, p, x, 0.0, T, abstol=1e-12, reltol=1e-12)
    zT = x @ expm(A * T).T
    assert np.abs(z - zT).max() < 1e-9
    assert np.abs(dl + np.trace(A) * T).max() < 1e-10
    logN = -0.5 * np.sum(zT ** 2, 1) - 0.5 * D * math.log(2 * math.pi)
    assert np.abs(lp - (logN + np.trace(A) * T)).max() < 1e-9
    # fixed probe: the estimate is e^T A e per sample
    e = rng.choice([-1.0, 1.0], size=(B, D))
    _, _, dle, _ = FF.forward(m, p, x, 0.0, T, probe=e, abstol=1e-12, reltol=1e-12)
    assert np.abs(dle + np.einsum("bi,ij,bj->b", e, A, e) * T).max() < 1e-10


def test_hutchinson_expectation_is_the_trace():
    rng = np.random.default_rng(1)
    D = 3
    m = MLP([D, 8, D], ["tanh", "identity"])
    p = m.glorot_init(rng)
    p[m.offsets[0][1]:m.offsets[0][2]] = rng.uniform(-0.5, 0.5, 8)
    z = rng.standard_normal((1, D))
    f = FF.hutchinson_rhs
    # brute-force Jacobian by central differences
    h = 1e-6
    J = np.zeros((D, D))
    for j in range(D):
        dz = np.zeros((1, D)); dz[0, j] = h
        J[:, j] = (m.forward(z + dz, p) - m.forward(z - dz, p))[0] / (2 * h)
    vals = []
    for k in range(2 ** D):
        e = np.array([[1.0 if (k >> i) & 1 else -1.0 for i in range(D)]])
        y = np.concatenate([z, [[0.0]]], 1)
        vals.append(-f(m, p, e)(y, 0.0)[0, D])
    assert abs(np.mean(vals) - np.trace(J)) < 1e-8
    ex = -FF.exact_trace_rhs(m, p, D)(np.concatenate([z, [[0.0]]], 1), 0.0)[0, D]
    assert abs(ex - np.trace(J)) < 1e-8


def test_time_concatenation_feeds_t():
    rng = np.random.default_rng(2)
    D = 2
    m = MLP([D + 1, 6, D], ["tanh", "identity"])
    p = m.glorot_init(rng)
    y = np.concatenate([rng.standard_normal((4, D)), np.zeros((4, 1))], 1)
    e = rng.choice([-1.0, 1.0], size=(4, D))
    f = FF.hutchinson_rhs(m, p, e, time_concat=True)
    t = 0.37
    want = m.forward(np.concatenate([y[:, :D], np.full((4, 1), t)], 1), p)
    assert np.allclose(f(y, t)[:, :D], want)
    assert not np.allclose(f(y, t)[:, :D], f(y, 0.0)[:, :D])
    planted = m.forward(np.concatenate([y[:, :D], np.zeros((4, 1))], 1), p)
    assert not np.allclose(f(y, t)[:, :D], planted)


# ---- FFJORD adjoint (second-order VJP of the Hutchinson term)
def _small_cnf(acts, tcat, seed=0):
    rng = np.random.default_rng(seed)
    D, H = 3, 8
    m = MLP([D + (1 if tcat else 0), H, H, D], acts)
    p = m.glorot_init(rng)
    for (w0, b0, e) in m.offsets:
        p[b0:e] = 0.1 * rng.standard_normal(e - b0)
    return m, p, D, rng


@pytest.mark.parametrize("acts", [["tanh", "tanh", "identity"], ["softplus", "sigmoid", "identity"],
                                  ["relu", "tanh", "identity"], ["tanh", "relu", "identity"]])
@pytest.mark.parametrize("tcat", [False, True])
def test_trace_vjp_matches_torch_double_backward(acts, tcat):
    import torch
    m, p, D, rng = _small_cnf(acts, tcat)
    B = 4
    x = rng.standard_normal((B, m.widths[0]))
    e = rng.choice([-1.0, 1.0], size=(B, D))
    lz = rng.standard_normal((B, D))
    ld = rng.standard_normal(B)
    gx, gp = FF.trace_vjp(m, p, x, e, lz, ld)
    xt = torch.tensor(x, requires_grad=True)
    pt = torch.tensor(p, requires_grad=True)
    acts_t = {"tanh": torch.tanh, "softplus": torch.nn.functional.softplus,
              "sigmoid": torch.sigmoid, "relu": torch.relu, "identity": lambda v: v}

    def f(xx):
        h = xx
        for (i, j), (w0, b0, e_), a in zip(zip(m.widths[:-1], m.widths[1:]), m.offsets, acts):
            h = acts_t[a](h @ pt[w0:b0].reshape(i, j) + pt[b0:e_])
        return h
    fx = f(xt)
    et = torch.tensor(e)
    vj = torch.autograd.grad(fx, xt, grad_outputs=et, create_graph=True)[0][:, :D]
    s = (vj * et).sum(1)
    obj = (torch.tensor(lz) * fx).sum() - (torch.tensor(ld) * s).sum()
    tgx, tgp = torch.autograd.grad(obj, [xt, pt])
    assert np.abs(gx - tgx.numpy()).max() < 1e-12 * max(1.0, np.abs(gx).max())
    assert np.abs(gp - tgp.numpy()).max() < 1e-12 * max(1.0, np.abs(gp).max())


def _nll(m, p, x, e, tcat, tol):
    lp, _, _, _ = FF.forward(m, p, x, 0.0, 0.7, probe=e, time_concat=tcat, **tol)
    return -float(np.mean(lp))


@pytest.mark.parametrize("tcat", [False, True])
def test_ffjord_adjoint_vs_central_fd(tcat):
    m, p, D, rng = _small_cnf(["tanh", "tanh", "identity"], tcat, seed=1)
    B = 3
    x = rng.standard_normal((B, D))
    e = rng.choice([-1.0, 1.0], size=(B, D))
    tol = dict(abstol=1e-11, reltol=1e-11)
    lp, z, dl, fwd = FF.forward(m, p, x, 0.0, 0.7, probe=e, time_concat=tcat, **tol)
    cz, cd = FF.nll_cotangents(z, B)
    gx, gp, _ = FF.backward(m, p, fwd, e, cz, cd, time_concat=tcat, **tol)
    h = 1e-6
    idx = np.random.default_rng(5).choice(m.nparams, 12, replace=False)
    fd = []
    for j in idx:
        d = np.zeros_like(p); d[j] = h
        fd.append((_nll(m, p + d, x, e, tcat, tol) - _nll(m, p - d, x, e, tcat, tol)) / (2 * h))
    fd = np.array(fd)
    assert np.abs(fd - gp[idx]).max() < 1e-6 * max(1.0, np.abs(fd).max())
    fx = []
    for b in range(B):
        for c in range(D):
            d = np.zeros_like(x); d[b, c] = h
            fx.append((_nll(m, p, x + d, e, tcat, tol) - _nll(m, p, x - d, e, tcat, tol)) / (2 * h))
    assert np.abs(np.array(fx).reshape(B, D) - gx).max() < 1e-6 * max(1.0, np.abs(gx).max())


def test_ffjord_adjoint_without_second_order_source_is_caught():
    """Dropping the σ'' source (a plausible mistake) fails the FD pin."""
    m, p, D, rng = _small_cnf(["tanh", "tanh", "identity"], False, seed=2)
    B = 3
    x = rng.standard_normal((B, D))
    e = rng.choice([-1.0, 1.0], size=(B, D))
    tol = dict(abstol=1e-11, reltol=1e-11)
    lp, z, dl, fwd = FF.forward(m, p, x, 0.0, 0.7, probe=e, **tol)
    cz, cd = FF.nll_cotangents(z, B)
    orig = FF._act_d2
    try:
        FF._act_d2 = lambda code, a: np.zeros_like(a)
        _, gbad, _ = FF.backward(m, p, fwd, e, cz, cd, **tol)
    finally:
        FF._act_d2 = orig
    _, gp, _ = FF.backward(m, p, fwd, e, cz, cd, **tol)
    assert np.abs(gbad - gp).max() > 1e-4 * np.abs(gp).max()


# ---- exact trace and regularisers (src/ffjord.jl:116-132, :149-169)
@pytest.mark.parametrize("monte_carlo", [True, False])
@pytest.mark.parametrize("regularize", [True, False])
@pytest.mark.parametrize("tcat", [False, True])
def test_rhs_vjp_matches_torch_double_backward(monte_carlo, regularize, tcat):
    """The VJP of the whole augmented right-hand side [f; -tr; ‖f‖²; ‖eᵀJ‖]
    (Hutchinson or exact trace, with and without the regularisers and with
    non-zero regulariser cotangents) against torch FP64 autograd of the same
    function built from the Jacobian torch computes itself."""
    import torch
    m, p, D, rng = _small_cnf(["tanh", "softplus", "identity"], tcat, seed=3)
    B = 4
    x = rng.standard_normal((B, m.widths[0]))
    e = rng.choice([-1.0, 1.0], size=(B, D))
    lz, ld = rng.standard_normal((B, D)), rng.standard_normal(B)
    l1 = rng.standard_normal(B) if regularize else None
    l2 = rng.standard_normal(B) if regularize else None
    gx, gp = FF.rhs_vjp(m, p, x, D, e, lz, ld, l1, l2, monte_carlo)
    xt = torch.tensor(x, requires_grad=True)
    pt = torch.tensor(p, requires_grad=True)
    acts_t = {1: torch.tanh, 2: torch.nn.functional.softplus, 0: lambda v: v}

    def f(xx):
        h = xx
        for (i, j), (w0, b0, e_), a in zip(zip(m.widths[:-1], m.widths[1:]), m.offsets, m.acts):
            h = acts_t[a](h @ pt[w0:b0].reshape(i, j) + pt[b0:e_])
        return h
    fx = f(xt)
    rows = [torch.autograd.grad(fx[:, d].sum(), xt, create_graph=True)[0][:, :D]
            for d in range(D)]                                   # J rows per sample
    J = torch.stack(rows, 1)                                     # (B, D, D): J[b, d, :]
    et = torch.tensor(e)
    tr = torch.diagonal(J, dim1=1, dim2=2).sum(1) if not monte_carlo else \
        torch.einsum("bi,bij,bj->b", et, J, et)
    obj = (torch.tensor(lz) * fx).sum() - (torch.tensor(ld) * tr).sum()
    if regularize:
        w = torch.einsum("bi,bij->bj", et, J)
        obj = obj + (torch.tensor(l1) * (fx * fx).sum(1)).sum() + \
            (torch.tensor(l2) * torch.linalg.norm(w, dim=1)).sum()
    tgx, tgp = torch.autograd.grad(obj, [xt, pt])
    assert np.abs(gx - tgx.numpy()).max() < 1e-11 * max(1.0, np.abs(gx).max())
    assert np.abs(gp - tgp.numpy()).max() < 1e-11 * max(1.0, np.abs(gp).max())
    if not regularize and monte_carlo:       # the pair form is trace_vjp
        hx, hp = FF.trace_vjp(m, p, x, e, lz, ld)
        assert np.abs(hx - gx).max() < 1e-13 and np.abs(hp - gp).max() < 1e-13


def test_regularisers_linear_flow_closed_forms():
    # f = A z: ‖eᵀJ‖ = ‖Aᵀe‖ constant, so r2(T) = T ‖Aᵀe‖; r1(T) = ∫‖A e^{At} x‖² dt
    from scipy.integrate import quad_vec
    rng = np.random.default_rng(4)
    D, B, T = 3, 4, 0.9
    m, p, A = linear_model(D, rng)
    x = rng.standard_normal((B, D))
    e = rng.choice([-1.0, 1.0], size=(B, D))
    tol = dict(abstol=1e-12, reltol=1e-12)
    for mc in (True, False):
        lp, z, dl, sol = FF.forward(m, p, x, 0.0, T, probe=e, regularize=True, monte_carlo=mc,
                                    **tol)
        assert np.abs(sol.lam2 - T * np.linalg.norm(e @ A, axis=1)).max() < 1e-10
        r1, _ = quad_vec(lambda t: np.sum((x @ expm(A * t).T @ A.T) ** 2, axis=1), 0.0, T,
                         epsabs=1e-13, epsrel=1e-13)
        assert np.abs(sol.lam1 - r1).max() < 1e-9
        tr = np.trace(A) if not mc else np.einsum("bi,ij,bj->b", e, A, e)
        assert np.abs(dl + tr * T).max() < 1e-10
        assert np.abs(z - x @ expm(A * T).T).max() < 1e-9


@pytest.mark.parametrize("monte_carlo", [True, False])
def test_regularised_exact_adjoint_vs_central_fd(monte_carlo):
    """dL/dp and dL/dx of L = -mean(logpx) + 0.3 mean(λ₁) + 0.2 mean(λ₂)
    through ``backward`` against central differences of ``forward``."""
    m, p, D, rng = _small_cnf(["tanh", "tanh", "identity"], False, seed=6)
    B = 3
    x = rng.standard_normal((B, D))
    e = rng.choice([-1.0, 1.0], size=(B, D))
    tol = dict(abstol=1e-11, reltol=1e-11)
    kw = dict(probe=e, regularize=True, monte_carlo=monte_carlo, **tol)

    def loss(pp, xx):
        lp, _, _, s = FF.forward(m, pp, xx, 0.0, 0.6, **kw)
        return -float(np.mean(lp)) + 0.3 * float(np.mean(s.lam1)) + 0.2 * float(np.mean(s.lam2))
    lp, z, dl, fwd = FF.forward(m, p, x, 0.0, 0.6, **kw)
    cz, cd = FF.nll_cotangents(z, B)
    gx, gp, _ = FF.backward(m, p, fwd, e, cz, cd, regularize=True, monte_carlo=monte_carlo,
                            dL_dlam1=np.full(B, 0.3 / B), dL_dlam2=np.full(B, 0.2 / B), **tol)
    h = 1e-6
    idx = np.random.default_rng(7).choice(m.nparams, 10, replace=False)
    fd = []
    for j in idx:
        d = np.zeros_like(p); d[j] = h
        fd.append((loss(p + d, x) - loss(p - d, x)) / (2 * h))
    assert np.abs(np.array(fd) - gp[idx]).max() < 1e-6 * max(1.0, np.abs(fd).max())
    fx = []
    for b in range(B):
        for c in range(D):
            d = np.zeros_like(x); d[b, c] = h
            fx.append((loss(p, x + d) - loss(p, x - d)) / (2 * h))
    assert np.abs(np.array(fx).reshape(B, D) - gx).max() < 1e-6 * max(1.0, np.abs(gx).max())


def test_sampling_linear_flow_and_round_trip():
    rng = np.random.default_rng(7)
    D, B, T = 3, 6, 0.9
    m, p, A = linear_model(D, rng)
    x = rng.standard_normal((B, D))
    z, sol = FF.sample(m, p, x, 0.0, T, abstol=1e-12, reltol=1e-12)
    want = x @ expm(-A * T).T
    assert np.abs(z - want).max() < 1e-9
    assert sol.t[-1] == 0.0
    # the slip of integrating forward in time lands on e^{+AT} x
    assert np.abs(x @ expm(A * T).T - want).max() > 1e-2
    # nonlinear time-concatenated flow: sampling inverts the forward map
    mn = MLP([D + 1, 12, D], ["tanh", "identity"])
    pn = mn.glorot_init(rng)
    e = rng.choice([-1.0, 1.0], size=(B, D))
    zs, _ = FF.sample(mn, pn, x, 0.0, T, probe=e, time_concat=True, abstol=1e-11, reltol=1e-11)
    _, xr, _, _ = FF.forward(mn, pn, zs, 0.0, T, probe=e, time_concat=True,
                             abstol=1e-11, reltol=1e-11)
    assert np.abs(xr - x).max() < 1e-8
    assert np.abs(zs - x).max() > 1e-2   # the flow moved the samples
