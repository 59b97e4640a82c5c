"""Pins the oracle MLP forward/VJP: direct evaluation, central FD, and an
independent torch-FP64 autograd."""
import numpy as np
import pytest
import torch

from oracle.mlp import MLP, MAP_CUBE, MAP_IDENTITY


@pytest.mark.parametrize("act", ["tanh", "softplus", "sigmoid", "relu", "identity"])
@pytest.mark.parametrize("imap", [MAP_IDENTITY, MAP_CUBE])
def test_vjp_vs_torch_autograd(act, imap):
    rng = np.random.default_rng(5)
    m = MLP([3, 7, 5, 3], [act, act, "identity"], input_map=imap)
    p = m.glorot_init(rng)
    p[m.offsets[0][1]:m.offsets[0][2]] = rng.standard_normal(7) * 0.1
    u = rng.standard_normal((4, 3))
    lam = rng.standard_normal((4, 3))
    gx, gp = m.vjp(u, p, lam)
    # torch restatement: Julia W (out x in) col-major == torch (in, out) view
    tu = torch.tensor(u, requires_grad=True)
    tp = torch.tensor(p, requires_grad=True)
    h = tu ** 3 if imap == MAP_CUBE else tu
    fn = {"tanh": torch.tanh, "softplus": torch.nn.functional.softplus,
          "sigmoid": torch.sigmoid, "relu": torch.relu, "identity": lambda z: z}
    for li, (i, j) in enumerate(zip(m.widths[:-1], m.widths[1:])):
        w0, b0, e = m.offsets[li]
        h = fn[m_act(m, li)](h @ tp[w0:b0].reshape(i, j) + tp[b0:e])
    (h * torch.tensor(lam)).sum().backward()
    if act == "relu":  # the gradient must see the kink: some units are off
        assert 0 < int((gp == 0).sum()) < gp.size
    assert np.allclose(gx, tu.grad.numpy(), rtol=1e-12, atol=1e-13)
    assert np.allclose(gp, tp.grad.numpy(), rtol=1e-12, atol=1e-13)


def m_act(m, li):
    inv = {0: "identity", 1: "tanh", 2: "softplus", 3: "relu", 4: "sigmoid"}
    return inv[m.acts[li]]


def test_vjp_vs_central_fd():
    rng = np.random.default_rng(6)
    m = MLP([2, 50, 2], ["tanh", "identity"], input_map=MAP_CUBE)
    p = m.glorot_init(rng)
    u = rng.standard_normal((3, 2))
    lam = rng.standard_normal((3, 2))
    gx, gp = m.vjp(u, p, lam)
    L = lambda uu, pp: float(np.sum(m.forward(uu, pp) * lam))
    h = 1e-6
    for j in rng.choice(m.nparams, 12, replace=False):
        e = np.zeros_like(p); e[j] = h
        assert abs((L(u, p + e) - L(u, p - e)) / (2 * h) - gp[j]) < 1e-7
    for idx in [(0, 0), (2, 1)]:
        e = np.zeros_like(u); e[idx] = h
        assert abs((L(u + e, p) - L(u - e, p)) / (2 * h) - gx[idx]) < 1e-7


def test_flat_layout_matches_lux_componentarray():
    # C1 (docs/src/examples/neural_ode.md:32): W2 [0,100), b2 [100,150),
    # W3 [150,250), b3 [250,252)
    m = MLP([2, 50, 2], ["tanh", "identity"], input_map=MAP_CUBE)
    assert m.nparams == 252
    assert m.offsets == [(0, 100, 150), (150, 250, 252)]
    # Julia W[o, i] (col-major) sits at i*out + o
    p = np.arange(252, dtype=float)
    Wt, b = m.unflatten(p)[0]
    assert Wt[1, 3] == 1 * 50 + 3 and b[0] == 100
