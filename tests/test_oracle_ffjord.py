"""Pins of the FFJORD oracle (oracle/ffjord.py, src/ffjord.jl:81-178).

* linear flow f = A z: z(T) = e^{AT} x (scipy expm), Δlogp(T) = -tr(A) T
  exactly with the exact trace and -eᵀAe T with a fixed probe; logpx is the
  change-of-variables density logN(z(T)) + tr(A) T;
* Hutchinson: the mean of eᵀJe over all 2^D Rademacher probes is exactly
  tr(J) (E[e eᵀ] = I), J the brute-force finite-difference Jacobian of a
  nonlinear MLP;
* time concatenation: the model sees [z; t] (a planted drop of t fails).
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
