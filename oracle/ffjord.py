"""FFJORD (continuous normalizing flow) forward, restated in numpy FP64.

Test infrastructure only (see oracle/__init__.py): the HIP path must never
import this module.

Follows ``src/ffjord.jl``:
  * ``__ffjord`` (:81-134): the augmented right-hand side
    ``d[z; Δlogp]/dt = [f(z); -tr(∂f/∂z)]`` with the Hutchinson estimate
    ``eᵀ (∂f/∂z) e`` from one VJP (``monte_carlo = true``, ``AutoZygote``,
    :95-115) or the exact trace of the batched Jacobian
    (``monte_carlo = false``, :116-126).  Here the probe ``e`` is an argument
    held fixed for the whole solve (the reference redraws ``randn!`` on every
    call, :96).  ``time_concat`` feeds ``[z; t]`` to the model (the
    time-concatenated network of the FFJORD workload); the reference passes
    only ``z`` (:88-90).
  * ``__forward_ffjord`` (:138-178): solve from ``[x; 0]`` saving only the end
    state, ``logpz = -|z|²/2 - D/2 log 2π`` (standard-normal base), and
    ``logpx = logpz - Δlogp``.
  * ``backward``: the InterpolatingAdjoint of that solve (``:143``) with the
    second-order VJP of the Hutchinson term (``trace_vjp``).  Pinned by torch
    FP64 double-backward of the same augmented right-hand side, central finite
    differences of ``forward``'s loss, and a planted mistake (the
    second-order source dropped) that the finite differences catch.
"""
import math

import numpy as np

from . import tsit5


def model_input(z, t, time_concat):
    if not time_concat:
        return z
    return np.concatenate([z, np.full((z.shape[0], 1), t, dtype=z.dtype)], axis=1)


def hutchinson_rhs(mlp, p, probe, time_concat=False):
    """RHS of the augmented state with a fixed probe (B, D)."""
    D = probe.shape[1]

    def f(y, t):
        z = y[:, :D]
        x = model_input(z, t, time_concat)
        fz = mlp.forward(x, p)
        ej, _ = mlp.vjp(x, p, probe.astype(y.dtype))   # e^T df/dx, (B, in)
        tr = np.sum(ej[:, :D] * probe, axis=1)
        return np.concatenate([fz, -tr[:, None]], axis=1)
    return f


def exact_trace_rhs(mlp, p, D, time_concat=False):
    """RHS with the exact trace: sum over the D unit probes."""
    def f(y, t):
        z = y[:, :D]
        x = model_input(z, t, time_concat)
        fz = mlp.forward(x, p)
        tr = np.zeros(z.shape[0], dtype=y.dtype)
        for d in range(D):
            e = np.zeros_like(z)
            e[:, d] = 1.0
            ej, _ = mlp.vjp(x, p, e)
            tr += ej[:, d]
        return np.concatenate([fz, -tr[:, None]], axis=1)
    return f


def forward(mlp, p, x, t0, t1, probe=None, time_concat=False, **solver_kw):
    """-> (logpx (B,), z(T) (B, D), Δlogp(T) (B,), Solution)."""
    B, D = x.shape
    f = (hutchinson_rhs(mlp, p, probe, time_concat) if probe is not None
         else exact_trace_rhs(mlp, p, D, time_concat))
    y0 = np.concatenate([x, np.zeros((B, 1), dtype=x.dtype)], axis=1)
    sol = tsit5.solve(f, y0, t0, t1, save_start=False, save_end=True, **solver_kw)
    yT = np.asarray(sol.u[-1])
    z, dl = yT[:, :D], yT[:, D]
    logpz = -0.5 * np.sum(z * z, axis=1) - 0.5 * D * math.log(2 * math.pi)
    return logpz - dl, z, dl, sol


# ---------------------------------------------------------------- adjoint
def _act_d2(code, a):
    """Second derivative of the activation at the pre-activation ``a``."""
    from .mlp import ACT_TANH, ACT_SOFTPLUS, ACT_SIGMOID
    if code == ACT_TANH:
        t = np.tanh(a)
        return -2.0 * t * (1.0 - t * t)
    if code == ACT_SOFTPLUS:
        s = 1.0 / (1.0 + np.exp(-a))
        return s * (1.0 - s)
    if code == ACT_SIGMOID:
        s = 1.0 / (1.0 + np.exp(-a))
        return s * (1.0 - s) * (1.0 - 2.0 * s)
    return np.zeros_like(a)


def trace_vjp(mlp, p, x, probe, lam_z, lam_d):
    """VJP of the augmented right-hand side ``g = [f(x); -eᵀ (∂f/∂z) e]`` at
    the model input ``x`` (B, in; ``z`` its first D columns) with cotangents
    ``lam_z`` (B, D) and ``lam_d`` (B,): returns (∂/∂x, ∂/∂p) of
    ``Σ_b lam_z·f - lam_d·s`` with ``s = eᵀ J e``, ``J = ∂f/∂z``.

    What Zygote's pullback through the Hutchinson term computes
    (``src/ffjord.jl:95-115`` differentiated by the InterpolatingAdjoint the
    layer selects at ``:143``): ``s`` is the tangent chain
    ``ṫ_{-1} = [e; 0]``, ``ȧ_l = ṫ_{l-1} W_l``, ``ṫ_l = σ'(a_l) ȧ_l``,
    ``s = e·ṫ_{L-1}``; its reverse gives per layer ``ǡ_l = ṫ̄_l σ'(a_l)``,
    ``W̄_l += ṫ_{l-1}ᵀ ǡ_l`` and the second-order source
    ``ṫ̄_l ȧ_l σ''(a_l)`` added to the primal backward chain's ``ā_l``."""
    from .mlp import act_fwd, act_deriv
    layers = mlp.unflatten(p)
    L = len(layers)
    D = probe.shape[1]
    h, hs, As, ts, adots = x, [x], [], [], []
    t = np.zeros_like(x)
    t[:, :D] = probe
    ts.append(t)
    for li, ((Wt, b), act) in enumerate(zip(layers, mlp.acts)):
        a = h @ Wt + b
        h = act_fwd(act, a)
        ad = ts[-1] @ Wt
        As.append(a)
        hs.append(h)
        adots.append(ad)
        ts.append(act_deriv(act, a, h) * ad)
    gp = np.zeros(mlp.nparams, dtype=np.result_type(x, p, lam_z))
    c = (-lam_d)[:, None]
    # tangent chain reverse, seeded by d(c s)/d t_{L-1} = c e
    tbar = c * probe
    src = [None] * L
    for li in range(L - 1, -1, -1):
        Wt, _ = layers[li]
        act = mlp.acts[li]
        abar_t = tbar * act_deriv(act, As[li], hs[li + 1])
        src[li] = tbar * adots[li] * _act_d2(act, As[li])
        w0, b0, _ = mlp.offsets[li]
        gp[w0:b0] += (ts[li].T @ abar_t).ravel()
        tbar = abar_t @ Wt.T
    # primal chain with the second-order sources
    g = lam_z
    for li in range(L - 1, -1, -1):
        Wt, _ = layers[li]
        act = mlp.acts[li]
        abar = g * act_deriv(act, As[li], hs[li + 1]) + src[li]
        w0, b0, e = mlp.offsets[li]
        gp[w0:b0] += (hs[li].T @ abar).ravel()
        gp[b0:e] += abar.sum(axis=0)
        g = abar @ Wt.T
    return g, gp


def backward(mlp, p, fwd, probe, dL_dz, dL_ddlogp, time_concat=False, **solver_kw):
    """InterpolatingAdjoint of the FFJORD solve (the layer's default
    ``sensealg``, ``src/ffjord.jl:143``): the adjoint state
    ``[λ_z; λ_Δ; μ]`` solves, from T back to t0,
    ``λ_z' = -(∂g/∂z)ᵀ λ``, ``λ_Δ' = 0``, ``μ' = -(∂g/∂p)ᵀ λ`` with
    ``λ(T) = [dL/dz(T); dL/dΔlogp(T)]`` and ``z(t)`` from the forward's dense
    interpolant.  Returns (dL/dx (B, D), dL/dp (P,), adjoint Solution)."""
    B, D = probe.shape
    P = mlp.nparams
    dtype = np.result_type(fwd.steps[0].y0, p)

    def f(y, t):
        lz = y[:B * D].reshape(B, D)
        ld = y[B * D:B * D + B]
        z = np.asarray(fwd.interp(t))[:, :D]
        gx, gp = trace_vjp(mlp, p, model_input(z, t, time_concat), probe.astype(dtype), lz, ld)
        return np.concatenate([-gx[:, :D].reshape(-1), np.zeros(B, dtype), -gp]).astype(dtype)

    y0 = np.concatenate([np.asarray(dL_dz, dtype).reshape(-1),
                         np.asarray(dL_ddlogp, dtype).reshape(-1), np.zeros(P, dtype)])
    kw = dict(solver_kw)
    adj = tsit5.solve(f, y0, fwd.t1, fwd.t0, saveat_interior=(), save_start=False,
                      save_end=True, save_everystep=False, record_dense=False, **kw)
    yT = adj.u[-1]
    return yT[:B * D].reshape(B, D), yT[B * D + B:], adj


def nll_cotangents(z, B):
    """Cotangents of the mean negative log-likelihood ``-mean(logpx)`` at T:
    ``dL/dz = z / B``, ``dL/dΔlogp = 1 / B`` (``logpx = logpz - Δlogp``)."""
    return z / B, np.full(z.shape[0], 1.0 / B)
