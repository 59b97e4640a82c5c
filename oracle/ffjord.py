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
