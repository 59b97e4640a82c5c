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
  * ``regularize`` (``:131-132``, ``:149-169``): two more state rows per
    sample, ``r1' = ‖f‖²`` and ``r2' = ‖eᵀ ∂f/∂z‖`` (``__norm_batched``), whose
    values at T are the layer's ``λ₁``, ``λ₂``.  (The reference selects them
    with ``selectdim(pred, N, ...)``, the batch dimension, which for a matrix
    input picks two sample columns rather than the two regulariser rows; the
    rows -- what the regularisers are -- are returned here.)
  * ``sample`` (``__backward_ffjord``, ``:183-214``): the same augmented
    solve with the time span reversed, from base samples to data space.
  * ``backward``: the InterpolatingAdjoint of that solve (``:143``) with the
    second-order VJP of the Hutchinson term (``trace_vjp``) or of the exact
    trace and the regularisers (``rhs_vjp``: bilinear forms ``aᵀ J b`` through
    one tangent chain per pair).  Pinned by torch FP64 double-backward of the
    same augmented right-hand side (every combination of ``monte_carlo`` and
    ``regularize``), central finite differences of ``forward``'s loss, closed
    forms on a linear flow (``Δlogp = -tr(A) T``, ``r2(T) = T ‖Aᵀe‖``,
    ``r1(T) = ∫‖A e^{At} x‖²``), and a planted mistake (the second-order
    source dropped) that the finite differences catch.
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


def ffjord_rhs(mlp, p, D, probe=None, time_concat=False, regularize=False, monte_carlo=True):
    """The augmented right-hand side of ``__ffjord`` (``src/ffjord.jl:81-134``)
    on the state ``[z; Δlogp]`` (``regularize``: ``[z; Δlogp; r1; r2]``):
    ``[f; -tr(∂f/∂z); (‖f‖²; ‖eᵀ ∂f/∂z‖)]`` per sample (rows of ``y``).
    ``monte_carlo``: Hutchinson ``eᵀJe`` with the fixed ``probe`` (``:95-115``);
    else the exact trace of the Jacobian (``:116-126``: the reference's batched
    Jacobian, here its D rows ``e_dᵀJ`` from D unit-probe VJPs), whose
    regulariser row ``eᵀJ`` uses ``probe`` as the reference's ``randn!`` draw."""
    def f(y, t):
        z = y[:, :D]
        x = model_input(z, t, time_concat)
        fz = mlp.forward(x, p)
        if monte_carlo:
            ej, _ = mlp.vjp(x, p, probe.astype(y.dtype))
            w = ej[:, :D]
            tr = np.sum(w * probe, axis=1)
        else:
            tr = np.zeros(z.shape[0], dtype=y.dtype)
            w = np.zeros_like(z)
            for d in range(D):
                e = np.zeros_like(z)
                e[:, d] = 1.0
                jd, _ = mlp.vjp(x, p, e)
                tr += jd[:, d]
                if regularize:
                    w += probe[:, d:d + 1].astype(y.dtype) * jd[:, :D]
        cols = [fz, -tr[:, None]]
        if regularize:
            cols += [np.sum(fz * fz, axis=1)[:, None], np.linalg.norm(w, axis=1)[:, None]]
        return np.concatenate(cols, axis=1)
    return f


def forward(mlp, p, x, t0, t1, probe=None, time_concat=False, regularize=False,
            monte_carlo=None, **solver_kw):
    """-> (logpx (B,), z(T) (B, D), Δlogp(T) (B,), Solution); with
    ``regularize`` the Solution carries ``lam1``, ``lam2`` (B,) (the
    regulariser rows at T).  ``monte_carlo`` defaults to ``probe is not None``
    (no probe: exact trace)."""
    B, D = x.shape
    if monte_carlo is None:
        monte_carlo = probe is not None
    f = ffjord_rhs(mlp, p, D, probe, time_concat, regularize, monte_carlo)
    extra = 3 if regularize else 1
    y0 = np.concatenate([x, np.zeros((B, extra), dtype=x.dtype)], axis=1)
    sol = tsit5.solve(f, y0, t0, t1, save_start=False, save_end=True, **solver_kw)
    yT = np.asarray(sol.u[-1])
    z, dl = yT[:, :D], yT[:, D]
    if regularize:
        sol.lam1, sol.lam2 = yT[:, D + 1], yT[:, D + 2]
    logpz = -0.5 * np.sum(z * z, axis=1) - 0.5 * D * math.log(2 * math.pi)
    return logpz - dl, z, dl, sol


def sample(mlp, p, x, t0, t1, probe=None, time_concat=False, regularize=False,
           monte_carlo=None, **solver_kw):
    """FFJORD sampling (``__backward_ffjord``, ``src/ffjord.jl:183-214``):
    base samples ``x`` (B, D) (the reference draws them from ``basedist``,
    standard normal by default, ``:186-191``) carried through the same
    augmented solve with the time span reversed (``reverse(n.tspan)``,
    ``:206``), from ``[x; 0]`` at ``t1`` to ``t0``; returns (z (B, D), Solution).
    ``Δlogp`` (and the regulariser rows) are integrated as in ``forward`` and
    dropped (``:211-213``).  Pinned by the linear flow ``z = e^{-AT} x`` and by
    ``forward(sample(x)) = x`` on a nonlinear flow (tests/test_oracle_ffjord.py)."""
    B, D = x.shape
    if monte_carlo is None:
        monte_carlo = probe is not None
    f = ffjord_rhs(mlp, p, D, probe, time_concat, regularize, monte_carlo)
    extra = 3 if regularize else 1
    y0 = np.concatenate([x, np.zeros((B, extra), dtype=x.dtype)], axis=1)
    sol = tsit5.solve(f, y0, t1, t0, save_start=False, save_end=True, **solver_kw)
    return np.asarray(sol.u[-1])[:, :D], sol


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


def bilinear_vjp(mlp, p, x, D, pairs, seed):
    """VJP of ``Σ_b seed_b·f(x_b) + Σ_pairs Σ_b a_bᵀ J_b c_b`` (``J = ∂f/∂z``,
    ``z`` the first D input columns) for per-sample vectors ``(a, c)``
    (B, D): each pair runs one tangent chain seeded with ``c`` and its
    reverse seeded with ``a`` (``trace_vjp`` is the pair ``(-λ_Δ e, e)``).
    Returns (∂/∂x (B, in), ∂/∂p (P,))."""
    from .mlp import act_fwd, act_deriv
    layers = mlp.unflatten(p)
    L = len(layers)
    h, hs, As = x, [x], []
    for (Wt, b), act in zip(layers, mlp.acts):
        a = h @ Wt + b
        h = act_fwd(act, a)
        As.append(a)
        hs.append(h)
    gp = np.zeros(mlp.nparams, dtype=np.result_type(x, p, seed))
    src = [np.zeros_like(a) for a in As]
    for left, right in pairs:
        t = np.zeros_like(x)
        t[:, :D] = right
        ts, adots = [t], []
        for li, ((Wt, _), act) in enumerate(zip(layers, mlp.acts)):
            ad = ts[-1] @ Wt
            adots.append(ad)
            ts.append(act_deriv(act, As[li], hs[li + 1]) * ad)
        tbar = left
        for li in range(L - 1, -1, -1):
            Wt, _ = layers[li]
            act = mlp.acts[li]
            abar_t = tbar * act_deriv(act, As[li], hs[li + 1])
            src[li] = src[li] + tbar * adots[li] * _act_d2(act, As[li])
            w0, b0, _ = mlp.offsets[li]
            gp[w0:b0] += (ts[li].T @ abar_t).ravel()
            tbar = abar_t @ Wt.T
    g = seed
    for li in range(L - 1, -1, -1):
        Wt, _ = layers[li]
        act = mlp.acts[li]
        abar = g * act_deriv(act, As[li], hs[li + 1]) + src[li]
        w0, b0, e = mlp.offsets[li]
        gp[w0:b0] += (hs[li].T @ abar).ravel()
        gp[b0:e] += abar.sum(axis=0)
        g = abar @ Wt.T
    return g, gp


def rhs_vjp(mlp, p, x, D, probe, lam_z, lam_d, lam_r1=None, lam_r2=None, monte_carlo=True):
    """VJP of the augmented right-hand side ``ffjord_rhs`` at the model input
    ``x`` with cotangents ``lam_z`` (B, D), ``lam_d`` (B,) and, with the
    regularisers, ``lam_r1``, ``lam_r2`` (B,): the trace term is the pair
    ``(-λ_Δ e, e)`` (Hutchinson) or ``(-λ_Δ e_d, e_d)`` for every unit probe
    (exact); ``λ_r1 ‖f‖²`` adds ``2 λ_r1 f`` to the primal seed; ``λ_r2 ‖w‖``
    with ``w = eᵀJ`` is the pair ``(e, λ_r2 w / ‖w‖)``."""
    B = x.shape[0]
    seed = np.array(lam_z, dtype=np.result_type(x, p, lam_z))
    c = -np.asarray(lam_d)[:, None]
    if monte_carlo:
        pairs = [(c * probe, probe)]
    else:
        pairs = []
        for d in range(D):
            e = np.zeros((B, D), dtype=seed.dtype)
            e[:, d] = 1.0
            pairs.append((c * e, e))
    if lam_r1 is not None and np.any(lam_r1 != 0):
        seed = seed + 2.0 * np.asarray(lam_r1)[:, None] * mlp.forward(x, p)
    if lam_r2 is not None and np.any(lam_r2 != 0):
        ej, _ = mlp.vjp(x, p, probe.astype(seed.dtype))
        w = ej[:, :D]
        nw = np.linalg.norm(w, axis=1)
        pairs.append((probe.astype(seed.dtype),
                      np.asarray(lam_r2)[:, None] * w / np.where(nw > 0, nw, 1.0)[:, None]))
    return bilinear_vjp(mlp, p, x, D, pairs, seed)


def backward(mlp, p, fwd, probe, dL_dz, dL_ddlogp, time_concat=False, regularize=False,
             monte_carlo=True, dL_dlam1=None, dL_dlam2=None, **solver_kw):
    """InterpolatingAdjoint of the FFJORD solve (the layer's default
    ``sensealg``, ``src/ffjord.jl:143``): the adjoint state
    ``[λ_z; λ_Δ; μ]`` solves, from T back to t0,
    ``λ_z' = -(∂g/∂z)ᵀ λ``, ``λ_Δ' = 0``, ``μ' = -(∂g/∂p)ᵀ λ`` with
    ``λ(T) = [dL/dz(T); dL/dΔlogp(T)]`` and ``z(t)`` from the forward's dense
    interpolant.  Returns (dL/dx (B, D), dL/dp (P,), adjoint Solution)."""
    B, D = probe.shape
    P = mlp.nparams
    dtype = np.result_type(fwd.steps[0].y0, p)
    nc = 3 if regularize else 1           # constant adjoint columns per sample
    z0 = np.zeros(B, dtype)
    l1 = np.asarray(dL_dlam1, dtype).reshape(-1) if dL_dlam1 is not None else z0
    l2 = np.asarray(dL_dlam2, dtype).reshape(-1) if dL_dlam2 is not None else z0

    def f(y, t):
        lz = y[:B * D].reshape(B, D)
        ld = y[B * D:B * D + B]
        z = np.asarray(fwd.interp(t))[:, :D]
        x = model_input(z, t, time_concat)
        if regularize or not monte_carlo:
            gx, gp = rhs_vjp(mlp, p, x, D, probe.astype(dtype), lz, ld,
                             y[B * D + B:B * D + 2 * B] if regularize else None,
                             y[B * D + 2 * B:B * D + 3 * B] if regularize else None,
                             monte_carlo)
        else:
            gx, gp = trace_vjp(mlp, p, x, probe.astype(dtype), lz, ld)
        return np.concatenate([-gx[:, :D].reshape(-1), np.zeros(nc * B, dtype),
                               -gp]).astype(dtype)

    cols = [np.asarray(dL_dz, dtype).reshape(-1), np.asarray(dL_ddlogp, dtype).reshape(-1)]
    if regularize:
        cols += [l1, l2]
    y0 = np.concatenate(cols + [np.zeros(P, dtype)])
    kw = dict(solver_kw)
    adj = tsit5.solve(f, y0, fwd.t1, fwd.t0, saveat_interior=(), save_start=False,
                      save_end=True, save_everystep=False, record_dense=False, **kw)
    yT = adj.u[-1]
    return yT[:B * D].reshape(B, D), yT[B * D + nc * B:], adj


def nll_cotangents(z, B):
    """Cotangents of the mean negative log-likelihood ``-mean(logpx)`` at T:
    ``dL/dz = z / B``, ``dL/dΔlogp = 1 / B`` (``logpx = logpz - Δlogp``)."""
    return z / B, np.full(z.shape[0], 1.0 / B)
