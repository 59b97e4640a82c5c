"""SciMLSensitivity ``InterpolatingAdjoint(autojacvec = ZygoteVJP())``, restated.

Third-party source restated: SciMLSensitivity >= 7.112 (``_concrete_solve_adjoint``
-> ``adjoint_sensitivities`` -> ``ODEAdjointProblem`` with
``ODEInterpolatingAdjointSensitivityFunction``), the default ``sensealg`` the
reference sets at ``src/neural_de.jl:57`` (also ``src/ffjord.jl:143``).

For ``L = sum_i g_i(u(t_i))`` over the saved times ``t_i`` with cotangents
``dL/du(t_i)``, the adjoint state ``y = [lambda; mu]`` (``lambda``: D x B,
``mu``: P) solves, backwards from ``T`` to ``t0``::

    lambda' = -(df/du)^T lambda,    mu' = -(df/dp)^T lambda,
    lambda(T) = 0, mu(T) = 0,  lambda += dL/du(t_i) at each saved t_i,

with the same ``Tsit5`` (adaptive, same abstol/reltol, error norm = RMS over
all ``D*B + P`` entries of ``y``), the saved times as tstops, each jump an
``IterativeCallback`` affect that marks ``u`` modified (FSAL re-evaluated;
the jump at ``T`` is the callback's initial affect, applied before the
initial-dt heuristic).  ``u(t)`` at each stage time comes from the forward
solution's dense interpolant (every forward step recorded).
``dL/du0 = lambda(t0)`` and ``dL/dp = mu(t0)``.

Pinned by (tests/test_oracle_adjoint.py): the linear-ODE closed forms
``dL/du0 = e^{W^T T} e^{W T} u0`` and the Van Loan block exponential for
``dL/dW`` (``L = |u(T)|^2 / 2``); central finite differences of the oracle's
own forward solve; and, for fixed dt, torch-FP64 autograd through the discrete
forward solver (the continuous adjoint agrees to O(dt^5)).  SciMLSensitivity's
exact adaptive adjoint step sequence needs Julia (absent): parity unpinned for
that sequence, pinned to solver tolerance.

``backward_backsolve`` restates ``BacksolveAdjoint(; checkpointing)`` (u
integrated back with the adjoint); pinned by the same linear closed forms and
saveat-jump closed form (with and without checkpointing), by central finite
differences and by agreement with ``backward`` on the spiral problem, and by a
planted mistake (u left un-integrated) that the closed form catches.
"""
import numpy as np

from . import tsit5


def forward(mlp, p, u0, t0, t1, saveat=None, save_everystep=None,
            save_start=None, save_end=None, **solver_kw):
    """NeuralODE forward (``src/neural_de.jl:47-61``): u' = model(u, p).

    ``u0`` is (B, D).  Returns the tsit5.Solution (dense record kept)."""
    interior, ss, se, sev = tsit5.save_plan(t0, t1, saveat, save_everystep,
                                            save_start, save_end)
    f = lambda y, t: mlp.forward(y, p)  # autonomous, tgrad = 0 (src/neural_de.jl:4,50)
    return tsit5.solve(f, u0, t0, t1, saveat_interior=interior, save_start=ss,
                       save_end=se, save_everystep=sev, **solver_kw)


def backward(mlp, p, fwd, dL_du, **solver_kw):
    """InterpolatingAdjoint backward pass.

    ``fwd``: the forward Solution (its ``t`` are the saved times);
    ``dL_du``: cotangents, one (B, D) array per saved time.
    Returns (dL/du0 (B, D), dL/dp (P,), adjoint Solution)."""
    u_shape = np.shape(fwd.u[0]) if fwd.u else np.shape(fwd.steps[0].y0)
    nlam = int(np.prod(u_shape))
    P = mlp.nparams
    ts = list(fwd.t)
    jumps = {}
    for t, g in zip(ts, dL_du):
        g = np.asarray(g).reshape(-1)
        jumps[t] = jumps.get(t, 0) + g
    dtype = np.result_type(fwd.steps[0].y0 if fwd.steps else fwd.u[0], p)

    def f(y, t):
        lam = y[:nlam].reshape(u_shape)
        u = fwd.interp(t)
        gx, gp = mlp.vjp(u, p, lam)
        return np.concatenate([-gx.reshape(-1), -gp]).astype(dtype, copy=False)

    def affect(y, t):
        y = y.copy()
        y[:nlam] += jumps[t].astype(dtype, copy=False)
        return y

    y0 = np.zeros(nlam + P, dtype=dtype)
    kw = dict(solver_kw)
    kw.pop("saveat", None)
    adj = tsit5.solve(f, y0, fwd.t1, fwd.t0, saveat_interior=(),
                      save_start=False, save_end=True, save_everystep=False,
                      callback_times=sorted(jumps), affect=affect,
                      record_dense=False, **kw)
    if not adj.u:  # the adjoint solve failed before T -> t0 (see adj.retcode)
        return np.full(u_shape, np.nan), np.full(P, np.nan), adj
    yT = adj.u[-1]
    return yT[:nlam].reshape(u_shape), yT[nlam:], adj


def backward_backsolve(mlp, p, fwd, dL_du, checkpointing=True, **solver_kw):
    """SciMLSensitivity ``BacksolveAdjoint(; checkpointing)``, restated
    (third-party: SciMLSensitivity >= 7.112, ``ODEBacksolveSensitivityFunction``
    and ``backsolve_checkpoint_callbacks``); selected by ``sensealg`` at the
    call in ``src/neural_de.jl:55-58`` (the reference's own test matrix runs it
    in ``test/BasicNeuralDE/neural_de_tests.jl:34-37``).

    Instead of interpolating the forward solution, ``u`` is integrated back
    from ``u(T)`` together with the adjoint: ``y = [u; lambda; mu]`` solves

        u' = f(u),  lambda' = -(df/du)^T lambda,  mu' = -(df/dp)^T lambda

    from ``T`` to ``t0`` with the same Tsit5 (error norm = RMS over all
    ``2 D B + P`` entries), the same cotangent jumps at the saved times, and,
    with ``checkpointing``, ``u`` reset to the forward solution at every saved
    time before ``T`` (a callback at the same times as the jumps).
    Returns (dL/du0 (B, D), dL/dp (P,), adjoint Solution)."""
    u_shape = np.shape(fwd.u[0]) if fwd.u else np.shape(fwd.steps[0].y0)
    nlam = int(np.prod(u_shape))
    P = mlp.nparams
    jumps, saved = {}, {}
    for t, g, u in zip(fwd.t, dL_du, fwd.u):
        g = np.asarray(g).reshape(-1)
        jumps[t] = jumps.get(t, 0) + g
        saved[t] = np.asarray(u).reshape(-1)
    dtype = np.result_type(fwd.steps[0].y0 if fwd.steps else fwd.u[0], p)

    def f(y, t):
        u = y[:nlam].reshape(u_shape)
        lam = y[nlam:2 * nlam].reshape(u_shape)
        fu = mlp.forward(u, p)
        gx, gp = mlp.vjp(u, p, lam)
        return np.concatenate([fu.reshape(-1), -gx.reshape(-1), -gp]).astype(dtype, copy=False)

    def affect(y, t):
        y = y.copy()
        y[nlam:2 * nlam] += jumps[t].astype(dtype, copy=False)
        if checkpointing and t != fwd.t1:
            y[:nlam] = saved[t].astype(dtype, copy=False)
        return y

    uT = np.asarray(fwd.interp(fwd.t1)).reshape(-1)
    y0 = np.concatenate([uT, np.zeros(nlam + P)]).astype(dtype)
    kw = dict(solver_kw)
    kw.pop("saveat", None)
    adj = tsit5.solve(f, y0, fwd.t1, fwd.t0, saveat_interior=(),
                      save_start=False, save_end=True, save_everystep=False,
                      callback_times=sorted(jumps), affect=affect,
                      record_dense=False, **kw)
    yT = adj.u[-1]
    return yT[nlam:2 * nlam].reshape(u_shape), yT[2 * nlam:], adj


def loss_mse_cotangent(pred, target):
    """d/dpred of mean((pred - target)^2) over all saved entries."""
    pred = np.asarray(pred)
    return 2.0 * (pred - np.asarray(target)) / pred.size
