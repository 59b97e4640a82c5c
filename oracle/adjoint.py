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


class Checkpoints:
    """``u(t)`` for ``InterpolatingAdjoint(checkpointing = true)``, restated
    (third-party: SciMLSensitivity >= 7.112, the checkpointed branch of
    ``ODEInterpolatingAdjointSensitivityFunction``; the flag is re-exported by
    the reference at ``src/DiffEqFlux.jl:53`` and reaches the solve through
    the ``sensealg`` keyword, ``src/neural_de.jl:57``).

    Checkpoints: ``t0``, the forward's saved times strictly inside the span,
    ``t1``, with their states (``u0``, the saved states, ``u(t1)``); nothing
    else of the forward solution is used.  The adjoint starts on the last
    interval, re-solved from its checkpoint state with the forward's
    tolerances (initial dt by the heuristic).  ``u(t)`` comes from the current
    interval's solution while ``t`` is not below the interval's start (at the
    start itself: the checkpoint state); a time below it moves to the interval
    holding ``t``, re-solved from its checkpoint with the last step of the
    previous re-solve as initial dt.  The initial-dt probe of the adjoint
    (``probe``) reads the current interval's interpolant wherever it falls.

    The exact checkpoint placement and re-solve options of SciMLSensitivity
    need Julia (absent): parity unpinned for those, pinned to solver
    tolerance (tests/test_oracle_adjoint.py: closed forms on linear ODEs,
    fixed-dt identity with the dense adjoint, agreement with it at tolerance,
    a planted wrong checkpoint).
    """

    def __init__(self, mlp, p, fwd, abstol=1e-6, reltol=1e-3, adaptive=True, dt=None,
                 maxiters=100000, **_):
        self.f = lambda y, t: mlp.forward(y, p)
        t0, t1 = fwd.t0, fwd.t1
        u0 = fwd.steps[0].y0 if fwd.steps else fwd.u[0]
        self.t = [t0] + [t for t in fwd.t if t0 < t < t1] + [t1]
        self.u = [u0] + [u for t, u in zip(fwd.t, fwd.u) if t0 < t < t1] + [fwd.interp(t1)]
        self.kw = dict(abstol=abstol, reltol=reltol, adaptive=adaptive, maxiters=maxiters)
        self.fixed_dt = None if adaptive else dt
        self.nf = 0
        self.resolves = []
        self.k = len(self.t) - 2
        self.cur = self._resolve(self.k, None)

    def _resolve(self, k, dt):
        if self.fixed_dt is not None:
            dt = self.fixed_dt
        s = tsit5.solve(self.f, self.u[k], self.t[k], self.t[k + 1], saveat_interior=(),
                        save_start=False, save_end=True, dt=dt, **self.kw)
        if s.retcode != tsit5.RET_SUCCESS:
            raise RuntimeError(f"checkpoint re-solve of [{self.t[k]}, {self.t[k + 1]}] failed")
        self.nf += s.nf
        self.resolves.append(k)
        return s

    def __call__(self, t, probe=False):
        if not probe and t < self.t[self.k]:
            last = abs(self.cur._times[-1] - self.cur._times[-2])
            while t < self.t[self.k]:
                self.k -= 1
            self.cur = self._resolve(self.k, last)
        return self.cur.interp(t)


def backward(mlp, p, fwd, dL_du, checkpointing=False, **solver_kw):
    """InterpolatingAdjoint backward pass.

    ``fwd``: the forward Solution (its ``t`` are the saved times);
    ``dL_du``: cotangents, one (B, D) array per saved time.
    ``checkpointing``: ``u(t)`` from interval re-solves (:class:`Checkpoints`)
    instead of the forward's dense record (forward times increasing only).
    Returns (dL/du0 (B, D), dL/dp (P,), adjoint Solution); with
    checkpointing the Solution carries ``ck_nf`` (re-solve RHS evaluations)
    and ``ck_resolves`` (interval indices in order)."""
    u_shape = np.shape(fwd.u[0]) if fwd.u else np.shape(fwd.steps[0].y0)
    nlam = int(np.prod(u_shape))
    P = mlp.nparams
    ts = list(fwd.t)
    jumps = {}
    for t, g in zip(ts, dL_du):
        g = np.asarray(g).reshape(-1)
        jumps[t] = jumps.get(t, 0) + g
    dtype = np.result_type(fwd.steps[0].y0 if fwd.steps else fwd.u[0], p)

    ck = Checkpoints(mlp, p, fwd, **solver_kw) if checkpointing else None

    def rhs(y, t, probe=False):
        lam = y[:nlam].reshape(u_shape)
        u = ck(t, probe) if ck is not None else fwd.interp(t)
        gx, gp = mlp.vjp(u, p, lam)
        return np.concatenate([-gx.reshape(-1), -gp]).astype(dtype, copy=False)

    def affect(y, t):
        y = y.copy()
        y[:nlam] += jumps[t].astype(dtype, copy=False)
        return y

    y0 = np.zeros(nlam + P, dtype=dtype)
    kw = dict(solver_kw)
    kw.pop("saveat", None)
    adj = tsit5.solve(rhs, y0, fwd.t1, fwd.t0, saveat_interior=(),
                      save_start=False, save_end=True, save_everystep=False,
                      callback_times=sorted(jumps), affect=affect,
                      record_dense=False, initdt_f=lambda y, t: rhs(y, t, True), **kw)
    if ck is not None:
        adj.ck_nf, adj.ck_resolves = ck.nf, ck.resolves
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
