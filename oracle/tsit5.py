"""The OrdinaryDiffEq ``Tsit5`` integrator loop, restated in numpy.

Third-party source restated: OrdinaryDiffEq 7 (``OrdinaryDiffEqTsit5`` +
``OrdinaryDiffEqCore`` integrator loop + ``initdt``), which the reference
reaches through ``solve(prob, Tsit5(); ...)`` at ``src/neural_de.jl:55-58``.
The keyword semantics exercised by the reference's tests
(``test/BasicNeuralDE/neural_de_tests.jl:27-42``: ``saveat`` range / scalar,
``save_everystep = false``, ``save_start = false``, ``abstol``/``reltol``) are
restated in :func:`save_plan`.

What is restated (upstream defaults for Tsit5):
  * perform_step: 7-stage FSAL explicit RK 5(4), 6 new RHS evals per step;
  * error: ``utilde = dt * sum(btilde_i k_i)``,
    ``EEst = rms(utilde ./ (abstol .+ max.(|uprev|, |u|) .* reltol))``;
  * PI controller: beta1 = 7/50, beta2 = 2/25, gamma = 9/10, qmin = 1/5,
    qmax = 10, qsteady = [1, 1], qoldinit = 1e-4; accept iff EEst <= 1;
  * initial dt: Hairer-Wanner starting-step heuristic (``ode_determine_initdt``);
  * tstops: dt clamped to land on the next tstop, t snapped to it within
    100 eps; ``dtmax`` defaults to ``tf - t0``;
  * saveat: values between steps from the free 4th-order interpolant with
    ``theta = (tsave - tprev) / dt``, exact ``u`` when ``tsave == t``;
  * retcodes: DtNaN, MaxIters, DtLessThanMin, Unstable (any NaN in ``u``).

Pinned by (tests/test_oracle_tsit5.py): fixed-dt global error slope 5 against
``scipy.linalg.expm``; adaptive saveat error against
``scipy.integrate.solve_ivp(DOP853, rtol=1e-13)`` proportional to ``rtol``;
exact landing on ``tf`` and on tstops; the controller's closed-form
behaviour on a scalar linear ODE (dt growth capped by qmax); host-forced
retcodes.  The exact adaptive dt sequence of OrdinaryDiffEq can only be pinned
by running it (Julia is absent): parity unpinned for that sequence, pinned to
solver tolerance only.
"""
import bisect
import math

import numpy as np

from . import tableau as tb

RET_DEFAULT, RET_SUCCESS, RET_TERMINATED, RET_DTNAN = 0, 1, 2, 3
RET_MAXITERS, RET_DTLESSTHANMIN, RET_UNSTABLE = 4, 5, 6

BETA1, BETA2, GAMMA, QMIN, QMAX = 7 / 50, 2 / 25, 9 / 10, 1 / 5, 10.0
QSTEADY_MIN, QSTEADY_MAX, QOLDINIT = 1.0, 1.0, 1e-4


def rms(x):
    """DiffEqBase default norm: sqrt(sum(abs2, x) / length(x)) (FP64 sum)."""
    x = np.asarray(x)
    if x.size == 0:
        return 0.0
    return math.sqrt(float(np.sum(np.square(x, dtype=np.float64))) / x.size)


def eps(t):
    return float(np.spacing(abs(float(t))))


def save_plan(t0, t1, saveat=None, save_everystep=None, save_start=None,
              save_end=None):
    """SciMLBase/OrdinaryDiffEq keyword semantics -> (interior saveat times,
    save_start, save_end, save_everystep).

    ``saveat`` may be None/empty, a number (step) or a sequence of times.
    """
    tdir = 1.0 if t1 >= t0 else -1.0
    is_num = saveat is not None and np.ndim(saveat) == 0
    seq = [] if (saveat is None or is_num) else [float(s) for s in saveat]
    empty = (not is_num) and len(seq) == 0
    if save_everystep is None:
        save_everystep = empty
    if is_num:
        interior = [s for s in float_range(t0, float(saveat), t1)[1:]
                    if tdir * t0 < tdir * s < tdir * t1]
    else:
        interior = sorted({s for s in seq if tdir * t0 < tdir * s < tdir * t1},
                          key=lambda s: tdir * s)
    if save_start is None:
        save_start = save_everystep or empty or is_num or (t0 in seq)
    if save_end is None:
        save_end = save_everystep or empty or is_num or (t1 in seq)
    return interior, bool(save_start), bool(save_end), bool(save_everystep)


def float_range(t0, step, t1):
    """Elements of Julia's ``t0:step:t1`` for floats.  Julia's range lifts
    decimal steps to exact rationals (``0.1`` is treated as ``1/10``), so each
    element is the correctly rounded ``t0 + i*step``; this does the same."""
    from fractions import Fraction
    fs = Fraction(step).limit_denominator(10 ** 9)
    f0 = Fraction(t0).limit_denominator(10 ** 9)
    if float(fs) != step or float(f0) != t0:
        fs, f0 = Fraction(step), Fraction(t0)
    f1 = Fraction(t1)
    n = int((f1 - f0) / fs) if fs != 0 else 0
    out = [float(f0 + i * fs) for i in range(n + 1)]
    if out and float(f0 + (n + 1) * fs) == t1:
        out.append(t1)
    return out


class Step:
    __slots__ = ("t0", "dt", "t1", "y0", "ks", "y1")

    def __init__(self, t0, dt, t1, y0, ks, y1):
        self.t0, self.dt, self.t1, self.y0, self.ks, self.y1 = t0, dt, t1, y0, ks, y1


class Solution:
    """Saved values + the dense (every-step) record used by the adjoint."""

    def __init__(self):
        self.t, self.u = [], []
        self.steps = []
        self.retcode = RET_DEFAULT
        self.nf = self.naccept = self.nreject = 0
        self.t0 = self.t1 = None

    def interp(self, t):
        """SciMLBase ``sol(t)`` over the dense record: exact stored state at a
        step time, else the Tsit5 interpolant on the bracketing step with
        ``dt = t_i - t_{i-1}``."""
        st = self.steps
        times = self._times
        tdir = 1.0 if self.t1 >= self.t0 else -1.0
        key = [tdir * s for s in times]
        i = bisect.bisect_left(key, tdir * t)
        if i < len(times) and times[i] == t:
            return st[i - 1].y1 if i > 0 else st[0].y0
        if i > 0 and times[i - 1] == t:
            return st[i - 2].y1 if i > 1 else st[0].y0
        i = min(max(i, 1), len(times) - 1)
        s = st[i - 1]
        dt = times[i] - times[i - 1]
        theta = (t - times[i - 1]) / dt
        return interpolant(theta, dt, s.y0, s.ks)

    def finalize_dense(self):
        if self.steps:
            self._times = [self.steps[0].t0] + [s.t1 for s in self.steps]
        else:
            self._times = [self.t0]


def interpolant(theta, dt, y0, ks):
    w = tb.interp_weights(theta)
    acc = ks[0] * w[0]
    for wi, k in zip(w[1:], ks[1:]):
        acc = acc + k * wi
    return y0 + dt * acc


def initdt(f, y0, f0, t, tdir, dtmax_abs, abstol, reltol, dtmin_opt):
    """ode_determine_initdt (out-of-place).  ``f0 = f(y0)`` is passed in (the
    reference evaluates it twice; the value is identical).  Returns (dt, n_new
    RHS evals)."""
    dtmin = float(np.nextafter(max(dtmin_opt, eps(t)), np.inf))
    smalldt = max(dtmin, 1e-6)
    sk = abstol + np.abs(y0) * reltol
    d0 = rms(y0 / sk)
    d1 = rms(f0 / sk)
    if d0 < 1e-5 or d1 < 1e-5:
        dt0 = smalldt
    else:
        dt0 = 0.01 * (d0 / d1)
    dt0 = min(dt0, dtmax_abs)
    if not (dt0 >= 10 * np.finfo(np.float64).eps):
        return tdir * smalldt, 0
    y1 = y0 + (tdir * dt0) * f0
    f1 = f(y1, t + tdir * dt0)
    if np.array_equal(f0, f1):
        return tdir * max(dtmin, 100 * dt0), 1
    d2 = rms((f1 - f0) / sk) / dt0
    m = max(d1, d2)
    if m <= 1e-15:
        dt1 = max(1e-6, dt0 * 1e-3)
    else:
        dt1 = 10.0 ** (-(2 + math.log10(m)) / tb.ORDER)
    return tdir * max(dtmin, min(100 * dt0, dt1, dtmax_abs)), 1


def solve(f, y0, t0, t1, *, abstol=1e-6, reltol=1e-3, adaptive=True, dt=None,
          dtmax=None, dtmin=None, maxiters=100000, saveat_interior=(),
          save_start=True, save_end=True, save_everystep=False, tstops=(),
          callback_times=(), affect=None, record_dense=True, initdt_f=None):
    """Integrate ``y' = f(y, t)`` from t0 to t1 with Tsit5.

    ``f(y, t)`` must return an array shaped like ``y``.  ``affect(y, t)`` is a
    discrete callback applied when the integrator lands on one of
    ``callback_times`` (each is also a tstop); it returns the new state and
    marks u as modified (FSAL re-evaluated), as SciMLSensitivity's
    IterativeCallback does for saveat cotangents.  ``initdt_f`` (default
    ``f``) evaluates the initial-dt heuristic's probe ``f(y0 + dt0 f0, t0 +
    dt0)``.
    """
    t0, t1 = float(t0), float(t1)
    tdir = 1.0 if t1 > t0 else -1.0
    sol = Solution()
    sol.t0, sol.t1 = t0, t1
    if dtmax is None:
        dtmax = t1 - t0
    dtmax = tdir * abs(dtmax)
    if dtmin is None:
        dtmin = max(eps(t0), eps(t1))
    dtmin = abs(dtmin)
    if not adaptive and dt is None:
        raise ValueError("fixed-step Tsit5 needs dt")
    dtcache = abs(dt) if dt is not None else 0.0

    stops = sorted({float(s) for s in tstops if tdir * t0 < tdir * s <= tdir * t1}
                   | {float(s) for s in callback_times if tdir * t0 < tdir * s <= tdir * t1}
                   | {t1}, key=lambda s: tdir * s)
    cb_set = set(float(s) for s in callback_times)
    saveq = sorted((float(s) for s in saveat_interior), key=lambda s: tdir * s)
    si = 0

    t = t0
    u = np.array(y0, copy=True)
    if save_start:
        sol.t.append(t)
        sol.u.append(u.copy())
    if affect is not None and t0 in cb_set:
        u = affect(u, t0)
    uprev = u
    fsalfirst = f(u, t)
    sol.nf += 1
    if adaptive and dt is None:
        dt, extra = initdt(initdt_f or f, u, fsalfirst, t, tdir, abs(dtmax), abstol, reltol,
                           dtmin)
        sol.nf += 1 + extra   # reference: f(u0) once in initialize!, once here
    else:
        dt = tdir * abs(dt)

    iter_ = 0
    accept_step = False
    qold, q11 = QOLDINIT, 1.0
    dtpropose = dt
    reeval_fsal = False
    fsallast = None
    tprev = t
    ks = None
    EEst = 0.0
    a = tb.A

    while stops:
        while tdir * t < tdir * stops[0]:
            # ---- loopheader
            if iter_ > 0:
                if adaptive and not accept_step:
                    dt = dt / min(1 / QMIN, q11 / GAMMA)
                else:
                    uprev = u
                    dt = dtpropose
                    if reeval_fsal:
                        fsalfirst = f(u, t)
                        sol.nf += 1
                    else:
                        fsalfirst = fsallast
                    reeval_fsal = False
            iter_ += 1
            if tdir > 0:
                dt = min(dtmax, dt)
            else:
                dt = max(dtmax, dt)
            dtmin_t = max(eps(t), dtmin)
            if abs(dt) < dtmin_t:
                dt = tdir * dtmin_t
            if adaptive:
                dt = tdir * min(abs(dt), abs(stops[0] - t))
            else:
                dt = tdir * min(dtcache, abs(stops[0] - t))
            # ---- check_error
            if math.isnan(dt):
                sol.retcode = RET_DTNAN
                return _finish(sol, record_dense)
            if iter_ > maxiters:
                sol.retcode = RET_MAXITERS
                return _finish(sol, record_dense)
            if adaptive and abs(dt) <= dtmin and (
                    not accept_step or tdir * (t + dt) < tdir * stops[0]):
                sol.retcode = RET_DTLESSTHANMIN
                return _finish(sol, record_dense)
            if accept_step and np.isnan(u).any():
                sol.retcode = RET_UNSTABLE
                return _finish(sol, record_dense)
            # ---- perform_step
            k = [fsalfirst]
            for s in range(1, 6):
                acc = k[0] * a[s][0]
                for j in range(1, s):
                    acc = acc + k[j] * a[s][j]
                k.append(f(uprev + dt * acc, t + tb.C[s] * dt))
            acc = k[0] * a[6][0]
            for j in range(1, 6):
                acc = acc + k[j] * a[6][j]
            unew = uprev + dt * acc
            k.append(f(unew, t + dt))
            sol.nf += 6
            fsallast = k[6]
            if adaptive:
                ut = k[0] * tb.BTILDE[0]
                for j in range(1, 7):
                    ut = ut + k[j] * tb.BTILDE[j]
                ut = dt * ut
                atmp = ut / (abstol + np.maximum(np.abs(uprev), np.abs(unew)) * reltol)
                EEst = rms(atmp)
            u = unew
            ks = k
            # ---- loopfooter
            ttmp = t + dt
            if adaptive:
                if EEst == 0.0:
                    q = 1 / QMAX
                else:
                    q11 = EEst ** BETA1
                    q = q11 / qold ** BETA2
                    q = max(1 / QMAX, min(1 / QMIN, q / GAMMA))
                if math.isnan(EEst):
                    sol.retcode = RET_DTNAN
                    return _finish(sol, record_dense)
                accept_step = EEst <= 1.0
                if accept_step:
                    sol.naccept += 1
                    if QSTEADY_MIN <= q <= QSTEADY_MAX:
                        q = 1.0
                    qold = max(EEst, QOLDINIT)
                    dtnew = dt / q
                    tprev = t
                    t = _snap(t, ttmp, stops[0])
                    dtnew = tdir * min(abs(dtmax), abs(dtnew))
                    dtpropose = tdir * max(abs(dtnew), dtmin)
                else:
                    sol.nreject += 1
                    u = uprev
                    continue
            else:
                sol.naccept += 1
                accept_step = True
                tprev = t
                t = _snap(t, ttmp, stops[0])
                dtpropose = dt
            # accepted: record, callbacks, saves
            if record_dense:
                sol.steps.append(Step(tprev, dt, t, uprev, ks, u))
            if affect is not None and t in cb_set:
                u = affect(u, t)
                reeval_fsal = True
            while si < len(saveq) and tdir * saveq[si] <= tdir * t:
                ts_ = saveq[si]
                si += 1
                if ts_ != t:
                    theta = (ts_ - tprev) / dt
                    sol.t.append(ts_)
                    sol.u.append(interpolant(theta, dt, uprev, ks))
                else:
                    sol.t.append(ts_)
                    sol.u.append(u.copy())
            if save_everystep:
                sol.t.append(t)
                sol.u.append(u.copy())
            if not stops:
                break
        # handle_tstop
        while stops and t == stops[0]:
            stops.pop(0)
    if save_end and (not sol.t or sol.t[-1] != t):
        sol.t.append(t)
        sol.u.append(u.copy())
    sol.retcode = RET_SUCCESS
    return _finish(sol, record_dense)


def _snap(t, ttmp, tstop):
    if abs(ttmp - tstop) < 100 * eps(max(t, tstop)):
        return tstop
    return ttmp


def _finish(sol, record_dense):
    if record_dense:
        sol.finalize_dense()
    return sol
