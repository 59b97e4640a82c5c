"""Pins the oracle's Tsit5 integrator (OrdinaryDiffEq restatement) by closed
forms and an independent high-accuracy integrator."""
import math

import numpy as np
import pytest
import scipy.linalg as sla
from scipy.integrate import solve_ivp

from oracle import tsit5
from oracle.problems import spiral_true_rhs


def _linear(D=4, seed=3):
    rng = np.random.default_rng(seed)
    W = rng.standard_normal((D, D)) * 0.6
    u0 = rng.standard_normal((2, D))
    return W, u0, (lambda y, t: y @ W.T)


def test_fixed_dt_order_five():
    W, u0, f = _linear()
    T = 1.0
    exact = u0 @ sla.expm(W * T).T
    errs = []
    for n in (8, 16, 32):
        sol = tsit5.solve(f, u0, 0.0, T, adaptive=False, dt=T / n,
                          save_start=False)
        assert sol.naccept == n and sol.t[-1] == T
        errs.append(np.abs(sol.u[-1] - exact).max())
    slopes = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(4.6 < s < 5.6 for s in slopes), slopes


def test_fixed_dt_last_step_clamped_to_tf():
    _, u0, f = _linear()
    sol = tsit5.solve(f, u0, 0.0, 1.0, adaptive=False, dt=0.3)
    assert sol.naccept == 4 and sol.t[-1] == 1.0
    assert abs(sol.steps[-1].dt - 0.1) < 1e-12


@pytest.mark.parametrize("rtol", [1e-4, 1e-6, 1e-8])
def test_adaptive_saveat_error_scales_with_tolerance(rtol):
    u0 = np.array([[2.0, 0.0]])
    ts = np.linspace(0, 1.5, 30)
    ref = solve_ivp(lambda t, y: spiral_true_rhs(y[None])[0], (0, 1.5),
                    u0[0], method="DOP853", rtol=1e-13, atol=1e-15, t_eval=ts)
    interior, ss, se, sev = tsit5.save_plan(0.0, 1.5, ts)
    sol = tsit5.solve(lambda y, t: spiral_true_rhs(y), u0, 0.0, 1.5,
                      abstol=rtol, reltol=rtol, saveat_interior=interior,
                      save_start=ss, save_end=se, save_everystep=sev)
    assert sol.retcode == tsit5.RET_SUCCESS
    assert np.allclose(sol.t, ts, rtol=0, atol=1e-15)
    err = np.abs(np.array(sol.u)[:, 0, :] - ref.y.T).max()
    assert err < 50 * rtol, err
    assert err > 1e-4 * rtol  # not trivially exact: the check has teeth


def test_saveat_exact_hit_copies_state():
    _, u0, f = _linear()
    sol = tsit5.solve(f, u0, 0.0, 1.0, adaptive=False, dt=0.25,
                      saveat_interior=[0.5, 0.6], save_start=True)
    assert sol.t == [0.0, 0.5, 0.6, 1.0]
    assert np.array_equal(sol.u[1], sol.steps[1].y1)  # t=0.5 is a step end


def test_interpolant_matches_saveat_and_dense():
    W, u0, f = _linear()
    sol = tsit5.solve(f, u0, 0.0, 1.0, abstol=1e-10, reltol=1e-10)
    for t in (0.013, 0.4, 0.77):
        assert np.abs(sol.interp(t) - u0 @ sla.expm(W * t).T).max() < 1e-8
    # step endpoints return stored states exactly
    s = sol.steps[2]
    assert np.array_equal(sol.interp(s.t1), s.y1)


def test_controller_growth_capped_by_qmax():
    # y' = 0 has EEst = 0 -> q = 1/qmax: dt grows 10x per accepted step
    f = lambda y, t: np.zeros_like(y)
    sol = tsit5.solve(f, np.ones((1, 1)), 0.0, 1e6, dt=1e-3, dtmax=1e9)
    dts = [s.dt for s in sol.steps]
    assert all(abs(dts[i + 1] / dts[i] - 10) < 1e-9 for i in range(len(dts) - 2))


def test_initdt_heuristic_closed_form():
    # y' = -y, y0 = 1: d0 = 1/(atol+rtol), d1 = d0, dt0 = 0.01, f1 = -(1-0.01)
    atol, rtol = 1e-6, 1e-3
    f = lambda y, t: -y
    y0 = np.ones(1)
    dt, _ = tsit5.initdt(f, y0, f(y0, 0), 0.0, 1.0, 10.0, atol, rtol, 1e-16)
    sk = atol + rtol
    d1 = 1 / sk
    d2 = (0.01 / sk) / 0.01
    expect = min(100 * 0.01, (0.01 / max(d1, d2)) ** (1 / 5))
    assert abs(dt - expect) < 1e-15


def _hairer_wanner_ii4(f, x0, y0, atol, rtol, order):
    """Starting step size, Hairer-Norsett-Wanner, Solving ODEs I, II.4 (the
    algorithm steps (a)-(f) as printed), with an explicit loop norm."""
    n = len(y0)
    sc = [atol + abs(v) * rtol for v in y0]
    norm = lambda v: math.sqrt(sum((v[i] / sc[i]) ** 2 for i in range(n)) / n)
    fx0 = f(y0, x0)
    d0, d1 = norm(y0), norm(fx0)                                   # (a)
    h0 = 1e-6 if (d0 < 1e-5 or d1 < 1e-5) else 0.01 * (d0 / d1)    # (b)
    y1 = [y0[i] + h0 * fx0[i] for i in range(n)]                   # (c)
    fx1 = f(y1, x0 + h0)
    d2 = norm([fx1[i] - fx0[i] for i in range(n)]) / h0            # (d)
    mx = max(d1, d2)
    h1 = max(1e-6, h0 * 1e-3) if mx <= 1e-15 else (0.01 / mx) ** (1.0 / order)  # (e)
    return min(100 * h0, h1)                                       # (f)


_INITDT_CASES = {
    # linear decay: the (d)-(e) branch binds, d2 = ||A f0|| whatever h0
    "fast_decay": (lambda y: y * np.array([-40.0, -3.0, -0.5]), [1.0, -2.0, 0.25], False),
    # near rest with a large forcing: ||y0|| << ||f(y0)||, so 100 h0 binds and
    # a swapped d0 / d1 changes the step by orders of magnitude
    "near_rest": (lambda y: np.stack([np.sin(3 * y[..., 1]) + 5.0, -y[..., 0] ** 2], -1),
                  [1e-3, 2e-3], True),
    "saturating": (lambda y: 50 * np.tanh(y * np.array([-40.0, -3.0, -0.5])) + 0.1 * y ** 2,
                   [1e-2, -2e-2, 2.5e-3], True),
}


@pytest.mark.parametrize("case", list(_INITDT_CASES))
def test_initdt_matches_hairer_wanner_with_distinct_norms(case):
    """Problems whose ||y0|| and ||f(y0)|| differ by orders of magnitude: the
    oracle's initial step equals the textbook algorithm (order 5 for Tsit5),
    and where the 100 h0 cap binds, a swapped d0 / d1 would not."""
    atol, rtol = 1e-6, 1e-3
    f1, y0l, discriminates = _INITDT_CASES[case]
    f = lambda y, t: f1(y)
    y0 = np.array(y0l)
    fl = lambda y, t: list(f1(np.asarray(y, float)))
    expect = _hairer_wanner_ii4(fl, 0.0, list(y0), atol, rtol, 5)
    dt, _ = tsit5.initdt(f, y0, f(y0, 0.0), 0.0, 1.0, 10.0, atol, rtol, 1e-16)
    assert abs(dt - expect) <= 1e-13 * expect
    sc = atol + np.abs(y0) * rtol
    d0 = float(np.sqrt(np.mean((y0 / sc) ** 2)))
    d1 = float(np.sqrt(np.mean((f(y0, 0.0) / sc) ** 2)))
    assert abs(math.log(d0 / d1)) > 1.0
    if discriminates:  # the same algorithm with d0 and d1 swapped lands elsewhere
        h0_sw = 0.01 * d1 / d0
        y1 = y0 + h0_sw * f(y0, 0.0)
        d2_sw = float(np.sqrt(np.mean(((f(y1, h0_sw) - f(y0, 0.0)) / sc) ** 2))) / h0_sw
        alt = min(100 * h0_sw, (0.01 / max(d1, d2_sw)) ** 0.2)
        assert abs(alt - expect) > 0.5 * expect


def test_retcode_maxiters():
    _, u0, f = _linear()
    sol = tsit5.solve(f, u0, 0.0, 1.0, maxiters=3)
    assert sol.retcode == tsit5.RET_MAXITERS and sol.naccept + sol.nreject == 3


def test_retcode_unstable_on_nan():
    f = lambda y, t: y * np.nan
    sol = tsit5.solve(f, np.ones((1, 2)), 0.0, 1.0, adaptive=False, dt=0.1)
    assert sol.retcode == tsit5.RET_UNSTABLE


def test_retcode_dtlessthanmin():
    _, u0, f = _linear()
    sol = tsit5.solve(f, u0, 0.0, 1.0, abstol=1e-14, reltol=1e-14, dtmin=0.2)
    assert sol.retcode == tsit5.RET_DTLESSTHANMIN


def test_save_plan_reference_kwargs():
    # test/BasicNeuralDE/neural_de_tests.jl:28,38,39
    assert tsit5.save_plan(0.0, 1.0, None, False, False) == ([], False, True, False)
    interior, ss, se, sev = tsit5.save_plan(0.0, 1.0, np.arange(0, 11) / 10)
    assert ss and se and not sev and len(interior) == 9
    assert tsit5.save_plan(0.0, 1.0, 0.1)[0] == [i / 10 for i in range(1, 10)]
    assert tsit5.save_plan(0.0, 1.0)[3] is True
