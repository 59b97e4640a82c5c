"""Failure retcodes forced from the host, GPU against the oracle
(``oracle/tsit5.py``, the restatement of OrdinaryDiffEq's ``check_error``):
tiny ``maxiters`` -> MaxIters, NaN parameters or cotangents -> DtNaN /
Unstable, ``dtmin`` above the step the tolerance needs -> DtLessThanMin; the
batched handle, the per-trajectory (ragged) path, and the adjoint.  A failed
forward cannot be differentiated (the reference's solve returns the retcode
and its adjoint needs a successful solve).  Every case is plain arithmetic
on NaN / a step-count limit: no kernel runs out of bounds."""
import numpy as np
import pytest

from oracle import adjoint as OA
from oracle import tsit5

from test_gpu_parity import setup_case

pytestmark = pytest.mark.gpu

RET = {"Success": 1, "DtNaN": 3, "MaxIters": 4, "DtLessThanMin": 5, "Unstable": 6}


def gpu_and_oracle(p_override=None, **kw):
    from diffeqflux_hip import Handle
    m, p, u0, h, (t0, t1) = setup_case("plain", "fp32")
    if p_override is not None:
        p = p_override(p)
        h.set_params(p.astype(np.float32))
    opts = Handle.make_opts(t0, t1, **kw)
    out, ts, st = h.forward(u0, opts, [])
    ref = OA.forward(m, p, u0.astype(np.float64), t0, t1, saveat=[t0, t1], **kw)
    print(f"  {kw}: gpu retcode {st['retcode']} ({st['naccept']}+{st['nreject']}), oracle "
          f"{ref.retcode} ({ref.naccept}+{ref.nreject})", flush=True)
    return h, st, ref, m, p, u0


def test_maxiters():
    h, st, ref, *_ = gpu_and_oracle(maxiters=3, abstol=1e-8, reltol=1e-8)
    assert ref.retcode == tsit5.RET_MAXITERS == RET["MaxIters"]
    assert st["retcode"] == ref.retcode
    assert st["naccept"] + st["nreject"] == 3


def test_nan_parameters_adaptive():
    def nan_bias(p):
        q = p.copy()
        q[-1] = np.nan
        return q
    h, st, ref, *_ = gpu_and_oracle(nan_bias, abstol=1e-6, reltol=1e-6)
    assert ref.retcode in (RET["DtNaN"], RET["Unstable"])
    assert st["retcode"] == ref.retcode


def test_nan_parameters_fixed_dt():
    def nan_bias(p):
        q = p.copy()
        q[-1] = np.nan
        return q
    h, st, ref, *_ = gpu_and_oracle(nan_bias, adaptive=False, dt=0.1)
    assert ref.retcode == RET["Unstable"]
    assert st["retcode"] == ref.retcode and st["naccept"] == ref.naccept


def test_dtmin_above_needed_step():
    h, st, ref, *_ = gpu_and_oracle(abstol=1e-12, reltol=1e-12, dtmin=0.2)
    assert ref.retcode == RET["DtLessThanMin"]
    assert st["retcode"] == ref.retcode


def test_failed_forward_cannot_be_differentiated():
    from diffeqflux_hip import NodeHipError
    h, st, ref, *_ = gpu_and_oracle(maxiters=3)
    assert st["retcode"] != RET["Success"]
    with pytest.raises(NodeHipError):
        h.backward(np.ones((2, h.B, h.D), np.float32))


def test_adjoint_with_nan_cotangent():
    """A NaN cotangent makes the adjoint's error estimate NaN: the backward
    solve stops with the oracle's retcode instead of looping."""
    h, st, ref, m, p, u0 = gpu_and_oracle(abstol=1e-6, reltol=1e-6)
    assert st["retcode"] == RET["Success"]
    c = np.zeros((2, h.B, h.D), np.float32)   # saves at t0 and T
    c[1, 0, 0] = np.nan
    _, _, sb = h.backward(c)
    _, _, adj = OA.backward(m, p, ref, [ci.astype(np.float64) for ci in c], abstol=1e-6,
                            reltol=1e-6)
    print(f"  adjoint: gpu retcode {sb['retcode']}, oracle {adj.retcode}", flush=True)
    assert adj.retcode in (RET["DtNaN"], RET["Unstable"])
    assert sb["retcode"] == adj.retcode


def test_ragged_per_trajectory_retcodes():
    """Per-trajectory solves report each trajectory's own retcode: one
    trajectory with a NaN initial state fails (oracle retcode), the others
    succeed; a tiny maxiters fails all of them with MaxIters."""
    from diffeqflux_hip import Handle
    from oracle.mlp import MLP
    widths, acts = [4, 32, 4], ["tanh", "identity"]
    m = MLP(widths, acts)
    p = m.glorot_init(np.random.default_rng(1)).astype(np.float32)
    h = Handle(widths, acts, 0, "fp32", 1)
    h.set_params(p)
    rng = np.random.default_rng(2)
    u0 = rng.standard_normal((5, 4)).astype(np.float32)
    u0[2, 1] = np.nan
    t1 = np.array([1.0, 1.5, 1.0, 0.5, 2.0])
    saves = [[float(t)] for t in t1]
    _, ret, ts, _ = h.ragged_forward(u0, t1, saves, abstol=1e-6, reltol=1e-6)
    want = []
    for i in range(5):
        r = OA.forward(m, p.astype(np.float64), u0[i:i + 1].astype(np.float64), 0.0, float(t1[i]),
                       saveat=[float(t1[i])], abstol=1e-6, reltol=1e-6)
        want.append(r.retcode)
    print(f"\n  ragged retcodes gpu {list(ret)} oracle {want}", flush=True)
    assert list(ret) == want and want[2] != RET["Success"] and want.count(RET["Success"]) == 4
    _, ret2, ts2, _ = h.ragged_forward(np.nan_to_num(u0), t1, saves, abstol=1e-10, reltol=1e-10,
                                       maxiters=2)
    assert list(ret2) == [RET["MaxIters"]] * 5
