"""The workloads of the configuration list at their own sizes on the GPU,
against the FP64 oracle on the same inputs:

* plain FP32 NeuralODE, D = 32, MLP 32-256-256-32 tanh (Glorot), B = 4096
  u0 ~ N(0, 1), tspan (0, 1), 20 saveat, rtol 1e-5, atol 1e-7, MSE loss
  against the saveat states of the seeded linear-plus-tanh teacher ODE:
  saved states within 20 (atol + rtol |u|) at every save, dL/dp and dL/du0
  within 1e-3 (norm-relative; FP32 state and error estimate: the step
  sequence may differ from the FP64 oracle's);
* the spiral of the docs example (B = 1, 30 saveat, abstol = reltol = 1e-6,
  Chain(x -> x.^3, Dense(2, 50, tanh), Dense(50, 2)), MSE against the true
  spiral's states): the same bars.
"""
import numpy as np
import pytest

from oracle import adjoint as OA
from oracle import problems as PR
from oracle import tsit5

pytestmark = pytest.mark.gpu


def _nrel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _run(pr, prec="fp32"):
    from diffeqflux_hip import Handle
    m, p, u0 = pr["model"], pr["p"], pr["u0"]
    ts = list(pr["saveat"])
    tgt = pr["targets"]
    p32, u32 = p.astype(np.float32), u0.astype(np.float32)
    B = u0.shape[0]
    h = Handle(m.widths, m.acts, m.input_map, prec, B)
    h.set_params(p32)
    tol = dict(abstol=pr["abstol"], reltol=pr["reltol"])
    opts = Handle.make_opts(*pr["tspan"], **tol)
    out, tg, st = h.forward(u32, opts, ts[1:-1])
    assert st["retcode"] == 1 and np.array_equal(tg, ts)
    cot = (2.0 / out.size) * (out - tgt.astype(np.float32))
    du0, dp, sb = h.backward(cot.astype(np.float32))
    assert sb["retcode"] == 1
    ref = OA.forward(m, p32.astype(np.float64), u32.astype(np.float64), *pr["tspan"], saveat=ts,
                     **tol)
    R = np.array(ref.u)
    rcot = [(2.0 / R.size) * (R[i] - tgt[i]) for i in range(len(ts))]
    gu, gp, adj = OA.backward(m, p32.astype(np.float64), ref, rcot, **tol)
    es = float(np.max(np.abs(out - R) / (pr["abstol"] + pr["reltol"] * np.abs(R))))
    ep, eu = _nrel(dp, gp), _nrel(du0, gu)
    print(f"  saves err {es:.2f} x tol; dL/dp {ep:.2e} dL/du0 {eu:.2e}; steps fwd "
          f"{st['naccept']}+{st['nreject']} (oracle {ref.naccept}+{ref.nreject}) adj "
          f"{sb['naccept']}+{sb['nreject']} (oracle {adj.naccept}+{adj.nreject})", flush=True)
    return es, ep, eu


def test_plain_workload_full_batch():
    print("\n[workload plain D=32 H=256x2 tanh B=4096 rtol 1e-5 atol 1e-7 fp32]", flush=True)
    pr = PR.plain(B=4096)
    pr["targets"] = PR.teacher_targets(pr["u0"], pr["saveat"])
    es, ep, eu = _run(pr)
    assert es < 20 and ep < 1e-3 and eu < 1e-3


def test_spiral_workload():
    print("\n[workload spiral B=1 30 saveat tol 1e-6 fp32]", flush=True)
    pr = PR.spiral()
    s = tsit5.solve(PR.spiral_true_rhs, pr["u0"], 0.0, 1.5, saveat_interior=list(pr["saveat"][1:-1]),
                    abstol=1e-12, reltol=1e-12, record_dense=False)
    pr["targets"] = np.array(s.u)
    es, ep, eu = _run(pr)
    assert es < 20 and ep < 1e-3 and eu < 1e-3
