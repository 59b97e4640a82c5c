"""GPU parity of FFJORD sampling (``__backward_ffjord``, src/ffjord.jl:183-214:
the augmented solve over the reversed time span) and of ``FFJORDDistribution``
(``:216-246``: ``rand`` / ``logpdf``) on the HIP path.

* the reversed solve (node_cnf_forward with t0 > t1) against the oracle's
  ``sample`` (FP64) with the same probe: z within 20 * (abstol + reltol*|z|);
* linear flow: z = e^{-AT} x (closed form) to 1e-5 relative;
* ``FFJORDDistribution``: ``rand`` from a seeded generator, then the forward
  flow of those samples returns the base draws (round trip; at the FFJORD
  workload shape in split BF16 too), ``logpdf`` equals the layer's logpx and
  ``_logpdf`` its first entry (the reference's ``first(first(...))``).
"""
import numpy as np
import pytest
from scipy.linalg import expm

from oracle import ffjord as FF
from oracle.mlp import MLP

pytestmark = pytest.mark.gpu

CASES = {
    # name: widths, acts, B, tol, precision
    "time_small": ([3, 16, 16, 2], ["tanh", "tanh", "identity"], 64, 1e-6, "fp32"),
    "z_only": ([2, 24, 2], ["tanh", "identity"], 37, 1e-6, "fp32"),
    "config5_shape": ([17, 512, 512, 16], ["tanh", "tanh", "identity"], 128, 1e-5, "bf16x3"),
}


@pytest.mark.parametrize("name", list(CASES))
def test_reversed_solve_matches_oracle_sample(name):
    from diffeqflux_hip import Handle, _lib
    widths, acts, B, tol, prec = CASES[name]
    D = widths[-1]
    m = MLP(widths, acts)
    rng = np.random.default_rng(3)
    p32 = m.glorot_init(rng).astype(np.float32)
    x = rng.standard_normal((B, D)).astype(np.float32)
    e = rng.choice(np.array([-1.0, 1.0], np.float32), size=(B, D))
    tc = widths[0] == D + 1
    print(f"\n[ffjord sampling {name} {prec} B={B}]", flush=True)
    h = Handle(widths, acts, _lib.MAP_CNF_TIME if tc else _lib.MAP_CNF, prec, B)
    h.set_params(p32)
    _, z, _, st = h.cnf_forward(x, e, 1.0, 0.0, abstol=tol, reltol=tol)
    assert st["retcode"] == 1
    rz, sol = FF.sample(m, p32.astype(np.float64), x.astype(np.float64), 0.0, 1.0,
                        probe=e.astype(np.float64), time_concat=tc, abstol=tol, reltol=tol)
    ez = np.abs(z - rz) / (tol + tol * np.abs(rz))
    moved = np.abs(rz - x).max()
    print(f"  steps gpu {st['naccept']}+{st['nreject']} oracle {sol.naccept}+{sol.nreject}; "
          f"scaled err z {ez.max():.3f}; flow moved samples by {moved:.3f}", flush=True)
    assert moved > 1e-2
    assert ez.max() < 20


def test_sampling_linear_flow_closed_form():
    from diffeqflux_hip import Handle, _lib
    D, B, T = 4, 16, 0.75
    m = MLP([D, D, D], ["identity", "identity"])
    rng = np.random.default_rng(5)
    W1 = 0.3 * rng.standard_normal((D, D)); W2 = 0.3 * rng.standard_normal((D, D))
    p = np.zeros(m.nparams, np.float32)
    (w0, b0, _), (w1, b1, _) = m.offsets
    p[w0:b0] = W1.T.ravel(); p[w1:b1] = W2.T.ravel()
    A = W2.astype(np.float32).astype(np.float64) @ W1.astype(np.float32).astype(np.float64)
    x = rng.standard_normal((B, D)).astype(np.float32)
    e = rng.choice(np.array([-1.0, 1.0], np.float32), size=(B, D))
    h = Handle([D, D, D], ["identity", "identity"], _lib.MAP_CNF, "fp32", B)
    h.set_params(p)
    _, z, _, st = h.cnf_forward(x, e, T, 0.0, abstol=1e-8, reltol=1e-8)
    want = x.astype(np.float64) @ expm(-A * T).T
    err = np.abs(z - want).max() / np.abs(want).max()
    print(f"\n  linear flow sampling: z rel err {err:.2e}", flush=True)
    assert st["retcode"] == 1
    assert err < 1e-5


def _layer(widths, acts, prec, tol, seed):
    from diffeqflux_hip import FFJORD, Chain, Dense
    D = widths[-1]
    layers = [Dense(i, o, a) for i, o, a in zip(widths[:-1], widths[1:], acts)]
    ffj = FFJORD(Chain(*layers), (0.0, 1.0), (D,), abstol=tol, reltol=tol, precision=prec)
    ps, st = ffj.setup(np.random.default_rng(seed))
    return ffj, np.asarray(ps, np.float32), st


@pytest.mark.parametrize("widths,acts,prec,tol,n,bar", [
    ([3, 16, 16, 2], ["tanh", "tanh", "identity"], "fp32", 1e-7, 64, 1e-4),
    ([17, 512, 512, 16], ["tanh", "tanh", "identity"], "bf16x3", 1e-6, 8192, 1e-3),
])
def test_distribution_rand_round_trip_and_logpdf(widths, acts, prec, tol, n, bar):
    from diffeqflux_hip import FFJORDDistribution
    ffj, ps, st = _layer(widths, acts, prec, tol, seed=11)
    D = widths[-1]
    st = dict(st, rng=np.random.default_rng(21))
    d = FFJORDDistribution(ffj, ps, st)
    print(f"\n[FFJORDDistribution {widths} {prec} n={n}]", flush=True)
    assert len(d) == D
    zs = d.rand(n, np.random.default_rng(5))
    assert zs.shape == (n, D) and np.isfinite(zs).all()
    x = np.random.default_rng(5).standard_normal((n, D)).astype(np.float32)   # the base draws
    # forward flow of the samples returns the base draws
    (lp, _, _), st2 = ffj(zs, ps, dict(st, rng=np.random.default_rng(1)))
    h = ffj._handle(n, False, True)
    _, xr, _, _ = h.cnf_forward(zs, np.ones((n, D), np.float32), 0.0, 1.0, abstol=tol, reltol=tol)
    err = np.abs(xr - x).max() / max(1.0, np.abs(x).max())
    moved = np.abs(zs - x).max()
    print(f"  round trip max err {err:.2e} (samples moved {moved:.3f})", flush=True)
    assert moved > 1e-2
    assert err < bar
    # logpdf of a batch is the layer's logpx; _logpdf its first entry; one sample alone
    lpd = d.logpdf(zs)
    st_p = dict(st, probe=np.ones((n, D), np.float32))
    (lp2, _, _), _ = ffj(zs, ps, st_p)
    d2 = FFJORDDistribution(ffj, ps, st_p)
    assert np.allclose(d2.logpdf(zs), np.asarray(lp2).reshape(-1), rtol=0, atol=0)
    assert d2._logpdf(zs) == float(np.asarray(lp2).reshape(-1)[0])
    assert np.isfinite(lpd).all() and lpd.shape == (n,)
    one = d.rand(None, np.random.default_rng(9))
    assert one.shape == (D,)
