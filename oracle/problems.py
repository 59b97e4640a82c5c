"""Workload definitions (synthetic data of each configuration's shape).

* spiral: the 2-D cubic spiral of ``docs/src/examples/neural_ode.md:18-33``
  and ``test/Layers/stiff_nested_ad_tests.jl:8-24``:
  ``du/dt = ((u.^3)' * [-0.1 2; -2 -0.1])'``, ``u0 = [2, 0]``,
  ``tspan = (0, 1.5)``, 30 uniform saveat points; model
  ``Chain(x -> x.^3, Dense(2, 50, tanh), Dense(50, 2))``.
* plain:  D=32, MLP 32-256-256-32 tanh, B=4096, u0 ~ N(0,1), tspan (0,1),
  20 saveat points, rtol 1e-5, atol 1e-7; targets from the teacher
  ``u' = A u + tanh(C u)`` with ``A = -I + 0.5 G/sqrt(D)``, ``C = G'/sqrt(D)``.
* mixed:  D=64, MLP 64-1024-1024-1024-64 softplus, B=4096, tspan (0,1),
  32 saveat points, rtol 1e-4 (atol 1e-6).
* ragged: 10,000 trajectories, D=16, MLP 16-128-128-16 tanh,
  T_i ~ U(0.1, 5), Nt_i ~ U{5..200}, rtol 1e-5.
* ffjord: D=16 Gaussian-mixture data, time-concatenated MLP 17-512-512-16.

Random streams are numpy PCG64 with the seeds given; Glorot-uniform weights,
zero biases (Lux's own init stream needs Julia and is not reproduced; parity
is always on identical parameters).
"""
import numpy as np

from .mlp import MLP, MAP_CUBE

SPIRAL_A = np.array([[-0.1, 2.0], [-2.0, -0.1]])


def spiral_true_rhs(u, t=None):
    """((u.^3)' * A)' on (B, 2) rows: each row u -> (u^3) @ A."""
    return (u ** 3) @ SPIRAL_A


def spiral(dtype=np.float64, seed=0):
    m = MLP([2, 50, 2], ["tanh", "identity"], input_map=MAP_CUBE)
    p = m.glorot_init(np.random.default_rng(seed), dtype)
    u0 = np.array([[2.0, 0.0]], dtype=dtype)
    ts = np.linspace(0.0, 1.5, 30)
    return dict(model=m, p=p, u0=u0, tspan=(0.0, 1.5), saveat=ts,
                abstol=1e-6, reltol=1e-6)


def teacher_rhs(D, seed):
    rng = np.random.default_rng(seed)
    A = -np.eye(D) + 0.5 * rng.standard_normal((D, D)) / np.sqrt(D)
    C = rng.standard_normal((D, D)) / np.sqrt(D)
    return lambda u, t=None: u @ A.T + np.tanh(u @ C.T)


def teacher_targets(u0, ts, seed=1, abstol=1e-12, reltol=1e-12):
    """Saved states of the teacher ``u' = A u + tanh(C u)`` (``teacher_rhs``
    with the given seed) at the times ``ts`` from ``u0`` (B, D), solved by the
    FP64 oracle Tsit5 at tight tolerance: (len(ts), B, D)."""
    from . import tsit5
    f = teacher_rhs(u0.shape[1], seed)
    s = tsit5.solve(f, np.asarray(u0, np.float64), ts[0], ts[-1], saveat_interior=list(ts[1:-1]),
                    save_start=True, save_end=True, abstol=abstol, reltol=reltol,
                    record_dense=False)
    return np.array(s.u)


def mlp_problem(D, hidden, n_hidden, act, B, n_save, reltol, abstol,
                dtype=np.float64, seed=0, u_seed=1):
    widths = [D] + [hidden] * n_hidden + [D]
    acts = [act] * n_hidden + ["identity"]
    m = MLP(widths, acts)
    p = m.glorot_init(np.random.default_rng(seed), dtype)
    u0 = np.random.default_rng(u_seed).standard_normal((B, D)).astype(dtype)
    ts = np.linspace(0.0, 1.0, n_save)
    return dict(model=m, p=p, u0=u0, tspan=(0.0, 1.0), saveat=ts,
                abstol=abstol, reltol=reltol)


def plain(B=4096, dtype=np.float64):
    return mlp_problem(32, 256, 2, "tanh", B, 20, 1e-5, 1e-7, dtype)


def mixed(B=4096, dtype=np.float64):
    return mlp_problem(64, 1024, 3, "softplus", B, 32, 1e-4, 1e-6, dtype)
