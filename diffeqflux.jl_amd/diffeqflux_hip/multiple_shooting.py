"""Direct multiple shooting on the HIP path (``src/multiple_shooting.jl``).

``multiple_shoot(p, ode_data, tsteps, prob, loss_function, [continuity_loss],
solver, group_size; continuity_term = 100, kwargs...)`` splits the time grid
into overlapping groups (``group_ranges``), solves the model on every group
from the data point at the group's start, and returns the summed group losses
plus ``continuity_term`` times the continuity losses between the last
prediction of a group and the data point that starts the next one
(``src/multiple_shooting.jl:36-84``); the ensemble form
(``src/multiple_shooting.jl:131-195``) does the same for every trajectory of
``ode_data``.

Every group (times every trajectory) is one trajectory of a single
per-trajectory solve (``node_ragged_forward``): its own initial state, time
span and save times in one persistent launch.  The right-hand side is
autonomous (``src/neural_de.jl:4``), so a group over ``[t_a, t_b]`` is solved
on ``[0, t_b - t_a]`` with its save times shifted by ``t_a``.  The gradient
with respect to the parameters runs through ``node_ragged_backward``.  Models
outside the per-trajectory kernels' limits (more than 4 Dense layers, D > 32,
hidden widths > 128) solve each group as one batched ``NeuralODE`` call over
the trajectories instead.

Array layout follows the rest of this package (the Julia arrays' bytes with
the axes reversed): ``ode_data`` is ``(T, D)`` for one trajectory and
``(ntraj, T, D)`` for the ensemble form; a group prediction is ``(len, D)``
resp. ``(ntraj, len, D)``; ``loss_function(u, û)`` receives the data and the
prediction of one group in that layout.
"""
import numpy as np

from ._lib import NodeHipError
from .neural_ode import NeuralODE, Tsit5

__all__ = ["group_ranges", "multiple_shoot", "EnsembleProblem", "EnsembleThreads",
           "EnsembleSerial", "EnsembleGPUArray"]

_SUCCESS = 1


def group_ranges(datasize, groupsize):
    """``group_ranges(datasize, groupsize)`` (``src/multiple_shooting.jl:231-239``)
    as 0-based Python ranges: ``range(s, min(datasize, s + groupsize))`` for
    ``s`` in ``0 : groupsize - 1 : datasize - 2``; consecutive groups share
    their boundary point and the last group holds what remains, e.g.
    ``group_ranges(10, 5) -> [range(0, 5), range(4, 9), range(8, 10)]``
    (Julia: ``1:5, 5:9, 9:10``)."""
    datasize, groupsize = int(datasize), int(groupsize)
    if not 2 <= groupsize <= datasize:
        raise ValueError(f"groupsize {groupsize}: datasize must be positive and groupsize "
                         "must be within [2, datasize]")
    return [range(s, min(datasize, s + groupsize)) for s in range(0, datasize - 1, groupsize - 1)]


def _default_continuity_loss(u_end, u_0):
    """``sum(abs, û_end - u_0)`` (``src/multiple_shooting.jl:243``)."""
    d = u_end - u_0
    return d.abs().sum() if hasattr(d, "abs") else np.abs(d).sum()


class EnsembleProblem:
    """``EnsembleProblem(prob)``: marks the ensemble form of ``multiple_shoot``
    (the trajectories are the leading axis of ``ode_data``)."""

    def __init__(self, prob, **kwargs):
        self.prob = prob
        self.kwargs = kwargs


class EnsembleSerial:
    pass


class EnsembleThreads:
    pass


class EnsembleGPUArray:
    pass


def _is_torch(x):
    try:
        import torch
        return isinstance(x, torch.Tensor)
    except ImportError:
        return False


def _parse(args):
    """``[continuity_loss,] solver, group_size`` (the reference's two methods)."""
    if len(args) == 2:
        return _default_continuity_loss, args[0], args[1]
    if len(args) == 3:
        return args[0], args[1], args[2]
    raise TypeError("expected ([continuity_loss,] solver, group_size)")


def multiple_shoot(p, ode_data, tsteps, prob, loss_function, *args, continuity_term=100,
                   **kwargs):
    """Returns ``(loss, group_predictions)``; ``loss`` is ``inf`` when a group
    solve does not succeed (``src/multiple_shooting.jl:66-67``).

    ``prob``: a ``NeuralODE`` (its model, precision and solver keywords; the
    time span is replaced per group) or an ``EnsembleProblem`` of one, in
    which case the next positional argument is the ensemble algorithm and
    ``trajectories`` must equal ``ode_data.shape[0]``.  ``p``: flat FP32
    parameters (a torch tensor with ``requires_grad`` makes the loss
    differentiable).  Remaining keywords are solver keywords (``abstol``,
    ``reltol``, ...)."""
    ensemble = isinstance(prob, EnsembleProblem)
    if ensemble:
        if not args:
            raise TypeError("ensemble form: (ensemblealg, loss_function, [continuity_loss,] "
                            "solver, group_size)")
        loss_function, args = args[0], args[1:]
    continuity_loss, solver, group_size = _parse(args)
    if not isinstance(solver, Tsit5):
        raise NotImplementedError(f"{solver!r}: only Tsit5() is on the HIP path")
    node = prob.prob if ensemble else prob
    if not isinstance(node, NeuralODE):
        raise NotImplementedError("multiple_shoot on the HIP path takes a NeuralODE problem")
    trajectories = kwargs.pop("trajectories", None)
    t = np.asarray(tsteps.detach().cpu().numpy() if _is_torch(tsteps) else tsteps, np.float64)
    data_t = ode_data if _is_torch(ode_data) else None
    data = np.asarray(ode_data.detach().cpu().numpy() if data_t is not None else ode_data,
                      np.float32)
    D = node.model.widths[0]
    if ensemble:
        if data.ndim != 3:
            raise ValueError("ensemble ode_data must be (trajectories, T, D)")
        if trajectories is not None and int(trajectories) != data.shape[0]:
            raise ValueError("trajectories must equal ode_data.shape[0]")
        datasize = data.shape[1]
    else:
        if data.ndim != 2:
            raise ValueError("ode_data must be (T, D)")
        datasize = data.shape[0]
    if data.shape[-1] != D:
        raise ValueError(f"ode_data must have D={D} entries per state")
    if datasize != t.size:
        raise ValueError("len(tsteps) must equal the number of data points")
    if not 2 <= int(group_size) <= datasize:
        raise ValueError(f"group_size {group_size} can't be < 2 or > number of data points")
    ranges = group_ranges(datasize, group_size)
    solve_kw = dict(node.kwargs)
    solve_kw.update(kwargs)
    solve_kw.pop("saveat", None)

    preds, ok = _solve_groups(node, p, data if ensemble else data[None], t, ranges, solve_kw)
    if not ensemble:
        preds = [g[0] for g in preds]
    if not ok:
        return float("inf"), preds
    u_all = data_t if data_t is not None else data
    if _is_torch(preds[0]) and not _is_torch(u_all):
        import torch
        u_all = torch.as_tensor(data, device=preds[0].device)
    loss = 0
    for i, rg in enumerate(ranges):
        sl = slice(rg.start, rg.stop)
        u = u_all[:, sl] if ensemble else u_all[sl]
        loss = loss + loss_function(u, preds[i])
        if i > 0:
            prev_end = preds[i - 1][:, -1] if ensemble else preds[i - 1][-1]
            first = u[:, 0] if ensemble else u[0]
            loss = loss + continuity_term * continuity_loss(prev_end, first)
    return loss, preds


def _segments(data, t, ranges):
    """Per (group, trajectory) solve set-up, group-major: initial states,
    shifted end times and the shifted save times as (concatenated times,
    offsets), the per-trajectory entry point's layout."""
    ntraj = data.shape[0]
    u0 = np.concatenate([data[:, rg.start] for rg in ranges]).astype(np.float32)
    ts = [t[rg.start:rg.stop] - t[rg.start] for rg in ranges]
    t1 = np.concatenate([np.full(ntraj, g[-1]) for g in ts])
    times = np.concatenate([np.tile(g, ntraj) for g in ts])
    lens = np.concatenate([np.full(ntraj, len(g), np.int64) for g in ts])
    off = np.zeros(lens.size + 1, np.int64)
    off[1:] = np.cumsum(lens)
    return u0, t1, (times, off)


def _solve_groups(node, p, data, t, ranges, kw):
    """Group predictions (list of (ntraj, len, D)) and whether every solve
    succeeded; differentiable in ``p`` when ``p`` is a torch tensor."""
    h = node._handle(1)
    try:
        return _solve_ragged(node, h, p, data, t, ranges, kw)
    except NodeHipError as e:
        if "per-trajectory solves support" not in str(e) and "LDS" not in str(e):
            raise
    return _solve_batched(node, p, data, t, ranges, kw)


_RAGGED_KW = ("abstol", "reltol", "adaptive", "dt", "dtmax", "dtmin", "maxiters")


def _ragged_args(kw):
    bad = set(kw) - set(_RAGGED_KW)
    if bad:
        raise NotImplementedError(f"solver keywords not on the per-trajectory path: {sorted(bad)}")
    return dict(abstol=kw.get("abstol", 1e-6), reltol=kw.get("reltol", 1e-3),
                adaptive=kw.get("adaptive", True), dt=kw.get("dt", 0.0),
                dtmax=kw.get("dtmax", 0.0), dtmin=kw.get("dtmin", 0.0),
                maxiters=kw.get("maxiters", 0))


def _split(flat, data, ranges):
    ntraj = data.shape[0]
    out, k = [], 0
    for rg in ranges:
        n = ntraj * len(rg)
        out.append(flat[k:k + n].reshape(ntraj, len(rg), data.shape[-1]))
        k += n
    return out


def _solve_ragged(node, h, p, data, t, ranges, kw):
    u0, t1, saves = _segments(data, t, ranges)
    rk = _ragged_args(kw)
    if _is_torch(p):
        flat, ok = _RaggedFn.apply(p, h, u0, t1, saves, rk)
        return _split(flat, data, ranges), ok
    h.set_params(np.asarray(p, np.float32))
    flat, ret, _, _ = h.ragged_forward(u0, t1, saves, flat_out=True, **rk)
    return _split(flat, data, ranges), bool(np.all(ret == _SUCCESS))


def _ragged_fn_factory():
    import torch

    class _Fn(torch.autograd.Function):
        """rrule of the per-trajectory solves: forward node_ragged_forward,
        backward node_ragged_backward (dL/dp; the initial states are data)."""

        @staticmethod
        def forward(ctx, p, h, u0, t1, saves, rk):
            h.set_params(p.detach().to(torch.float32).cpu().numpy())
            out, ret, _, _ = h.ragged_forward(u0, t1, saves, flat_out=True, **rk)
            ctx.h, ctx.pdev, ctx.pdtype, ctx.gen = h, p.device, p.dtype, h.rg_gen
            ctx.ok = bool(np.all(ret == _SUCCESS))
            flat = torch.from_numpy(out).to(p.device)
            return flat, ctx.ok

        @staticmethod
        def backward(ctx, g, _ok=None):
            if not ctx.ok:
                raise NodeHipError("multiple_shoot: gradient of a failed group solve")
            if ctx.h.rg_gen != ctx.gen:
                raise RuntimeError("multiple_shoot backward: the handle ran another "
                                   "per-trajectory solve since this one (it keeps one tape)")
            _, dp, _, _, _, _ = ctx.h.ragged_backward(g.detach().cpu().numpy())
            return (torch.from_numpy(dp).to(ctx.pdev, ctx.pdtype), None, None, None, None, None)

    return _Fn


class _LazyFn:
    fn = None

    def apply(self, *a):
        if self.fn is None:
            self.fn = _ragged_fn_factory()
        return self.fn.apply(*a)


_RaggedFn = _LazyFn()


def _solve_batched(node, p, data, t, ranges, kw):
    """One batched NeuralODE solve per group over the trajectories; each
    group keeps its own handles (a handle holds the tape of its last forward,
    which the group's backward adjoints)."""
    ntraj = data.shape[0]
    preds, ok = [], True
    cache = node.__dict__.setdefault("_group_handles", {})
    for gi, rg in enumerate(ranges):
        ta, tb = float(t[rg.start]), float(t[rg.stop - 1])
        g = NeuralODE(node.model, (ta, tb), Tsit5(), precision=node.precision,
                      wgrad_precision=node.wgrad_precision,
                      saveat=list(t[rg.start:rg.stop]), **kw)
        g._handles = cache.setdefault(gi, {})
        x = data[:, rg.start]
        if _is_torch(p):
            import torch
            x = torch.as_tensor(x, device=p.device)
        sol, _ = g(x, p)
        ok = ok and sol.retcode == "Success"
        arr = sol.array
        # (len, ntraj, D) -> (ntraj, len, D)
        preds.append(arr.transpose(0, 1) if _is_torch(arr) else np.swapaxes(arr, 0, 1))
    return preds, ok
