"""Python mirror of DiffEqFlux's ``NeuralODE`` Lux layer over the HIP library.

Mirrors ``src/neural_de.jl:35-61``::

    node = NeuralODE(model, tspan, Tsit5(); saveat=..., sensealg=..., kwargs...)
    ps, st = setup(rng, node)
    sol, st = node(u0, ps, st)

with the same argument meaning: ``model`` is a ``Chain`` of ``Dense`` layers
(optionally preceded by the parameter-free ``x -> x.^3`` map), the RHS is
autonomous (``dudt(u, p, t) = model(u, p)``, ``tgrad = 0``), solver keywords
follow OrdinaryDiffEq (``saveat`` range or scalar, ``save_everystep``,
``save_start``, ``save_end``, ``abstol``, ``reltol``, ``dt``, ``adaptive``,
``dtmax``, ``dtmin``, ``maxiters``), and the default ``sensealg`` is
``InterpolatingAdjoint(autojacvec=ZygoteVJP())`` with a user ``sensealg``
overriding it (``src/neural_de.jl:55-58``: the rightmost keyword wins).

Gradients: when ``u0`` or ``ps`` are torch tensors that require grad, the
solve is a ``torch.autograd.Function`` (the counterpart of the ChainRulesCore
``rrule`` SciMLSensitivity attaches to ``solve``), whose backward runs the
InterpolatingAdjoint on the GPU.

Layouts are Julia's: ``u0`` is ``(B, D)`` (a ``D x B`` column-major matrix) or
``(D,)``; ``Array(sol)`` is ``(Nt, B, D)`` (``D x B x Nt``).
"""
import math

import numpy as np

from .handle import Handle
from . import _lib


# ---------------------------------------------------------------- algorithms
class Tsit5:
    """Tsitouras 5(4) with FSAL, embedded error and PI step control."""

    def __repr__(self):
        return "Tsit5()"


class ZygoteVJP:
    def __repr__(self):
        return "ZygoteVJP()"


class InterpolatingAdjoint:
    def __init__(self, autojacvec=None, checkpointing=False):
        self.autojacvec = autojacvec if autojacvec is not None else ZygoteVJP()
        self.checkpointing = checkpointing

    def __repr__(self):
        return (f"InterpolatingAdjoint(autojacvec={self.autojacvec!r}, "
                f"checkpointing={self.checkpointing})")


class _Unsupported:
    def __init__(self, *a, **k):
        pass

    def __repr__(self):
        return f"{type(self).__name__}()"


class BacksolveAdjoint:
    """``BacksolveAdjoint(; autojacvec, checkpointing = true)``: the adjoint
    integrates ``u`` back from ``u(T)`` together with ``[lambda; mu]``
    (checkpointing: ``u`` reset to the forward's saved state at each saved
    time).  On the HIP path with ``autojacvec`` ``ZygoteVJP()`` or unset."""

    def __init__(self, autojacvec=None, checkpointing=True, noisemixing=False):
        self.autojacvec = autojacvec
        self.checkpointing = bool(checkpointing)
        self.noisemixing = noisemixing

    def __repr__(self):
        return f"BacksolveAdjoint(checkpointing={self.checkpointing})"


class TrackerAdjoint(_Unsupported):
    pass


class QuadratureAdjoint(_Unsupported):
    pass


class GaussAdjoint(_Unsupported):
    pass


# ---------------------------------------------------------------- layers
class Dense:
    """Lux ``Dense(in => out, activation)``: ``act.(W * x .+ b)``."""

    def __init__(self, in_dims, out_dims=None, activation="identity"):
        if out_dims is None:  # Dense((in, out)) / Dense(in => out) forms
            in_dims, out_dims = in_dims
        if callable(activation):
            activation = getattr(activation, "__name__", str(activation))
        if activation not in _lib.ACT:
            raise ValueError(f"unsupported activation {activation!r}")
        self.in_dims, self.out_dims, self.activation = int(in_dims), int(out_dims), activation

    @property
    def nparams(self):
        return self.in_dims * self.out_dims + self.out_dims


class Cube:
    """The parameter-free ``x -> x.^3`` WrappedFunction layer."""
    nparams = 0


class Chain:
    def __init__(self, *layers):
        self.layers = list(layers)
        dense = [l for l in self.layers if isinstance(l, Dense)]
        for l in self.layers:
            if not isinstance(l, (Dense, Cube)):
                raise NotImplementedError(f"layer {l!r} is not on the HIP path")
        if any(isinstance(l, Cube) for l in self.layers[1:]):
            raise NotImplementedError("x.^3 is supported only as the first layer")
        self.input_map = _lib.MAP_CUBE if isinstance(self.layers[0], Cube) else _lib.MAP_IDENTITY
        self.widths = [dense[0].in_dims] + [d.out_dims for d in dense]
        for a, b in zip(dense[:-1], dense[1:]):
            if a.out_dims != b.in_dims:
                raise ValueError("Dense widths do not chain")
        self.acts = [d.activation for d in dense]
        self.dense = dense

    @property
    def nparams(self):
        return sum(d.nparams for d in self.dense)

    def init_params(self, rng, init="pytorch"):
        """Flat FP32 parameters in ComponentArray order.  ``init='pytorch'``
        is Lux 1.x's default Dense init (U(+-1/sqrt(in)) for weight and bias);
        ``init='glorot'`` gives Glorot-uniform weights and zero biases."""
        p = []
        for d in self.dense:
            if init == "glorot":
                lim = math.sqrt(6.0 / (d.in_dims + d.out_dims))
                w = rng.uniform(-lim, lim, size=d.in_dims * d.out_dims)
                b = np.zeros(d.out_dims)
            else:
                lim = 1.0 / math.sqrt(d.in_dims)
                w = rng.uniform(-lim, lim, size=d.in_dims * d.out_dims)
                b = rng.uniform(-lim, lim, size=d.out_dims)
            p += [w, b]
        return np.concatenate(p).astype(np.float32)


# ---------------------------------------------------------------- solution
class ODESolution:
    """What downstream code uses of SciMLBase's ODESolution (returned at
    ``src/neural_de.jl:54-60``): ``.u`` (one state per saved time, e.g.
    ``sol.u[-1]`` as ``sol.u[end]`` in docs/src/examples/mnist_neural_ode.md:43),
    ``.t``, ``.retcode``, ``.stats``, ``Array(sol)`` (``np.asarray`` /
    ``.array``), ``len``, indexing, ``sum``, and ``sol(t)``: the state at any
    time of the span by linear interpolation between the saved states (what
    SciMLBase gives for a solution saved at ``saveat`` times, which carries no
    dense output)."""

    def __init__(self, array, t, retcode, stats):
        self.array = array            # (Nt, B, D) numpy or torch
        self.t = np.asarray(t)
        self.retcode = _lib.RET_NAMES.get(retcode, str(retcode))
        self.stats = stats

    @property
    def u(self):
        return [self.array[i] for i in range(len(self.t))]

    def __len__(self):
        return len(self.t)

    def __getitem__(self, i):
        return self.array[i]

    def __array__(self, dtype=None, copy=None):
        a = self.array
        if hasattr(a, "detach"):
            a = a.detach().cpu().numpy()
        return np.asarray(a, dtype=dtype)

    def sum(self):
        return self.array.sum()

    def __call__(self, t):
        ts = self.t
        if len(ts) == 0:
            raise ValueError("solution holds no saved states")
        t = float(t)
        lo, hi = (ts[0], ts[-1]) if ts[-1] >= ts[0] else (ts[-1], ts[0])
        if not lo <= t <= hi:
            raise ValueError(f"t = {t} lies outside the saved times [{lo}, {hi}]: "
                             "the solution cannot extrapolate")
        i = int(np.searchsorted(ts, t, side="right")) - 1 if ts[-1] >= ts[0] else \
            int(np.searchsorted(-ts, -t, side="right")) - 1
        i = min(max(i, 0), len(ts) - 1)
        if ts[i] == t or i == len(ts) - 1:
            return self.array[i]
        w = (t - ts[i]) / (ts[i + 1] - ts[i])
        return self.array[i] + w * (self.array[i + 1] - self.array[i])


def save_plan(t0, t1, saveat=None, save_everystep=None, save_start=None, save_end=None):
    """OrdinaryDiffEq's saving keywords -> (interior times, start, end, everystep).

    ``saveat`` number -> the range ``t0:saveat:t1`` (elements correctly rounded,
    as Julia's float ranges are); a sequence -> those times; empty ->
    every step (``save_everystep`` defaults to ``isempty(saveat)``)."""
    tdir = 1.0 if t1 >= t0 else -1.0
    is_num = saveat is not None and np.ndim(saveat) == 0
    seq = [] if (saveat is None or is_num) else [float(s) for s in np.asarray(saveat).ravel()]
    empty = (not is_num) and not seq
    if save_everystep is None:
        save_everystep = empty
    if is_num:
        interior = [s for s in _julia_range(t0, float(saveat), t1)[1:]
                    if tdir * t0 < tdir * s < tdir * t1]
    else:
        interior = sorted({s for s in seq if tdir * t0 < tdir * s < tdir * t1},
                          key=lambda s: tdir * s)
    if save_start is None:
        save_start = bool(save_everystep or empty or is_num or any(s == t0 for s in seq))
    if save_end is None:
        save_end = bool(save_everystep or empty or is_num or any(s == t1 for s in seq))
    return interior, bool(save_start), bool(save_end), bool(save_everystep)


def _julia_range(t0, step, t1):
    from fractions import Fraction
    fs = Fraction(step).limit_denominator(10 ** 9)
    f0 = Fraction(t0).limit_denominator(10 ** 9)
    if float(fs) != step or float(f0) != t0:
        fs, f0 = Fraction(step), Fraction(t0)
    n = int((Fraction(t1) - f0) / fs) if fs else 0
    out = [float(f0 + i * fs) for i in range(n + 1)]
    if out and float(f0 + (n + 1) * fs) == t1:
        out.append(t1)
    return out


# ---------------------------------------------------------------- the layer
_SOLVER_KW = {"saveat", "save_everystep", "save_start", "save_end", "abstol", "reltol",
              "dt", "adaptive", "dtmax", "dtmin", "maxiters", "sensealg"}


class NeuralODE:
    """``NeuralODE(model, tspan, alg=Tsit5(); kwargs...)`` (src/neural_de.jl:42-45).

    Extra keywords (no counterpart in the reference): ``precision`` ('fp32',
    'bf16' or 'bf16x3') selects the GEMM precision of the right-hand side and
    its VJPs on the GPU; ``wgrad_precision`` (default: ``precision``) that of
    the adjoint's weight-gradient products ('bf16' on a 'bf16x3' layer)."""

    def __init__(self, model, tspan, alg=None, *args, precision="fp32", wgrad_precision=None,
                 **kwargs):
        if alg is None:
            alg = Tsit5()
        if not isinstance(alg, Tsit5):
            raise NotImplementedError(f"{alg!r}: only Tsit5() is on the HIP path")
        bad = set(kwargs) - _SOLVER_KW
        if bad:
            raise NotImplementedError(f"solver keywords not on the HIP path: {sorted(bad)}")
        if not isinstance(model, Chain):
            model = Chain(*model) if isinstance(model, (list, tuple)) else model
        self.model = model
        self.tspan = (float(tspan[0]), float(tspan[1]))
        self.alg = alg
        self.args = args
        self.kwargs = dict(kwargs)
        self.precision = precision
        self.wgrad_precision = wgrad_precision
        self._handles = {}

    @property
    def sensealg(self):
        return self.kwargs.get("sensealg", InterpolatingAdjoint(autojacvec=ZygoteVJP()))

    def setup(self, rng, init="pytorch"):
        """``Lux.setup(rng, node)``: (flat parameters, empty state)."""
        return self.model.init_params(rng, init), {}

    def _handle(self, B):
        h = self._handles.get(B)
        if h is None:
            h = Handle(self.model.widths, self.model.acts, self.model.input_map,
                       self.precision, B)
            if self.wgrad_precision is not None:
                h.set_option("wgrad_precision", self.wgrad_precision)
            self._handles[B] = h
        return h

    def _opts(self):
        kw = self.kwargs
        t0, t1 = self.tspan
        interior, ss, se, sev = save_plan(t0, t1, kw.get("saveat"), kw.get("save_everystep"),
                                          kw.get("save_start"), kw.get("save_end"))
        adaptive = kw.get("adaptive", True)
        o = Handle.make_opts(t0, t1, abstol=kw.get("abstol", 1e-6), reltol=kw.get("reltol", 1e-3),
                             adaptive=adaptive, dt=kw.get("dt", 0.0), dtmax=kw.get("dtmax", 0.0),
                             dtmin=kw.get("dtmin", 0.0), maxiters=kw.get("maxiters", 0),
                             save_start=ss, save_end=se, save_everystep=sev)
        return o, interior

    def _sense_mode(self):
        sens = self.sensealg
        vjp_ok = sens.autojacvec is None or isinstance(sens.autojacvec, ZygoteVJP) \
            if isinstance(sens, (InterpolatingAdjoint, BacksolveAdjoint)) else False
        if isinstance(sens, InterpolatingAdjoint) and isinstance(sens.autojacvec, ZygoteVJP):
            return "interpolating_ckpt" if sens.checkpointing else "interpolating"
        if isinstance(sens, BacksolveAdjoint) and vjp_ok and not sens.noisemixing:
            return "backsolve" if sens.checkpointing else "backsolve_nockpt"
        raise NotImplementedError(f"sensealg {sens!r} is not on the HIP path")

    def __call__(self, x, p, st=None):
        mode = self._sense_mode()
        vec = np.ndim(x) == 1
        D = self.model.widths[0]
        B = 1 if vec else int(x.shape[0])
        if (x.shape[-1] if not vec else x.shape[0]) != D:
            raise ValueError(f"u0 must have D={D} entries per trajectory")
        h = self._handle(B)
        h.set_option("sensealg", "interpolating" if mode == "interpolating_ckpt" else mode)
        h.set_option("checkpointing", mode == "interpolating_ckpt")
        opts, interior = self._opts()
        try:
            import torch
            diff = isinstance(x, torch.Tensor) or isinstance(p, torch.Tensor)
        except ImportError:
            diff = False
        if diff:
            arr, t, stats = _NeuralODEFn.apply(x, p, h, opts, interior)
        else:
            h.set_params(p)
            arr, t, stats = h.forward(np.asarray(x, np.float32).reshape(B, D), opts, interior)
        if vec:
            arr = arr.reshape(arr.shape[0], D)
        return ODESolution(arr, t, stats["retcode"], stats), st


def _NeuralODEFn_factory():
    import torch

    class _Fn(torch.autograd.Function):
        """rrule of the solve: forward = HIP Tsit5 solve (tape kept in the
        handle), backward = HIP InterpolatingAdjoint."""

        @staticmethod
        def forward(ctx, x, p, h, opts, interior):
            xt = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x, np.float32))
            pt = p if isinstance(p, torch.Tensor) else torch.as_tensor(np.asarray(p, np.float32))
            h.set_params(pt.detach().cuda() if xt.is_cuda else pt.detach().cpu().numpy())
            u0 = xt.detach().reshape(h.B, h.D)
            if u0.is_cuda:
                out, t, stats = h.forward(u0, opts, interior)
            else:
                o, t, stats = h.forward(u0.cpu().numpy(), opts, interior)
                out = torch.from_numpy(np.ascontiguousarray(o))
            ctx.h = h
            ctx.gen = h.fwd_gen
            ctx.xshape = xt.shape
            ctx.pdtype, ctx.xdtype = pt.dtype, xt.dtype
            ctx.pdev, ctx.xdev = pt.device, xt.device
            ctx.mark_non_differentiable  # noqa: B018
            return out, t, stats

        @staticmethod
        def backward(ctx, g, _gt=None, _gs=None):
            h = ctx.h
            if h.fwd_gen != ctx.gen:
                raise RuntimeError("NeuralODE backward: the layer's handle ran another forward "
                                   "solve since this one (the handle keeps one tape); call "
                                   "backward before the next forward of the same layer and batch")
            if g.is_cuda:
                du0, dp, _ = h.backward(g.contiguous())
            else:
                a, b, _ = h.backward(g.detach().cpu().numpy())
                du0, dp = torch.from_numpy(a), torch.from_numpy(b)
            du0 = du0.to(ctx.xdev, ctx.xdtype).reshape(ctx.xshape)
            dp = dp.to(ctx.pdev, ctx.pdtype)
            return du0, dp, None, None, None

    return _Fn


class _Lazy:
    _fn = None

    def apply(self, *a):
        if self._fn is None:
            self._fn = _NeuralODEFn_factory()
        out, t, stats = self._fn.apply(*a)
        return out, t, stats


_NeuralODEFn = _Lazy()


# ---------------------------------------------------------------- wrappers
class AugmentedNDELayer:
    """``AugmentedNDELayer(nde, adim)`` (src/neural_de.jl:372-386): append
    ``adim`` zero rows to the state (``__augment``: ``vcat(x, zeros)`` for a
    vector, ``cat(x, zeros; dims = ndims(x) - 1)`` for a ``D x B`` matrix) and
    run the wrapped NeuralODE on the augmented state.  The zeros carry no
    gradient (``@ignore_derivatives``)."""

    def __init__(self, nde, adim):
        if not isinstance(nde, NeuralODE):
            raise NotImplementedError("AugmentedNDELayer wraps a NeuralODE on the HIP path")
        self.nde = nde
        self.adim = int(adim)

    def setup(self, rng, init="pytorch"):
        return self.nde.setup(rng, init)

    def augment(self, x):
        try:
            import torch
            if isinstance(x, torch.Tensor):
                z = torch.zeros(x.shape[:-1] + (self.adim,), dtype=x.dtype, device=x.device)
                return torch.cat([x, z], dim=-1)
        except ImportError:
            pass
        x = np.asarray(x)
        return np.concatenate([x, np.zeros(x.shape[:-1] + (self.adim,), x.dtype)], axis=-1)

    def __call__(self, x, p, st=None):
        return self.nde(self.augment(x), p, st)


class DimMover:
    """``DimMover(from, to)`` (src/neural_de.jl:397-412): ``cat(eachslice(x;
    dims = from)...; dims = to)`` with negative dims counted from the end in
    Julia's column-major order.  Arrays here are the C-order mirror of the
    Julia array (Julia dim ``k`` of an ``n``-dim array is numpy axis
    ``n - k``), e.g. the ``(Nt, B, D)`` saved states of a ``D x B x Nt``
    solution."""

    def __init__(self, from_=-2, to=-1):
        if from_ == 0 or to == 0:
            raise ValueError("dims are 1-based (negative counts from the end)")
        self.from_, self.to = from_, to

    def __call__(self, x, ps=None, st=None):
        n = x.ndim
        f = self.from_ if self.from_ > 0 else n + 1 + self.from_
        t = self.to if self.to > 0 else n + 1 + self.to
        ax_f, ax_t = n - f, n - t          # numpy axes of the Julia dims
        slices = [x.take(i, axis=ax_f) if hasattr(x, "take") and not _is_torch(x)
                  else x.select(ax_f, i) for i in range(x.shape[ax_f])]
        cat_ax = ax_t if ax_t < ax_f else ax_t - 1   # axis inside a slice
        if _is_torch(x):
            import torch
            return torch.cat(slices, dim=cat_ax), st
        return np.concatenate(slices, axis=cat_ax), st


def _is_torch(x):
    try:
        import torch
        return isinstance(x, torch.Tensor)
    except ImportError:
        return False
