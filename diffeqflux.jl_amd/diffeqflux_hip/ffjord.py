"""Python mirror of DiffEqFlux's ``FFJORD`` continuous normalizing flow over
the HIP library (``src/ffjord.jl:55-178``)::

    ffjord = FFJORD(model, tspan, input_dims, Tsit5(); kwargs...)
    ps, st = ffjord.setup(rng)
    (logpx, λ₁, λ₂), st = ffjord(x, ps, st)

``x`` is ``(B, D)`` (a Julia ``D x B`` matrix).  ``st`` carries
``regularize`` / ``monte_carlo`` as in the reference (``:60-65``):
``monte_carlo = false`` takes the exact trace (D unit-probe passes per
right-hand side on the GPU), ``regularize = true`` integrates ``‖f‖²`` and
``‖eᵀ ∂f/∂z‖`` as two more state rows and returns them as ``λ₁``, ``λ₂``
(without it ``λ₁ = λ₂ = Δlogp``, ``:166-167``).  The probe (Hutchinson, and
the regulariser's ``e``) is drawn once per call as a Rademacher vector per
sample (``st["rng"]``, a numpy Generator, or ``st["probe"]`` to pass one) and
held fixed for the whole solve; the reference redraws ``randn!`` on every
right-hand-side call (``:96``).  The regularisers carry no gradient on the
HIP path: a loss that differentiates ``λ₁`` / ``λ₂`` raises.  A model whose input width is ``D + 1`` gets ``[z; t]``
(time-concatenated network); a ``D``-wide input gets ``z`` as in the
reference (``:88``).

Differentiation: with ``x`` or ``ps`` a ``torch.Tensor`` the solve is a
``torch.autograd.Function`` whose backward is the HIP InterpolatingAdjoint of
the augmented solve (the reference's ``sensealg``, ``:143``) with the same
probe, including the second-order VJP of the Hutchinson term; ``logpx`` is
then assembled from ``z(T)`` and ``Δlogp`` in torch.

``FFJORDDistribution`` (``:216-246``) samples by the reversed-time solve and
evaluates densities by the forward solve, on the same handles.
"""
import numpy as np

from . import _lib
from .handle import Handle
from .neural_ode import Chain, Tsit5, _SOLVER_KW


class FFJORD:
    def __init__(self, model, tspan, input_dims, alg=None, *args, basedist=None, ad=None,
                 precision="fp32", **kwargs):
        if alg is None:
            alg = Tsit5()
        if not isinstance(alg, Tsit5):
            raise NotImplementedError(f"{alg!r}: only Tsit5() is on the HIP path")
        if basedist is not None:
            raise NotImplementedError("only the standard-normal base distribution is on the HIP path")
        bad = set(kwargs) - _SOLVER_KW
        if bad:
            raise NotImplementedError(f"solver keywords not on the HIP path: {sorted(bad)}")
        if not isinstance(model, Chain):
            raise NotImplementedError("FFJORD on the HIP path takes a Chain of Dense layers")
        D = model.widths[-1]
        if model.widths[0] not in (D, D + 1):
            raise ValueError("model input width must be D or D + 1 (time-concatenated)")
        if model.acts[-1] != "identity":
            raise NotImplementedError("the last Dense layer must be linear")
        self.model, self.tspan = model, (float(tspan[0]), float(tspan[1]))
        self.input_dims = tuple(np.atleast_1d(input_dims).tolist())
        self.D = D
        self.time_concat = model.widths[0] == D + 1
        self.kwargs = dict(kwargs)
        self.precision = precision
        self._handles = {}

    def setup(self, rng, init="pytorch"):
        return self.model.init_params(rng, init), {"regularize": False, "monte_carlo": True}

    def _handle(self, B, reg=False, mc=True):
        key = (B, bool(reg), bool(mc))
        h = self._handles.get(key)
        if h is None:
            imap = _lib.MAP_CNF_TIME if self.time_concat else _lib.MAP_CNF
            if reg:
                imap |= _lib.MAP_CNF_REGULARIZE
            if not mc:
                imap |= _lib.MAP_CNF_EXACT_TRACE
            h = Handle(self.model.widths, self.model.acts, imap, self.precision, B)
            h.cnf_reg = bool(reg)
            self._handles[key] = h
        return h

    def __call__(self, x, ps, st=None):
        st = dict(st or {"regularize": False, "monte_carlo": True})
        reg = bool(st.get("regularize", False))
        mc = bool(st.get("monte_carlo", True))
        x_in = x
        x = x.detach().cpu().numpy() if hasattr(x, "detach") else x
        x = np.asarray(x, np.float32)
        if x.ndim == 1:
            x = x[None, :]
            if hasattr(x_in, "detach"):
                x_in = x_in.reshape(1, -1)
        B = x.shape[0]
        probe = _probe(st, B, self.D)
        h = self._handle(B, reg, mc)
        try:
            import torch
            diff = isinstance(x_in, torch.Tensor) or isinstance(ps, torch.Tensor)
        except ImportError:
            diff = False
        if diff:
            z, dl, l1, l2, stats = _FFJORDFn.apply(x_in, ps, self, h, probe)
            D = self.D
            logpz = -0.5 * (z * z).sum(dim=1) - 0.5 * D * float(np.log(2 * np.pi))
            lp = (logpz - dl).reshape(1, B)
            dl = dl.reshape(1, B)
            st["stats"] = stats
            if reg:
                return (lp, l1.reshape(1, B), l2.reshape(1, B)), st
            return (lp, dl, dl), st
        h.set_params(np.asarray(ps, np.float32))
        lp, z, dl, l1, l2, stats = self._solve(h, x, probe)
        st["stats"] = stats
        dl = dl.reshape(1, B)
        if reg:
            return (lp.reshape(1, B), l1.reshape(1, B), l2.reshape(1, B)), st
        return (lp.reshape(1, B), dl, dl), st

    def _solve(self, h, x, probe, reverse=False):
        kw = self.kwargs
        t0, t1 = self.tspan[::-1] if reverse else self.tspan
        lp, z, dl, stats = h.cnf_forward(x, probe, t0, t1,
                                         abstol=kw.get("abstol", 1e-6), reltol=kw.get("reltol", 1e-3),
                                         adaptive=kw.get("adaptive", True), dt=kw.get("dt", 0.0),
                                         dtmax=kw.get("dtmax", 0.0), dtmin=kw.get("dtmin", 0.0),
                                         maxiters=kw.get("maxiters", 0))
        if stats["retcode"] != 1:
            raise RuntimeError(f"FFJORD solve failed: {_lib.RET_NAMES.get(stats['retcode'])}")
        l1 = l2 = None
        if getattr(h, "cnf_reg", False):
            l1, l2 = h.cnf_regularizers()
        return lp, z, dl, l1, l2, stats


def _probe(st, B, D):
    probe = st.get("probe")
    if probe is None:
        rng = st.get("rng") or np.random.default_rng()
        probe = rng.choice(np.array([-1.0, 1.0], np.float32), size=(B, D))
    return probe


class FFJORDDistribution:
    """``FFJORDDistribution(model, ps, st)`` (``src/ffjord.jl:216-246``): the
    flow as a distribution over the data space.

    * ``rand(n, rng)`` draws ``n`` standard-normal base samples and carries
      them through the flow backwards in time (``__backward_ffjord``,
      ``:183-214``: the same augmented solve over ``reverse(tspan)``), one
      batched HIP solve; returns ``(n, D)`` (a Julia ``D x n`` matrix), or
      ``(D,)`` when ``n`` is None.
    * ``_logpdf(x)`` is the reference's ``first(first(__forward_ffjord(...)))``
      (``:235-240``): the first sample's log-likelihood; ``logpdf(x)`` gives
      every sample's (a vector for a ``(B, D)`` batch, as Distributions'
      ``logpdf`` over the columns of a matrix), from one batched solve.
    ``st`` as for ``FFJORD`` (``regularize``, ``monte_carlo``, ``rng`` /
    ``probe``)."""

    def __init__(self, model, ps, st=None):
        self.model, self.ps = model, ps
        self.st = dict(st or {"regularize": False, "monte_carlo": True})

    def __len__(self):
        return int(np.prod(self.model.input_dims))

    @property
    def dtype(self):
        return np.asarray(self.ps).dtype

    def logpdf(self, x):
        x = np.asarray(x.detach().cpu().numpy() if hasattr(x, "detach") else x, np.float32)
        (lp, _, _), _ = self.model(x, np.asarray(self.ps, np.float32), self.st)
        lp = np.asarray(lp).reshape(-1)
        return float(lp[0]) if x.ndim == 1 else lp

    def _logpdf(self, x):
        return float(np.asarray(self.logpdf(x)).reshape(-1)[0])

    def rand(self, n=None, rng=None):
        ffj = self.model
        B = 1 if n is None else int(n)
        rng = rng if rng is not None else (self.st.get("rng") or np.random.default_rng())
        x = rng.standard_normal((B, ffj.D)).astype(np.float32)
        reg = bool(self.st.get("regularize", False))
        mc = bool(self.st.get("monte_carlo", True))
        probe = _probe(self.st, B, ffj.D) if (mc or reg) else None
        h = ffj._handle(B, reg, mc)
        h.set_params(np.asarray(self.ps, np.float32))
        _, z, _, _, _, stats = ffj._solve(h, x, probe, reverse=True)
        self.st["stats"] = stats
        return z[0] if n is None else z


def _FFJORDFn_factory():
    import torch

    class _Fn(torch.autograd.Function):
        """rrule of the FFJORD solve: forward = HIP Tsit5 solve of [z; Δlogp]
        (tape kept in the handle), backward = HIP InterpolatingAdjoint with
        the second-order VJP of the Hutchinson term."""

        @staticmethod
        def forward(ctx, x, p, ffj, h, probe):
            xt = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x, np.float32))
            pt = p if isinstance(p, torch.Tensor) else torch.as_tensor(np.asarray(p, np.float32))
            h.set_params(pt.detach().cpu().numpy().astype(np.float32))
            xn = xt.detach().cpu().numpy().astype(np.float32)
            _, z, dl, l1, l2, stats = ffj._solve(h, xn, probe)
            if l1 is None:
                l1 = l2 = np.zeros_like(dl)
            ctx.h, ctx.gen = h, h.fwd_gen
            ctx.xshape, ctx.xdtype, ctx.xdev = xt.shape, xt.dtype, xt.device
            ctx.pdtype, ctx.pdev = pt.dtype, pt.device
            dev = xt.device
            return (torch.from_numpy(z).to(dev), torch.from_numpy(dl).to(dev),
                    torch.from_numpy(l1).to(dev), torch.from_numpy(l2).to(dev), stats)

        @staticmethod
        def backward(ctx, gz, gdl, gl1=None, gl2=None, _gs=None):
            h = ctx.h
            if h.fwd_gen != ctx.gen:
                raise RuntimeError("FFJORD backward: the layer's handle ran another solve since "
                                   "this one (the handle keeps one tape); call backward before "
                                   "the next forward of the same layer and batch")
            for gl in (gl1, gl2):
                if gl is not None and bool((gl != 0).any()):
                    raise NotImplementedError("FFJORD: gradients through the regularisers "
                                              "(λ₁, λ₂) are not on the HIP path")
            gz = torch.zeros(ctx.xshape) if gz is None else gz
            gdl = torch.zeros(ctx.xshape[0]) if gdl is None else gdl
            dx, dp, _ = h.cnf_backward(gz.detach().cpu().numpy(), gdl.detach().cpu().numpy())
            dx = torch.from_numpy(dx).to(ctx.xdev, ctx.xdtype).reshape(ctx.xshape)
            dp = torch.from_numpy(dp).to(ctx.pdev, ctx.pdtype)
            return dx, dp, None, None, None

    return _Fn


class _LazyFn:
    _fn = None

    def apply(self, *a):
        if self._fn is None:
            self._fn = _FFJORDFn_factory()
        return self._fn.apply(*a)


_FFJORDFn = _LazyFn()
