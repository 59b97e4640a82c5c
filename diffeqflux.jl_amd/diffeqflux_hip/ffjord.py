"""Python mirror of DiffEqFlux's ``FFJORD`` continuous normalizing flow over
the HIP library (``src/ffjord.jl:55-178``)::

    ffjord = FFJORD(model, tspan, input_dims, Tsit5(); kwargs...)
    ps, st = ffjord.setup(rng)
    (logpx, λ₁, λ₂), st = ffjord(x, ps, st)

``x`` is ``(B, D)`` (a Julia ``D x B`` matrix).  ``st`` carries
``regularize`` / ``monte_carlo`` as in the reference (``:60-65``); the HIP
path runs ``monte_carlo = true`` without regularisation.  The Hutchinson
probe is drawn once per call as a Rademacher vector per sample (``st["rng"]``,
a numpy Generator, or ``st["probe"]`` to pass one) and held fixed for the
whole solve; the reference redraws ``randn!`` on every right-hand-side call
(``:96``).  A model whose input width is ``D + 1`` gets ``[z; t]``
(time-concatenated network); a ``D``-wide input gets ``z`` as in the
reference (``:88``).
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

    def _handle(self, B):
        h = self._handles.get(B)
        if h is None:
            imap = _lib.MAP_CNF_TIME if self.time_concat else _lib.MAP_CNF
            h = Handle(self.model.widths, self.model.acts, imap, self.precision, B)
            self._handles[B] = h
        return h

    def __call__(self, x, ps, st=None):
        st = dict(st or {"regularize": False, "monte_carlo": True})
        if st.get("regularize", False):
            raise NotImplementedError("regularize = true is not on the HIP path")
        if not st.get("monte_carlo", True):
            raise NotImplementedError("the exact-trace (monte_carlo = false) form is not on the HIP path")
        x = np.asarray(x, np.float32)
        if x.ndim == 1:
            x = x[None, :]
        B = x.shape[0]
        probe = st.get("probe")
        if probe is None:
            rng = st.get("rng") or np.random.default_rng()
            probe = rng.choice(np.array([-1.0, 1.0], np.float32), size=(B, self.D))
        h = self._handle(B)
        h.set_params(np.asarray(ps, np.float32))
        kw = self.kwargs
        lp, z, dl, stats = h.cnf_forward(x, probe, self.tspan[0], self.tspan[1],
                                         abstol=kw.get("abstol", 1e-6), reltol=kw.get("reltol", 1e-3),
                                         adaptive=kw.get("adaptive", True), dt=kw.get("dt", 0.0),
                                         dtmax=kw.get("dtmax", 0.0), dtmin=kw.get("dtmin", 0.0),
                                         maxiters=kw.get("maxiters", 0))
        if stats["retcode"] != 1:
            raise RuntimeError(f"FFJORD solve failed: {_lib.RET_NAMES.get(stats['retcode'])}")
        st["stats"] = stats
        dl = dl.reshape(1, B)
        return (lp.reshape(1, B), dl, dl), st
