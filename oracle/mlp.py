"""Lux ``Chain`` of ``Dense`` layers: forward, VJP and the flat parameter layout.

Restates the RHS that ``NeuralODE`` builds at ``src/neural_de.jl:48-50``
(``dudt(u, p, t) = model(u, p)``) for the models the reference uses:
``Chain(x -> x.^3, Dense(2, 50, tanh), Dense(50, 2))``
(``docs/src/examples/neural_ode.md:32``) and
``Chain(Dense(2 => 50, tanh), Dense(50 => 2))``
(``test/BasicNeuralDE/neural_de_tests.jl:17``).  The numerics are Lux/LuxLib
(third-party, absent from the reference tree): ``Dense`` computes
``act.(W * x .+ b)`` with ``W`` an ``out x in`` column-major matrix, and the
pullback is Zygote's (``ZygoteVJP``, ``src/neural_de.jl:57``).

Array convention: a Julia ``D x B`` column-major matrix (one column per
trajectory) has the same bytes as a C-order numpy array of shape ``(B, D)``;
this module works on ``(B, D)`` arrays.  Likewise Julia's ``W`` (``out x in``,
column-major) is stored as a C-order ``(in, out)`` array ``Wt``.

Flat parameters follow ``ComponentArray(ps)`` order
(``test/BasicNeuralDE/neural_de_tests.jl:45``): depth first over layers,
``weight`` (column-major) then ``bias``; the parameter-free ``x -> x.^3``
layer contributes nothing.

Pinned by (tests/test_oracle_mlp.py): direct evaluation, central finite
differences of the forward, and an independent torch-FP64 autograd.
"""
import numpy as np

ACT_IDENTITY, ACT_TANH, ACT_SOFTPLUS, ACT_RELU, ACT_SIGMOID = 0, 1, 2, 3, 4
MAP_IDENTITY, MAP_CUBE = 0, 1

ACT_NAMES = {"identity": ACT_IDENTITY, "tanh": ACT_TANH,
             "softplus": ACT_SOFTPLUS, "relu": ACT_RELU,
             "sigmoid": ACT_SIGMOID}


def act_fwd(code, z):
    if code == ACT_IDENTITY:
        return z
    if code == ACT_TANH:
        return np.tanh(z)
    if code == ACT_SOFTPLUS:  # NNlib: log1p(exp(-|x|)) + relu(x)
        return np.log1p(np.exp(-np.abs(z))) + np.maximum(z, 0)
    if code == ACT_RELU:
        return np.maximum(z, 0)
    if code == ACT_SIGMOID:
        return 1.0 / (1.0 + np.exp(-z))
    raise ValueError(f"unknown activation {code}")


def act_deriv(code, z, y):
    """d act / dz given pre-activation z and output y."""
    if code == ACT_IDENTITY:
        return np.ones_like(z)
    if code == ACT_TANH:
        return 1 - y * y
    if code == ACT_SOFTPLUS:
        return 1.0 / (1.0 + np.exp(-z))
    if code == ACT_RELU:
        return (z > 0).astype(z.dtype)
    if code == ACT_SIGMOID:
        return y * (1 - y)
    raise ValueError(f"unknown activation {code}")


class MLP:
    """A Dense chain ``widths[0] -> ... -> widths[-1]`` with per-layer
    activations and an optional parameter-free input map (``x.^3``)."""

    def __init__(self, widths, acts, input_map=MAP_IDENTITY):
        assert len(acts) == len(widths) - 1
        self.widths = [int(w) for w in widths]
        self.acts = [ACT_NAMES[a] if isinstance(a, str) else int(a)
                     for a in acts]
        self.input_map = int(input_map)
        offs, o = [], 0
        for i, j in zip(self.widths[:-1], self.widths[1:]):
            offs.append((o, o + i * j, o + i * j + j))
            o += i * j + j
        self.offsets = offs          # (w_start, b_start, end) per layer
        self.nparams = o

    @property
    def D(self):
        return self.widths[0]

    def unflatten(self, p):
        """Flat Lux vector -> list of (Wt (in,out), b (out,))."""
        out = []
        for (i, j), (w0, b0, e) in zip(zip(self.widths[:-1], self.widths[1:]),
                                       self.offsets):
            out.append((p[w0:b0].reshape(i, j), p[b0:e]))
        return out

    def glorot_init(self, rng, dtype=np.float64):
        """Glorot-uniform weights U(+-sqrt(6/(in+out))), zero biases."""
        p = np.zeros(self.nparams, dtype=dtype)
        for (i, j), (w0, b0, e) in zip(zip(self.widths[:-1], self.widths[1:]),
                                       self.offsets):
            lim = np.sqrt(6.0 / (i + j))
            p[w0:b0] = rng.uniform(-lim, lim, size=i * j).astype(dtype)
        return p

    def _map(self, u):
        return u * u * u if self.input_map == MAP_CUBE else u

    def forward(self, u, p):
        """u: (B, D) -> f(u): (B, D)."""
        h = self._map(u)
        for (Wt, b), act in zip(self.unflatten(p), self.acts):
            h = act_fwd(act, h @ Wt + b)
        return h

    def vjp(self, u, p, lam):
        """Returns (lam^T df/du, lam^T df/dp) for lam: (B, D_out).

        The parameter gradient is summed over the batch (``dL/dp`` of the whole
        ``D x B`` state, as Zygote's pullback of the batched Chain gives).
        """
        layers = self.unflatten(p)
        x0 = self._map(u)
        hs, zs = [x0], []
        h = x0
        for (Wt, b), act in zip(layers, self.acts):
            z = h @ Wt + b
            h = act_fwd(act, z)
            zs.append(z)
            hs.append(h)
        gp = np.zeros(self.nparams, dtype=np.result_type(u, p, lam))
        g = lam
        for li in range(len(layers) - 1, -1, -1):
            Wt, _ = layers[li]
            delta = g * act_deriv(self.acts[li], zs[li], hs[li + 1])
            w0, b0, e = self.offsets[li]
            gp[w0:b0] = (hs[li].T @ delta).ravel()
            gp[b0:e] = delta.sum(axis=0)
            g = delta @ Wt.T
        if self.input_map == MAP_CUBE:
            g = g * (3 * u * u)
        return g, gp
