"""Python owner of one ``node_handle_t`` (the C ABI of ``include/node_hip.h``).

Arrays follow the ABI's Julia layout: a ``D x B`` state is a C-order
``(B, D)`` array, saved states are ``(Nt, B, D)`` (Julia ``D x B x Nt``), and
parameters are the flat Lux/ComponentArray FP32 vector.  numpy arrays go
through the host entry points; torch CUDA tensors go through the ``*_device``
entry points without host copies.
"""
import ctypes

import numpy as np

from . import _lib


def _is_torch_cuda(x):
    try:
        import torch
        return isinstance(x, torch.Tensor) and x.is_cuda
    except ImportError:
        return False


def _f32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32))


def _ptr(a):
    if _is_torch_cuda(a):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


class Handle:
    """A model (Dense chain) + batch size bound to device buffers."""

    def __init__(self, widths, acts, input_map=0, precision="fp32", batch=1):
        self.lib = _lib.load()
        self.desc = _lib.make_desc(widths, acts, input_map, precision, batch)
        self.widths = list(widths)
        self.D = int(widths[0])
        self.B = int(batch)
        self.precision = precision
        self.P = int(self.lib.node_num_params(ctypes.byref(self.desc)))
        h = ctypes.c_void_p()
        rc = self.lib.node_create(ctypes.byref(self.desc), ctypes.byref(h))
        if rc != 0:
            raise _lib.NodeHipError(f"node_create failed ({_lib.ERR_NAMES.get(rc, rc)})")
        self.h = h
        self.last_stats = None
        self.saved_times = None
        # count of forward solves (batched / per-trajectory): a backward pass
        # adjoints the handle's most recent forward of its kind only
        self.fwd_gen = 0
        self.rg_gen = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.node_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        _lib.check(self.lib, self.h, rc, what)

    # ------------------------------------------------------------ params
    def set_params(self, p):
        if _is_torch_cuda(p):
            p = p.detach().to(dtype=__import__("torch").float32).contiguous()
            __import__("torch").cuda.current_stream(p.device).synchronize()
            self._check(self.lib.node_set_params_device(self.h, _ptr(p), p.numel()),
                        "node_set_params_device")
        else:
            p = _f32(p).reshape(-1)
            self._check(self.lib.node_set_params(self.h, _ptr(p), p.size), "node_set_params")

    # ------------------------------------------------------------ solves
    @staticmethod
    def make_opts(t0, t1, abstol=1e-6, reltol=1e-3, adaptive=True, dt=0.0,
                  dtmax=0.0, dtmin=0.0, maxiters=0, save_start=True, save_end=True,
                  save_everystep=False):
        o = _lib.SolveOpts()
        o.t0, o.t1 = float(t0), float(t1)
        o.abstol, o.reltol = float(abstol), float(reltol)
        o.adaptive = 1 if adaptive else 0
        o.dt, o.dtmax, o.dtmin = float(dt or 0.0), float(dtmax or 0.0), float(dtmin or 0.0)
        o.maxiters = int(maxiters or 0)
        o.save_start, o.save_end = int(bool(save_start)), int(bool(save_end))
        o.save_everystep = int(bool(save_everystep))
        return o

    def forward(self, u0, opts, saveat_interior=()):
        """Returns (states (Nt, B, D), times (Nt,), stats dict)."""
        self.fwd_gen += 1
        sa = np.ascontiguousarray(np.asarray(saveat_interior, dtype=np.float64).reshape(-1))
        sa_p = sa.ctypes.data_as(ctypes.c_void_p) if sa.size else None
        st = _lib.Stats()
        n_fixed = sa.size + int(opts.save_start) + int(opts.save_end)
        if _is_torch_cuda(u0):
            import torch
            u0 = u0.detach().to(torch.float32).contiguous()
            cap = 0 if opts.save_everystep else n_fixed
            out = torch.empty((max(cap, 1), self.B, self.D), device=u0.device, dtype=torch.float32)
            torch.cuda.current_stream(u0.device).synchronize()  # library runs on its own stream
            rc = self.lib.node_forward_device(self.h, ctypes.byref(opts), _ptr(u0), sa_p,
                                              sa.size, _ptr(out) if cap else None, cap,
                                              ctypes.byref(st))
            self._check(rc, "node_forward_device")
            if opts.save_everystep:
                host = np.empty((st.nsaved, self.B, self.D), dtype=np.float32)
                self._check(self.lib.node_copy_saved(self.h, _ptr(host), st.nsaved),
                            "node_copy_saved")
                out = torch.from_numpy(host).to(u0.device)
            else:
                out = out[:st.nsaved]
        else:
            u0 = _f32(u0).reshape(self.B, self.D)
            cap = 0 if opts.save_everystep else n_fixed
            out = np.empty((max(cap, 1), self.B, self.D), dtype=np.float32)
            rc = self.lib.node_forward(self.h, ctypes.byref(opts), _ptr(u0), sa_p, sa.size,
                                       _ptr(out) if cap else None, cap, ctypes.byref(st))
            self._check(rc, "node_forward")
            if opts.save_everystep:
                out = np.empty((st.nsaved, self.B, self.D), dtype=np.float32)
                self._check(self.lib.node_copy_saved(self.h, _ptr(out), st.nsaved),
                            "node_copy_saved")
            else:
                out = out[:st.nsaved]
        ts = np.empty(max(st.nsaved, 1), dtype=np.float64)
        self._check(self.lib.node_saved_times(self.h, _ptr(ts), ts.size), "node_saved_times")
        self.saved_times = ts[:st.nsaved]
        self.last_stats = st.as_dict()
        return out, self.saved_times, self.last_stats

    def backward(self, dLdout):
        """dLdout: (Nt, B, D) -> (dL/du0 (B, D) fp32, dL/dp (P,) fp64, stats)."""
        st = _lib.Stats()
        if _is_torch_cuda(dLdout):
            import torch
            g = dLdout.detach().to(torch.float32).contiguous()
            du0 = torch.empty((self.B, self.D), device=g.device, dtype=torch.float32)
            dp = torch.empty((self.P,), device=g.device, dtype=torch.float64)
            torch.cuda.current_stream(g.device).synchronize()
            rc = self.lib.node_backward_device(self.h, _ptr(g), _ptr(du0), _ptr(dp),
                                               ctypes.byref(st))
            self._check(rc, "node_backward_device")
        else:
            g = _f32(dLdout)
            du0 = np.empty((self.B, self.D), dtype=np.float32)
            dp = np.empty((self.P,), dtype=np.float64)
            rc = self.lib.node_backward(self.h, _ptr(g), _ptr(du0), _ptr(dp), ctypes.byref(st))
            self._check(rc, "node_backward")
        return du0, dp, st.as_dict()

    def cnf_forward(self, x, probe, t0, t1, abstol=1e-6, reltol=1e-3, adaptive=True, dt=0.0,
                    dtmax=0.0, dtmin=0.0, maxiters=0):
        """FFJORD forward on a CNF handle.  x, probe: (B, D).  Returns
        (logpx (B,), z(T) (B, D), delta_logp (B,), stats dict)."""
        D = self.widths[-1]
        x = _f32(x).reshape(self.B, D)
        e = _f32(probe).reshape(self.B, D) if probe is not None else None
        z = np.empty((self.B, D), np.float32)
        dl = np.empty(self.B, np.float32)
        lp = np.empty(self.B, np.float32)
        o = self.make_opts(t0, t1, abstol=abstol, reltol=reltol, adaptive=adaptive, dt=dt,
                           dtmax=dtmax, dtmin=dtmin, maxiters=maxiters)
        st = _lib.Stats()
        self.fwd_gen += 1
        self._check(self.lib.node_cnf_forward(self.h, ctypes.byref(o), _ptr(x),
                                              _ptr(e) if e is not None else None, _ptr(z),
                                              _ptr(dl), _ptr(lp), ctypes.byref(st)),
                    "node_cnf_forward")
        return lp, z, dl, st.as_dict()

    def cnf_regularizers(self):
        """(r1(T), r2(T)) per sample of the last ``cnf_forward`` on a
        regularised CNF handle (the layer's λ₁, λ₂)."""
        l1 = np.empty(self.B, np.float32)
        l2 = np.empty(self.B, np.float32)
        self._check(self.lib.node_cnf_regularizers(self.h, _ptr(l1), _ptr(l2)),
                    "node_cnf_regularizers")
        return l1, l2

    def cnf_backward(self, dL_dz, dL_ddlogp):
        """InterpolatingAdjoint of the last ``cnf_forward`` (same probe).
        dL_dz: (B, D), dL_ddlogp: (B,).  Returns (dL/dx (B, D) fp32,
        dL/dp (P,) fp64, stats dict)."""
        D = self.widths[-1]
        gz = _f32(dL_dz).reshape(self.B, D)
        gd = _f32(dL_ddlogp).reshape(self.B)
        dx = np.empty((self.B, D), np.float32)
        dp = np.empty((self.P,), np.float64)
        st = _lib.Stats()
        self._check(self.lib.node_cnf_backward(self.h, _ptr(gz), _ptr(gd), _ptr(dx), _ptr(dp),
                                               ctypes.byref(st)), "node_cnf_backward")
        return dx, dp, st.as_dict()

    def ragged_forward(self, u0, t1, save_times, t0=0.0, abstol=1e-6, reltol=1e-3, adaptive=True,
                       dt=0.0, dtmax=0.0, dtmin=0.0, maxiters=0, flat_out=False):
        """Per-trajectory solves.  u0: (n, D); t1: (n,); save_times: list of n
        increasing sequences, or a pair (times, offsets) of the concatenated
        save times and the n + 1 offsets of each trajectory's run (the C
        entry point's own layout).  Returns (list of (len_i, D) saved states
        -- or, with flat_out, the (offsets[n], D) array they are views of --,
        retcodes (n,), per-trajectory stats (n, 4) [nf, naccept, nreject,
        records], totals dict)."""
        self.rg_gen += 1
        u0 = _f32(u0).reshape(-1, self.D)
        n = u0.shape[0]
        t1 = np.ascontiguousarray(np.asarray(t1, dtype=np.float64).reshape(n))
        if isinstance(save_times, tuple):
            st, off = save_times
            off = np.ascontiguousarray(np.asarray(off, np.int64).reshape(n + 1))
            st = np.ascontiguousarray(np.asarray(st, np.float64).reshape(-1)
                                      if off[-1] else np.zeros(1))
        else:
            lens = [len(s) for s in save_times]
            off = np.zeros(n + 1, np.int64)
            off[1:] = np.cumsum(lens)
            st = np.ascontiguousarray(np.concatenate([np.asarray(s, np.float64) for s in save_times])
                                      if off[-1] else np.zeros(1))
        out = np.zeros((max(int(off[-1]), 1), self.D), np.float32)
        ret = np.zeros(n, np.int32)
        ts = np.zeros((n, 4), np.int32)
        o = self.make_opts(t0, t0, abstol=abstol, reltol=reltol, adaptive=adaptive, dt=dt,
                           dtmax=dtmax, dtmin=dtmin, maxiters=maxiters)
        rs = _lib.RaggedStats()
        self._check(self.lib.node_ragged_forward(self.h, ctypes.byref(o), n, _ptr(t1), _ptr(u0),
                                                 _ptr(off), _ptr(st), _ptr(out), _ptr(ret),
                                                 _ptr(ts), ctypes.byref(rs)),
                    "node_ragged_forward")
        self.ragged_offsets = off
        if flat_out:
            return out[:off[-1]], ret, ts, rs.as_dict()
        return [out[off[i]:off[i + 1]] for i in range(n)], ret, ts, rs.as_dict()

    def ragged_backward(self, dLdout, per_trajectory=False):
        """Adjoint of the last ragged_forward.  dLdout: list of (len_i, D)
        cotangents (or one stacked (total, D) array).  Returns (dL/du0 (n, D),
        dL/dp (P,) FP64, per-trajectory dL/dp (n, P) FP32 or None, retcodes,
        per-trajectory stats (n, 4), totals)."""
        if isinstance(dLdout, (list, tuple)):
            dLdout = np.concatenate([np.asarray(x, np.float32).reshape(-1, self.D) for x in dLdout])
        g = _f32(dLdout).reshape(-1, self.D)
        if g.shape[0] == 0:
            g = np.zeros((1, self.D), np.float32)
        n = len(self.ragged_offsets) - 1
        du0 = np.zeros((n, self.D), np.float32)
        dp = np.zeros(self.P, np.float64)
        rows = np.zeros((n, self.P), np.float32) if per_trajectory else None
        ret = np.zeros(n, np.int32)
        ts = np.zeros((n, 4), np.int32)
        rs = _lib.RaggedStats()
        self._check(self.lib.node_ragged_backward(self.h, _ptr(g), _ptr(du0), _ptr(dp),
                                                  _ptr(rows) if rows is not None else None,
                                                  _ptr(ret), _ptr(ts), ctypes.byref(rs)),
                    "node_ragged_backward")
        return du0, dp, rows, ret, ts, rs.as_dict()

    def rhs(self, u):
        u = _f32(u).reshape(self.B, self.D)
        du = np.empty_like(u)
        self._check(self.lib.node_rhs(self.h, _ptr(u), _ptr(du)), "node_rhs")
        return du

    def vjp(self, u, lam):
        u = _f32(u).reshape(self.B, self.D)
        lam = _f32(lam).reshape(self.B, self.D)
        du = np.empty_like(u)
        dp = np.empty((self.P,), dtype=np.float64)
        self._check(self.lib.node_vjp(self.h, _ptr(u), _ptr(lam), _ptr(du), _ptr(dp)),
                    "node_vjp")
        return du, dp

    def set_option(self, key, value):
        """key: 'graphs' (bool), 'graph_attempts' (1..16), 'autotune' (bool),
        'wgrad_precision' ('fp32' / 'bf16' / 'bf16x3') or 'sensealg'
        ('interpolating' / 'backsolve' / 'backsolve_nockpt'), 'checkpointing'
        (bool: InterpolatingAdjoint(checkpointing = true), the forward keeps
        checkpoint states only and the adjoint re-solves each interval),
        'tape_budget' (bytes of forward tape, 0 = unlimited) or
        'ragged_vjp_precision' ('fp32' / 'bf16': the per-trajectory adjoint's
        VJP products on BF16 operands)."""
        k = {"graphs": _lib.OPT_GRAPHS, "graph_attempts": _lib.OPT_GRAPH_ATTEMPTS,
             "autotune": _lib.OPT_AUTOTUNE, "wgrad_precision": _lib.OPT_WGRAD_PRECISION,
             "sensealg": _lib.OPT_SENSEALG, "checkpointing": _lib.OPT_CHECKPOINTING,
             "tape_budget": _lib.OPT_TAPE_BUDGET,
             "ragged_vjp_precision": _lib.OPT_RAGGED_VJP_PRECISION}[key]
        if key in ("wgrad_precision", "ragged_vjp_precision") and isinstance(value, str):
            value = _lib.PREC[value]
        if key == "sensealg" and isinstance(value, str):
            value = _lib.SENSE[value]
        if key == "sensealg" and getattr(self, "_sense", None) == int(value):
            return  # unchanged: keep the captured adjoint graph
        self._check(self.lib.node_set_option(self.h, k, int(value)), "node_set_option")
        if key == "sensealg":
            self._sense = int(value)

    def plans(self):
        """{(role name, layer): (tile rows, tile cols, ksplit)} currently in use."""
        out = {}
        L = len(self.widths) - 1
        for role, name in self.ROLES.items():
            for layer in range(L):
                if role in (0, 2) and layer == L - 1:
                    continue
                if role == 1 and layer != L - 1:
                    continue
                if role == 3 and layer == 0:
                    continue
                if role == 5 and layer != 0:
                    continue
                r, c, k = _lib.c_i32(), _lib.c_i32(), _lib.c_i32()
                if self.lib.node_get_plan(self.h, role, layer, ctypes.byref(r), ctypes.byref(c),
                                          ctypes.byref(k)) == 0:
                    out[(name, layer + 1)] = (r.value, c.value, k.value)
        return out

    # ------------------------------------------------------------ profiling
    ROLES = {0: "fwd_hidden", 1: "fwd_stage", 2: "adj_hidden", 3: "bwd_data",
             4: "wgrad", 5: "bwd_last"}

    def prof_enable(self, keys=None):
        """Record HIP-event timings around GEMM launches.  ``keys``: iterable
        of (role, layer) or None for all; ``[]`` disables."""
        if keys is None:
            mask = (1 << 64) - 1
        else:
            mask = 0
            for role, layer in keys:
                mask |= 1 << (role * 8 + layer)
        self._check(self.lib.node_prof_enable(self.h, ctypes.c_uint64(mask)), "node_prof_enable")

    def prof_collect(self, cap=1 << 20):
        """-> list of ((role, layer), ms) per recorded launch."""
        keys = np.empty(cap, np.int32)
        ms = np.empty(cap, np.float32)
        n = ctypes.c_int64(0)
        self._check(self.lib.node_prof_collect(self.h, _ptr(keys), _ptr(ms), cap,
                                               ctypes.byref(n)), "node_prof_collect")
        k = min(n.value, cap)
        return [((int(keys[i]) // 8, int(keys[i]) % 8), float(ms[i])) for i in range(k)]

