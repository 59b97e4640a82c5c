"""ctypes binding of the C ABI in ``include/node_hip.h``.

The product path: every numeric call goes through ``libnode_hip.so``.  There
is no CPU fallback -- ``load()`` raises when the library is missing.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnode_hip.so")

MAX_LAYERS = 8
ACT = {"identity": 0, "tanh": 1, "softplus": 2, "relu": 3, "sigmoid": 4}
MAP_IDENTITY, MAP_CUBE, MAP_CNF, MAP_CNF_TIME = 0, 1, 2, 3
MAP_CNF_REGULARIZE, MAP_CNF_EXACT_TRACE = 0x10, 0x20     # flags on MAP_CNF / MAP_CNF_TIME
OPT_GRAPHS, OPT_GRAPH_ATTEMPTS, OPT_AUTOTUNE, OPT_WGRAD_PRECISION, OPT_SENSEALG = 1, 2, 3, 4, 5
OPT_CHECKPOINTING, OPT_TAPE_BUDGET, OPT_RAGGED_VJP_PRECISION = 6, 7, 8
SENSE = {"interpolating": 0, "backsolve": 1, "backsolve_nockpt": 2}
PREC = {"fp32": 0, "bf16": 1, "bf16x3": 2}
RET_NAMES = {0: "Default", 1: "Success", 2: "Terminated", 3: "DtNaN",
             4: "MaxIters", 5: "DtLessThanMin", 6: "Unstable"}
ERR_NAMES = {-1: "NODE_ERR_INVALID", -2: "NODE_ERR_HIP", -3: "NODE_ERR_STATE",
             -4: "NODE_ERR_CAPACITY", -5: "NODE_ERR_UNSUPPORTED"}

c_i32, c_i64, c_f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
P_f32 = ctypes.POINTER(ctypes.c_float)
P_f64 = ctypes.POINTER(ctypes.c_double)


class NodeDesc(ctypes.Structure):
    _fields_ = [("n_layers", c_i32), ("widths", c_i32 * (MAX_LAYERS + 1)),
                ("acts", c_i32 * MAX_LAYERS), ("input_map", c_i32),
                ("precision", c_i32), ("batch", c_i64)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("t0", c_f64), ("t1", c_f64), ("abstol", c_f64), ("reltol", c_f64),
                ("adaptive", c_i32), ("save_start", c_i32), ("save_end", c_i32),
                ("save_everystep", c_i32), ("dt", c_f64), ("dtmax", c_f64),
                ("dtmin", c_f64), ("maxiters", c_i64)]


class Stats(ctypes.Structure):
    _fields_ = [("nf", c_i64), ("naccept", c_i64), ("nreject", c_i64),
                ("nsaved", c_i64), ("retcode", c_i32), ("reserved", c_i32),
                ("t_final", c_f64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class RaggedStats(ctypes.Structure):
    _fields_ = [("nf", c_i64), ("naccept", c_i64), ("nreject", c_i64), ("records", c_i64),
                ("n_success", c_i64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# (name, restype, argtypes) of every entry point declared in node_hip.h
SIGNATURES = [
    ("node_abi_version", c_i32, []),
    ("node_num_params", c_i64, [ctypes.POINTER(NodeDesc)]),
    ("node_create", c_i32, [ctypes.POINTER(NodeDesc), ctypes.POINTER(ctypes.c_void_p)]),
    ("node_destroy", c_i32, [ctypes.c_void_p]),
    ("node_last_error", ctypes.c_char_p, [ctypes.c_void_p]),
    ("node_set_params", c_i32, [ctypes.c_void_p, ctypes.c_void_p, c_i64]),
    ("node_set_params_device", c_i32, [ctypes.c_void_p, ctypes.c_void_p, c_i64]),
    ("node_forward", c_i32, [ctypes.c_void_p, ctypes.POINTER(SolveOpts), ctypes.c_void_p,
                             ctypes.c_void_p, c_i64, ctypes.c_void_p, c_i64,
                             ctypes.POINTER(Stats)]),
    ("node_forward_device", c_i32, [ctypes.c_void_p, ctypes.POINTER(SolveOpts),
                                    ctypes.c_void_p, ctypes.c_void_p, c_i64,
                                    ctypes.c_void_p, c_i64, ctypes.POINTER(Stats)]),
    ("node_cnf_forward", c_i32, [ctypes.c_void_p, ctypes.POINTER(SolveOpts), ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.POINTER(Stats)]),
    ("node_cnf_backward", c_i32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Stats)]),
    ("node_saved_times", c_i32, [ctypes.c_void_p, ctypes.c_void_p, c_i64]),
    ("node_copy_saved", c_i32, [ctypes.c_void_p, ctypes.c_void_p, c_i64]),
    ("node_backward", c_i32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_void_p, ctypes.POINTER(Stats)]),
    ("node_backward_device", c_i32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.POINTER(Stats)]),
    ("node_ragged_forward", c_i32, [ctypes.c_void_p, ctypes.POINTER(SolveOpts), c_i64,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.POINTER(RaggedStats)]),
    ("node_ragged_backward", c_i32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.POINTER(RaggedStats)]),
    ("node_cnf_regularizers", c_i32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("node_rhs", c_i32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("node_rhs_device_noalloc", c_i32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        c_i64]),
    ("node_vjp", c_i32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                         ctypes.c_void_p, ctypes.c_void_p]),
    ("node_set_option", c_i32, [ctypes.c_void_p, c_i32, c_i64]),
    ("node_get_plan", c_i32, [ctypes.c_void_p, c_i32, c_i32, ctypes.POINTER(c_i32),
                              ctypes.POINTER(c_i32), ctypes.POINTER(c_i32)]),
    ("node_prof_enable", c_i32, [ctypes.c_void_p, ctypes.c_uint64]),
    ("node_prof_collect", c_i32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, c_i64,
                                  ctypes.POINTER(c_i64)]),
]

_LIB = None


class NodeHipError(RuntimeError):
    pass


def load(path=None):
    """Load ``libnode_hip.so`` (built in-tree by ``diffeqflux_hip.build``)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise NodeHipError(
            f"{p} not found: build it with `python -m diffeqflux_hip.build` "
            "(there is no CPU fallback for the NeuralODE hot path)")
    lib = ctypes.CDLL(p)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.node_abi_version() != 1:
        raise NodeHipError("libnode_hip ABI version mismatch")
    if path is None:
        _LIB = lib
    return lib


def make_desc(widths, acts, input_map=MAP_IDENTITY, precision="fp32", batch=1):
    d = NodeDesc()
    L = len(widths) - 1
    if not 1 <= L <= MAX_LAYERS:
        raise ValueError("1..8 Dense layers supported")
    d.n_layers = L
    for i, w in enumerate(widths):
        d.widths[i] = int(w)
    for i, a in enumerate(acts):
        d.acts[i] = ACT[a] if isinstance(a, str) else int(a)
    d.input_map = int(input_map)
    d.precision = PREC[precision] if isinstance(precision, str) else int(precision)
    d.batch = int(batch)
    return d


def check(lib, handle, rc, what):
    if rc != 0:
        msg = lib.node_last_error(handle) if handle else b""
        raise NodeHipError(f"{what} failed ({ERR_NAMES.get(rc, rc)}): "
                           f"{msg.decode() if msg else ''}")
