#!/usr/bin/env python3
"""Forward + adjoint RHS evaluations per second of the NeuralODE hot path.

Workload (the headline configuration): state D=64, MLP 64-1024-1024-1024-64
with softplus, B=4096 trajectories u0 ~ N(0,1), tspan (0,1) with 32 saveat
points, Tsit5 rtol=1e-4 (atol 1e-6), MSE loss against the saveat states of a
random linear-plus-tanh teacher ODE, gradient by InterpolatingAdjoint.  GEMMs
and VJPs on BF16 MFMA (FP32 accumulation): by default split BF16 (hi + lo
operand planes, three passes; --wgrad-precision bf16 runs the adjoint's
weight-gradient products in one BF16 pass); state, stage sums and adjoint
state in FP32; time / controller / dL/dp quadrature in FP64.

One "step" = one forward solve + loss cotangent + one adjoint solve, all on
inputs resident in HBM.  ``value`` = batched RHS evaluations (forward RHS +
adjoint VJP, each over all B trajectories) per second over the timed steps.

Also reported: the same rate for the FP32 GPU path (the 5x comparison), the
dominant GEMM's roofline from HIP events on the library's stream (a second
pass over the same steps: the event hooks make the library launch kernels one
by one instead of replaying its captured graphs), and
(``cpu_baseline`` leg, rank 0 only) the oracle restated in numpy FP32 timed on
a bounded sample, plus the BF16/FP32 gradient error against the FP64 oracle on
that sample.
"""
import argparse
import json
import math
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "diffeqflux.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = ("fwd+adjoint RHS evals/sec, D=64 H=1024 B=4096; BF16-MFMA vs FP32 at grad rel err")
PEAK = {"bf16": 2500.0, "bf16x3": 2500.0, "fp32": 157.3}   # MI355X dense TFLOP/s
ROLE_NAMES = {0: "fwd_hidden", 1: "fwd_stage", 2: "adj_hidden", 3: "bwd_data", 4: "wgrad",
              5: "bwd_last"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------- workload
D, H, NH = 64, 1024, 3
WIDTHS = [D] + [H] * NH + [D]
ACTS = ["softplus"] * NH + ["identity"]
TSPAN = (0.0, 1.0)
NSAVE = 32
RTOL, ATOL = 1e-4, 1e-6


def glorot_params(seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for i, o in zip(WIDTHS[:-1], WIDTHS[1:]):
        lim = math.sqrt(6.0 / (i + o))
        out += [rng.uniform(-lim, lim, i * o), np.zeros(o)]
    return np.concatenate(out).astype(np.float32)


def teacher_targets(u0, ts, seed=1, substeps=8):
    """Saveat states of u' = A u + tanh(C u), A = -I + 0.5 G/sqrt(D),
    C = G'/sqrt(D), by classical RK4 (synthetic targets only)."""
    rng = np.random.default_rng(seed)
    A = -np.eye(D) + 0.5 * rng.standard_normal((D, D)) / math.sqrt(D)
    C = rng.standard_normal((D, D)) / math.sqrt(D)
    f = lambda u: u @ A.T + np.tanh(u @ C.T)
    u = u0.astype(np.float64)
    out = [u.copy()]
    for a, b in zip(ts[:-1], ts[1:]):
        h = (b - a) / substeps
        for _ in range(substeps):
            k1 = f(u); k2 = f(u + h / 2 * k1); k3 = f(u + h / 2 * k2); k4 = f(u + h * k3)
            u = u + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        out.append(u.copy())
    return np.stack(out).astype(np.float32)


def make_inputs(B, seed=1):
    u0 = np.random.default_rng(seed).standard_normal((B, D)).astype(np.float32)
    ts = np.linspace(TSPAN[0], TSPAN[1], NSAVE)
    return u0, ts, teacher_targets(u0, ts)


def flops_per_launch(role, layer, B):
    """Algorithmic FLOPs of one launch of the (role, layer) GEMM: 2*B*in*out
    (the work the Dense chain or its pullback asks for, whatever the number of
    MFMA passes the precision mode issues for it)."""
    i, o = WIDTHS[layer], WIDTHS[layer + 1]
    return 2.0 * B * i * o


def passes(role, precision, wgrad_precision=None):
    """BF16 MFMA passes issued per algorithmic product: split BF16 runs three
    (hi*hi + lo*hi + hi*lo); role 4 (weight gradients) runs in
    ``wgrad_precision`` when that is given."""
    prec = wgrad_precision if (role == 4 and wgrad_precision) else precision
    return 3 if prec == "bf16x3" else 1


# ---------------------------------------------------------------- GPU legs
class Runner:
    def __init__(self, precision, B, p, u0, ts, targets, wgrad_precision=None):
        import torch
        from diffeqflux_hip import Handle
        self.torch = torch
        self.h = Handle(WIDTHS, ACTS, 0, precision, B)
        if wgrad_precision:
            self.h.set_option("wgrad_precision", wgrad_precision)
        self.wgrad_precision = wgrad_precision
        self.h.set_params(p)
        self.opts = Handle.make_opts(TSPAN[0], TSPAN[1], abstol=ATOL, reltol=RTOL)
        self.interior = list(ts[1:-1])
        self.u0 = torch.from_numpy(u0).cuda()
        self.tgt = torch.from_numpy(targets).cuda()
        self.B = B

    def step(self):
        out, _, sf = self.h.forward(self.u0, self.opts, self.interior)
        cot = (2.0 / out.numel()) * (out - self.tgt)
        self.torch.cuda.current_stream().synchronize()
        du0, dp, sa = self.h.backward(cot)
        return sf, sa, du0, dp

    def timed(self, steps, warmup):
        torch = self.torch
        for _ in range(warmup):
            self.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nf = nfa = 0
        last = None
        for _ in range(steps):
            last = self.step()
            nf += last[0]["nf"]
            nfa += last[1]["nf"]
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        return dt, nf, nfa, last


def gemm_shape(role, layer, B):
    """(M, N, K) of the GEMM of a (role, layer) launch."""
    i, o = WIDTHS[layer], WIDTHS[layer + 1]
    if role in (0, 1, 2):      # forward / recompute of layer l: B x out, K = in
        return (B, o, i)
    if role in (3, 5):         # data gradient of layer l: B x in, K = out
        return (B, i, o)
    return (i, o, B)           # weight gradient: in x out, K = batch


# HBM bytes per launch of the dominant kernel come from the committed PMC
# summary of this bench's command (two rocprofv3 --pmc runs, FETCH_SIZE then
# WRITE_SIZE in KiB per dispatch, each scaled by what the counter reads for a
# plain device copy of known size); the summary carries one machine-readable
# line "traffic_bytes_per_launch: <precision> <role> <M>x<N>x<K> <bytes>".
PROFILES = os.path.join(ROOT, "profiles")


def measured_traffic():
    out = {}
    if not os.path.isdir(PROFILES):
        return out
    for fn in sorted(os.listdir(PROFILES)):
        if not fn.startswith("traffic_"):
            continue
        with open(os.path.join(PROFILES, fn)) as f:
            for line in f:
                mt = re.match(r"\s*traffic_bytes_per_launch:\s*(\S+)\s+(\d+)\s+(\d+)x(\d+)x(\d+)"
                              r"\s+(\d+)", line)
                if mt:
                    shape = tuple(int(mt.group(k)) for k in (3, 4, 5))
                    out[(mt.group(1), int(mt.group(2)), shape)] = (int(mt.group(6)), fn)
    return out


# Whole-solve counters of the headline step from the committed PMC summary:
# "step_profile: <precision> hbm_bytes_per_solve <B> mfma_busy_cycles_per_solve <C>"
# (C summed over the chip's SIMDs).
def measured_step_profile():
    out = {}
    if not os.path.isdir(PROFILES):
        return out
    for fn in sorted(os.listdir(PROFILES)):
        if not fn.startswith("pmc_"):
            continue
        with open(os.path.join(PROFILES, fn)) as f:
            for line in f:
                mt = re.match(r"\s*step_profile:\s*(\S+)\s+hbm_bytes_per_solve\s+(\d+)\s+"
                              r"mfma_busy_cycles_per_solve\s+(\d+)", line)
                if mt:
                    out[mt.group(1)] = (int(mt.group(2)), int(mt.group(3)), fn)
    return out


def pmc_summary(root, timed=3):
    """Per-solve totals of rocprofv3 --pmc runs of this bench's headline leg
    (``--steps 3 --warmup 0``, csv output).  ``root`` holds one run per
    counter set (SQ_VALU_MFMA_BUSY_CYCLES/, FETCH_SIZE/, WRITE_SIZE/) and
    cFETCH_SIZE/, cWRITE_SIZE/ (a 256 MiB device copy, 3 dispatches).  The
    dispatch stream is cut at every k_fwd_init (one per solve: the warm solve
    with the autotuner, the timed solves, then the roofline passes); the
    figures are the mean over the ``timed`` solves after the warm one."""
    import csv
    import glob

    def kname(nm):
        m = re.search(r"gemm_tn<.*?, (\d+), (\d+), (\d+), (\d+), \d+, nodehip::(\w+)>", nm)
        if m:
            return f"gemm_tn<{m.group(1)}x{m.group(2)},{m.group(3)}x{m.group(4)},{m.group(5)}>"
        return re.sub(r"^void |<.*|\(.*", "", nm).split("::")[-1]

    def solves(tag):
        fs = glob.glob(os.path.join(root, tag, "*_counter_collection.csv"))
        disp = {}
        with open(fs[0]) as f:
            for r in csv.DictReader(f):
                d = disp.setdefault(int(r["Dispatch_Id"]), {"k": kname(r["Kernel_Name"])})
                d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        segs, cur = [], None
        for did in sorted(disp):
            if disp[did]["k"] == "k_fwd_init":
                cur = []
                segs.append(cur)
            if cur is not None:
                cur.append(disp[did])
        return segs[1:1 + timed]

    def per_solve(tag, ctr):
        segs = solves(tag)
        out = {}
        for seg in segs:
            for d in seg:
                out[d["k"]] = out.get(d["k"], 0.0) + d.get(ctr, 0.0) / len(segs)
        return out

    def copy_kib(tag, ctr):
        fs = glob.glob(os.path.join(root, tag, "*_counter_collection.csv"))
        with open(fs[0]) as f:
            v = [float(r["Counter_Value"]) for r in csv.DictReader(f)
                 if "copyBuffer" in r["Kernel_Name"] and r["Counter_Name"] == ctr]
        return sum(v) / len(v)

    fscale = 262144.0 / copy_kib("cFETCH_SIZE", "FETCH_SIZE")
    wscale = 262144.0 / copy_kib("cWRITE_SIZE", "WRITE_SIZE")
    fe, wr = per_solve("FETCH_SIZE", "FETCH_SIZE"), per_solve("WRITE_SIZE", "WRITE_SIZE")
    mb = per_solve("SQ_VALU_MFMA_BUSY_CYCLES", "SQ_VALU_MFMA_BUSY_CYCLES")
    mi = per_solve("SQ_VALU_MFMA_BUSY_CYCLES", "SQ_INSTS_MFMA")
    ks = set(fe) | set(wr) | set(mb)
    rows = [{"kernel": k, "fetch_bytes": fe.get(k, 0.0) * fscale * 1024,
             "write_bytes": wr.get(k, 0.0) * wscale * 1024,
             "mfma_busy_cycles": mb.get(k, 0.0), "mfma_insts": mi.get(k, 0.0)} for k in ks]
    rows.sort(key=lambda r: -(r["fetch_bytes"] + r["write_bytes"]))
    tot = {c: sum(r[c] for r in rows) for c in ("fetch_bytes", "write_bytes", "mfma_busy_cycles",
                                                 "mfma_insts")}
    return {"fetch_scale": fscale, "write_scale": wscale, "solves": timed, "per_kernel": rows,
            "per_solve": tot}


def profile_dominant(run):
    """One profiled step with every GEMM timed.  Launches are grouped by
    (role, GEMM shape) -- the layers of one shape run the same kernel -- and
    the group with the largest total time is the dominant kernel."""
    run.h.prof_enable(None)
    run.step()
    recs = run.h.prof_collect()
    run.h.prof_enable([])
    tot, grp = {}, {}
    for k, ms in recs:
        tot[k] = tot.get(k, 0.0) + ms
        g = (k[0], gemm_shape(k[0], k[1], run.B))
        grp[g] = grp.get(g, 0.0) + ms
    dom = max(grp, key=grp.get)
    keys = sorted(k for k in tot if (k[0], gemm_shape(k[0], k[1], run.B)) == dom)
    return dom, keys, tot


def roofline_for(run, dom, keys, precision, durations):
    """achieved = algorithmic FLOPs per launch / average duration of the
    launches that did work (launches gated off on the device -- FSAL stage
    without re-evaluation, attempts enqueued after the end -- return at once
    and are excluded by duration).  Durations come from HIP event pairs
    around each launch on the library's stream; they include the event
    commands themselves (a few microseconds above rocprof's kernel-only
    duration of the same launches)."""
    d = np.array(durations, dtype=np.float64)
    if d.size == 0:
        return None, {}
    thr = 0.25 * np.percentile(d, 90)
    real = d[d >= thr]
    avg_ms = float(real.mean())
    role, layer = keys[0]
    fl = flops_per_launch(role, layer, run.B)
    npass = passes(role, precision, run.wgrad_precision)
    achieved = fl / (avg_ms * 1e-3) / 1e12
    peak = PEAK[precision]
    M, N, K = dom[1]
    info = {"kernel": f"{ROLE_NAMES[role]}[layers {','.join(str(k[1] + 1) for k in keys)}] "
                      f"{M}x{N}, K={K}",
            "launches": int(d.size), "launches_with_work": int(real.size),
            "avg_ms": avg_ms, "flop_per_launch": fl, "mfma_passes": npass,
            "issued_tflops": round(achieved * npass, 2),
            "issued_frac": round(achieved * npass / peak, 4)}
    prec_key = run.wgrad_precision if (role == 4 and run.wgrad_precision) else precision
    tr = measured_traffic().get((prec_key, role, dom[1]))
    traffic = None
    if tr is not None:
        traffic = tr[0]
        info["traffic_source"] = (f"profiles/{tr[1]}: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE per "
                                  "dispatch, copy-kernel calibrated")
    return {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak,
            "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
            "traffic": traffic}, info


# ---------------------------------------------------------------- ragged leg
RAGGED_WIDTHS = [16, 128, 128, 16]
RAGGED_ACTS = ["tanh", "tanh", "identity"]


def ragged_inputs(n, seed=4):
    """Per-trajectory workload: u0 ~ N(0,1), t1 ~ U(0.1, 5), save count ~
    U{5..200} on a uniform grid over [0, t1] (endpoints included)."""
    rng = np.random.default_rng(seed)
    u0 = rng.standard_normal((n, RAGGED_WIDTHS[0])).astype(np.float32)
    t1 = rng.uniform(0.1, 5.0, n)
    saves = [np.linspace(0.0, t1[i], int(rng.integers(5, 201))) for i in range(n)]
    return u0, t1, saves


def ragged_leg(n=10000, reps=3, cpu_sample=8):
    """Per-trajectory solves (10,000 trajectories, D=16, MLP 16-128-128-16
    tanh, own t1 / save times / adaptive steps, rtol 1e-5, atol 1e-6) forward
    + adjoint of an MSE-style loss, FP32 MFMA; the oracle on a few of the same
    trajectories (batch 1 each) as the CPU baseline."""
    from diffeqflux_hip import Handle
    from oracle.mlp import MLP
    m = MLP(RAGGED_WIDTHS, RAGGED_ACTS)
    p = m.glorot_init(np.random.default_rng(2)).astype(np.float32)
    u0, t1, saves = ragged_inputs(n)
    # the entry point's own layout: concatenated save times + offsets in, the
    # flat saved-state array out (no per-trajectory Python lists timed)
    lens = np.array([len(s) for s in saves], np.int64)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum(lens)
    times = np.concatenate(saves)
    den = np.repeat((0.5 * lens * n).astype(np.float32), lens)[:, None]
    h = Handle(RAGGED_WIDTHS, RAGGED_ACTS, 0, "fp32", 1)
    h.set_params(p)
    kw = dict(abstol=1e-6, reltol=1e-5)
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        flat, ret, ts, tot = h.ragged_forward(u0, t1, (times, off), flat_out=True, **kw)
        t1f = time.perf_counter()
        cots = flat / den                                   # d/du of sum ||u - u/2||^2 / N
        t1c = time.perf_counter()
        du0, dp, _, aret, ats, atot = h.ragged_backward(cots)
        t2 = time.perf_counter()
        rec = (t2 - t0 - (t1c - t1f), t1f - t0, t2 - t1c, tot, atot)
        if best is None or rec[0] < best[0]:
            best = rec
    # the BF16-operand VJP option on the same forward and cotangents
    # (best of two: the first also loads that kernel)
    h.set_option("ragged_vjp_precision", "bf16")
    tb = float("inf")
    for _ in range(2):
        t0 = time.perf_counter()
        du0b, dpb, _, aretb, _, atotb = h.ragged_backward(cots)
        tb = min(tb, time.perf_counter() - t0)
    h.set_option("ragged_vjp_precision", "fp32")
    wall, tf, ta, tot, atot = best
    evals = tot["nf"] + atot["nf"]
    res = {"workload": "10,000 trajectories, D=16, MLP 16-128-128-16 tanh, t1~U(0.1,5), "
                       "saveat count~U{5..200}, rtol 1e-5 atol 1e-6, per-trajectory adaptive dt, "
                       "FP32 MFMA",
           "traj_per_s": round(n / wall, 1), "rhs_evals_per_s_traj": round(evals / wall, 1),
           "fwd_ms": round(1e3 * tf, 2), "adj_ms": round(1e3 * ta, 2),
           "fwd_steps": tot["naccept"], "fwd_rejects": tot["nreject"],
           "adj_steps": atot["naccept"], "adj_rejects": atot["nreject"],
           "all_success": bool(np.all(ret == 1) and np.all(aret == 1)),
           "timing": "host entry points incl. PCIe copies of inputs/outputs (flat save-time / "
                     "saved-state layout)",
           "bf16_vjp_option": {
               "adj_ms": round(1e3 * tb, 2), "adj_steps": atotb["naccept"],
               "adj_rejects": atotb["nreject"], "all_success": bool(np.all(aretb == 1)),
               "dLdp_rel_err_vs_fp32": float(np.abs(dpb - dp).max() / np.abs(dp).max()),
               "dLdu0_rel_err_vs_fp32": float(np.abs(du0b - du0).max() / np.abs(du0).max())}}
    # CPU baseline: the oracle (FP32) on the first cpu_sample trajectories
    from oracle import adjoint as OA
    tc = time.perf_counter()
    ev = 0
    for i in range(cpu_sample):
        f = OA.forward(m, p, u0[i:i + 1], 0.0, float(t1[i]), saveat=list(saves[i]), **kw)
        c = [x / (0.5 * len(saves[i]) * n) for x in np.array(f.u)]
        _, _, a = OA.backward(m, p, f, c, **kw)
        ev += f.nf + a.nf
    cs = time.perf_counter() - tc
    res["cpu_baseline"] = {"value": round(cpu_sample / cs, 3), "unit": "trajectories/s",
                           "cores": 1, "kind": "port",
                           "sample": f"oracle FP32, trajectories 0..{cpu_sample - 1} solved one "
                                     f"by one (batch 1), forward+adjoint, {ev} evals in {cs:.1f} s"}
    return res


# ---------------------------------------------------------------- multiple shooting leg
def spiral_ensemble(ntraj, t, seed=8, substeps=64):
    """The multiple-shooting docs problem (u' = ((u.^3)' A)', A = [-0.1 2; -2
    -0.1]) from ntraj initial states on radius 1.5..2.5, integrated by
    classical RK4 with `substeps` steps per data interval (synthetic data)."""
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0, 2 * math.pi, ntraj)
    rad = rng.uniform(1.5, 2.5, ntraj)
    u = np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1)
    A = np.array([[-0.1, 2.0], [-2.0, -0.1]])
    f = lambda x: (x ** 3) @ A                                    # noqa: E731
    out = [u.copy()]
    for k in range(len(t) - 1):
        h = (t[k + 1] - t[k]) / substeps
        for _ in range(substeps):
            k1 = f(u); k2 = f(u + 0.5 * h * k1); k3 = f(u + 0.5 * h * k2); k4 = f(u + h * k3)
            u = u + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        out.append(u.copy())
    return np.stack(out, 1).astype(np.float32)                    # (ntraj, T, 2)


def multiple_shoot_leg(ntraj=1024, reps=3, cpu_traj=2):
    """Ensemble multiple shooting (src/multiple_shooting.jl:131-195) on the
    docs example: Chain(x -> x.^3, Dense(2,16,tanh), Dense(16,2)), 30 points
    on (0,5), group size 3, continuity term 200, ntraj trajectories: one loss
    + dL/dp evaluation = 15 x ntraj segment solves forward + adjoint (one
    per-trajectory launch each way); the oracle on the segments of cpu_traj
    trajectories as the CPU baseline."""
    import torch
    from diffeqflux_hip import (Chain, Cube, Dense, NeuralODE, Tsit5, multiple_shoot,
                                EnsembleProblem, EnsembleThreads, group_ranges)
    from oracle.mlp import MLP, MAP_CUBE
    m = MLP([2, 16, 2], ["tanh", "identity"], input_map=MAP_CUBE)
    p = m.glorot_init(np.random.default_rng(3)).astype(np.float32)
    t = np.linspace(0.0, 5.0, 30)
    data = spiral_ensemble(ntraj, t)
    node = NeuralODE(Chain(Cube(), Dense(2, 16, "tanh"), Dense(16, 2)), (0.0, 5.0), Tsit5(),
                     saveat=list(t))
    loss_fn = lambda u, v: ((u - v) ** 2).sum()                   # noqa: E731
    nseg = len(group_ranges(30, 3)) * ntraj
    best = None
    for _ in range(reps):
        pt = torch.tensor(p, requires_grad=True)
        t0 = time.perf_counter()
        loss, _ = multiple_shoot(pt, data, t, EnsembleProblem(node), EnsembleThreads(), loss_fn,
                                 Tsit5(), 3, continuity_term=200, trajectories=ntraj)
        loss.backward()
        dt = time.perf_counter() - t0
        if best is None or dt < best[0]:
            best = (dt, float(loss.detach()))
    wall, lval = best
    res = {"workload": f"ensemble multiple shooting, {ntraj} trajectories x 30 points of the docs "
                       "spiral, group size 3 ({nseg} segments), Chain(x.^3, Dense(2,16,tanh), "
                       "Dense(16,2)), Tsit5 reltol 1e-3 abstol 1e-6, loss + dL/dp".format(nseg=nseg),
           "segments_per_s": round(nseg / wall, 1), "evals_per_s": round(1.0 / wall, 3),
           "ms_per_eval": round(1e3 * wall, 2), "loss": lval,
           "timing": "host entry point incl. PCIe copies and torch loss assembly"}
    from oracle import adjoint as OA
    tc = time.perf_counter()
    nsc = 0
    for j in range(cpu_traj):
        for rg in group_ranges(30, 3):
            s = [float(x) for x in t[rg.start:rg.stop]]
            f = OA.forward(m, p, data[j, rg.start][None], s[0], s[-1], saveat=s, abstol=1e-6,
                           reltol=1e-3)
            c = [2.0 * (np.array(f.u)[k] - data[j, rg.start + k][None]) for k in range(len(s))]
            OA.backward(m, p, f, c, abstol=1e-6, reltol=1e-3)
            nsc += 1
    cs = time.perf_counter() - tc
    res["cpu_baseline"] = {"value": round(nsc / cs, 2), "unit": "segments/s", "cores": 1,
                           "kind": "port",
                           "sample": f"oracle FP32, the {nsc} segments of {cpu_traj} trajectories "
                                     f"solved one by one, forward+adjoint, in {cs:.1f} s"}
    return res


# ---------------------------------------------------------------- FFJORD leg
def gaussian_ring(n, D=16, seed=6):
    """8-component Gaussian mixture, means on a radius-4 ring in the first two
    coordinates, std 0.5; N(0,1) in the other coordinates."""
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 8, n)
    x = rng.standard_normal((n, D)).astype(np.float32)
    ang = 2 * math.pi * k / 8
    x[:, 0] = 4 * np.cos(ang) + 0.5 * rng.standard_normal(n)
    x[:, 1] = 4 * np.sin(ang) + 0.5 * rng.standard_normal(n)
    return x


def ffjord_leg(B=8192, reps=3, precision="bf16x3", cpu_batch=8192):
    """FFJORD log-likelihood forward (D=16, time-concatenated MLP 17-512-512-16
    tanh, Hutchinson trace with one Rademacher probe per sample, tspan (0,1),
    Tsit5 rtol 1e-5 atol 1e-5) on the Gaussian-ring data, and the training
    step (forward + InterpolatingAdjoint of -mean(logpx) with the second-order
    VJP of the Hutchinson term); the oracle (FP32) on a slice of the batch as
    the CPU baselines."""
    from diffeqflux_hip import Handle, _lib
    from oracle.mlp import MLP
    widths, acts = [17, 512, 512, 16], ["tanh", "tanh", "identity"]
    m = MLP(widths, acts)
    p = m.glorot_init(np.random.default_rng(3)).astype(np.float32)
    x = gaussian_ring(B)
    e = np.random.default_rng(7).choice(np.array([-1.0, 1.0], np.float32), size=(B, 16))
    h = Handle(widths, acts, _lib.MAP_CNF_TIME, precision, B)
    h.set_params(p)
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        lp, z, dl, st = h.cnf_forward(x, e, 0.0, 1.0, abstol=1e-5, reltol=1e-5)
        dt = time.perf_counter() - t0
        if best is None or dt < best[0]:
            best = (dt, st, lp)
    dt, st, lp = best
    # training step: forward + InterpolatingAdjoint of -mean(logpx) (second-order
    # VJP of the Hutchinson term), same probe
    tbest = None
    for _ in range(reps):
        t0 = time.perf_counter()
        _, z, _, st2 = h.cnf_forward(x, e, 0.0, 1.0, abstol=1e-5, reltol=1e-5)
        dx, dp, sb = h.cnf_backward(z / B, np.full(B, 1.0 / B, np.float32))
        tt = time.perf_counter() - t0
        if tbest is None or tt < tbest[0]:
            tbest = (tt, st2, sb)
    fl = 2.0 * B * sum(i * o for i, o in zip(widths[:-1], widths[1:])) * 2  # f + probe VJP
    res = {"workload": "FFJORD forward, B=8192, D=16, MLP 17-512-512-16 tanh (time-concat), "
                       "Hutchinson 1 Rademacher probe, tspan (0,1), rtol 1e-5 atol 1e-5",
           "precision": precision, "ms": round(1e3 * dt, 2), "nf": st["nf"],
           "steps": f"{st['naccept']}+{st['nreject']}",
           "rhs_evals_per_s": round(st["nf"] / dt, 1), "samples_per_s": round(B / dt, 1),
           "tflops": round(st["nf"] * fl / dt / 1e12, 1),
           "tflops_issued": round(st["nf"] * fl * (3 if precision == "bf16x3" else 1) / dt / 1e12, 1),
           "mean_logpx": float(np.mean(lp)),
           "train_step_ms": round(1e3 * tbest[0], 2),
           "train_evals": f"fwd {tbest[1]['nf']} adj {tbest[2]['nf']} "
                          f"(adjoint steps {tbest[2]['naccept']}+{tbest[2]['nreject']})",
           "train_samples_per_s": round(B / tbest[0], 1),
           "timing": "host entry points incl. PCIe copies of x / probe / cotangents / results"}
    from oracle import ffjord as FF
    tc = time.perf_counter()
    rlp, _, _, sol = FF.forward(m, p, x[:cpu_batch], 0.0, 1.0, probe=e[:cpu_batch],
                                time_concat=True, abstol=1e-5, reltol=1e-5)
    cs = time.perf_counter() - tc
    res["cpu_baseline"] = {"value": round(cpu_batch / cs, 2), "unit": "samples/s",
                           "cores": blas_cores(),
                           "kind": "port",
                           "sample": f"oracle FP32 FFJORD forward on {cpu_batch} of the samples "
                                     f"(same probe), {sol.nf} evals in {cs:.1f} s"}
    tc = time.perf_counter()
    _, rz, _, sol = FF.forward(m, p, x[:cpu_batch], 0.0, 1.0, probe=e[:cpu_batch],
                               time_concat=True, abstol=1e-5, reltol=1e-5)
    gz, gd = FF.nll_cotangents(rz, cpu_batch)
    _, _, adj = FF.backward(m, p, sol, e[:cpu_batch], gz.astype(np.float32),
                            gd.astype(np.float32), time_concat=True, abstol=1e-5, reltol=1e-5)
    cs = time.perf_counter() - tc
    res["cpu_baseline_train"] = {"value": round(cpu_batch / cs, 2), "unit": "samples/s",
                                 "cores": blas_cores(), "kind": "port",
                                 "sample": f"oracle FP32 FFJORD forward + adjoint on {cpu_batch} of "
                                           f"the samples, {sol.nf}+{adj.nf} evals in {cs:.1f} s"}
    return res


# ---------------------------------------------------------------- C1 / C2 legs
SPIRAL_A = np.array([[-0.1, 2.0], [-2.0, -0.1]])


def rk4_states(f, u0, ts, substeps=32):
    """States of u' = f(u) at the times ts by classical RK4 (synthetic data)."""
    u = np.asarray(u0, np.float64)
    out = [u.copy()]
    for a, b in zip(ts[:-1], ts[1:]):
        h = (b - a) / substeps
        for _ in range(substeps):
            k1 = f(u); k2 = f(u + h / 2 * k1); k3 = f(u + h / 2 * k2); k4 = f(u + h * k3)
            u = u + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        out.append(u.copy())
    return np.stack(out).astype(np.float32)


def glorot(widths, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i, o in zip(widths[:-1], widths[1:]):
        lim = math.sqrt(6.0 / (i + o))
        out += [rng.uniform(-lim, lim, i * o), np.zeros(o)]
    return np.concatenate(out).astype(np.float32)


class SmallRunner:
    """forward + MSE cotangent + adjoint of one workload on the GPU (device
    tensors, FP32 handle)."""

    def __init__(self, widths, acts, imap, B, p, u0, ts, tgt, tol, precision="fp32"):
        import torch
        from diffeqflux_hip import Handle
        self.torch = torch
        self.h = Handle(widths, acts, imap, precision, B)
        self.h.set_params(p)
        self.opts = Handle.make_opts(ts[0], ts[-1], **tol)
        self.interior = list(ts[1:-1])
        self.u0 = torch.from_numpy(u0).cuda()
        self.tgt = torch.from_numpy(tgt).cuda()

    def step(self):
        out, _, sf = self.h.forward(self.u0, self.opts, self.interior)
        cot = (2.0 / out.numel()) * (out - self.tgt)
        self.torch.cuda.current_stream().synchronize()
        du0, dp, sa = self.h.backward(cot)
        return sf, sa, du0, dp

    def time(self, reps, warmup=1):
        for _ in range(warmup):
            self.step()
        self.torch.cuda.synchronize()
        t0 = time.perf_counter()
        nf = 0
        for _ in range(reps):
            sf, sa, _, _ = self.step()
            nf += sf["nf"] + sa["nf"]
        self.torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, nf / reps, sf, sa


def spiral_leg(reps=20):
    """The docs example on its own terms (B = 1): u0 = [2, 0], tspan (0, 1.5),
    30 saveat, Chain(x -> x.^3, Dense(2, 50, tanh), Dense(50, 2)) with Glorot
    weights (seed 0), Tsit5 abstol = reltol = 1e-6, MSE against the true
    spiral's states; forward + adjoint on the GPU (FP32, latency-bound at one
    trajectory) and the oracle in numpy FP32 on the host (one core: the whole
    problem is one 2-vector)."""
    widths, acts = [2, 50, 2], ["tanh", "identity"]
    p = glorot(widths, 0)
    u0 = np.array([[2.0, 0.0]], np.float32)
    ts = np.linspace(0.0, 1.5, 30)
    tgt = rk4_states(lambda u: (u ** 3) @ SPIRAL_A, u0, ts)
    tol = dict(abstol=1e-6, reltol=1e-6)
    r = SmallRunner(widths, acts, 1, 1, p, u0, ts, tgt, tol)
    sec, nf, sf, sa = r.time(reps, warmup=2)
    res = {"workload": "spiral docs example, B=1, 30 saveat, Chain(x.^3, Dense(2,50,tanh), "
                       "Dense(50,2)), Tsit5 abstol=reltol=1e-6, MSE vs the true spiral, FP32",
           "ms_per_loss_grad": round(1e3 * sec, 3), "rhs_evals_per_s": round(nf / sec, 1),
           "steps_fwd_adj": f"{sf['naccept']}+{sf['nreject']}/{sa['naccept']}+{sa['nreject']}",
           "timing": "device-resident u0/targets, host-synchronous entry points"}
    from oracle import adjoint as OA
    from oracle.mlp import MLP, MAP_CUBE
    m = MLP(widths, acts, input_map=MAP_CUBE)
    _, _, du0, dp = r.step()

    def cpu(dtype):
        f = OA.forward(m, p.astype(dtype), u0.astype(dtype), 0.0, 1.5, saveat=ts, **tol)
        pr = np.array(f.u)
        cot = (2.0 / pr.size) * (pr - tgt.astype(dtype))
        gu, gp, a = OA.backward(m, p.astype(dtype), f, list(cot), **tol)
        return f, a, gp
    cpu(np.float32)
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < 3.0 or n < 3:
        f, a, _ = cpu(np.float32)
        n += 1
    cs = (time.perf_counter() - t0) / n
    gp64 = cpu(np.float64)[2]
    res["grad_rel_err"] = float(f"{np.linalg.norm(dp.cpu().numpy() - gp64) / np.linalg.norm(gp64):.3e}")
    res["cpu_baseline"] = {"value": round((f.nf + a.nf) / cs, 1), "unit": "RHS evals/s",
                           "ms_per_loss_grad": round(1e3 * cs, 2), "cores": 1, "kind": "port",
                           "sample": f"oracle numpy FP32, the whole problem, mean of {n} runs "
                                     "after one warm-up"}
    # the same problem through the per-trajectory entry points (n = 1: one
    # persistent launch per direction instead of a launch sequence per step)
    from diffeqflux_hip import Handle, _lib
    hr = Handle(widths, acts, _lib.MAP_CUBE, "fp32", 1)
    hr.set_params(p)
    off = np.array([0, ts.size], np.int64)

    def rg_step():
        flat, _, _, tot = hr.ragged_forward(u0, [1.5], (ts, off), flat_out=True, **tol)
        cot = ((2.0 / flat.size) * (flat - tgt.reshape(flat.shape))).astype(np.float32)
        _, dpr, _, _, _, atot = hr.ragged_backward(cot)
        return dpr, tot, atot
    for _ in range(2):
        rg_step()
    t0 = time.perf_counter()
    for _ in range(reps):
        dpr, tot, atot = rg_step()
    sr = (time.perf_counter() - t0) / reps
    res["per_trajectory"] = {
        "ms_per_loss_grad": round(1e3 * sr, 3),
        "steps_fwd_adj": f"{tot['naccept']}+{tot['nreject']}/{atot['naccept']}+{atot['nreject']}",
        "grad_rel_err": float(f"{np.linalg.norm(dpr - gp64) / np.linalg.norm(gp64):.3e}"),
        "timing": "node_ragged_forward / node_ragged_backward with n = 1, host pointers"}
    return res


def plain_leg(reps=5, cpu_batch=256):
    """Batched plain FP32 NeuralODE: D = 32, MLP 32-256-256-32 tanh (Glorot,
    seed 0), B = 4096 u0 ~ N(0, 1), tspan (0, 1), 20 saveat, Tsit5 rtol 1e-5
    atol 1e-7, MSE against the saveat states of the seeded linear-plus-tanh
    teacher (RK4); forward + adjoint on the GPU in FP32, and the oracle in
    numpy FP32 on a batch sample as the CPU baseline."""
    D, B = 32, 4096
    widths, acts = [32, 256, 256, 32], ["tanh", "tanh", "identity"]
    p = glorot(widths, 0)
    ts = np.linspace(0.0, 1.0, 20)
    tol = dict(abstol=1e-7, reltol=1e-5)

    def data(n):
        u0 = np.random.default_rng(1).standard_normal((n, D)).astype(np.float32)
        rng = np.random.default_rng(1)
        A = -np.eye(D) + 0.5 * rng.standard_normal((D, D)) / math.sqrt(D)
        C = rng.standard_normal((D, D)) / math.sqrt(D)
        return u0, rk4_states(lambda u: u @ A.T + np.tanh(u @ C.T), u0, ts, substeps=8)
    u0, tgt = data(B)
    r = SmallRunner(widths, acts, 0, B, p, u0, ts, tgt, tol)
    sec, nf, sf, sa = r.time(reps)
    res = {"workload": "D=32, MLP 32-256-256-32 tanh, B=4096, tspan (0,1), 20 saveat, Tsit5 "
                       "rtol 1e-5 atol 1e-7, MSE vs teacher ODE, FP32 MFMA",
           "rhs_evals_per_s": round(nf / sec, 1), "traj_per_s": round(B / sec, 1),
           "ms_per_step": round(1e3 * sec, 2),
           "steps_fwd_adj": f"{sf['naccept']}+{sf['nreject']}/{sa['naccept']}+{sa['nreject']}",
           "timing": "inputs resident in HBM"}
    from oracle import adjoint as OA
    from oracle.mlp import MLP
    m = MLP(widths, acts)
    us, tgs = data(cpu_batch)
    rs = SmallRunner(widths, acts, 0, cpu_batch, p, us, ts, tgs, tol)
    _, _, du0, dp = rs.step()

    def cpu(dtype):
        f = OA.forward(m, p.astype(dtype), us.astype(dtype), 0.0, 1.0, saveat=ts, **tol)
        pr = np.array(f.u)
        cot = (2.0 / pr.size) * (pr - tgs.astype(dtype))
        gu, gp, a = OA.backward(m, p.astype(dtype), f, list(cot), **tol)
        return f, a, gp
    t0 = time.perf_counter()
    f, a, _ = cpu(np.float32)
    cs = time.perf_counter() - t0
    gp64 = cpu(np.float64)[2]
    res["grad_rel_err"] = float(f"{np.linalg.norm(dp.cpu().numpy() - gp64) / np.linalg.norm(gp64):.3e}")
    res["grad_rel_err_sample"] = f"B={cpu_batch} sample vs FP64 oracle"
    res["cpu_baseline"] = {"value": round((f.nf + a.nf) / cs * cpu_batch / B, 2),
                           "unit": "RHS evals/s (full-batch equivalent)", "cores": blas_cores(),
                           "kind": "port",
                           "sample": f"one FP32 forward+adjoint solve of the oracle at B={cpu_batch}, "
                                     f"{f.nf + a.nf} evals in {cs:.1f} s, rate scaled by "
                                     f"{cpu_batch}/{B}"}
    return res


# ---------------------------------------------------------------- C3 finite differences
FD_B, FD_DIRS, FD_TOL = 8, 64, dict(abstol=1e-12, reltol=1e-9)


def fd_direction(i, P):
    """Unit direction i of the finite-difference check: N(0, I) / norm, seeded by (2, i)."""
    v = np.random.default_rng([2, i]).standard_normal(P)
    return v / np.linalg.norm(v)


_FD = {}


def _fd_init():
    try:
        from threadpoolctl import threadpool_limits
        threadpool_limits(1)
    except Exception:
        pass
    from oracle.mlp import MLP
    u0, ts, tgt = make_inputs(FD_B)
    _FD.update(m=MLP(WIDTHS, ACTS), p=glorot_params().astype(np.float64),
               u0=u0.astype(np.float64), ts=ts, tgt=tgt.astype(np.float64))


def _fd_loss(task):
    """MSE loss of the FP64 oracle solve at p + s * eps * v_i (tight tolerance)."""
    from oracle import adjoint as OA
    i, s, eps = task
    m, p = _FD["m"], _FD["p"]
    pp = p + (s * eps) * fd_direction(i, p.size) if s else p
    f = OA.forward(m, pp, _FD["u0"], TSPAN[0], TSPAN[1], saveat=_FD["ts"], **FD_TOL)
    return float(np.mean((np.array(f.u) - _FD["tgt"]) ** 2))


def fd_leg(gpu_grad):
    """C3's gradient against central finite differences of the FP64 oracle
    along FD_DIRS random unit directions of parameter space, on a batch
    sample (B = FD_B of the same u0 / targets): the FP64 loss is solved at a
    tight tolerance (rtol 1e-9) so that its finite differences follow the
    exact ODE's loss; eps is picked by a sweep on direction 0 (the value whose
    central difference changes least when eps is halved).  Reported: the GPU
    gradient's directional derivatives (its own tolerance, rtol 1e-4) against
    the differences, and the FP64 oracle adjoint at the tight tolerance
    against the same differences (which pins them)."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    from oracle import adjoint as OA
    from oracle.mlp import MLP
    P = int(gpu_grad.size)
    nproc = max(1, min(16, len(os.sched_getaffinity(0))))
    t0 = time.perf_counter()
    # fresh interpreters (this process holds a GPU context); a worker that
    # dies raises here instead of being replaced forever
    with ProcessPoolExecutor(nproc, mp_context=mp.get_context("spawn"),
                             initializer=_fd_init) as pool:
        sweep = [1e-2, 5e-3, 1e-3, 5e-4, 1e-4, 5e-5]
        lv = list(pool.map(_fd_loss, [(0, s, e) for e in sweep for s in (1, -1)]))
        fd0 = [(lv[2 * k] - lv[2 * k + 1]) / (2 * e) for k, e in enumerate(sweep)]
        k = int(np.argmin([abs(fd0[k] - fd0[k + 1]) for k in range(len(sweep) - 1)]))
        eps = sweep[k]
        lv = list(pool.map(_fd_loss, [(i, s, eps) for i in range(FD_DIRS) for s in (1, -1)]))
    fd = np.array([(lv[2 * i] - lv[2 * i + 1]) / (2 * eps) for i in range(FD_DIRS)])
    t_fd = time.perf_counter() - t0
    # FP64 oracle adjoint at the tight tolerance on the same sample
    m = MLP(WIDTHS, ACTS)
    u0, ts, tgt = make_inputs(FD_B)
    p64 = glorot_params().astype(np.float64)
    f = OA.forward(m, p64, u0.astype(np.float64), TSPAN[0], TSPAN[1], saveat=ts, **FD_TOL)
    pr = np.array(f.u)
    cot = (2.0 / pr.size) * (pr - tgt.astype(np.float64))
    _, g64, _ = OA.backward(m, p64, f, list(cot), **FD_TOL)
    V = [fd_direction(i, P) for i in range(FD_DIRS)]
    dg = np.array([float(np.dot(gpu_grad, v)) for v in V])
    do = np.array([float(np.dot(g64, v)) for v in V])
    scale = float(np.abs(fd).max())
    return {"sample": f"B={FD_B} of the headline inputs, {FD_DIRS} unit directions "
                      f"N(0,I)/norm (seeds (2,i)), FP64 oracle loss at rtol 1e-9 atol 1e-12",
            "eps": eps, "eps_sweep": {f"{e:g}": float(f"{v:.6e}") for e, v in zip(sweep, fd0)},
            "gpu_max_rel_err": float(f"{np.abs(dg - fd).max() / scale:.3e}"),
            "gpu_median_rel_err": float(f"{np.median(np.abs(dg - fd) / np.abs(fd)):.3e}"),
            "gpu_cosine": float(f"{np.dot(dg, fd) / np.linalg.norm(dg) / np.linalg.norm(fd):.8f}"),
            "oracle_adjoint_max_rel_err": float(f"{np.abs(do - fd).max() / scale:.3e}"),
            "cpu_s": round(t_fd, 1), "processes": nproc}


# ---------------------------------------------------------------- CPU leg
def blas_cores():
    """Threads the oracle's numpy BLAS runs on (the CPU baselines' cores)."""
    try:
        from threadpoolctl import threadpool_info
        return max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:
        return len(os.sched_getaffinity(0))


def cpu_baseline_leg(p, Bs, gpu_grads):
    """The oracle (numpy FP32 restatement of the reference algorithm) timed on
    a bounded sample: one forward + adjoint solve at Bs trajectories (same D,
    H, tolerances, saveat, loss).  Also the FP64 oracle gradient on the same
    sample, against which the GPU gradients of that sample are checked."""
    from oracle import adjoint as OA
    from oracle.mlp import MLP
    cores = blas_cores()
    m = MLP(WIDTHS, ACTS)
    u0, ts, tgt = make_inputs(Bs)

    def solve(dtype):
        pp, uu = p.astype(dtype), u0.astype(dtype)
        f = OA.forward(m, pp, uu, TSPAN[0], TSPAN[1], saveat=ts, abstol=ATOL, reltol=RTOL)
        pred = np.array(f.u)
        cot = (2.0 / pred.size) * (pred - tgt.astype(dtype))
        gu, gp, a = OA.backward(m, pp, f, list(cot), abstol=ATOL, reltol=RTOL)
        return f, a, gu, gp

    t0 = time.perf_counter()
    f32, a32, _, _ = solve(np.float32)
    cpu_s = time.perf_counter() - t0
    evals = f32.nf + a32.nf
    _, _, gu64, gp64 = solve(np.float64)
    errs = {}
    for name, (du0, dp) in gpu_grads.items():
        errs[name] = float(np.linalg.norm(np.asarray(dp, np.float64) - gp64) / np.linalg.norm(gp64))
        errs[name + "_du0"] = float(np.linalg.norm(np.asarray(du0, np.float64) - gu64)
                                    / np.linalg.norm(gu64))
    return {"value": round(evals / cpu_s * Bs / 4096.0, 3),
            "unit": "RHS evals/s (full-batch equivalent)", "cores": int(cores), "kind": "port",
            "sample": f"one FP32 forward+adjoint solve of the oracle at B={Bs} (same D/H/tol/"
                      f"saveat/loss), {evals} batched evals in {cpu_s:.1f} s, rate scaled by "
                      f"{Bs}/4096"}, errs


# ---------------------------------------------------------------- main
# MI355X: 256 CUs x 4 SIMDs, peak engine clock 2.4 GHz
N_SIMD, CLOCK_HZ = 1024, 2.4e9


def solve_counters(precision, ms_per_solve):
    """Whole-solve HBM rate and MFMA busy fraction: per-solve bytes and MFMA
    busy cycles from the committed PMC summary of this command, over this
    run's solve time (busy fraction at the peak clock)."""
    sp = measured_step_profile().get(precision)
    if sp is None:
        return {}
    nbytes, busy, fn = sp
    t = ms_per_solve * 1e-3
    return {"solve_hbm_gbs": round(nbytes / t / 1e9, 1),
            "solve_hbm_frac": round(nbytes / t / 8e12, 4),
            "solve_mfma_busy": round(busy / (N_SIMD * CLOCK_HZ * t), 4),
            "solve_counters_source": f"profiles/{fn}: rocprofv3 --pmc per solve (FETCH_SIZE + "
                                     "WRITE_SIZE copy-calibrated, SQ_VALU_MFMA_BUSY_CYCLES)"}


def main():
    if len(sys.argv) == 3 and sys.argv[1] == "--pmc-summary":
        print(json.dumps(pmc_summary(sys.argv[2]), indent=1))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--precision", default="bf16x3", choices=["bf16x3", "bf16", "fp32"])
    ap.add_argument("--wgrad-precision", default="same", choices=["bf16x3", "bf16", "fp32", "same"],
                    help="precision of the adjoint weight-gradient products ('bf16' is allowed "
                         "on a bf16x3 run; 'same' = --precision)")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 comparison leg")
    ap.add_argument("--no-bf16", action="store_true", help="skip the pure-BF16 leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-batch", type=int, default=256)
    ap.add_argument("--no-ragged", action="store_true",
                    help="skip the per-trajectory (ragged) workload leg")
    ap.add_argument("--no-ffjord", action="store_true", help="skip the FFJORD workload leg")
    ap.add_argument("--no-multiple-shoot", action="store_true",
                    help="skip the multiple-shooting workload leg")
    ap.add_argument("--no-spiral", action="store_true", help="skip the spiral (B=1) leg")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain FP32 D=32 B=4096 workload leg")
    ap.add_argument("--no-fd", action="store_true",
                    help="skip the finite-difference check of the headline gradient")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the NeuralODE library runs on one GPU only (no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    p = glorot_params()
    u0, ts, tgt = make_inputs(args.batch)
    wgp = None if args.wgrad_precision in ("same", args.precision) else args.wgrad_precision
    if wgp and args.precision != "bf16x3":
        wgp = None
    log(f"[bench] {torch.cuda.get_device_name(0)} B={args.batch} {args.precision}"
        f" (weight gradients {wgp or args.precision})")

    run = Runner(args.precision, args.batch, p, u0, ts, tgt, wgp)
    run.step()  # warm the allocator / code objects; the handle autotunes its tiles here
    plans = {f"{k[0]}{k[1]}": f"{v[0]}x{v[1]}" + (f"/k{v[2]}" if v[2] > 1 else "")
             for k, v in run.h.plans().items()}
    log("[bench] autotuned tiles:", plans)
    # timed region: step attempts replayed from captured HIP graphs, no hooks
    dt, nf, nfa, last = run.timed(args.steps, args.warmup)
    value = (nf + nfa) / dt
    S = sum(i * o for i, o in zip(WIDTHS[:-1], WIDTHS[1:]))
    solve_flops = 2.0 * args.batch * S * nf + 6.0 * args.batch * S * nfa
    # roofline pass: the same steps again with HIP events around every launch
    # of the dominant GEMM (event hooks launch kernels one by one)
    dom, keys, tot = profile_dominant(run)
    log("[bench] per-role GEMM time (ms, one step, eager launches):",
        {f"{ROLE_NAMES[k[0]]}{k[1] + 1}": round(v, 2) for k, v in sorted(tot.items())})
    run.h.prof_enable(keys)
    dtp, _, _, _ = run.timed(args.steps, 0)
    durs = [ms for _, ms in run.h.prof_collect()]
    run.h.prof_enable([])
    roof, rinfo = roofline_for(run, dom, keys, args.precision, durs)
    rinfo["pass_ms_per_step"] = round(1e3 * dtp / args.steps, 2)
    sf, sa = last[0], last[1]
    log(f"[bench] {args.precision}: {value:.1f} evals/s, {1e3 * dt / args.steps:.2f} ms/step, "
        f"fwd {sf['naccept']}+{sf['nreject']} steps nf {sf['nf']}, adj {sa['naccept']}+"
        f"{sa['nreject']} steps nf {sa['nf']}; dominant {rinfo}")

    extra = {}
    legs = []
    if not args.no_fp32 and args.precision != "fp32":
        legs.append(("fp32", max(2, args.steps // 5)))
    if not args.no_bf16 and args.precision != "bf16":
        legs.append(("bf16", 1))
    for prec, nsteps in legs:
        r2 = Runner(prec, args.batch, p, u0, ts, tgt)
        dt2, nf2, nfa2, l2 = r2.timed(nsteps, 1)
        v2 = (nf2 + nfa2) / dt2
        extra[f"{prec}_evals_per_s"] = round(v2, 2)
        extra[f"{prec}_ms_per_step"] = round(1e3 * dt2 / nsteps, 2)
        extra[f"{prec}_steps_fwd_adj"] = (f"{l2[0]['naccept']}+{l2[0]['nreject']}/"
                                          f"{l2[1]['naccept']}+{l2[1]['nreject']}")
        log(f"[bench] {prec}: {v2:.1f} evals/s, {1e3 * dt2 / nsteps:.1f} ms/step "
            f"(steps fwd/adj {extra[f'{prec}_steps_fwd_adj']})")
        del r2
    if "fp32_evals_per_s" in extra:
        extra["speedup_vs_fp32"] = round(value / extra["fp32_evals_per_s"], 3)
        extra["time_speedup_vs_fp32"] = round(extra["fp32_ms_per_step"] / (1e3 * dt / args.steps), 3)

    cpu = None
    if not args.no_cpu:
        Bs = args.cpu_batch
        us, tss, tgs = make_inputs(Bs)
        grads = {}
        variants = [("bf16x3", "bf16x3", None), ("bf16x3_wgrad_bf16", "bf16x3", "bf16"),
                    ("bf16", "bf16", None), ("fp32", "fp32", None)]
        for name, prec, wg in variants:
            rs = Runner(prec, Bs, p, us, tss, tgs, wg)
            _, _, du0, dp = rs.step()
            grads[name] = (du0.cpu().numpy(), dp.cpu().numpy())
        cpu, errs = cpu_baseline_leg(p, Bs, grads)
        for prec, _, _ in variants:
            extra[f"grad_rel_err_{prec}"] = float(f"{errs[prec]:.3e}")
            extra[f"grad_rel_err_du0_{prec}"] = float(f"{errs[prec + '_du0']:.3e}")
        extra["grad_rel_err_sample"] = f"||dL/dp - FP64 oracle|| / ||oracle||, B={Bs} sample"
        head = "bf16x3_wgrad_bf16" if wgp == "bf16" else args.precision
        extra["grad_rel_err"] = extra[f"grad_rel_err_{head}"]
        log(f"[bench] cpu baseline {cpu}; grad rel err {errs}")

    if not args.no_fd:
        rf = Runner(args.precision, FD_B, p, *make_inputs(FD_B), wgp)
        _, _, _, dpf = rf.step()
        extra["fd64"] = fd_leg(dpf.cpu().numpy())
        log(f"[bench] fd64: {extra['fd64']}")
        del rf
    if not args.no_spiral:
        extra["spiral"] = spiral_leg()
        log(f"[bench] spiral: {extra['spiral']}")
    if not args.no_plain:
        extra["plain"] = plain_leg()
        log(f"[bench] plain: {extra['plain']}")
    if not args.no_ragged:
        extra["ragged"] = ragged_leg()
        log(f"[bench] ragged: {extra['ragged']}")
    if not args.no_ffjord:
        extra["ffjord"] = ffjord_leg()
        log(f"[bench] ffjord: {extra['ffjord']}")
    if not args.no_multiple_shoot:
        extra["multiple_shoot"] = multiple_shoot_leg()
        log(f"[bench] multiple_shoot: {extra['multiple_shoot']}")

    rec = {
        "metric": METRIC, "value": round(value, 2), "unit": "RHS evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.precision,
        "wgrad_dtype": wgp or args.precision,
        "dtype_note": {"bf16x3": "BF16 MFMA operands split hi+lo (3 passes), FP32 accumulate; "
                                 "FP32 state, FP64 time/controller/dL/dp quadrature",
                       "bf16": "BF16 MFMA operands, FP32 accumulate; FP32 state",
                       "fp32": "FP32 MFMA"}[args.precision],
        "data": "synthetic (u0 ~ N(0,1), teacher-ODE targets, Glorot weights)",
        "config": {"workload": "NeuralODE fwd+InterpolatingAdjoint, D=64, MLP "
                               "64-1024-1024-1024-64 softplus, B=4096, tspan (0,1), 32 saveat, "
                               "Tsit5 rtol=1e-4 atol=1e-6, MSE loss",
                   "batch": args.batch, "precision": args.precision,
                   "wgrad_precision": wgp or args.precision},
        "traj_per_s": round(args.batch * args.steps / dt, 2),
        # whole solve: algorithmic FLOPs (forward RHS 2BS, adjoint VJP 6BS,
        # S = sum in*out) per second of the timed steps, against the peak
        "solve_tflops": round(solve_flops / dt / 1e12, 2),
        "solve_mfma_frac": round(solve_flops / dt / 1e12 / PEAK[args.precision], 4),
        **solve_counters(args.precision, 1e3 * dt / args.steps),
        "nf_fwd": sf["nf"], "nf_adj": sa["nf"],
        "steps_fwd": f"{sf['naccept']}+{sf['nreject']}",
        "steps_adj": f"{sa['naccept']}+{sa['nreject']}",
        "roofline": roof, "roofline_kernel": rinfo, "tiles": plans, "cpu_baseline": cpu,
    }
    rec.update(extra)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
