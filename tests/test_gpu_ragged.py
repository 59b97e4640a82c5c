"""GPU parity of the per-trajectory ("ragged") solves: every trajectory of one
launch against the oracle solving that trajectory alone (batch 1), on the
same FP32 parameters.

Tolerances: saved states within 20 * (abstol + reltol*|u|) of the FP64 oracle
at every save time (the FP32 error estimate may take a different, equally
valid step sequence); accepted-step counts within max(2, n/5) of the oracle
run in FP32 (the reference's own precision).
"""
import numpy as np
import pytest

from oracle import adjoint as OA
from oracle.mlp import MLP, MAP_CUBE

pytestmark = pytest.mark.gpu


def ragged_case(widths, acts, imap, n, seed, tmax=3.0, nsave=(3, 12)):
    m = MLP(widths, acts, input_map=imap)
    rng = np.random.default_rng(seed)
    p32 = m.glorot_init(rng).astype(np.float32)
    p32[m.offsets[0][1]:m.offsets[0][2]] = rng.uniform(-0.1, 0.1, widths[1])
    u0 = rng.standard_normal((n, widths[0])).astype(np.float32)
    if imap == MAP_CUBE:
        u0 *= 0.7
    t1 = rng.uniform(0.1, tmax, n)
    saves = []
    for i in range(n):
        k = int(rng.integers(nsave[0], nsave[1] + 1))
        s = np.linspace(0.0, t1[i], k)
        if i % 3 == 1:
            s = s[1:]          # t0 not saved
        if i % 4 == 2:
            s = s[:-1]         # t1 not saved
        saves.append(s)
    return m, p32, u0, t1, saves


@pytest.mark.parametrize("case", ["config4", "spiral", "identity_last2"])
def test_ragged_forward_matches_per_trajectory_oracle(case):
    from diffeqflux_hip import Handle
    if case == "config4":
        widths, acts, imap, n, tol = [16, 128, 128, 16], ["tanh", "tanh", "identity"], 0, 24, 1e-5
    elif case == "spiral":
        widths, acts, imap, n, tol = [2, 50, 2], ["tanh", "identity"], MAP_CUBE, 20, 1e-6
    else:
        widths, acts, imap, n, tol = [32, 64, 32], ["softplus", "identity"], 0, 17, 1e-4
    m, p32, u0, t1, saves = ragged_case(widths, acts, imap, n, seed=3)
    h = Handle(widths, acts, imap, "fp32", 1)
    h.set_params(p32)
    outs, ret, ts, tot = h.ragged_forward(u0, t1, saves, abstol=tol, reltol=tol)
    print(f"\n[ragged {case} n={n}] totals {tot}", flush=True)
    assert np.all(ret == 1), ret
    p = p32.astype(np.float64)
    worst = 0.0
    for i in range(n):
        s = list(saves[i])
        ref = OA.forward(m, p, u0[i:i + 1].astype(np.float64), 0.0, float(t1[i]), saveat=s,
                         abstol=tol, reltol=tol)
        ref32 = OA.forward(m, p32, u0[i:i + 1], 0.0, float(t1[i]), saveat=s, abstol=tol,
                           reltol=tol)
        R = np.array(ref.u)[:, 0, :]
        assert outs[i].shape == R.shape, (i, outs[i].shape, R.shape)
        err = np.abs(outs[i] - R) / (tol + tol * np.abs(R))
        worst = max(worst, float(err.max()))
        assert err.max() < 20, (i, err.max())
        slack = max(2, ref32.naccept // 5)
        assert abs(int(ts[i, 1]) - ref32.naccept) <= slack, (i, ts[i], ref32.naccept)
        assert ts[i, 3] == ts[i, 1] + 1
    print(f"  worst scaled err {worst:.3f}; steps gpu {ts[:6, 1].tolist()} ...", flush=True)
    assert tot["naccept"] == int(ts[:, 1].sum()) and tot["n_success"] == n


def test_ragged_fixed_dt_matches_oracle_closely():
    """Fixed dt: same step sequence as the oracle, so FP32 rounding only."""
    from diffeqflux_hip import Handle
    widths, acts = [16, 128, 128, 16], ["tanh", "tanh", "identity"]
    m, p32, u0, t1, saves = ragged_case(widths, acts, 0, 9, seed=5)
    t1 = np.round(t1 * 8) / 8 + 0.25          # multiples of dt
    saves = [np.linspace(0.0, t, 5) for t in t1]
    h = Handle(widths, acts, 0, "fp32", 1)
    h.set_params(p32)
    outs, ret, ts, _ = h.ragged_forward(u0, t1, saves, adaptive=False, dt=1 / 32)
    assert np.all(ret == 1)
    p = p32.astype(np.float64)
    for i in range(len(t1)):
        ref = OA.forward(m, p, u0[i:i + 1].astype(np.float64), 0.0, float(t1[i]),
                         saveat=list(saves[i]), adaptive=False, dt=1 / 32)
        R = np.array(ref.u)[:, 0, :]
        e = np.abs(outs[i] - R).max() / np.abs(R).max()
        assert e < 2e-5, (i, e)
        assert ts[i, 1] == ref.naccept


def test_ragged_rejects_unsupported_and_invalid():
    from diffeqflux_hip import Handle, NodeHipError
    h = Handle([64, 256, 64], ["tanh", "identity"], 0, "fp32", 1)
    h.set_params(np.zeros(h.P, np.float32))
    with pytest.raises(NodeHipError, match="D <= 32"):
        h.ragged_forward(np.zeros((2, 64)), [1.0, 1.0], [[0.5], [0.5]])
    h2 = Handle([16, 32, 16], ["tanh", "identity"], 0, "fp32", 1)
    h2.set_params(np.zeros(h2.P, np.float32))
    with pytest.raises(NodeHipError, match="t1 > t0"):
        h2.ragged_forward(np.zeros((2, 16)), [1.0, 0.0], [[0.5], [0.0]])
    with pytest.raises(NodeHipError, match="increase"):
        h2.ragged_forward(np.zeros((1, 16)), [1.0], [[0.5, 0.2]])


@pytest.mark.parametrize("case", ["config4", "spiral", "padded_rows"])
def test_ragged_adjoint_matches_per_trajectory_oracle(case):
    """dL/du0 and each trajectory's dL/dp against the oracle's
    InterpolatingAdjoint of that trajectory alone; dL/dp summed over
    trajectories in FP64.  The cases cover the parameter rows' tile orders:
    rows along the outputs (16-128-128-16 first layers), transposed rows
    along the inputs (its last layer), padding tiles in both (2-50-2), and
    partial tiles whose padding features carry nonzero activations
    (3-40-40-3 with sigmoid: sigmoid(0) = 1/2 sits in the padding rows)."""
    from diffeqflux_hip import Handle
    if case == "config4":
        widths, acts, imap, n, tol = [16, 128, 128, 16], ["tanh", "tanh", "identity"], 0, 10, 1e-5
    elif case == "spiral":
        widths, acts, imap, n, tol = [2, 50, 2], ["tanh", "identity"], MAP_CUBE, 8, 1e-6
    else:
        widths, acts, imap, n, tol = [3, 40, 40, 3], ["sigmoid", "tanh", "identity"], 0, 6, 1e-6
    m, p32, u0, t1, saves = ragged_case(widths, acts, imap, n, seed=7, nsave=(2, 9))
    h = Handle(widths, acts, imap, "fp32", 1)
    h.set_params(p32)
    outs, ret, ts, _ = h.ragged_forward(u0, t1, saves, abstol=tol, reltol=tol)
    assert np.all(ret == 1)
    rng = np.random.default_rng(11)
    cots = [rng.standard_normal(o.shape).astype(np.float32) for o in outs]
    du0, dp, rows, aret, ats, tot = h.ragged_backward(cots, per_trajectory=True)
    print(f"\n[ragged adjoint {case} n={n}] totals {tot}", flush=True)
    assert np.all(aret == 1), aret
    p = p32.astype(np.float64)
    gp_sum = np.zeros_like(p)
    worst_u = worst_p = 0.0
    for i in range(n):
        s = list(saves[i])
        ref = OA.forward(m, p, u0[i:i + 1].astype(np.float64), 0.0, float(t1[i]), saveat=s,
                         abstol=tol, reltol=tol)
        c = [cots[i][k:k + 1].astype(np.float64) for k in range(len(s))]
        gu0, gp, adj = OA.backward(m, p, ref, c, abstol=tol, reltol=tol)
        gp_sum += gp
        eu = float(np.abs(du0[i] - gu0[0]).max() / max(np.abs(gu0).max(), 1e-30))
        ep = float(np.abs(rows[i] - gp).max() / max(np.abs(gp).max(), 1e-30))
        worst_u, worst_p = max(worst_u, eu), max(worst_p, ep)
        print(f"  traj {i}: saves {len(s)} adj steps gpu {ats[i, 1]}+{ats[i, 2]} oracle "
              f"{adj.naccept}+{adj.nreject}; rel err du0 {eu:.2e} dp {ep:.2e}", flush=True)
        assert eu < 50 * tol + 1e-5 and ep < 50 * tol + 1e-5, (i, eu, ep)
    e = float(np.abs(dp - gp_sum).max() / np.abs(gp_sum).max())
    print(f"  worst du0 {worst_u:.2e} dp {worst_p:.2e}; summed dL/dp rel err {e:.2e}", flush=True)
    assert e < 50 * tol + 1e-5
    assert np.allclose(dp, rows.astype(np.float64).sum(0), rtol=1e-6, atol=1e-9)


def test_ragged_adjoint_bf16_vjp_option():
    """NODE_OPT_RAGGED_VJP_PRECISION = BF16: the adjoint's VJP products on
    BF16 operands (FP32 accumulation, everything else FP32) on the
    16-128-128-16 case, against the per-trajectory FP64 oracle at a BF16 bar
    (operand rounding 2^-9 relative, through three layers and the adjoint's
    error control); the FP32 setting restored gives the FP32 results again."""
    from diffeqflux_hip import Handle, NodeHipError
    widths, acts, tol, n = [16, 128, 128, 16], ["tanh", "tanh", "identity"], 1e-5, 10
    m, p32, u0, t1, saves = ragged_case(widths, acts, 0, n, seed=7, nsave=(2, 9))
    h = Handle(widths, acts, 0, "fp32", 1)
    h.set_params(p32)
    outs, ret, _, _ = h.ragged_forward(u0, t1, saves, abstol=tol, reltol=tol)
    assert np.all(ret == 1)
    rng = np.random.default_rng(11)
    cots = [rng.standard_normal(o.shape).astype(np.float32) for o in outs]
    du0_32, dp_32, _, _, _, tot32 = h.ragged_backward(cots)
    with pytest.raises(NodeHipError):
        h.set_option("ragged_vjp_precision", "bf16x3")
    h.set_option("ragged_vjp_precision", "bf16")
    du0, dp, rows, aret, ats, tot = h.ragged_backward(cots, per_trajectory=True)
    assert np.all(aret == 1), aret
    p = p32.astype(np.float64)
    gp_sum, gu = np.zeros_like(p), []
    for i in range(n):
        s = list(saves[i])
        ref = OA.forward(m, p, u0[i:i + 1].astype(np.float64), 0.0, float(t1[i]), saveat=s,
                         abstol=tol, reltol=tol)
        c = [cots[i][k:k + 1].astype(np.float64) for k in range(len(s))]
        gu0, gp, _ = OA.backward(m, p, ref, c, abstol=tol, reltol=tol)
        gp_sum += gp
        gu.append(gu0[0])
    gu = np.array(gu)
    eu = float(np.abs(du0 - gu).max() / np.abs(gu).max())
    ep = float(np.abs(dp - gp_sum).max() / np.abs(gp_sum).max())
    print(f"\n[ragged adjoint, BF16 VJP operands] adjoint steps {tot['naccept']}+{tot['nreject']} "
          f"(FP32 {tot32['naccept']}+{tot32['nreject']}); rel err du0 {eu:.2e} dL/dp {ep:.2e}", flush=True)
    assert eu < 2e-2 and ep < 2e-2
    h.set_option("ragged_vjp_precision", "fp32")
    du0_b, dp_b, _, _, _, _ = h.ragged_backward(cots)
    assert np.array_equal(dp_b, dp_32) and np.array_equal(du0_b, du0_32)


def test_ragged_many_trajectories_take_queue_positions():
    """More trajectories than the launch's slots (16 per wave, 4 waves per CU:
    16,384 on 256 CUs): the rest start from the queue as slots free up.  10
    distinct trajectories repeated 2,000 times: every copy's saved states,
    dL/du0 and own dL/dp equal its first copy's bit for bit (whichever slot
    or queue position ran it), and the 10 against the FP64 oracle."""
    from diffeqflux_hip import Handle
    widths, acts, tol, k, reps = [2, 16, 2], ["tanh", "identity"], 1e-6, 10, 2000
    m, p32, u0s, t1s, saves_s = ragged_case(widths, acts, 0, k, seed=13, nsave=(2, 4))
    n = k * reps
    u0 = np.tile(u0s, (reps, 1))
    t1 = np.tile(t1s, reps)
    saves = [saves_s[i % k] for i in range(n)]
    h = Handle(widths, acts, 0, "fp32", 1)
    h.set_params(p32)
    outs, ret, _, tot = h.ragged_forward(u0, t1, saves, abstol=tol, reltol=tol)
    assert np.all(ret == 1)
    rng = np.random.default_rng(17)
    cots_s = [rng.standard_normal(o.shape).astype(np.float32) for o in outs[:k]]
    cots = [cots_s[i % k] for i in range(n)]
    du0, dp, rows, aret, _, atot = h.ragged_backward(cots, per_trajectory=True)
    print(f"\n[ragged n={n}] forward {tot['naccept']} / adjoint {atot['naccept']} steps", flush=True)
    assert np.all(aret == 1)
    for i in range(k, n):
        j = i % k
        assert np.array_equal(outs[i], outs[j]), i
    du0r, rowsr = du0.reshape(reps, k, -1), rows.reshape(reps, k, -1)
    assert np.array_equal(du0r, np.broadcast_to(du0r[:1], du0r.shape))
    assert np.array_equal(rowsr, np.broadcast_to(rowsr[:1], rowsr.shape))
    p = p32.astype(np.float64)
    for i in range(k):
        s = list(saves[i])
        ref = OA.forward(m, p, u0[i:i + 1].astype(np.float64), 0.0, float(t1[i]), saveat=s,
                         abstol=tol, reltol=tol)
        R = np.array(ref.u)[:, 0, :]
        assert np.all(np.abs(outs[i] - R) <= 20 * (tol + tol * np.abs(R))), i
        c = [cots[i][q:q + 1].astype(np.float64) for q in range(len(s))]
        gu0, gp, _ = OA.backward(m, p, ref, c, abstol=tol, reltol=tol)
        assert np.abs(du0[i] - gu0[0]).max() <= 50 * tol * max(np.abs(gu0).max(), 1e-3), i
        assert np.abs(rows[i] - gp).max() <= 50 * tol * max(np.abs(gp).max(), 1e-3), i


def test_ragged_flat_layout_forms_match_lists():
    """save_times as (times, offsets) and flat_out=True give exactly the
    per-trajectory list forms' results (the entry point's own layout)."""
    from diffeqflux_hip import Handle
    widths, acts, tol, n = [2, 50, 2], ["tanh", "identity"], 1e-6, 12
    m, p32, u0, t1, saves = ragged_case(widths, acts, MAP_CUBE, n, seed=21)
    h = Handle(widths, acts, MAP_CUBE, "fp32", 1)
    h.set_params(p32)
    outs, ret, ts, _ = h.ragged_forward(u0, t1, saves, abstol=tol, reltol=tol)
    cots = [np.full(o.shape, 0.5, np.float32) for o in outs]
    du0, dp, _, _, _, _ = h.ragged_backward(cots)
    lens = np.array([len(s) for s in saves], np.int64)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum(lens)
    flat, ret2, ts2, _ = h.ragged_forward(u0, t1, (np.concatenate(saves), off), flat_out=True,
                                          abstol=tol, reltol=tol)
    assert flat.shape == (off[-1], 2)
    assert np.array_equal(flat, np.concatenate(outs))
    assert np.array_equal(ret, ret2) and np.array_equal(ts, ts2)
    du0f, dpf, _, _, _, _ = h.ragged_backward(np.full(flat.shape, 0.5, np.float32))
    assert np.array_equal(du0, du0f) and np.array_equal(dp, dpf)
