"""Host logic of multiple shooting (``src/multiple_shooting.jl``): the group
partition, the per-group solve set-up and the argument forms.  No GPU."""
import numpy as np
import pytest

from diffeqflux_hip import multiple_shooting as MS
from diffeqflux_hip import (Chain, Dense, Cube, NeuralODE, Tsit5, group_ranges, multiple_shoot,
                            EnsembleProblem, EnsembleThreads)


def test_group_ranges_docstring_example():
    # src/multiple_shooting.jl:222-228: group_ranges(10, 5) == [1:5, 5:9, 9:10]
    assert group_ranges(10, 5) == [range(0, 5), range(4, 9), range(8, 10)]


def test_group_ranges_docs_example_sizes():
    # docs/src/examples/multiple_shooting.md: datasize 30, group_size 3
    g = group_ranges(30, 3)
    assert len(g) == 15 and g[0] == range(0, 3) and g[-1] == range(28, 30)
    assert group_ranges(30, 30) == [range(0, 30)]     # single shooting
    assert group_ranges(2, 2) == [range(0, 2)]


@pytest.mark.parametrize("bad", [(10, 1), (10, 11), (1, 2), (5, 0)])
def test_group_ranges_rejects_bad_sizes(bad):
    with pytest.raises(ValueError):
        group_ranges(*bad)


def test_group_ranges_partition_properties():
    """Every index covered, consecutive groups share exactly their boundary
    point, no group longer than groupsize, all but the last exactly that long
    (brute force over small sizes)."""
    for n in range(2, 41):
        for g in range(2, n + 1):
            rs = group_ranges(n, g)
            assert rs[0].start == 0 and rs[-1].stop == n
            for a, b in zip(rs, rs[1:]):
                assert b.start == a.stop - 1
                assert len(a) == g
            assert 2 <= len(rs[-1]) <= g
            assert sorted(set(i for r in rs for i in r)) == list(range(n))


def test_segments_shift_each_group_to_zero():
    rng = np.random.default_rng(0)
    data = rng.standard_normal((3, 11, 2)).astype(np.float32)   # (ntraj, T, D)
    t = np.linspace(0.5, 3.0, 11)
    rs = group_ranges(11, 4)
    u0, t1, (times, off) = MS._segments(data, t, rs)
    assert u0.shape == (3 * len(rs), 2) and u0.dtype == np.float32
    assert off.shape == (3 * len(rs) + 1,) and off[0] == 0 and off[-1] == times.size
    for gi, rg in enumerate(rs):
        for j in range(3):
            k = gi * 3 + j                                       # group-major
            sk = times[off[k]:off[k + 1]]
            assert np.array_equal(u0[k], data[j, rg.start])
            assert len(sk) == len(rg)
            assert sk[0] == 0.0 and sk[-1] == t1[k]
            assert np.allclose(sk + t[rg.start], t[rg.start:rg.stop], rtol=0, atol=1e-15)


def test_split_restores_group_shapes():
    data = np.zeros((2, 7, 3), np.float32)
    rs = group_ranges(7, 3)
    flat = np.arange(sum(2 * len(r) for r in rs) * 3, dtype=np.float32).reshape(-1, 3)
    parts = MS._split(flat, data, rs)
    assert [p.shape for p in parts] == [(2, len(r), 3) for r in rs]
    assert np.array_equal(np.concatenate([p.reshape(-1, 3) for p in parts]), flat)


def test_argument_forms_and_errors():
    nn = Chain(Cube(), Dense(2, 16, "tanh"), Dense(16, 2))
    node = NeuralODE(nn, (0.0, 1.0), Tsit5())
    data = np.zeros((5, 2), np.float32)
    t = np.linspace(0, 1, 5)
    loss = lambda u, v: ((u - v) ** 2).sum()                      # noqa: E731
    assert MS._parse((Tsit5(), 3))[0] is MS._default_continuity_loss
    cl = lambda a, b: 0.0                                         # noqa: E731
    assert MS._parse((cl, Tsit5(), 3))[0] is cl
    with pytest.raises(TypeError):
        MS._parse((3,))
    with pytest.raises(ValueError):      # group_size out of range, before any solve
        multiple_shoot(None, data, t, node, loss, Tsit5(), 6)
    with pytest.raises(ValueError):      # tsteps / data length mismatch
        multiple_shoot(None, data, t[:4], node, loss, Tsit5(), 3)
    with pytest.raises(NotImplementedError):
        multiple_shoot(None, data, t, node, loss, "Vern7", 3)
    with pytest.raises(ValueError):      # ensemble data must be 3-D
        multiple_shoot(None, data, t, EnsembleProblem(node), EnsembleThreads(), loss, Tsit5(), 3)
    with pytest.raises(ValueError):      # trajectories keyword must match
        multiple_shoot(None, np.zeros((2, 5, 2), np.float32), t, EnsembleProblem(node),
                       EnsembleThreads(), loss, Tsit5(), 3, trajectories=3)


def test_default_continuity_loss_is_sum_abs():
    a = np.array([1.0, -2.0, 0.5]); b = np.array([0.0, 1.0, 0.5])
    assert MS._default_continuity_loss(a, b) == 4.0
