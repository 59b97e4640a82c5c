"""Host-side layer logic that needs no GPU: DimMover and the augmentation of
AugmentedNDELayer (test/BasicNeuralDE/neural_de_tests.jl:235-241 checks
DimMover by permutation equality)."""
import numpy as np
import torch

from diffeqflux_hip import AugmentedNDELayer, Chain, Dense, DimMover, NeuralODE


def julia_dimmover(x_julia, frm, to):
    """Direct restatement on a column-major (Fortran-order) array."""
    n = x_julia.ndim
    f = frm if frm > 0 else n + 1 + frm
    t = to if to > 0 else n + 1 + to
    slices = [np.take(x_julia, i, axis=f - 1) for i in range(x_julia.shape[f - 1])]
    return np.concatenate(slices, axis=t - 1 if t < f else t - 2)


def test_dimmover_matches_column_major_semantics():
    rng = np.random.default_rng(0)
    xj = np.asfortranarray(rng.standard_normal((3, 4, 5)))       # Julia D x B x Nt
    xc = np.ascontiguousarray(np.transpose(xj, (2, 1, 0)))       # our (Nt, B, D)
    for frm, to in [(-2, -1), (1, 3), (2, 1), (-1, -3)]:
        want = julia_dimmover(xj, frm, to)
        got, _ = DimMover(frm, to)(xc)
        assert np.array_equal(np.transpose(got, tuple(range(got.ndim))[::-1]), want)
        gt, _ = DimMover(frm, to)(torch.tensor(xc))
        assert np.array_equal(gt.numpy(), got)


def test_augment_appends_zero_rows():
    node = NeuralODE(Chain(Dense(4, 8, "tanh"), Dense(8, 4)), (0.0, 1.0))
    a = AugmentedNDELayer(node, 2)
    x = np.ones((3, 2), np.float32)            # Julia 2 x 3
    y = a.augment(x)
    assert y.shape == (3, 4) and np.all(y[:, 2:] == 0) and np.all(y[:, :2] == 1)
    xv = torch.ones(2, requires_grad=True)
    yv = a.augment(xv)
    assert yv.shape == (4,)
    yv.sum().backward()
    assert torch.equal(xv.grad, torch.ones(2))


def test_odesolution_consumers():
    # the ODESolution consumers of the reference docs: Array(sol)
    # (docs/src/examples/neural_ode.md:37), reduce(hcat, sol.u)
    # (docs/src/examples/GPUs.md:107), sol.u[end] (mnist_neural_ode.md:43),
    # last / sum (test/CUDA/cuda_tests.jl:52, test/BasicNeuralDE/neural_de_tests.jl:46)
    from diffeqflux_hip import ODESolution
    rng = np.random.default_rng(0)
    arr = rng.standard_normal((5, 3, 2)).astype(np.float32)     # Julia 2 x 3 x 5
    ts = np.array([0.0, 0.25, 0.5, 0.75, 1.0])
    sol = ODESolution(arr, ts, 1, {"nf": 42, "naccept": 7, "nreject": 0})
    assert sol.retcode == "Success" and sol.stats["nf"] == 42 and len(sol) == 5
    assert np.array_equal(np.asarray(sol), arr)
    assert np.array_equal(np.concatenate(sol.u, axis=0), arr.reshape(15, 2))   # reduce(hcat, u)
    assert np.array_equal(sol.u[-1], arr[-1]) and np.array_equal(sol[2], arr[2])
    assert np.isclose(sol.sum(), arr.sum())
    # sol(t): the saved state at a saved time, linear in between
    assert np.array_equal(sol(0.5), arr[2])
    assert np.allclose(sol(0.6), arr[2] + 0.4 * (arr[3] - arr[2]), atol=1e-7)
    assert np.array_equal(sol(1.0), arr[4])
    import pytest
    with pytest.raises(ValueError):
        sol(1.5)
    st = ODESolution(torch.tensor(arr), ts, 1, {})
    assert torch.allclose(st(0.6), torch.tensor(arr[2] + 0.4 * (arr[3] - arr[2])), atol=1e-6)
