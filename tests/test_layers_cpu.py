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
