"""Pins the Tsit5 tableau by the mathematics that fixes it (no reference
fixture exists for this path): RK order conditions, embedded-pair order,
dense-output conditions."""
import itertools

import numpy as np
import pytest

from oracle import tableau as tb

A = tb.a_matrix()
c = np.array(tb.C)
b = np.array(tb.B)
bt = np.array(tb.BTILDE)


def test_row_sums_equal_c():
    assert np.allclose(A.sum(axis=1), c, atol=1e-14, rtol=0)


def _tree_conditions(w, order):
    """Order conditions sum w_i Phi_i(tree) = 1/gamma(tree) for all rooted
    trees up to `order` (explicit RK, c = row sums)."""
    Ac = A @ c
    Ac2 = A @ c ** 2
    AAc = A @ Ac
    conds = [
        (w @ np.ones(7), 1.0),
        (w @ c, 1 / 2),
        (w @ c ** 2, 1 / 3), (w @ Ac, 1 / 6),
        (w @ c ** 3, 1 / 4), (w @ (c * Ac), 1 / 8), (w @ Ac2, 1 / 12),
        (w @ AAc, 1 / 24),
    ]
    if order >= 5:
        conds += [
            (w @ c ** 4, 1 / 5), (w @ (c ** 2 * Ac), 1 / 10),
            (w @ (c * Ac2), 1 / 15), (w @ (c * AAc), 1 / 30),
            (w @ (Ac * Ac), 1 / 20), (w @ (A @ c ** 3), 1 / 20),
            (w @ (A @ (c * Ac)), 1 / 40), (w @ (A @ Ac2), 1 / 60),
            (w @ (A @ AAc), 1 / 120),
        ]
    return conds


def test_fifth_order_conditions():
    conds = _tree_conditions(b, 5)
    assert len(conds) == 17
    for lhs, rhs in conds:
        assert abs(lhs - rhs) < 1e-14


def test_btilde_annihilates_low_order_trees():
    # b - btilde must be a 4th-order method: btilde orthogonal to all trees <= 4
    for lhs, _ in _tree_conditions(bt, 4):
        assert abs(lhs) < 1e-14
    # but not to every order-5 tree (otherwise there is no error estimate)
    assert max(abs(lhs) for lhs, _ in _tree_conditions(bt, 5)[8:]) > 1e-6


def test_dense_output_endpoints():
    assert np.allclose(tb.interp_weights(1.0), b, atol=1e-14)
    assert np.allclose(tb.interp_weights(0.0), 0.0)


@pytest.mark.parametrize("theta", [0.1, 0.3, 0.5, 0.7, 0.95])
def test_dense_output_continuous_order4(theta):
    w = np.array(tb.interp_weights(theta))
    Ac = A @ c
    # continuous order conditions: sum b_i(t) Phi_i = t^k / gamma
    assert abs(w.sum() - theta) < 1e-14
    assert abs(w @ c - theta ** 2 / 2) < 1e-14
    assert abs(w @ c ** 2 - theta ** 3 / 3) < 1e-14
    assert abs(w @ Ac - theta ** 3 / 6) < 1e-14
    assert abs(w @ c ** 3 - theta ** 4 / 4) < 1e-13
    assert abs(w @ (c * Ac) - theta ** 4 / 8) < 1e-13
    assert abs(w @ (A @ c ** 2) - theta ** 4 / 12) < 1e-13
    assert abs(w @ (A @ Ac) - theta ** 4 / 24) < 1e-13


def test_a_typo_would_be_caught():
    """A plausible mistake (a transposed pair of coefficients) breaks the
    order conditions: the checks above have teeth."""
    bad = b.copy()
    bad[[2, 3]] = bad[[3, 2]]
    errs = [abs(lhs - rhs) for lhs, rhs in _tree_conditions(bad, 5)]
    assert max(errs) > 1e-3
