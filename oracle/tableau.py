"""Tsit5 tableau (Tsitouras 2011, Comput. Math. Appl. 62:770-775).

Restates OrdinaryDiffEq's ``Tsit5ConstantCacheActual`` / ``Tsit5Interp``
coefficients (third-party package OrdinaryDiffEq 7, absent from the reference
tree; the reference calls it at ``src/neural_de.jl:55-56`` with
``alg = Tsit5()`` from ``test/BasicNeuralDE/neural_de_tests.jl:43``).

Pinned by (tests/test_oracle_tableau.py):
  * row sums of A equal c;
  * all 17 order-5 conditions on b (rooted trees of order <= 5);
  * btilde annihilates every tree of order <= 4 (b - btilde is 4th order);
  * dense output b_i(1) = b_i, b_i(0) = 0 and the continuous order-4
    conditions at interior theta;
  * global error slope 5 against scipy.linalg.expm on a linear ODE.
"""
import numpy as np

C = (0.0, 0.161, 0.327, 0.9, 0.9800255409045097, 1.0, 1.0)

# A[s] holds the coefficients a_{s+1, 1..s} (0-based stage s produces k_{s+1})
A = (
    (),
    (0.161,),
    (-0.008480655492356989, 0.335480655492357),
    (2.897153057105493, -6.359448489975075, 4.3622954328695815),
    (5.325864828439257, -11.748883564062828, 7.4955393428898365,
     -0.09249506636175525),
    (5.86145544294642, -12.92096931784711, 8.159367898576159,
     -0.071584973281401, -0.028269050394068383),
    (0.09646076681806523, 0.01, 0.4798896504144996, 1.379008574103742,
     -3.290069515436081, 2.324710524099774),
)

# 5th-order weights (FSAL: b = last row of A, b7 = 0)
B = A[6] + (0.0,)

# embedded error weights: utilde = dt * sum(BTILDE_i k_i)
BTILDE = (-0.00178001105222577714, -0.0008164344596567469,
          0.007880878010261995, -0.1447110071732629, 0.5823571654525552,
          -0.45808210592918697, 0.015151515151515152)

# free 4th-order dense output: b_1(t) = t(r11 + t(r12 + t(r13 + t r14)))
# and b_i(t) = t^2 (ri2 + t(ri3 + t ri4)) for i >= 2
R1 = (1.0, -2.763706197274826, 2.9132554618219126, -1.0530884977290216)
RI = (
    (0.13169999999999998, -0.2234, 0.1017),
    (3.9302962368947516, -5.941033872131505, 2.490627285651253),
    (-12.411077166933676, 30.33818863028232, -16.548102889244902),
    (37.50931341651104, -88.1789048947664, 47.37952196281928),
    (-27.896526289197286, 65.09189467479366, -34.8706578614966),
    (1.5, -4.0, 2.5),
)

ORDER = 5


def interp_weights(theta):
    """Dense-output weights b_i(theta), i = 1..7 (Python floats)."""
    t = float(theta)
    t2 = t * t
    w = [t * (R1[0] + t * (R1[1] + t * (R1[2] + t * R1[3])))]
    for r2, r3, r4 in RI:
        w.append(t2 * (r2 + t * (r3 + t * r4)))
    return w


def a_matrix():
    """Full 7x7 A as a numpy array (row s = coefficients of stage s)."""
    a = np.zeros((7, 7))
    for s, row in enumerate(A):
        a[s, :len(row)] = row
    return a
