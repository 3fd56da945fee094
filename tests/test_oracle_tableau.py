"""DOP54 tableau of the oracle: exact rationals, order conditions, row sums,
and the float slot layout of Zandpack/mpi_RK4pars_dev.py:121-147."""
from fractions import Fraction as Fr

import numpy as np

from oracle import zand_oracle as O


def test_order_conditions_exact():
    c, a, b5, b4 = O.dop54_rationals()
    # row sums c_i = sum_j a_ij  (the reference's own check, mpi_RK4pars_dev.py:203-205)
    for i in range(1, 7):
        assert sum(a[i], Fr(0)) == c[i]
    for b, order in ((b5, 5), (b4, 4)):
        assert sum(b) == 1
        for q in range(1, order):
            assert sum(bj * cj ** q for bj, cj in zip(b, c)) == Fr(1, q + 1)
    # order-4 condition sum_i b_i c_i sum_j a_ij c_j = 1/8, and b5 = last row (FSAL)
    s = sum(b5[i] * sum(a[i][j] * c[j] for j in range(i)) * c[i] for i in range(7))
    assert s == Fr(1, 8)
    assert list(a[6]) + [Fr(0)] == list(b5)


def test_float_layout():
    t = O.dop54_tableau()
    B, CT, CH = t["B"], t["CT"], t["CH"]
    assert B.shape == (7, 8) and CH.shape == (8,) and CT.shape == (7,)
    assert np.all(B[1:, 7] == 1.0) and B[0].sum() == 0
    assert abs(CT.sum()) < 1e-15 and CT[1] == 0.0
    assert CH[7] == 1.0 and abs(CH[:7].sum() - 1) < 1e-15
    np.testing.assert_allclose(t["A"], B[:, :7].sum(axis=1), atol=1e-15)


def test_hnew():
    assert O.hnew(1.0, 1.0, 1.0) == 0.9
    assert abs(O.hnew(0.1, 32e-6, 1e-6) - 0.9 * 0.1 * 0.5) < 1e-15
