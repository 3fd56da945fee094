"""DOP54 tableau of the oracle: exact rationals, order conditions, row sums,
and the float slot layout of Zandpack/mpi_RK4pars_dev.py:121-147."""
from fractions import Fraction as Fr

import numpy as np

from oracle import zand_oracle as O
from zandpack_amd import builder


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


def test_rkf45_order_conditions():
    """Fehlberg 4(5) of the in-process driver: row sums, fifth-order conditions of
    the step weights CH, fourth-order ones of the embedded C."""
    from fractions import Fraction as Fr
    c, a, ch, c4 = O.rkf45_rationals()
    for i in range(1, 6):
        assert sum(a[i]) == c[i]
    for q in range(5):
        assert sum(w * ci ** q for w, ci in zip(ch, c)) == Fr(1, q + 1)
    for q in range(4):
        assert sum(w * ci ** q for w, ci in zip(c4, c)) == Fr(1, q + 1)
    # the order-3 condition Σ b_i Σ_j a_ij c_j = 1/6
    s = sum(ch[i] * sum(a[i][j] * c[j] for j in range(i)) for i in range(6))
    assert s == Fr(1, 6)


def test_rkf45_von_neumann():
    """ξ = 0: the RKF45 loop follows σ(t) = U σ0 U† to its error level."""
    p = builder.config_small(seed=2)
    p.pulse = builder.Pulse()
    c = p.consts()
    c = {k: (np.zeros_like(v) if k in ("xi", "Ixi") else v) for k, v in c.items()}
    rng = np.random.default_rng(0)
    n = p.dims[1]
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    s0 = (A @ A.conj().T)[None] / n
    y0 = (s0, np.zeros_like(p.xi), p.zero_state()[2])
    out = O.rkf45_propagate(c, y0, 0.0, 0.5, 0.01, 1e-9, p.H0, O.zero_dH, p.bias, 2)
    T = float(np.sum(out["h_accepted"]))
    e, V = np.linalg.eigh(p.H0[0])
    U = (V * np.exp(-1j * e * T / O.HBAR)) @ V.conj().T
    ref = U @ s0[0] @ U.conj().T
    assert np.abs(out["y"][0][0] - ref).max() < 1e-7 * np.abs(ref).max()
    assert len(out["times"]) > 3 and out["times"][-1] <= 0.5


def test_scaled_error_norm():
    """The scaled RMS norm: with rtol = 0 it is the absolute norm over atol sqrt(N),
    and an entry-by-entry loop with max(|y0|, |y_new|) gives the same value."""
    p = builder.config_small(seed=8)
    rng = np.random.default_rng(8)
    y0 = tuple(0.05 * (rng.normal(size=a.shape) + 1j * rng.normal(size=a.shape))
               for a in p.zero_state())
    args = (p.consts(), y0, 0.1, 0.01, p.H0, p.dH, p.bias, 2)
    ks, te_abs, _, y_new = O.dop54_attempt(*args)
    N = sum(a.size for a in y0)
    _, te_a, _, _ = O.dop54_attempt(*args, rtol=0.0, atol=1e-7)
    assert abs(te_a - te_abs / (1e-7 * N ** 0.5)) < 1e-12 * te_a
    rtol, atol = 1e-3, 1e-6
    _, te_s, _, _ = O.dop54_attempt(*args, rtol=rtol, atol=atol)
    ct = O.dop54_tableau()["CT"]
    acc = 0.0
    for part in range(3):
        e = sum(ct[j] * ks[j][part] for j in range(7)).ravel()
        a, b = y0[part].ravel(), y_new[part].ravel()
        for i in range(e.size):
            sc = atol + rtol * max(abs(a[i]), abs(b[i]))
            acc += abs(e[i]) ** 2 / sc ** 2
    assert abs(te_s - (acc / N) ** 0.5) < 1e-12 * te_s
