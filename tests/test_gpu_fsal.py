"""First same as last (zh_set_fsal): an attempt after an accepted one takes that one's
k_7 = h f(t0 + h, y_new) as its k_1 (scaled to its own step) where f is the same there,
i.e. the stage parameters at its t0 equal those of that k_7 bit for bit.  Against the
same attempts with it off (to rounding; bitwise where it is not taken), the oracle's
adaptive propagation, and the cases that must evaluate the first stage: a new state, a
rejected step, a time-dependent bias at another time, instances deciding each alone."""
import sys

import numpy as np
import pytest

from oracle import zand_oracle as O
from oracle_side import dH_of
from zandpack_amd import builder

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _state(p, seed, scale=0.02):
    rng = np.random.default_rng(seed)
    NK, n, NL, NX, NOT = p.dims
    M = p.M
    psi = scale * (rng.normal(size=p.xi.shape) + 1j * rng.normal(size=p.xi.shape))
    A = rng.normal(size=(NK, M, M)) + 1j * rng.normal(size=(NK, M, M))
    om = scale * (A - A.conj().transpose(0, 2, 1))
    S = rng.normal(size=(NK, n, n)) + 1j * rng.normal(size=(NK, n, n))
    sig = p.sigma0 + 0.02 * (S + S.conj().transpose(0, 2, 1))
    return sig, psi, om.reshape(NK, NL, NX, NOT, NL, NX, NOT)


def _pulsed(p, kind, amp=0.2):
    p.pulse = builder.Pulse(kind, amp, 0.1, 0.05, (0.5, -0.5))
    return p


CASES = {
    "step": lambda: _pulsed(builder.config_small(seed=3), "step"),
    "gauss": lambda: _pulsed(builder.config_small(seed=4), "gauss"),
    "spin_hubbard": lambda: _pulsed(builder.config_small(NK=2, hubbard_U=2.0, seed=5), "step"),
    "c4_100": lambda: builder.config_c4_instance(100, seed=6),
}

STEPS = [0.010, 0.012, 0.009, 0.011, 0.010, 0.013, 0.008, 0.010]


def _run(p, precision, mode, t_next, seed=11, eps=1.0):
    """Eight forced attempts with varying steps from a random start: their results, the
    final state and how many took k_7."""
    from zandpack_amd.engine import Propagator
    a = Propagator(p, precision, state=_state(p, seed))
    a.set_fsal(mode)
    t, out = 0.05, []
    for h in STEPS:
        r = a.attempt(t, h, eps, force_accept=True)
        out.append(r)
        t = t_next(t, h)
    st, info = a.get_state(), a.fsal_info()
    a.close()
    return out, st, info


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", list(CASES))
def test_fsal_matches_full_evaluation(native, precision, case):
    """Chained steps (t += h): the seventh stage's time is the next start, so k_7 is taken
    at every attempt after the first, pulse or not; results agree with every stage
    evaluated (FP64 1e-12, mixed 1e-5 relative; error norms 1e-9, mixed 1e-3: the norm is
    a difference of nearly equal stage sums, and the reused k_1 and a new one differ in
    their FP32 rounding, which it shows at ~1e-4)."""
    p = CASES[case]()
    say("fsal on/off", case, precision, p.dims, p.M)
    on, s_on, i_on = _run(p, precision, "on", lambda t, h: t + h)
    off, s_off, i_off = _run(p, precision, "off", lambda t, h: t + h)
    assert i_on["on"] and not i_off["on"]
    assert i_off["taken"] == 0
    assert i_on["taken"] == len(STEPS) - 1
    tol_y, tol_te = (1e-12, 1e-9) if precision == "fp64" else (1e-5, 1e-3)
    for r1, r0 in zip(on, off):
        assert abs(r1["te"] - r0["te"]) <= tol_te * r0["te"], (r1["te"], r0["te"])
        assert rel(r1["J"], r0["J"]) < tol_y
    for x, y in zip(s_on, s_off):
        assert rel(x, y) < tol_y


def test_fsal_not_taken_where_f_differs(native):
    """A Gaussian bias and starts that skip ahead of the seventh stage's time (t += 1.5 h):
    the parameters differ, k_1 is evaluated, and the results are bitwise those with it
    off.  A constant bias (step after its onset) takes it at the same starts."""
    jump = lambda t, h: t + 1.5 * h  # noqa: E731
    on, s_on, i_on = _run(CASES["gauss"](), "fp64", "on", jump)
    off, s_off, _ = _run(CASES["gauss"](), "fp64", "off", jump)
    assert i_on["taken"] == 0
    for r1, r0 in zip(on, off):
        assert r1["te"] == r0["te"] and np.array_equal(r1["J"], r0["J"])
    for x, y in zip(s_on, s_off):
        assert np.array_equal(x, y)
    _, _, i_step = _run(CASES["step"](), "fp64", "on", jump)
    assert i_step["taken"] >= len(STEPS) - 2  # all but the start that crosses the onset


def test_fsal_after_new_state_and_rejection(native):
    """A new state and a rejected step: the next attempt evaluates its first stage; the
    one after an accepted step takes k_7 again.  Against the oracle (FP64 1e-12)."""
    from zandpack_amd.engine import Propagator
    p = CASES["spin_hubbard"]()
    y0 = _state(p, 21)
    a = Propagator(p, "fp64", state=y0)
    r = a.attempt(0.1, 0.01, 1.0)
    assert r["accepted"] and a.fsal_info()["taken"] == 0
    a.put_state(*y0)
    tA = 0.11
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, tA, 0.01, p.H0, dH_of(p), p.bias, 2)
    r = a.attempt(tA, 0.01, 1.0)
    assert a.fsal_info()["taken"] == 0  # a new state: k_1 evaluated
    assert abs(r["te"] - TE) <= 1e-10 * TE
    for got, w in zip(a.get_state(), y_new):
        assert rel(got, w) < 1e-12
    # the next one takes k_7, against the oracle from the same state
    y1 = a.get_state()
    tB = tA + 0.01  # the seventh stage's time of the step before
    ks, TE, J1, y2 = O.dop54_attempt(p.consts(), y1, tB, 0.007, p.H0, dH_of(p), p.bias, 2)
    r = a.attempt(tB, 0.007, 1.0)
    assert a.fsal_info()["taken"] == 1
    assert abs(r["te"] - TE) <= 1e-10 * TE and rel(r["J"], J1) < 1e-12
    yB = a.get_state()
    for got, w in zip(yB, y2):
        assert rel(got, w) < 1e-12
    # rejected (eps tiny): state kept, the retry evaluates its first stage
    tC = tB + 0.007
    r = a.attempt(tC, 0.007, 1e-30)
    assert not r["accepted"] and a.fsal_info()["taken"] == 2
    y3 = a.get_state()
    for got, w in zip(y3, yB):
        assert np.array_equal(got, w)
    ks, TE, J3, y4 = O.dop54_attempt(p.consts(), y3, tC, 0.005, p.H0, dH_of(p), p.bias, 2)
    r = a.attempt(tC, 0.005, 1.0)
    assert a.fsal_info()["taken"] == 2
    assert abs(r["te"] - TE) <= 1e-10 * TE and rel(r["J"], J3) < 1e-12
    for got, w in zip(a.get_state(), y4):
        assert rel(got, w) < 1e-12
    a.close()


@pytest.mark.parametrize("case,time_update", [("chain20", "correct"), ("step", "reference")])
def test_fsal_propagation_against_oracle(native, case, time_update):
    """The adaptive loop: config 1's chain (its bias ramps in, so only the corrected time
    update, t += h, starts where the last k_7 was formed) and a step bias with zand's
    update (t += h_next; constant after the onset).  k_7 is taken after most accepted
    steps, and the step sizes, times and currents are the oracle's (1e-10)."""
    from zandpack_amd.engine import Propagator
    p = builder.config_chain() if case == "chain20" else CASES["step"]()
    t1 = 2.0 if case == "chain20" else 1.0
    ref = O.propagate(p.consts(), p.zero_state(), 0.0, t1, 0.01, 1e-8, p.H0, dH_of(p), p.bias, 2,
                      time_update=time_update)
    prop = Propagator(p, "fp64")
    out = prop.propagate(0.0, t1, 0.01, 1e-8, time_update=time_update)
    info = prop.fsal_info()
    say("fsal propagate", case, time_update, out["n_steps"], out["n_rejected"], info)
    assert info["taken"] >= out["n_steps"] // 2
    assert out["n_steps"] == len(ref["times"]) - 1 and out["n_rejected"] == ref["n_reject"]
    np.testing.assert_allclose(out["times"], ref["times"], rtol=1e-10)
    np.testing.assert_allclose(out["h"], ref["h_accepted"], rtol=1e-10)
    assert rel(out["J"], ref["J"]) < 1e-10
    prop.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fsal_instances_decide_alone(native, precision):
    """Three instances (Gaussian biases, t += h), one rejecting now and then: each takes
    k_7 exactly when it would alone, and every result is bitwise that of the instance
    alone."""
    from zandpack_amd.engine import Instances, Propagator
    ps = [CASES["c4_100"]() for _ in range(3)]
    ys = [_state(p, 40 + g) for g, p in enumerate(ps)]
    grp = Instances(ps, precision, states=ys)
    alone = [Propagator(p, precision, state=y) for p, y in zip(ps, ys)]
    t = np.array([0.0, 0.1, 0.2])
    h = np.array([0.01, 0.02, 0.015])
    eps = np.array([1e-2, 1e-12, 1e-2])
    for it in range(6):
        rg = grp.attempt(t, h, eps)
        for g in range(3):
            ra = alone[g].attempt(t[g], h[g], eps[g])
            assert ra["te"] == rg[g]["te"] and np.array_equal(ra["J"], rg[g]["J"])
            assert ra["accepted"] == rg[g]["accepted"]
            if ra["accepted"]:
                t[g] += h[g]  # the step just taken: the next start is the k_7's time
            h[g] = ra["h_next"]
        eps[1] = 1e-12 if it % 2 else 1e-2
    taken = sum(a.fsal_info()["taken"] for a in alone)
    assert taken > 0 and grp.fsal_info()["taken"] == taken
    full = grp.get_state()
    for g in range(3):
        for x, y in zip(alone[g].get_state(), grp.instance_state(g, full)):
            assert np.array_equal(x, y)
        alone[g].close()
    grp.close()
