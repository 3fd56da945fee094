"""The in-process drivers' run files, resume and tolerance-error branch
(Zandpack/TimedependentTransport.py:3226-3235, 3258-3281, 3343-3390), driven by an
``f`` over the oracle's right-hand side so that the host logic runs without a GPU."""
import numpy as np
import pytest

from oracle import zand_oracle as O
from zandpack_amd import builder, drop_in


def oracle_f(p):
    consts = p.consts()

    def f(t, sig, psi, omg, dH, delta_variant, dH_given=True):
        Ht = p.H0 + dH(t, sig) if dH_given else dH(t, sig)
        delta = np.array([delta_variant(t, a) for a in range(2)])
        f.last_J = O.currents(consts, psi)
        return O.rhs(consts, sig, psi, omg, Ht, delta)

    f.last_J = None
    return f


@pytest.fixture(scope="module")
def case():
    return builder.config_small(seed=3)


@pytest.mark.parametrize("driver", ["AdaptiveRK4", "AdaptiveDOP"])
def test_run_files(driver, case, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    p = case
    drv = getattr(drop_in, driver)
    seen = []

    def write_func(fh, t, s, ps, om):
        seen.append(t)
        fh.write(f"hook {t}\n")

    times, data = drv(oracle_f(p), *p.zero_state(), 1e-7, 0.0, 0.12, p.dH, p.bias, p.Ixi,
                      h_guess=0.01, name="run", print_step=2, write_func=write_func)
    n = len(times)
    assert n >= 4
    # snapshots are taken after accepted steps 0, 2, 4, ... (the last one at step n_last)
    n_last = 2 * ((n - 1) // 2)
    t_saved = np.load("_times.npy")
    np.testing.assert_array_equal(t_saved, times[: n_last + 1])
    ne = np.load("_#electrons_device.npy")
    tr = np.trace(data["density matrix"][: n_last + 1], axis1=2, axis2=3)
    np.testing.assert_allclose(ne, tr, rtol=0, atol=0)
    assert float(np.load("run_last_time.npy")) == times[n_last]
    if n_last == n - 1:
        np.testing.assert_array_equal(np.load("run_last_sig.npy"), data["state"][0])
    stems = ("_current_left", "_current_right") if driver == "AdaptiveRK4" else ("_JL", "_JR")
    for stem, key in zip(stems, ("current_left", "current_right")):
        np.testing.assert_array_equal(np.load(stem + ".npy"), np.array(data[key][: n_last + 1]))
    log = open("run.txt").read()
    assert log.startswith("\n\n\n\n\nStart (Wait for compilation)\n")
    assert log.count(" %\n") == n_last // 2 + 1 == len(seen)
    assert log.count("current timestep: ") == n_last // 2 + 1
    assert "100%\nRuntime: " in log
    # the hook receives the start time of the step just taken
    for t in seen:
        assert np.min(np.abs(times - t)) < 1e-12


def test_no_files_when_disabled(case, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    p = case
    drop_in.AdaptiveRK4(oracle_f(p), *p.zero_state(), 1e-7, 0.0, 0.03, p.dH, p.bias, p.Ixi,
                        h_guess=0.01, print_to_file=False)
    assert list(tmp_path.iterdir()) == []


def test_resume_from_last_files(case, tmp_path, monkeypatch):
    """A missing state is read back from <name>_last_*.npy and a missing t0 from
    <name>_last_time.npy (the start time of the last saved step)."""
    monkeypatch.chdir(tmp_path)
    p = case
    f = oracle_f(p)
    times, data = drop_in.AdaptiveRK4(f, *p.zero_state(), 1e-7, 0.0, 0.05, p.dH, p.bias,
                                      p.Ixi, h_guess=0.01, name="r", print_step=1)
    s_last = np.load("r_last_sig.npy")
    np.testing.assert_array_equal(s_last, data["state"][0])
    t2, d2 = drop_in.AdaptiveRK4(f, None, None, None, 1e-7, None, 0.08, p.dH, p.bias, p.Ixi,
                                 h_guess=0.01, name="r", print_step=1)
    assert t2[0] == times[-1]
    np.testing.assert_array_equal(d2["density matrix"][0], s_last)


def test_tolerance_error_branch(case, tmp_path, monkeypatch):
    """With atol/rtol the error is 1 ** max(f) = 1: no step is accepted for eps < 1,
    and for eps >= 1 every attempt is accepted with h <- 0.9 h eps^(1/5)."""
    monkeypatch.chdir(tmp_path)
    p = case
    f = oracle_f(p)
    tol = dict(atol=(1e-9, 1e-9, 1e-9), rtol=(1e-6, 1e-6, 1e-6))
    assert drop_in._tolerance_error([np.full(3, np.nan)] * 3, [np.zeros(3)] * 3,
                                    (0, 0, 0), (1, 1, 1)) == 1.0
    with pytest.raises(RuntimeError):
        drop_in.AdaptiveRK4(f, *p.zero_state(), 0.5, 0.0, 0.05, p.dH, p.bias, p.Ixi,
                            h_guess=0.01, print_to_file=False, max_attempts=5, **tol)
    times, _ = drop_in.AdaptiveRK4(f, *p.zero_state(), 2.0, 0.0, 0.05, p.dH, p.bias, p.Ixi,
                                   h_guess=0.01, print_to_file=False, **tol)
    h = 0.01
    expect = [0.0]
    while True:
        h = 0.9 * h * 2.0 ** 0.2
        if expect[-1] + h > 0.05:
            break
        expect.append(expect[-1] + h)
    np.testing.assert_allclose(times, expect, rtol=1e-12)
