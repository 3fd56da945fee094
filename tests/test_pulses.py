"""Host pulse library (the device evaluates the same forms): closed-form limits of
the generic and stairs pulses, and the data-driven pulses through their measured
samples (tests/golden/*_pulse.npy are the reference's ExpPulses tables, kept as
data)."""
import os

import numpy as np
import pytest

from zandpack_amd import builder

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_generic_pulse_limits():
    q = builder.Pulse("generic", 0.7, 0.4, 1.3, (0.5, -0.5), 0.25, 2.0)  # s, f, phase, d
    # at the delay: amp sin(-phase); a few widths later the envelope has decayed
    assert np.isclose(q.value(2.0), -0.7 * np.sin(0.25))
    assert abs(q.value(2.0 + 5 * 0.4)) < 1e-9
    # a zero of the carrier: 2π f (t-d) = phase
    t0 = 2.0 + 0.25 / (2 * np.pi * 1.3)
    assert abs(q.value(t0)) < 1e-14


def test_stairs_limits():
    p = builder.Pulse("stairs", 0.2, 1.5, 5, (0.5, -0.5), 0.05)      # w, n, soft
    assert abs(p.value(-10.0)) < 1e-12
    assert np.isclose(p.value(100.0), 0.2 * 4)
    # half of the i-th stair at its centre, the earlier ones complete
    assert np.isclose(p.value(3.0), 0.2 * (1 + 0.5), atol=1e-12)


@pytest.mark.parametrize("name", ["air_photonics", "toptica"])
def test_experimental_pulse_through_samples(name):
    tab = np.load(os.path.join(GOLD, f"{name}_pulse.npy"))
    p = builder.experimental_pulse(name, tab, amp=2.0)
    spec = builder.EXPERIMENTAL_PULSES[name]
    lo, hi = spec["rows"]
    t = tab[lo:hi, 0] * 1000.0
    v = tab[lo:hi, 1] / np.abs(tab[lo:hi, 1]).max()
    t = t - t.min()
    a, b, ts = spec["window"]
    win = lambda x: (np.tanh((x - a) / ts) - np.tanh((x - b) / ts)) / 2
    for tk, vk in zip(t[:-1], v[:-1]):
        assert np.isclose(p.value(tk), 2.0 * vk * win(tk), rtol=1e-12, atol=1e-15)
    assert p.value(t[0] - 1e-3) == 0.0 and p.value(t[-1] + 1e-3) == 0.0
    # inside an interval: the cubic is scipy's spline
    from scipy.interpolate import CubicSpline
    cs = CubicSpline(t, v)
    tm = 0.5 * (t[3] + t[4])
    assert np.isclose(p.value(tm), 2.0 * cs(tm) * win(tm), rtol=1e-12)
