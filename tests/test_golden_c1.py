"""Config-1 fixture (tests/golden/c1_chain20.npz, written by make_c1_fixture.py):
the oracle must keep reproducing it (a regression pin of the whole driver), and
its physics is checked against invariants that do not come from the oracle."""
import os

import numpy as np

from oracle import zand_oracle as O
from zandpack_amd import builder

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    return np.load(os.path.join(HERE, "golden", "c1_chain20.npz"))


def test_fixture_invariants():
    f = load()
    t, J, h = f["times"], f["J"], f["h"]
    assert len(t) == len(J) + 1 == len(h) + 1 and t[-1] >= f["t_end"]
    # the reference advances t by the step proposed next (t_{k+1} - t_k = h_{k+1})
    np.testing.assert_allclose(np.diff(t)[:-1], h[1:], rtol=1e-11)
    assert np.all(f["TE"] < f["eps"])
    # σ stays Hermitian with its trace changed only by the lead currents
    for s in f["sigma_every"]:
        np.testing.assert_allclose(s, s.conj().transpose(0, 2, 1), atol=1e-12)
    p = builder.config_chain()
    n0 = np.trace(p.sigma0[0]).real
    ntr = np.array([np.trace(s[0]).real for s in f["sigma_every"]])
    assert np.all(np.abs(ntr - n0) < 1.0)
    # with V_R = 0 and V_L > 0 ramping up, charge flows in from the left lead
    assert J[-1, 0, 0] > 0


def test_oracle_reproduces_fixture():
    f = load()
    p = builder.config_chain()
    out = O.propagate(p.consts(), p.zero_state(), 0.0, float(f["t_end"]), float(f["h0"]),
                      float(f["eps"]), p.H0, p.dH, p.bias, 2)
    assert len(out["times"]) == len(f["times"])
    np.testing.assert_allclose(out["times"], f["times"], rtol=1e-12)
    np.testing.assert_allclose(out["J"], f["J"], rtol=0, atol=1e-10 * np.abs(f["J"]).max())
    np.testing.assert_allclose(out["sigma"][::int(f["every"])], f["sigma_every"], atol=1e-10)
