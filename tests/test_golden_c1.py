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


def start(f):
    return f["sigma0"], f["psi0"], f["omega0"]


def test_fixture_invariants():
    f = load()
    t, J, h = f["times"], f["J"], f["h"]
    assert len(t) == len(J) + 1 == len(h) + 1 and t[-1] >= f["t_end"]
    # the reference advances t by the step proposed next, h_next = 0.9 h (eps/TE)^(1/5)
    # of the accepted attempt (mpi/zand:396-407, 427), whether or not the next attempt
    # is then rejected
    np.testing.assert_allclose(np.diff(t), O.hnew(h, f["TE"], float(f["eps"])), rtol=1e-12)
    assert int(f["n_reject"]) > 0  # the run exercises rejections
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
    # the run starts in equilibrium: no current at t = 0, σ₀ a density (occupations in
    # [0, 1]) that the equations of motion leave unchanged at zero bias
    assert np.abs(J[0]).max() < 1e-10 * np.abs(J).max()
    e = np.linalg.eigvalsh(f["sigma0"][0])
    assert e.min() > -1e-10 and e.max() < 1 + 1e-10
    ds, dp, do = O.rhs(p.consts(), *start(f), p.H0, np.zeros(2))
    assert np.abs(ds).max() < 1e-12 and np.abs(dp).max() < 1e-12 * max(1, np.abs(f["psi0"]).max())


def test_start_is_the_equilibrium_steady_state():
    """σ₀ is the equilibrium density of the chain's own Fermi expansion and Ψ₀, Ω₀
    psinought's steady state for it (the SCF and psinought equivalents)."""
    f = load()
    p = builder.config_chain()
    sig = O.equilibrium_density(p.consts(), p.H0, float(f["kT"]), int(f["N_F"]))
    np.testing.assert_allclose(f["sigma0"], sig, atol=1e-13)
    psi, omg = O.steady_state_dense(p.consts(), np.asarray(p.H0, complex), sig, np.zeros(2))
    np.testing.assert_allclose(f["psi0"], psi, atol=1e-12)


def test_oracle_reproduces_fixture():
    f = load()
    p = builder.config_chain()
    out = O.propagate(p.consts(), start(f), 0.0, float(f["t_end"]), float(f["h0"]),
                      float(f["eps"]), p.H0, O.zero_dH, p.bias, 2)
    assert len(out["times"]) == len(f["times"])
    np.testing.assert_allclose(out["times"], f["times"], rtol=1e-12)
    np.testing.assert_allclose(out["J"], f["J"], rtol=0, atol=1e-10 * np.abs(f["J"]).max())
    np.testing.assert_allclose(out["sigma"][::int(f["every"])], f["sigma_every"], atol=1e-10)
