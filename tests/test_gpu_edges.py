"""Edge shapes through the C ABI against the oracle: one lead, three leads (one
sharing orbitals), a single mode per lead, no coupling at all (ξ = 0: von
Neumann evolution, J = 0), odd sizes, and argument errors.  Bars as in the
parity tests: FP64 1e-12 for one evaluation, mixed 2e-5."""
import sys

import numpy as np
import pytest

from oracle import zand_oracle as O
from oracle_side import dH_of, diss_of  # noqa: F401
from zandpack_amd import builder

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _leads(p, idx, weights):
    """The problem restricted / extended to the leads ``idx`` of p (repeats allowed)."""
    sel = lambda a: np.ascontiguousarray(np.asarray(a)[:, idx])
    q = builder.AMEProblem(H0=p.H0, sigma0=p.sigma0, xi=sel(p.xi), Ixi=sel(p.Ixi),
                           GG_P=sel(p.GG_P), GL_P=sel(p.GL_P), GG_M=sel(p.GG_M),
                           GL_M=sel(p.GL_M), Xpp=sel(p.Xpp), Xpm=sel(p.Xpm),
                           pulse=builder.Pulse("gauss", 0.3, 1.0, 0.8, tuple(weights)))
    return q


def _uncoupled(p):
    q = _leads(p, [0, 1], (0.5, -0.5))
    q.xi = np.zeros_like(q.xi)
    q.Ixi = np.zeros_like(q.Ixi)
    return q


CASES = {
    "one_lead": lambda: _leads(builder.config_small(seed=1), [0], (0.7,)),
    "three_leads": lambda: _leads(builder.config_small(seed=2), [0, 1, 0], (0.5, -0.5, 0.2)),
    "one_channel": lambda: builder.config_small(N_L=1, slots=2, NOT=1, seed=3),
    "odd_sizes": lambda: builder.config_small(n_w=3, n_l=3, N_L=3, slots=2, NOT=3, seed=4),
    "uncoupled": lambda: _uncoupled(builder.config_small(seed=5)),
}


def _rand_state(p, seed, scale=0.05):
    rng = np.random.default_rng(seed)
    NK, n, NL, NX, NOT = p.dims
    psi = scale * (rng.normal(size=p.xi.shape) + 1j * rng.normal(size=p.xi.shape))
    shp = (NK, NL, NX, NOT, NL, NX, NOT)
    omg = scale * (rng.normal(size=shp) + 1j * rng.normal(size=shp))
    A = rng.normal(size=(NK, n, n)) + 1j * rng.normal(size=(NK, n, n))
    sig = p.sigma0 + 0.05 * (A + A.conj().transpose(0, 2, 1))
    return sig, psi, omg


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_edge_rhs(native, case, precision):
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    sig, psi, omg = _rand_state(p, 7)
    NL = p.dims[2]
    delta = np.linspace(0.05, -0.05, NL)
    say("edge rhs", case, precision, p.dims)
    ref = O.rhs(p.consts(), sig, psi, omg, p.H0, delta)
    prop = Propagator(p, precision)
    got = prop.rhs(p.H0, delta, sig, psi, omg)
    tol = 1e-12 if precision == "fp64" else 2e-5
    for g, w in zip(got[:3], ref):
        assert rel(g, w) < tol, (case, rel(g, w))
    Jref = O.currents(p.consts(), psi)
    if case == "uncoupled":
        assert np.abs(got[3]).max() == 0.0
    else:
        assert rel(got[3], Jref) < (1e-12 if precision == "fp64" else 1e-5)
    prop.close()


@pytest.mark.parametrize("case", list(CASES))
def test_edge_attempt(native, case):
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    y0 = _rand_state(p, 11, scale=0.02)
    say("edge attempt", case, p.dims)
    prop = Propagator(p, "fp64", state=y0)
    t0, h, eps = 0.2, 0.02, 1e-3
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, t0, h, p.H0, dH_of(p), p.bias, p.dims[2])
    r = prop.attempt(t0, h, eps)
    assert abs(r["te"] - TE) <= 1e-10 * TE
    want = y_new if r["accepted"] else y0
    for g, w in zip(prop.get_state(), want):
        assert rel(g, w) < 1e-12
    prop.close()


@pytest.mark.parametrize("case", list(CASES))
def test_edge_attempt_mixed(native, case):
    """The same attempt in mixed precision (odd orbital counts take the one-element
    stage combination and error-norm kernels in complex64, E6 included), then a second
    one that reuses k_7 (t += h): error norms within 1e-3, the new state within 2e-5
    of the oracle's."""
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    y0 = _rand_state(p, 11, scale=0.02)
    say("edge attempt mixed", case, p.dims)
    prop = Propagator(p, "mixed", state=y0)
    t0, h = 0.2, 0.02
    NL = p.dims[2]
    ks, TE, J0, y1 = O.dop54_attempt(p.consts(), y0, t0, h, p.H0, dH_of(p), p.bias, NL)
    r = prop.attempt(t0, h, 1.0, force_accept=True)
    assert abs(r["te"] - TE) <= 1e-3 * TE, (r["te"], TE)
    for g, w in zip(prop.get_state(), y1):
        assert rel(g, w) < 2e-5
    ks, TE2, J1, y2 = O.dop54_attempt(p.consts(), y1, t0 + h, h, p.H0, dH_of(p), p.bias, NL)
    r = prop.attempt(t0 + h, h, 1.0, force_accept=True)
    assert abs(r["te"] - TE2) <= 1e-3 * TE2, (r["te"], TE2)
    if TE2 > 0:
        for g, w in zip(prop.get_state(), y2):
            assert rel(g, w) < 2e-5
    prop.close()


def test_uncoupled_is_von_neumann(native):
    """ξ = 0: σ(t) = U σ₀ U† with U = exp(-iHt/ħ), J = 0, Ψ and Ω stay zero."""
    from zandpack_amd.engine import Propagator
    p = CASES["uncoupled"]()
    p.pulse = builder.Pulse()
    rng = np.random.default_rng(0)
    n = p.dims[1]
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    s0 = (A @ A.conj().T)[None] / n
    say("edge von Neumann")
    prop = Propagator(p, "fp64", state=(s0, np.zeros_like(p.xi), p.zero_state()[2]))
    out = prop.propagate(0.0, 1.0, 0.01, 1e-10)
    e, V = np.linalg.eigh(p.H0[0])
    # the integrated time is the sum of the accepted steps (the reported times follow
    # zand's t += h_next bookkeeping)
    t = float(np.sum(out["h"]))
    U = (V * np.exp(-1j * e * t / p.hbar)) @ V.conj().T
    err = rel(prop.sigma()[0], U @ s0[0] @ U.conj().T)
    assert err < 1e-8, err
    assert np.abs(out["J"]).max() == 0.0
    s, ps, om = prop.get_state()
    assert np.abs(ps).max() == 0.0 and np.abs(om).max() == 0.0
    prop.close()


def test_argument_errors(native):
    from zandpack_amd.engine import Propagator
    from zandpack_amd import _native as N
    p = builder.config_small(seed=1)
    say("edge errors")
    prop = Propagator(p, "fp64")
    with pytest.raises(Exception, match="step size"):
        prop.attempt(0.0, 0.0, 1e-6)
    with pytest.raises(Exception, match="step size"):
        prop.attempt(0.0, float("nan"), 1e-6)
    with pytest.raises(Exception):
        prop.propagate(0.0, 1.0, -0.01, 1e-6)
    prop.close()
    q = _leads(builder.config_small(seed=1), [0] * 9, (0.1,) * 9)
    with pytest.raises(Exception):
        Propagator(q, "fp64")
