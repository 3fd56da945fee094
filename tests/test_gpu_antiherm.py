"""Ω held as its block upper triangle (zh_set_antihermitian): the Ω equation
(Zandpack/mpi/zand:784-809) keeps Ω anti-Hermitian when every mode has
ΔM_j Ixi_j = -conj(ΔP_j ξ_j) and χ⁻ = conj χ⁺, so the device computes the 64 x 64
blocks on and above the diagonal only.  Against the oracle (which computes every
entry): one evaluation on anti-Hermitian states (FP64 1e-12, mixed 2e-5), full attempts
(FP64 1e-12, TE 1e-10), adaptive propagations from the zero state (step counts, times
1e-10, J 1e-10), the steady state; against the same problem with the mode off; the
mode's selection and refusals.  Problems with several 64-blocks of device modes."""
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


def _complex_h(seed=8):
    p = builder.config_small(NK=2, seed=seed)
    rng = np.random.default_rng(seed)
    n = p.H0.shape[-1]
    A = rng.normal(size=(n, n))
    p.H0 = p.H0 + 0.2j * (A - A.T)[None]
    return p


CASES = {
    "small_nk2": lambda: builder.config_small(NK=2, seed=3),
    "complex_H": _complex_h,
    "chain20": builder.config_chain,
    "tile_aligned": lambda: builder.config_small(n_w=8, n_l=3, N_L=4, slots=4, NOT=8, seed=6),
    "c4_100": lambda: builder.config_c4_instance(100, seed=3),
}


def _anti_state(p, seed, scale=0.05):
    rng = np.random.default_rng(seed)
    NK, n, NL, NX, NOT = p.dims
    M = p.M
    psi = scale * (rng.normal(size=p.xi.shape) + 1j * rng.normal(size=p.xi.shape))
    A = rng.normal(size=(NK, M, M)) + 1j * rng.normal(size=(NK, M, M))
    om = scale * (A - A.conj().transpose(0, 2, 1))
    S = rng.normal(size=(NK, n, n)) + 1j * rng.normal(size=(NK, n, n))
    sig = p.sigma0 + 0.02 * (S + S.conj().transpose(0, 2, 1))
    return sig, psi, om.reshape(NK, NL, NX, NOT, NL, NX, NOT)


def _antiherm_err(om, p):
    o = om.reshape(p.dims[0], p.M, p.M)
    return np.abs(o + o.conj().transpose(0, 2, 1)).max() / max(np.abs(o).max(), 1e-300)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("layout", ["reference", "pruned"])
def test_rhs_antihermitian(native, case, precision, layout):
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    sig, psi, omg = _anti_state(p, 5)
    delta = np.array([0.05, -0.03])
    say("antiherm rhs", case, precision, layout, p.dims, p.M)
    ref = O.rhs(p.consts(), sig, psi, omg, p.H0, delta)
    Jref = O.currents(p.consts(), psi)
    a = Propagator(p, precision, layout=layout)
    info = a.antihermitian_info()
    assert info["constants"] and info["active"], info  # σ0 with zero Ψ, Ω: active
    ds, dp, do, J = a.rhs(p.H0, delta, sig, psi, omg)
    a.close()
    b = Propagator(p, precision, layout=layout, antihermitian="off")
    assert not b.antihermitian_info()["active"]
    full = b.rhs(p.H0, delta, sig, psi, omg)
    b.close()
    tol = 1e-12 if precision == "fp64" else 2e-5
    for got, want in zip((ds, dp, do), ref):
        assert rel(got, want) < tol, (case, rel(got, want))
    assert rel(J, Jref) < (1e-12 if precision == "fp64" else 1e-5)
    # dΩ comes back in full and anti-Hermitian (the diagonal 64-blocks are computed in
    # full, so their two triangles agree to rounding); the full evaluation agrees
    assert _antiherm_err(do, p) < (1e-13 if precision == "fp64" else 1e-5)
    for got, want in zip((ds, dp, do), full[:3]):
        assert rel(got, want) < tol


def test_rhs_non_antihermitian_input_runs_in_full(native):
    """An input Ω without the symmetry is evaluated in full, whatever the mode."""
    from zandpack_amd.engine import Propagator
    p = CASES["chain20"]()
    sig, psi, omg = _anti_state(p, 6)
    rng = np.random.default_rng(1)
    omg = omg + 0.01 * rng.normal(size=omg.shape)
    delta = np.array([0.02, -0.01])
    say("antiherm rhs non-anti input")
    ref = O.rhs(p.consts(), sig, psi, omg, p.H0, delta)
    a = Propagator(p, "fp64")
    out = a.rhs(p.H0, delta, sig, psi, omg)
    assert a.antihermitian_info()["active"]  # the resident state is untouched
    a.close()
    for got, want in zip(out[:3], ref):
        assert rel(got, want) < 1e-12


@pytest.mark.parametrize("case", ["small_nk2", "complex_H", "chain20", "tile_aligned", "c4_100"])
@pytest.mark.parametrize("layout", ["reference", "pruned"])
def test_attempt_antihermitian_fp64(native, case, layout):
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    p.pulse = builder.Pulse("step", amp=0.2, p1=0.0, weights=(0.5, -0.5))
    y0 = _anti_state(p, 11, scale=0.01)
    say("antiherm attempt", case, layout)
    prop = Propagator(p, "fp64", state=y0, layout=layout, antihermitian="on")
    t0, h = 0.3, 0.01
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, t0, h, p.H0, dH_of(p), p.bias, p.dims[2])
    r = prop.attempt(t0, h, 1.0, force_accept=True)
    assert abs(r["te"] - TE) <= 1e-10 * TE
    assert rel(r["J"], J0) < 1e-12
    for got, w in zip(prop.get_state(), y_new):
        assert rel(got, w) < 1e-12
    assert _antiherm_err(prop.get_state()[2], p) < 1e-13
    prop.close()


@pytest.mark.parametrize("case", ["small_nk2", "chain20", "c4_100"])
def test_attempt_antihermitian_mixed(native, case):
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    y0 = _anti_state(p, 12, scale=0.01)
    say("antiherm attempt mixed", case)
    t0, h = 0.3, 0.01
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, t0, h, p.H0, dH_of(p), p.bias, p.dims[2])
    prop = Propagator(p, "mixed", state=y0, antihermitian="on")
    r = prop.attempt(t0, h, 1.0, force_accept=True)
    got = prop.get_state()
    prop.close()
    assert abs(r["te"] - TE) <= 2e-3 * TE
    assert rel(r["J"], J0) < 1e-5
    for g, w in zip(got, y_new):
        assert rel(g, w) < 1e-5


@pytest.mark.parametrize("case", ["chain20", "c4_100"])
def test_propagate_antihermitian_vs_oracle_and_full(native, case):
    """From the zero state (where auto selects the mode) over 0.5 fs: the oracle's
    accepted steps, times and currents; and the same run with the mode off."""
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    p.pulse = builder.Pulse("step", amp=0.2, p1=0.0, weights=(0.5, -0.5))
    y0 = p.zero_state()
    say("antiherm propagate", case)
    ref = O.propagate(p.consts(), y0, 0.0, 0.5, 0.01, 1e-8, p.H0, dH_of(p), p.bias, p.dims[2])
    outs = []
    for mode in ("auto", "off"):
        prop = Propagator(p, "fp64", antihermitian=mode)
        assert prop.antihermitian_info()["active"] == (mode == "auto")
        out = prop.propagate(0.0, 0.5, 0.01, 1e-8)
        s, ps, om = prop.get_state()
        prop.close()
        assert out["n_steps"] == len(ref["times"]) - 1
        assert out["n_rejected"] == ref["n_reject"]
        np.testing.assert_allclose(out["times"], ref["times"], rtol=1e-10)
        assert rel(out["J"], ref["J"]) < 1e-10
        for g, w in zip((s, ps, om), ref["y"]):
            assert rel(g, w) < 1e-10
        outs.append(out)
    # the two forms round differently (Ωξ in two passes, the error norm's order)
    np.testing.assert_allclose(outs[0]["J"], outs[1]["J"], rtol=0, atol=1e-11 * np.abs(outs[1]["J"]).max())


def test_switching_off_fills_the_lower_triangle(native):
    """Steps with the mode on, then off (the resident lower triangle filled in on the
    device), then on again: the run continues as the full one does."""
    from zandpack_amd.engine import Propagator
    p = CASES["c4_100"]()
    y0 = _anti_state(p, 21, scale=0.01)
    say("antiherm switch")
    a = Propagator(p, "fp64", state=y0, antihermitian="on")
    b = Propagator(p, "fp64", state=y0, antihermitian="off")
    t, h = 0.0, 0.01
    for i in range(6):
        if i == 2:
            a.set_antihermitian("off")
            assert not a.antihermitian_info()["active"]
        if i == 4:
            a.set_antihermitian("on")
            assert a.antihermitian_info()["active"]
        ra = a.attempt(t, h, 1.0, force_accept=True)
        rb = b.attempt(t, h, 1.0, force_accept=True)
        assert abs(ra["te"] - rb["te"]) <= 1e-10 * rb["te"]
        assert rel(ra["J"], rb["J"]) < 1e-12
        t += h
    for g, w in zip(a.get_state(), b.get_state()):
        assert rel(g, w) < 1e-12
    a.close()
    b.close()


def test_selection_and_refusals(native):
    from zandpack_amd._native import ZHError
    from zandpack_amd.engine import Propagator
    p = CASES["chain20"]()
    sig, psi, omg = _anti_state(p, 3)
    say("antiherm selection")
    prop = Propagator(p, "fp64")
    assert prop.antihermitian_info() == {"active": True, "constants": True, "state": True}
    # a state without the symmetry: auto falls back to the full form, "on" is refused
    bad = omg + 0.01
    prop.put_state(sig, psi, bad)
    assert prop.antihermitian_info() == {"active": False, "constants": True, "state": False}
    with pytest.raises(ZHError, match="not anti-Hermitian"):
        prop.set_antihermitian("on")
    np.testing.assert_array_equal(prop.get_state()[2], bad)
    prop.put_state(sig, psi, omg)
    assert prop.antihermitian_info()["active"]
    prop.set_antihermitian("on")
    with pytest.raises(ZHError, match="not anti-Hermitian"):
        prop.put_state(sig, psi, bad)
    prop.close()
    # constants without the symmetry (one active mode's Ixi not conj(ξ))
    q = CASES["chain20"]()
    act = np.argwhere(q.dGP[0] != 0)[0]
    Ixi = q.Ixi.copy()
    Ixi[(0, *act)] *= 1.5
    q.Ixi = Ixi
    prop = Propagator(q, "fp64")
    assert prop.antihermitian_info() == {"active": False, "constants": False, "state": True}
    with pytest.raises(ZHError, match="constants"):
        prop.set_antihermitian("on")
    prop.close()


def test_steady_state_antihermitian(native):
    """psinought's Ψ₀ and Ω₀ with Ω held as its triangle agree with the full form."""
    from zandpack_amd.engine import Propagator
    p = CASES["c4_100"]()
    say("antiherm steady")
    out = []
    for mode in ("auto", "off"):
        prop = Propagator(p, "fp64", antihermitian=mode)
        assert prop.antihermitian_info()["active"] == (mode == "auto")
        prop.steady_state(0.0)
        out.append(prop.get_state())
        prop.close()
    for g, w in zip(out[0], out[1]):
        assert rel(g, w) < 1e-10
    assert _antiherm_err(out[0][2], p) < 1e-13


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["chain20", "c4_100"])
def test_every_tile_candidate_antihermitian(native, precision, case):
    """Every enabled tile candidate (the mirrored X ξ_c product has its own list) gives
    the oracle's attempt with Ω held as its triangle; the autotuner's choice is one of
    them."""
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    y0 = _anti_state(p, 31, scale=0.02)
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, 0.1, 0.01, p.H0, dH_of(p), p.bias, 2)
    tol = 1e-11 if precision == "fp64" else 1e-4
    ncand = 2 if precision == "fp64" else 3
    prop = Propagator(p, precision, state=y0, antihermitian="on")
    tiles = prop.autotune(force=True)
    assert prop.antihermitian_info()["active"]
    assert all(0 <= v < ncand for v in tiles.values())
    for c in range(ncand):
        say("antiherm tiles", case, precision, c)
        prop.put_state(*y0)
        prop.set_tiles({k: c for k in tiles})
        r = prop.attempt(0.1, 0.01, 1.0, force_accept=True)
        assert abs(r["te"] - TE) <= tol * TE, (c, r["te"], TE)
        for got, w in zip(prop.get_state(), y_new):
            assert rel(got, w) < (1e-12 if precision == "fp64" else 1e-5), c
    prop.close()
