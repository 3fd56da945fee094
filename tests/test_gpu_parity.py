"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle on the
same seeded inputs.  FP64 bar: 1e-12 relative (max-norm) for one evaluation,
1e-10 for propagated currents; mixed precision (complex64 Ψ/Ω, FP32 MFMA) bar:
1e-5 for one evaluation, 1e-6 relative on J(t) over a propagation."""
import os
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


def _rand_state(p, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    NK, n, NL, NX, NOT = p.dims
    psi = scale * (rng.normal(size=p.xi.shape) + 1j * rng.normal(size=p.xi.shape))
    shp = (NK, NL, NX, NOT, NL, NX, NOT)
    omg = scale * (rng.normal(size=shp) + 1j * rng.normal(size=shp))
    A = rng.normal(size=(NK, n, n)) + 1j * rng.normal(size=(NK, n, n))
    sig = p.sigma0 + 0.05 * (A + A.conj().transpose(0, 2, 1))
    return sig, psi, omg


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("mixed", 3e-6)])
@pytest.mark.parametrize("shape", [(37, 53, 29), (64, 64, 16), (130, 70, 200), (5, 3, 1)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_building_block(native, precision, tol, shape, ta, tb):
    from zandpack_amd.engine import selftest_gemm
    M, N, K = shape
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    A = rng.normal(size=(M, K)) + 1j * rng.normal(size=(M, K))
    B = rng.normal(size=(K, N)) + 1j * rng.normal(size=(K, N))
    say("gemm", precision, shape, ta, tb)
    C = selftest_gemm(A.T.copy() if ta else A, B.T.copy() if tb else B, bool(ta), bool(tb),
                      precision)
    ref = A @ B
    assert np.abs(C - ref).max() <= tol * np.sqrt(K) * np.abs(A).max() * np.abs(B).max() * 4


def _complex_h(seed=8):
    """Small case with a complex Hermitian Hamiltonian (Peierls-like phases), so
    the general complex H products run instead of the real-H fast path."""
    p = builder.config_small(seed=seed)
    rng = np.random.default_rng(seed)
    n = p.H0.shape[-1]
    A = rng.normal(size=(n, n))
    p.H0 = p.H0 + 0.2j * (A - A.T)[None]
    return p


CASES = {
    "small": lambda: builder.config_small(seed=1),
    "complex_H": _complex_h,
    # 2 leads x 4 Lorentzians x 8 channels = 64 active modes: auto picks the
    # active-first order with no padding
    "tile_aligned": lambda: builder.config_small(n_w=8, n_l=3, N_L=4, slots=4, NOT=8, seed=6),
    "small_spin_hubbard": lambda: builder.config_small(NK=2, hubbard_U=2.0, seed=2),
    "chain20": lambda: builder.config_chain(),
    "wide": lambda: builder.config_small(n_w=6, n_l=7, N_L=5, slots=6, NOT=3, seed=4),
    # 72 coupled orbitals: the 80-column tiles of the Ωξ and Π products
    "wide_edges": lambda: builder.config_small(n_w=36, n_l=3, N_L=2, slots=2, NOT=4, seed=9),
    # a one-column ribbon: both leads couple the same orbitals (no lead-aligned tiles)
    "shared_orbitals": lambda: builder.config_small(n_w=4, n_l=1, seed=13),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("layout", ["reference", "pruned"])
@pytest.mark.parametrize("hproducts", ["dense", "sparse"])
def test_rhs_parity(native, case, precision, layout, hproducts):
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    sig, psi, omg = _rand_state(p, 7)
    Ht = p.H0 + dH_of(p)(0.2, sig)
    delta = np.array([p.bias(0.2, a) for a in range(2)]) + np.array([0.05, -0.03])
    ref = O.rhs(p.consts(), sig, psi, omg, Ht, delta)
    Jref = O.currents(p.consts(), psi)
    say("rhs", case, precision, layout, hproducts, p.dims)
    prop = Propagator(p, precision, layout=layout, hproducts=hproducts)
    ds, dp, do, J = prop.rhs(Ht, delta, sig, psi, omg)
    assert prop.hamiltonian_info()["sparse"] == (hproducts == "sparse")
    tol = 1e-12 if precision == "fp64" else 2e-5
    for got, want in zip((ds, dp, do), ref):
        assert rel(got, want) < tol, (case, precision, rel(got, want))
    assert rel(J, Jref) < (1e-12 if precision == "fp64" else 1e-5)
    prop.close()


@pytest.mark.parametrize("case", ["small", "complex_H", "small_spin_hubbard", "chain20",
                                  "tile_aligned", "shared_orbitals"])
@pytest.mark.parametrize("layout", ["reference", "pruned"])
@pytest.mark.parametrize("zero_ff", [False, True])
def test_attempt_parity(native, case, layout, zero_ff):
    """One attempt from a random state; with zero_ff the Fermi x Fermi block of Ω
    starts at zero, which the active-first order skips."""
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    y0 = _rand_state(p, 11, scale=0.02)
    if zero_ff:
        passive = (p.dGP == 0) & (p.dGM == 0)
        pm = passive.all(axis=0).reshape(-1)
        om = y0[2].reshape(p.dims[0], p.M, p.M)
        om[:, pm[:, None] & pm[None, :]] = 0
    say("attempt", case, layout, zero_ff)
    prop = Propagator(p, "fp64", state=y0, layout=layout)
    info = prop.layout_info()
    assert info["active_first"] == (layout == "pruned")
    assert info["skip_passive_block"] == (layout == "pruned" and zero_ff)
    t0, h, eps = 0.3, 0.02, 1e-3
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, t0, h, p.H0, dH_of(p), p.bias, 2)
    r = prop.attempt(t0, h, eps)
    assert abs(r["te"] - TE) <= 1e-10 * TE
    assert r["accepted"] == (TE < eps)
    np.testing.assert_allclose(r["h_next"], O.hnew(h, TE, eps), rtol=1e-10)
    assert rel(r["J"], J0) < 1e-12
    s, ps, om = prop.get_state()
    want = y_new if r["accepted"] else y0
    for got, w in zip((s, ps, om), want):
        assert rel(got, w) < 1e-12
    prop.close()


@pytest.mark.parametrize("case", ["small", "complex_H", "small_spin_hubbard", "chain20"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_attempt_sparse_h(native, case, precision):
    """One attempt with the H products as gathers over the row-compressed H."""
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    y0 = _rand_state(p, 13, scale=0.02)
    say("attempt sparse-H", case, precision)
    prop = Propagator(p, precision, state=y0, hproducts="sparse")
    assert prop.hamiltonian_info()["sparse"]
    t0, h, eps = 0.3, 0.02, 1e-3
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, t0, h, p.H0, dH_of(p), p.bias, 2)
    r = prop.attempt(t0, h, eps)
    tol = 1e-10 if precision == "fp64" else 1e-4
    assert abs(r["te"] - TE) <= tol * TE
    assert rel(r["J"], J0) < (1e-12 if precision == "fp64" else 1e-5)
    if precision == "fp64":
        assert r["accepted"] == (TE < eps)
        s, ps, om = prop.get_state()
        want = y_new if r["accepted"] else y0
        for got, w in zip((s, ps, om), want):
            assert rel(got, w) < 1e-12
    prop.close()


def test_sparse_h_choice(native):
    """Auto takes the row-compressed H for a sparse tight-binding H and the dense
    products for a dense one; forcing it on a row with more than 64 entries fails."""
    from zandpack_amd import _native as N
    from zandpack_amd.engine import Propagator
    p = builder.config_c2()
    say("sparse choice", p.dims)
    prop = Propagator(p, "mixed")
    info = prop.hamiltonian_info()
    assert info["real"] and info["sparse"] and 1 <= info["row_entries"] <= 8
    prop.set_hproducts("dense")
    assert not prop.hamiltonian_info()["sparse"]
    prop.close()
    q = builder.config_small(n_w=10, n_l=10, seed=3)
    rng = np.random.default_rng(3)
    A = rng.normal(size=q.H0.shape[1:])
    q.H0 = q.H0 + 0.01 * (A + A.T)[None]
    prop = Propagator(q, "fp64")
    assert not prop.hamiltonian_info()["sparse"]
    with pytest.raises(N.ZHError):
        prop.set_hproducts("sparse")
    assert not prop.hamiltonian_info()["sparse"]
    prop.close()


def test_sparse_dense_propagation_agree(native):
    """Adaptive propagation of config 1 with either form of the H products."""
    from zandpack_amd.engine import Propagator
    outs = []
    for mode in ("dense", "sparse"):
        p = builder.config_chain()
        say("propagate", mode)
        prop = Propagator(p, "fp64", hproducts=mode)
        outs.append((prop.propagate(0.0, 2.0, 0.01, 1e-8), prop.get_state()))
        prop.close()
    (a, sa), (b, sb) = outs
    assert a["n_steps"] == b["n_steps"] and a["n_rejected"] == b["n_rejected"]
    np.testing.assert_allclose(a["times"], b["times"], rtol=1e-10)
    assert rel(a["J"], b["J"]) < 1e-10
    for x, y in zip(sa, sb):
        assert rel(x, y) < 1e-10


@pytest.mark.parametrize("case", ["small", "small_spin_hubbard"])
@pytest.mark.parametrize("kind", ["const", "step"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_attempt_dissipator(native, case, kind, precision):
    """One attempt with dissipator(t, σ) = η(t)(σ - σ_eq) (mpi/zand:343)."""
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    rng = np.random.default_rng(21)
    NK, n = p.dims[0], p.dims[1]
    A = rng.normal(size=(NK, n, n)) + 1j * rng.normal(size=(NK, n, n))
    p.dissipator = builder.Dissipator(kind, eta0=2.5, t0=0.31, width=0.02,
                                      sig_eq=p.sigma0 + 0.1 * (A + A.conj().transpose(0, 2, 1)))
    y0 = _rand_state(p, 14, scale=0.02)
    say("attempt dissipator", case, kind, precision)
    prop = Propagator(p, precision, state=y0)
    t0, h, eps = 0.3, 0.02, 1e-3
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, t0, h, p.H0, dH_of(p), p.bias, 2,
                                        dissipator=diss_of(p))
    r = prop.attempt(t0, h, eps, force_accept=True)
    tol = 1e-10 if precision == "fp64" else 1e-4
    assert abs(r["te"] - TE) <= tol * TE
    s = prop.sigma()
    assert rel(s, y_new[0]) < (1e-12 if precision == "fp64" else 1e-6)
    # without the dissipator the step differs by far more than that
    ks0, TE0, _, y_nd = O.dop54_attempt(p.consts(), y0, t0, h, p.H0, dH_of(p), p.bias, 2)
    assert rel(y_nd[0], y_new[0]) > 1e-4
    prop.close()


def test_attempt_reject_keeps_state(native):
    from zandpack_amd.engine import Propagator
    p = CASES["small"]()
    y0 = _rand_state(p, 12, scale=0.02)
    prop = Propagator(p, "fp64", state=y0)
    r = prop.attempt(0.0, 0.05, 1e-14)
    assert not r["accepted"]
    s, ps, om = prop.get_state()
    assert rel(s, y0[0]) == 0 and rel(ps, y0[1]) == 0 and rel(om, y0[2]) == 0
    prop.close()


@pytest.mark.parametrize("graph,layout", [(True, "auto"), (False, "auto"), (True, "pruned")])
def test_propagate_chain20(native, graph, layout):
    """Config 1 (n=20, M=96): adaptive DOP54 with zand's control semantics, 0 -> 3 fs."""
    from zandpack_amd.engine import Propagator
    p = builder.config_chain()
    y0 = p.zero_state()
    say("propagate chain20 graph", graph, layout)
    ref = O.propagate(p.consts(), y0, 0.0, 3.0, 0.01, 1e-8, p.H0, dH_of(p), p.bias, 2)
    prop = Propagator(p, "fp64", graph=graph, layout=layout)
    out = prop.propagate(0.0, 3.0, 0.01, 1e-8)
    assert out["n_steps"] == len(ref["times"]) - 1
    assert out["n_rejected"] == ref["n_reject"]
    np.testing.assert_allclose(out["times"], ref["times"], rtol=1e-10)
    np.testing.assert_allclose(out["h"], ref["h_accepted"], rtol=1e-10)
    assert rel(out["J"], ref["J"]) < 1e-10
    assert rel(prop.sigma(), ref["y"][0]) < 1e-10
    prop.close()


def test_graph_replays_match_direct(native):
    """hipGraph replays (both parities, several times) equal direct launches bit for bit."""
    from zandpack_amd.engine import Propagator
    p = CASES["small_spin_hubbard"]()
    y0 = _rand_state(p, 21, scale=0.02)
    res = []
    for graph in (True, False):
        prop = Propagator(p, "fp64", graph=graph, state=y0)
        Js = [prop.attempt(0.1 * i, 0.01, 1.0)["J"] for i in range(5)]
        res.append((np.array(Js), prop.get_state()))
        prop.close()
    np.testing.assert_array_equal(res[0][0], res[1][0])
    for a, b in zip(res[0][1], res[1][1]):
        np.testing.assert_array_equal(a, b)


def test_layouts_agree_over_propagation(native):
    """Reference order vs active-first order with the passive block skipped: the
    same adaptive run (steps, rejections) and currents to rounding."""
    from zandpack_amd.engine import Propagator
    p = CASES["tile_aligned"]()
    outs = []
    for layout in ("reference", "pruned"):
        prop = Propagator(p, "fp64", layout=layout)
        outs.append((prop.propagate(0.0, 1.0, 0.01, 1e-8), prop.get_state(), prop.layout_info()))
        prop.close()
    (a, sa, ia), (b, sb, ib) = outs
    assert not ia["active_first"] and ib["active_first"] and ib["skip_passive_block"]
    assert a["n_steps"] == b["n_steps"] and a["n_rejected"] == b["n_rejected"]
    assert rel(b["J"], a["J"]) < 1e-10
    for x, y in zip(sb, sa):
        assert rel(x, y) < 1e-10


@pytest.mark.parametrize("driver", ["AdaptiveRK4", "AdaptiveRK4_Opti"])
def test_drop_in_rkf45(native, driver):
    """The in-process RKF45 drivers over make_f against the oracle's restatement."""
    from zandpack_amd import drop_in
    from zandpack_amd.engine import Propagator
    p = CASES["small_spin_hubbard"]()
    prop = Propagator(p, "fp64")
    f = drop_in.make_f(prop)
    y0 = p.zero_state()
    say("rkf45", driver)
    # (with the driver's default first step, (t1-t0)/1000, the first error norms are at
    # the rounding level and the step sequence is not reproducible to 1e-10)
    times, data = getattr(drop_in, driver)(f, *y0, 1e-7, 0.0, 0.3, dH_of(p), p.bias, p.Ixi,
                                           h_guess=0.01, print_to_file=False)
    ref = O.rkf45_propagate(p.consts(), y0, 0.0, 0.3, 0.01, 1e-7, p.H0, dH_of(p), p.bias, 2)
    assert len(times) == len(ref["times"])
    np.testing.assert_allclose(times, ref["times"], rtol=1e-10)
    J = np.array([data["current_left"], data["current_right"]]).transpose(1, 0, 2)
    assert rel(J, ref["J"]) < 1e-10
    assert rel(data["state"][0], ref["y"][0]) < 1e-10
    assert rel(data["density matrix"], ref["sigma"]) < 1e-10
    prop.close()


@pytest.mark.parametrize("case,precision", [("small", "fp64"), ("small", "mixed"),
                                            ("tile_aligned", "fp64")])
def test_scaled_norm_attempt(native, case, precision):
    """Attempts under the scaled RMS norm (zh_set_tolerance) against the oracle: TE,
    accept and the committed state; (0, 0) restores the absolute norm."""
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    y0 = _rand_state(p, 17, scale=0.02)
    rtol, atol = 1e-4, 1e-8
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, 0.1, 0.01, p.H0, dH_of(p), p.bias, 2,
                                         rtol=rtol, atol=atol)
    # mixed: the weights reach 1/atol² on near-zero entries, where the complex64
    # rounding of the stored stage slots is not small against atol
    tol = 1e-11 if precision == "fp64" else 1e-2
    prop = Propagator(p, precision, state=y0)
    prop.set_tolerance(rtol, atol)
    say("scaled norm", case, precision, TE)
    r = prop.attempt(0.1, 0.01, 1.0)
    assert abs(r["te"] - TE) <= tol * TE, (r["te"], TE)
    assert r["accepted"] == (TE < 1.0)
    if r["accepted"]:
        s, ps, om = prop.get_state()
        for got, w in zip((s, ps, om), y_new):
            assert rel(got, w) < (1e-12 if precision == "fp64" else 1e-5)
    prop.put_state(*y0)
    prop.set_tolerance(0.0, 0.0)
    _, TE_abs, _, _ = O.dop54_attempt(p.consts(), y0, 0.1, 0.01, p.H0, dH_of(p), p.bias, 2)
    r = prop.attempt(0.1, 0.01, 1.0)
    assert abs(r["te"] - TE_abs) <= (1e-11 if precision == "fp64" else 1e-4) * TE_abs
    prop.close()


def test_drop_in_scipy_ode(native):
    """solve_ivp over the GPU f against solve_ivp over the oracle's RHS."""
    from scipy.integrate import solve_ivp
    from zandpack_amd import drop_in
    from zandpack_amd.engine import Propagator
    p = CASES["small"]()
    prop = Propagator(p, "fp64")
    f = drop_in.make_f(prop)
    y0 = p.zero_state()
    t_eval = np.array([0.1, 0.2, 0.3])
    say("scipy_ode")
    t, rho, cl = drop_in.scipy_ode(f, *y0, 0.0, t_eval, dH_of(p), p.bias, p.Ixi, atol=1e-8, rtol=1e-6)
    shapes = [a.shape for a in y0]
    sizes = [a.size for a in y0]

    def F(tt, y):
        parts = np.split(y, np.cumsum(sizes)[:-1])
        s, ps, om = (v.reshape(sh) for v, sh in zip(parts, shapes))
        delta = np.array([p.bias(tt, a) for a in range(2)])
        d = O.rhs(p.consts(), s, ps, om, p.H0 + dH_of(p)(tt, s), delta)
        return np.concatenate([v.ravel() for v in d])

    y = np.concatenate([v.ravel().astype(complex) for v in y0])
    tp = 0.0
    for i, te in enumerate(t_eval):
        y = solve_ivp(F, (tp, te), y, method="RK45", t_eval=[te], rtol=1e-6, atol=1e-8).y[:, -1]
        parts = np.split(y, np.cumsum(sizes)[:-1])
        s, ps = parts[0].reshape(shapes[0]), parts[1].reshape(shapes[1])
        assert rel(rho[i], s[0]) < 1e-9
        assert rel(cl[i], O.currents(p.consts(), ps)[0]) < 1e-8
        tp = te
    prop.close()


def test_drop_in_f_and_adaptive_dop(native):
    """The reference call surface: make_f + AdaptiveDOP vs the oracle driver."""
    from zandpack_amd.drop_in import AdaptiveDOP, make_f
    from zandpack_amd.engine import Propagator
    p = CASES["small_spin_hubbard"]()
    prop = Propagator(p, "fp64")
    f = make_f(prop)
    y0 = p.zero_state()
    times, data = AdaptiveDOP(f, *y0, 1e-7, 0.0, 0.3, dH_of(p), p.bias, p.Ixi, h_guess=0.01,
                              print_to_file=False)
    ref = O.propagate(p.consts(), y0, 0.0, 0.3, 0.01, 1e-7, p.H0, dH_of(p), p.bias, 2)
    np.testing.assert_allclose(times, ref["times"][:-1], rtol=1e-10)
    J = np.array([data["current_left"], data["current_right"]]).transpose(1, 0, 2)
    assert rel(J, ref["J"]) < 1e-10
    assert rel(data["state"][0], ref["y"][0]) < 1e-10
    prop.close()


def test_mixed_precision_currents(native):
    """Mixed precision follows the FP64 run's accepted (t, h) sequence (replay
    mode); J(t) relative error (normalised by max |J|) stays below 1e-6."""
    from zandpack_amd.engine import Propagator
    p = builder.config_small(n_w=6, n_l=7, N_L=5, slots=6, NOT=3, seed=4)
    say("mixed replay")
    a = Propagator(p, "fp64")
    out = a.propagate(0.0, 2.0, 0.01, 1e-7)
    b = Propagator(p, "mixed")
    Jm = []
    # attempt k starts at times[k] (the reference bookkeeping advanced t by h_next)
    for t, h in zip(out["times"][:-1], out["h"]):
        r = b.attempt(t, h, 1.0, force_accept=True)
        Jm.append(r["J"])
    Jm = np.array(Jm)
    err = np.abs(Jm - out["J"]).max() / np.abs(out["J"]).max()
    say("mixed J rel err", err)
    assert err < 1e-6
    a.close()
    b.close()


def test_c2_size_properties(native):
    """Config 2 shapes (n=400, M=1600): one evaluation against the oracle, plus
    the size-independent invariants (dσ Hermitian, Tr dσ = Σ J)."""
    from zandpack_amd.engine import Propagator
    p = builder.config_c2()
    sig, psi, omg = _rand_state(p, 3, scale=0.01)
    Ht = p.H0.copy()
    delta = np.array([0.25, -0.25])
    say("c2 rhs", p.dims, p.M)
    prop = Propagator(p, "fp64")
    ds, dp, do, J = prop.rhs(Ht, delta, sig, psi, omg)
    ref = O.rhs(p.consts(), sig, psi, omg, Ht, delta)
    for got, want in zip((ds, dp, do), ref):
        assert rel(got, want) < 1e-12
    np.testing.assert_allclose(ds, ds.conj().transpose(0, 2, 1), atol=1e-12 * np.abs(ds).max())
    tr = np.trace(ds, axis1=1, axis2=2).real
    np.testing.assert_allclose(tr, J.sum(axis=0), rtol=1e-10, atol=1e-12 * np.abs(ds).max())
    prop.close()


_HUNG = ("hung once inside zh_attempt at tile candidate 2 in graph mode (FP64, complex H, "
         "n = 12, M = 28); no cause was found in the kernels' index arithmetic or the "
         "capture's fork/join, so FP64 candidate 2 is switched off in the engine and this "
         "case stays off the GPU until a cause is shown")


@pytest.mark.parametrize("case,precision", [
    ("wide_edges", "fp64"), ("wide_edges", "mixed"),
    pytest.param("complex_H", "fp64", marks=pytest.mark.skip(reason=_HUNG)),
    ("complex_H", "mixed")])
def test_every_tile_candidate_matches_oracle(native, precision, case):
    """The autotuner may pick any enabled tile shape: every candidate gives the
    same attempt as the oracle (FP64: candidates 0 and 1; candidate 2 is off)."""
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    y0 = _rand_state(p, 31, scale=0.02)
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, 0.1, 0.01, p.H0, dH_of(p), p.bias, 2)
    # the error norm is a difference of nearly equal stage sums: its FP32 relative
    # error is larger than the state's
    tol = 1e-11 if precision == "fp64" else 1e-4
    ncand = 2 if precision == "fp64" else 3
    prop = Propagator(p, precision, state=y0, autotune=True)
    tiles = prop.autotune(force=True)
    assert all(0 <= v < ncand for v in tiles.values())
    for c in range(ncand):
        say("tiles", case, precision, c)
        prop.put_state(*y0)
        prop.set_tiles({k: c for k in tiles})
        r = prop.attempt(0.1, 0.01, 1.0)
        assert abs(r["te"] - TE) <= tol * TE, (c, r["te"], TE)
        s, ps, om = prop.get_state()
        for got, w in zip((s, ps, om), y_new):
            assert rel(got, w) < (1e-12 if precision == "fp64" else 1e-5), c
    prop.close()


def test_fp64_candidate2_refused(native):
    """FP64 tile candidate 2 is switched off: set_tiles refuses it (nothing runs)."""
    from zandpack_amd._native import ZHError
    from zandpack_amd.engine import Propagator
    p = CASES["small"]()
    prop = Propagator(p, "fp64")
    with pytest.raises(ZHError, match="switched off"):
        prop.set_tiles({k: 2 for k in ("psi", "hsig", "omega", "omgxi", "pi")})
    prop.close()


def test_config1_scaled_tolerances(native):
    """Config 1 as specified (DOP54, rtol 1e-8, atol 1e-10; here to 4 fs): the
    device propagation under the scaled norm takes the oracle's accepted steps, with
    its currents and final state (FP64 ≤ 1e-10)."""
    from zandpack_amd.engine import Propagator
    p = builder.config_chain()
    y0 = p.zero_state()
    rtol, atol = 1e-8, 1e-10
    say("c1 scaled tolerances")
    ref = O.propagate(p.consts(), y0, 0.0, 4.0, 0.01, 1.0, p.H0, dH_of(p), p.bias, 2,
                      rtol=rtol, atol=atol)
    prop = Propagator(p, "fp64")
    prop.set_tolerance(rtol, atol)
    out = prop.propagate(0.0, 4.0, 0.01, 1.0)
    assert out["n_steps"] == len(ref["times"]) - 1
    np.testing.assert_allclose(out["times"], ref["times"], rtol=1e-10, atol=1e-12)
    assert rel(out["J"], ref["J"]) < 1e-10
    assert rel(prop.sigma(), ref["y"][0]) < 1e-10
    prop.close()


def test_config1_golden_fixture(native):
    """Config 1 to 50 fs from its equilibrium steady state against
    tests/golden/c1_chain20.npz: replay of the fixture's (t, h) sequence reproduces J(t)
    and σ(t) (FP64 ≤ 1e-10; mixed J within one FP32 rounding of the lead sums, below),
    and the free-running adaptive run takes the same steps."""
    import os
    from zandpack_amd.engine import Propagator
    f = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_chain20.npz"))
    p = builder.config_chain()
    ts, hs = f["times"][:-1], f["h"]
    every = int(f["every"])
    y0 = (f["sigma0"], f["psi0"], f["omega0"])  # the equilibrium steady state
    # From equilibrium J is a small difference of large lead sums (J(0) = 0 while
    # Σ|ψ₀·Ixi| 2/ħ² is ~110 max|J|): complex64 storage of Ψ bounds the mixed error at
    # one FP32 rounding of that sum relative to max|J| (6.6e-6 here), not at 1e-6.
    S = 2.0 * np.abs(f["psi0"] * p.Ixi).sum() / builder.HBAR ** 2
    mixed_tol = max(1e-6, 2.0 ** -24 * S / np.abs(f["J"]).max())
    for precision, tol in (("fp64", 1e-10), ("mixed", mixed_tol)):
        say("c1 replay", precision)
        prop = Propagator(p, precision, state=y0)
        J, sig = [], []
        for i, (t, h) in enumerate(zip(ts, hs)):
            if i % every == 0:
                sig.append(prop.sigma())
            J.append(prop.attempt(t, h, 1.0, force_accept=True)["J"])
        J = np.array(J)
        assert np.abs(J - f["J"]).max() / np.abs(f["J"]).max() < tol
        if precision == "fp64":
            assert np.abs(np.array(sig) - f["sigma_every"][:len(sig)]).max() < 1e-10
        prop.close()
    prop = Propagator(p, "fp64", state=y0)
    out = prop.propagate(0.0, float(f["t_end"]), float(f["h0"]), float(f["eps"]))
    assert out["n_steps"] == len(f["times"]) - 1 and out["n_rejected"] == int(f["n_reject"])
    # free running, each step size comes from an error estimate that cancels to ~1e-8 of
    # the stage derivatives, so rounding-level differences (summation orders) move h, the
    # times and J at ~1e-9; the replay above holds the FP64 bar
    assert np.abs(out["times"] - f["times"]).max() < 1e-7 * float(f["t_end"])
    assert rel(out["J"], f["J"]) < 1e-7
    prop.close()


def _pulse_cases():
    tab = np.load(os.path.join(os.path.dirname(__file__), "golden", "toptica_pulse.npy"))
    knots = np.linspace(-0.2, 2.5, 9)
    return {
        "generic": builder.Pulse("generic", 0.4, 0.5, 1.1, (0.5, -0.5), 0.3, 0.8),
        "stairs": builder.Pulse("stairs", 0.15, 0.4, 4, (0.5, -0.5), 0.05),
        "spline": builder.spline_pulse(knots, np.sin(3 * knots), (0.1, 1.8, 0.2), amp=0.3),
        "toptica": builder.experimental_pulse("toptica", tab, amp=0.5),
    }


@pytest.mark.parametrize("kind", ["generic", "stairs", "spline", "toptica"])
def test_device_pulses_propagate(native, kind):
    """The device-evaluated bias of each pulse kind against the host form (Δ(t) at
    every stage time) over an adaptive propagation, 0 -> 2 fs."""
    from zandpack_amd.engine import Propagator
    p = builder.config_chain()
    p.pulse = _pulse_cases()[kind]
    y0 = p.zero_state()
    say("pulse", kind)
    ref = O.propagate(p.consts(), y0, 0.0, 2.0, 0.01, 1e-8, p.H0, dH_of(p), p.bias, 2)
    prop = Propagator(p, "fp64")
    out = prop.propagate(0.0, 2.0, 0.01, 1e-8)
    assert out["n_steps"] == len(ref["times"]) - 1
    np.testing.assert_allclose(out["times"], ref["times"], rtol=1e-10)
    assert rel(out["J"], ref["J"]) < 1e-10
    assert np.abs(ref["J"]).max() > 1e-6
    prop.close()


@pytest.mark.parametrize("case", ["small", "tile_aligned", "small_spin_hubbard", "wide_edges",
                                  "shared_orbitals", "chain20"])
@pytest.mark.parametrize("layout", ["reference", "pruned"])
@pytest.mark.parametrize("zero_ff", [False, True])
def test_fused_omega_matches_unfused(native, case, layout, zero_ff):
    """Mixed precision: the attempt with the Ωξ product fused into the Ω stage kernel
    against the oracle and against the unfused attempt, over three attempts (the y0 ξ_c
    carried from one accepted attempt to the next), for every chunk width."""
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    y0 = _rand_state(p, 41, scale=0.02)
    if zero_ff:
        passive = (p.dGP == 0) & (p.dGM == 0)
        pm = passive.all(axis=0).reshape(-1)
        om = y0[2].reshape(p.dims[0], p.M, p.M)
        om[:, pm[:, None] & pm[None, :]] = 0
    say("fused", case, layout, zero_ff)
    ref_state = y0
    seq = [(0.1, 0.01), (0.11, 0.012), (0.122, 0.008)]
    refs = []
    for t, h in seq:
        ks, TE, J0, y_new = O.dop54_attempt(p.consts(), ref_state, t, h, p.H0, dH_of(p), p.bias, 2)
        refs.append((TE, J0, y_new))
        ref_state = y_new
    for c in range(3):
        a = Propagator(p, "mixed", state=y0, layout=layout)
        b = Propagator(p, "mixed", state=y0, layout=layout)
        a.set_fusion("on")
        b.set_fusion("off")
        tiles = {k: 0 for k in ("psi", "hsig", "omega", "omgxi", "pi")}
        tiles["omega"] = c
        a.set_tiles(tiles)
        for (t, h), (TE, J0, y_new) in zip(seq, refs):
            ra = a.attempt(t, h, 1.0, force_accept=True)
            rb = b.attempt(t, h, 1.0, force_accept=True)
            say("  te fused/unfused/oracle", c, ra["te"], rb["te"], TE)
            # the error norm is a difference of nearly equal stage sums: FP32 rounding of
            # the stages reaches it at up to ~1e-3 on small steps, on either path alike
            assert abs(ra["te"] - TE) <= 2e-3 * TE, (c, ra["te"], TE)
            assert abs(ra["te"] - rb["te"]) <= 2e-3 * TE
            assert rel(ra["J"], J0) < 1e-5
            sa, pa, oa = a.get_state()
            for got, w in zip((sa, pa, oa), y_new):
                assert rel(got, w) < 1e-5, c
        a.close()
        b.close()
