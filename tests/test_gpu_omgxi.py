"""The streaming Ωξ kernel (zh_set_omgxi): the Ω ξ_c product of the Ψ equation
(Zandpack/mpi/zand:771-778) with Ω and ξ_c staged through LDS by direct global-to-LDS
loads, one lead's columns per tile.  Against the oracle: one evaluation (FP64 1e-12,
mixed 2e-5) and full attempts (TE 1e-10, state 1e-12; mixed TE 1e-4, state 1e-5) with Ω
in full or held as its triangle, with and without the skipped passive block; against the
tiled GEMM on the same problem; the fallback where the layout does not allow it."""
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
    "small": lambda: builder.config_small(seed=1),
    "complex_H": _complex_h,
    "small_spin_hubbard": lambda: builder.config_small(NK=2, hubbard_U=2.0, seed=2),
    "chain20": builder.config_chain,
    "tile_aligned": lambda: builder.config_small(n_w=8, n_l=3, N_L=4, slots=4, NOT=8, seed=6),
    "wide": lambda: builder.config_small(n_w=6, n_l=7, N_L=5, slots=6, NOT=3, seed=4),
    "wide_edges": lambda: builder.config_small(n_w=36, n_l=3, N_L=2, slots=2, NOT=4, seed=9),
    "c4_100": lambda: builder.config_c4_instance(100, seed=3),
}


def _state(p, seed, anti, zero_ff, scale=0.03):
    """σ near σ0, random Ψ, Ω anti-Hermitian or not, its passive x passive block zero
    or not."""
    rng = np.random.default_rng(seed)
    NK, n, NL, NX, NOT = p.dims
    M = p.M
    psi = scale * (rng.normal(size=p.xi.shape) + 1j * rng.normal(size=p.xi.shape))
    A = rng.normal(size=(NK, M, M)) + 1j * rng.normal(size=(NK, M, M))
    om = scale * ((A - A.conj().transpose(0, 2, 1)) if anti else A)
    if zero_ff:
        pm = ((p.dGP == 0) & (p.dGM == 0)).all(axis=0).reshape(-1)
        om[:, pm[:, None] & pm[None, :]] = 0
    S = rng.normal(size=(NK, n, n)) + 1j * rng.normal(size=(NK, n, n))
    sig = p.sigma0 + 0.02 * (S + S.conj().transpose(0, 2, 1))
    return sig, psi, om.reshape(NK, NL, NX, NOT, NL, NX, NOT)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("anti,zero_ff", [(False, False), (True, False), (False, True),
                                          (True, True)])
def test_stream_attempt(native, case, precision, anti, zero_ff):
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    y0 = _state(p, 17, anti, zero_ff)
    t0, h, eps = 0.2, 0.02, 1e-3
    say("omgxi stream attempt", case, precision, anti, zero_ff, p.dims, p.M)
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, t0, h, p.H0, dH_of(p), p.bias, 2)
    res = {}
    for mode in ("stream", "gemm"):
        a = Propagator(p, precision, state=y0, layout="pruned")
        a.set_omgxi(mode)
        info = a.omgxi_info()
        assert info["possible"], info
        assert info["stream"] == (mode == "stream")
        assert a.antihermitian_info()["active"] == (anti and a.antihermitian_info()["constants"])
        r = a.attempt(t0, h, eps)
        res[mode] = (r, a.get_state())
        a.close()
    r, st = res["stream"]
    tol_te, tol_y = (1e-10, 1e-12) if precision == "fp64" else (1e-4, 1e-5)
    assert abs(r["te"] - TE) <= tol_te * TE, (r["te"], TE)
    assert rel(r["J"], J0) < (1e-12 if precision == "fp64" else 1e-5)
    if precision == "fp64":
        assert r["accepted"] == (TE < eps)
    want = y_new if r["accepted"] else y0
    for got, w in zip(st, want):
        assert rel(got, w) < tol_y
    # the tiled GEMM on the same problem: the same attempt to rounding
    rg, sg = res["gemm"]
    assert abs(r["te"] - rg["te"]) <= (1e-12 if precision == "fp64" else 1e-5) * rg["te"]
    for got, w in zip(st, sg):
        assert rel(got, w) < (1e-13 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("case", ["small", "complex_H", "wide_edges", "c4_100"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("layout", ["reference", "pruned"])
def test_stream_rhs(native, case, precision, layout):
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    sig, psi, omg = _state(p, 5, anti=False, zero_ff=False)
    Ht = p.H0 + dH_of(p)(0.2, sig)
    delta = np.array([0.05, -0.03])
    say("omgxi stream rhs", case, precision, layout)
    ref = O.rhs(p.consts(), sig, psi, omg, Ht, delta)
    a = Propagator(p, precision, layout=layout)
    a.set_omgxi("stream")
    assert a.omgxi_info()["stream"]
    out = a.rhs(Ht, delta, sig, psi, omg)
    a.close()
    tol = 1e-12 if precision == "fp64" else 2e-5
    for got, want in zip(out[:3], ref):
        assert rel(got, want) < tol


def test_stream_propagation_matches_gemm(native):
    """An adaptive FP64 propagation (tile_aligned, 40 steps) takes the same steps and
    currents with either kernel."""
    from zandpack_amd.engine import Propagator
    p = CASES["tile_aligned"]()
    y0 = _state(p, 3, anti=True, zero_ff=True, scale=0.0)
    out = {}
    for mode in ("stream", "gemm"):
        a = Propagator(p, "fp64", state=y0)
        a.set_omgxi(mode)
        say("omgxi propagate", mode)
        out[mode] = a.propagate(0.0, 2.0, 0.01, 1e-6, cap=64)
        a.close()
    s, g = out["stream"], out["gemm"]
    assert s["n_steps"] == g["n_steps"] and s["n_steps"] > 10
    np.testing.assert_allclose(s["times"], g["times"], rtol=1e-12)
    assert rel(s["J"], g["J"]) < 1e-10


def test_shared_orbitals_fall_back(native):
    """Both leads couple the same orbitals: no lead-aligned column tiles, the streaming
    kernel does not apply, and forcing it runs the GEMM."""
    from zandpack_amd.engine import Propagator
    p = builder.config_small(n_w=4, n_l=1, seed=13)
    y0 = _state(p, 9, anti=False, zero_ff=False)
    a = Propagator(p, "fp64", state=y0)
    a.set_omgxi("stream")
    assert a.omgxi_info() == {"stream": False, "possible": False}
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, 0.1, 0.01, p.H0, dH_of(p), p.bias, 2)
    r = a.attempt(0.1, 0.01, 1.0)
    assert abs(r["te"] - TE) <= 1e-10 * TE
    a.close()
