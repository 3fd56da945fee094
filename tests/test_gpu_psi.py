"""One-pass Ψ right-hand side with a row-compressed H (zh_set_psi_kernel): the stage-Ψ
rows of a tile staged in LDS over the orbitals its H rows reach, the ħ ΔP (σ ξ) part
formed beside the gather (Zandpack/mpi/zand:746-778).  Against the oracle (FP64: TE
1e-10, state 1e-12; mixed: TE 1e-4, state 1e-5) and against the split form on the same
problem; the split fallback when H reaches too far; instances."""
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


def _state(p, seed, scale=0.02, zero_ff=False):
    rng = np.random.default_rng(seed)
    NK, n, NL, NX, NOT = p.dims
    M = p.M
    psi = scale * (rng.normal(size=p.xi.shape) + 1j * rng.normal(size=p.xi.shape))
    A = rng.normal(size=(NK, M, M)) + 1j * rng.normal(size=(NK, M, M))
    om = scale * (A - A.conj().transpose(0, 2, 1))
    if zero_ff:
        pm = ((p.dGP == 0) & (p.dGM == 0)).all(axis=0).reshape(-1)
        om[:, pm[:, None] & pm[None, :]] = 0
    S = rng.normal(size=(NK, n, n)) + 1j * rng.normal(size=(NK, n, n))
    sig = p.sigma0 + 0.02 * (S + S.conj().transpose(0, 2, 1))
    return sig, psi, om.reshape(NK, NL, NX, NOT, NL, NX, NOT)


CASES = {
    "small": lambda: builder.config_small(seed=1),
    "small_spin_hubbard": lambda: builder.config_small(NK=2, hubbard_U=2.0, seed=2),
    "chain20": builder.config_chain,
    "tile_aligned": lambda: builder.config_small(n_w=8, n_l=3, N_L=4, slots=4, NOT=8, seed=6),
    "wide_edges": lambda: builder.config_small(n_w=36, n_l=3, N_L=2, slots=2, NOT=4, seed=9),
    "c4_100": lambda: builder.config_c4_instance(100, seed=3),
    "c4_200": lambda: builder.config_c4_instance(200, seed=4),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("zero_ff", [False, True])
def test_fused_psi_attempt(native, case, precision, zero_ff):
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    y0 = _state(p, 23, zero_ff=zero_ff)
    t0, h, eps = 0.2, 0.02, 1e-3
    say("psi fused attempt", case, precision, zero_ff, p.dims, p.M)
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, t0, h, p.H0, dH_of(p), p.bias, 2)
    out = {}
    for mode in ("fused", "split"):
        a = Propagator(p, precision, state=y0, hproducts="sparse", layout="pruned")
        a.set_psi_kernel(mode)
        assert a.psi_kernel_info() == mode
        r = a.attempt(t0, h, eps)
        out[mode] = (r, a.get_state())
        a.close()
    r, st = out["fused"]
    tol_te, tol_y = (1e-10, 1e-12) if precision == "fp64" else (1e-4, 1e-5)
    assert abs(r["te"] - TE) <= tol_te * TE, (r["te"], TE)
    assert rel(r["J"], J0) < (1e-12 if precision == "fp64" else 1e-5)
    want = y_new if r["accepted"] else y0
    for got, w in zip(st, want):
        assert rel(got, w) < tol_y
    # the split form on the same problem; the error norm is a difference of nearly
    # equal stage sums, so its FP32 rounding shows at 1e-5 of it
    rs, ss = out["split"]
    assert abs(r["te"] - rs["te"]) <= (1e-12 if precision == "fp64" else 1e-4) * rs["te"]
    for got, w in zip(st, ss):
        assert rel(got, w) < (1e-13 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_psi_rhs(native, precision):
    from zandpack_amd.engine import Propagator
    p = CASES["c4_200"]()
    sig, psi, omg = _state(p, 5)
    delta = np.array([0.05, -0.03])
    ref = O.rhs(p.consts(), sig, psi, omg, p.H0, delta)
    a = Propagator(p, precision, hproducts="sparse")
    assert a.psi_kernel_info() == "fused"
    out = a.rhs(p.H0, delta, sig, psi, omg)
    a.close()
    for got, want in zip(out[:3], ref):
        assert rel(got, want) < (1e-12 if precision == "fp64" else 2e-5)


def test_far_reaching_h_falls_back(native):
    """H rows reaching across the device (here a coupling between the first and the last
    orbital of a 200-orbital device): auto runs the split form, forcing one pass fails."""
    from zandpack_amd._native import ZHError
    from zandpack_amd.engine import Propagator
    p = CASES["c4_200"]()
    n = p.H0.shape[-1]
    p.H0 = p.H0.copy()
    p.H0[:, 0, n - 1] = p.H0[:, n - 1, 0] = -0.3
    y0 = _state(p, 9)
    a = Propagator(p, "fp64", state=y0, hproducts="sparse")
    assert a.hamiltonian_info()["sparse"] and a.psi_kernel_info() == "split"
    with pytest.raises(ZHError, match="reach"):
        a.set_psi_kernel("fused")
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, 0.1, 0.01, p.H0, dH_of(p), p.bias, 2)
    r = a.attempt(0.1, 0.01, 1.0)
    assert abs(r["te"] - TE) <= 1e-10 * TE
    a.close()


def test_fused_psi_instances_equal_alone(native):
    from zandpack_amd.engine import Instances, Propagator
    ps = [builder.config_c4_instance(100, seed=s) for s in (1, 2, 3)]
    ys = [_state(p, 30 + i) for i, p in enumerate(ps)]
    grp = Instances(ps, "mixed", states=ys)
    assert grp.psi_kernel_info() == "fused"
    rg = grp.attempt([0.0, 0.1, 0.2], [0.01, 0.005, 0.008], 1e-3)
    for g, p in enumerate(ps):
        a = Propagator(p, "mixed", state=ys[g])
        ra = a.attempt([0.0, 0.1, 0.2][g], [0.01, 0.005, 0.008][g], 1e-3)
        assert ra["te"] == rg[g]["te"] and np.array_equal(ra["J"], rg[g]["J"])
        for x, y in zip(a.get_state(), grp.instance_state(g)):
            assert np.array_equal(x, y)
        a.close()
    grp.close()
