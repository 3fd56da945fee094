"""Instances (zh_problem_create_instances): several propagations of one shape stacked on
the k axis, every kernel of an attempt launched once for all of them, each instance
with its own pulse, time, step, tolerance, error norm and acceptance.  Against the same
problems alone (bitwise, same tiles), the oracle (FP64 1e-12), mixed acceptance (the
rejected instance keeps its state bitwise), held instances (h = 0, forced), and the
batched workload's shapes."""
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
    """σ near σ0, random Ψ, anti-Hermitian Ω (the Ω triangle applies)."""
    rng = np.random.default_rng(seed)
    NK, n, NL, NX, NOT = p.dims
    M = p.M
    psi = scale * (rng.normal(size=p.xi.shape) + 1j * rng.normal(size=p.xi.shape))
    A = rng.normal(size=(NK, M, M)) + 1j * rng.normal(size=(NK, M, M))
    om = scale * (A - A.conj().transpose(0, 2, 1))
    S = rng.normal(size=(NK, n, n)) + 1j * rng.normal(size=(NK, n, n))
    sig = p.sigma0 + 0.02 * (S + S.conj().transpose(0, 2, 1))
    return sig, psi, om.reshape(NK, NL, NX, NOT, NL, NX, NOT)


def _pulsed(p, amp, kind="gauss"):
    p.pulse = builder.Pulse(kind, amp, 0.3, 0.2, (0.5, -0.5))
    return p


FAMILIES = {
    "small": lambda s: _pulsed(builder.config_small(seed=s), 0.1 + 0.05 * s),
    "spin_hubbard": lambda s: _pulsed(builder.config_small(NK=2, hubbard_U=2.0, seed=s),
                                      0.05 * s, "step"),
    "c4_100": lambda s: builder.config_c4_instance(100, seed=s),
}


def _alone_and_grouped(fam, precision, G=3, seed0=1):
    from zandpack_amd.engine import Instances, Propagator
    ps = [FAMILIES[fam](seed0 + g) for g in range(G)]
    ys = [_state(p, 40 + g) for g, p in enumerate(ps)]
    alone = [Propagator(p, precision, state=y) for p, y in zip(ps, ys)]
    grp = Instances(ps, precision, states=ys)
    return ps, ys, alone, grp


@pytest.mark.parametrize("fam", list(FAMILIES))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_instances_equal_alone(native, fam, precision):
    """Three instances attempted together over five attempts with per-instance t, h
    and eps (one of them rejects on the first attempt): every result and state equal
    to the instance alone, bit for bit."""
    ps, ys, alone, grp = _alone_and_grouped(fam, precision)
    say("instances vs alone", fam, precision, grp.dims if hasattr(grp, "dims") else ps[0].dims)
    G = len(ps)
    t = np.array([0.0, 0.1, 0.2])
    h = np.array([0.01, 0.02, 0.015])
    eps = np.array([1e-2, 1e-12, 1e-2])  # instance 1 rejects at first
    for it in range(5):
        rg = grp.attempt(t, h, eps)
        for g in range(G):
            ra = alone[g].attempt(t[g], h[g], eps[g])
            assert ra["te"] == rg[g]["te"], (it, g, ra["te"], rg[g]["te"])
            assert ra["h_next"] == rg[g]["h_next"]
            assert ra["accepted"] == rg[g]["accepted"]
            assert np.array_equal(ra["J"], rg[g]["J"])
            if ra["accepted"]:
                t[g] += ra["h_next"]
            h[g] = ra["h_next"]
        if it == 0:
            assert not rg[1]["accepted"] and rg[0]["accepted"]
        eps[1] = 1e-2
    full = grp.get_state()
    for g in range(G):
        for a, b in zip(alone[g].get_state(), grp.instance_state(g, full)):
            assert np.array_equal(a, b), g
        alone[g].close()
    grp.close()


def test_instances_against_oracle(native):
    """One FP64 attempt of two spin-polarised Hubbard instances against the oracle."""
    from zandpack_amd.engine import Instances
    ps = [FAMILIES["spin_hubbard"](s) for s in (3, 4)]
    ys = [_state(p, 7 + s) for s, p in enumerate(ps)]
    grp = Instances(ps, "fp64", states=ys)
    t0 = [0.3, 0.1]
    h = [0.02, 0.01]
    r = grp.attempt(t0, h, 1.0)
    full = grp.get_state()
    for g, p in enumerate(ps):
        say("instances oracle", g)
        ks, TE, J0, y_new = O.dop54_attempt(p.consts(), ys[g], t0[g], h[g], p.H0, dH_of(p),
                                            p.bias, 2)
        assert abs(r[g]["te"] - TE) <= 1e-10 * TE
        assert rel(r[g]["J"], J0) < 1e-12
        for got, w in zip(grp.instance_state(g, full), y_new):
            assert rel(got, w) < 1e-12
    grp.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_held_instance_keeps_state(native, precision):
    """h = 0 with force_accept holds an instance (a finished one) bitwise while the
    others advance."""
    from zandpack_amd.engine import Instances
    ps = [FAMILIES["small"](s) for s in (1, 2)]
    ys = [_state(p, 3 + s) for s, p in enumerate(ps)]
    grp = Instances(ps, precision, states=ys)
    before = grp.instance_state(1)
    for _ in range(3):
        r = grp.attempt([0.0, 0.0], [0.01, 0.0], [1.0, 1.0], force_accept=[False, True])
        assert r[0]["accepted"] and r[1]["accepted"]
    after = grp.instance_state(1)
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    assert not np.array_equal(grp.instance_state(0)[1], np.asarray(ys[0][1]))
    grp.close()


def test_single_instance_calls_refused(native):
    from zandpack_amd._native import ZHError
    from zandpack_amd.engine import Instances
    ps = [FAMILIES["small"](s) for s in (1, 2)]
    grp = Instances(ps, "fp64")
    with pytest.raises(ZHError, match="several instances"):
        grp.propagate(0.0, 0.1, 0.01, 1e-6)
    with pytest.raises(ZHError, match="one instance"):
        grp.rhs(ps[0].H0, np.zeros(2), *[np.concatenate([a, a]) for a in _state(ps[0], 1)])
    grp.close()


def test_batched_shapes_equal_alone(native):
    """Eight instances of the batched workload's n = 200 shape (M = 800), mixed
    precision, autotuned once: two attempts each equal to the instance alone on the
    group's tiles."""
    from zandpack_amd.engine import Instances, Propagator
    ps = [builder.config_c4_instance(200, seed=10 + g) for g in range(8)]
    grp = Instances(ps, "mixed", autotune=True)
    tiles = grp.tiles
    say("c4 group tiles", tiles, grp.omgxi_info())
    xi_mode = "stream" if grp.omgxi_info()["stream"] else "gemm"
    res = [grp.attempt(0.0, 0.005, 1e-6) for _ in range(2)]
    for g in (0, 3, 7):
        a = Propagator(ps[g], "mixed")
        a.set_tiles(tiles)
        a.set_omgxi(xi_mode)
        for it in range(2):
            ra = a.attempt(0.0, 0.005, 1e-6)
            assert ra["te"] == res[it][g]["te"] and ra["accepted"] == res[it][g]["accepted"]
            assert ra["h_next"] == res[it][g]["h_next"]
            assert np.array_equal(ra["J"], res[it][g]["J"])
        a.close()
    grp.close()
