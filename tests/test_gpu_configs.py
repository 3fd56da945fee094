"""The GPU path at the benchmark configurations, against the oracle.

* C3 (n = 1000, M = 4000 per k, NK = 2 spin, Hubbard U = 2 eV, step bias): one
  right-hand side in FP64 and mixed precision against the oracle's zand form
  (orbital-pruned, the reference's own contraction set), the size-independent
  invariants (dσ Hermitian, Tr dσ = Σ_α J_α), one full FP64 DOP54 attempt against
  the oracle's attempt, and mixed precision replayed on 300 accepted FP64 steps
  with the current error bounded at every step.
* C4 (uneven batched instances, n = 100..800, M = 4n): the batched attempts equal
  the single attempts bitwise, and one instance's attempt equals the oracle's.
* C5 (serving): a Poisson stream with n in [50, 1200], M in [2n, 5n]; one served
  FP64 request's currents against the oracle's propagation of the same problem.
"""
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


def _c3_state(p, seed, scale=0.01, zero_ff=True):
    rng = np.random.default_rng(seed)
    NK, n, NL, NX, NOT = p.dims
    M = p.M
    psi = scale * (rng.standard_normal(p.xi.shape) + 1j * rng.standard_normal(p.xi.shape))
    om = scale * (rng.standard_normal((NK, M, M)) + 1j * rng.standard_normal((NK, M, M)))
    if zero_ff:
        pm = ((p.dGP == 0) & (p.dGM == 0)).all(axis=0).reshape(-1)
        om[:, pm[:, None] & pm[None, :]] = 0
    A = rng.standard_normal((NK, n, n)) + 1j * rng.standard_normal((NK, n, n))
    sig = p.sigma0 + 0.01 * (A + A.conj().transpose(0, 2, 1))
    return sig, psi, om.reshape((NK, NL, NX, NOT, NL, NX, NOT))


@pytest.fixture(scope="module")
def c3():
    return builder.config_c3(seed=0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c3_rhs(native, c3, precision):
    from zandpack_amd.engine import Propagator
    p = c3
    sig, psi, omg = _c3_state(p, 1, zero_ff=False)
    t = 0.7
    Ht = p.H0 + dH_of(p)(t, sig)
    delta = np.array([p.bias(t, a) for a in range(p.dims[2])])
    say("c3 rhs", precision, p.dims, p.M)
    ref = O.rhs(p.consts(), sig, psi, omg, Ht, delta, prune=True)
    Jref = O.currents(p.consts(), psi)
    prop = Propagator(p, precision)
    ds, dp, do, J = prop.rhs(Ht, delta, sig, psi, omg)
    prop.close()
    tol = 1e-12 if precision == "fp64" else 2e-5
    errs = [rel(g, w) for g, w in zip((ds, dp, do), ref)]
    say("c3 rhs rel err", precision, errs, rel(J, Jref))
    assert max(errs) < tol
    assert rel(J, Jref) < (1e-12 if precision == "fp64" else 1e-5)
    # invariants: dσ Hermitian, Tr dσ/dt = Σ_α J_α (no dissipator)
    scale = np.abs(ds).max()
    assert np.abs(ds - ds.conj().transpose(0, 2, 1)).max() < 1e-12 * scale
    tr = np.trace(ds, axis1=1, axis2=2)
    np.testing.assert_allclose(tr.real, J.sum(axis=0), rtol=1e-9, atol=1e-12 * scale)


def test_c3_attempt_fp64(native, c3):
    """One full DOP54 attempt at C3 (zero passive block: the device skips it) against
    the oracle: error estimate, controller, stage-0 currents and the new state."""
    from zandpack_amd.engine import Propagator
    p = c3
    y0 = _c3_state(p, 2, scale=0.002)
    t0, h, eps = 0.4, 0.002, 1e-6
    say("c3 attempt fp64")
    prop = Propagator(p, "fp64", state=y0)
    assert prop.layout_info()["skip_passive_block"]
    r = prop.attempt(t0, h, eps, force_accept=True)
    s, ps, om = prop.get_state()
    prop.close()
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, t0, h, p.H0, dH_of(p), p.bias, 2)
    del ks
    errs = [rel(g, w) for g, w in zip((s, ps, om), y_new)]
    say("c3 attempt: TE", r["te"], TE, "state errs", errs, "J", rel(r["J"], J0))
    assert abs(r["te"] - TE) <= 1e-10 * TE
    np.testing.assert_allclose(r["h_next"], O.hnew(h, TE, eps), rtol=1e-10)
    assert rel(r["J"], J0) < 1e-12
    assert max(errs) < 1e-12


def test_c3_mixed_replay_300_steps(native, c3):
    """Mixed precision forced onto 300 accepted FP64 steps of the C3 propagation
    (from σ0 with Ψ = Ω = 0, eps = 1e-6, absolute norm): the current error
    max_α,k |J_mixed - J_fp64| / max |J_fp64| stays at or below 1e-6 at every
    step, and the final σ agrees to 1e-6."""
    from zandpack_amd.engine import Propagator
    p = c3
    steps = 300
    a = Propagator(p, "fp64")
    seq, J64 = [], []
    t, h = 0.0, 0.002
    while len(seq) < steps:
        r = a.attempt(t, h, 1e-6)
        h_old, h = h, r["h_next"]
        if r["accepted"]:
            seq.append((t, h_old))
            J64.append(r["J"])
            t += h
    s64 = a.sigma()
    a.close()
    b = Propagator(p, "mixed")
    Jm = np.array([b.attempt(ts, hs, 1.0, force_accept=True)["J"] for ts, hs in seq])
    sm = b.sigma()
    b.close()
    J64 = np.array(J64)
    scale = np.abs(J64).max()
    per_step = np.abs(Jm - J64).max(axis=(1, 2)) / scale
    serr = np.linalg.norm(sm - s64) / np.linalg.norm(s64)
    say(f"c3 replay: {steps} steps to t = {t:.3f} fs, max J err {per_step.max():.3e}, "
        f"final sigma err {serr:.3e}, max |J| {scale:.3e}")
    assert per_step.max() <= 1e-6
    assert serr < 1e-6


# ------------------------------------------------------------------------- C4
def test_c4_batch_equals_single(native):
    """Eight config-4 instances, n = 100..800 (M = 4n), mixed precision: three batched
    attempts equal the single attempts bitwise."""
    from zandpack_amd.engine import Propagator, batch_attempt
    ns = list(range(100, 801, 100))
    probs = [builder.config_c4_instance(n, seed=500 + i) for i, n in enumerate(ns)]
    say("c4 batch", ns)
    a = [Propagator(q, "mixed") for q in probs]
    b = [Propagator(q, "mixed") for q in probs]
    t = [0.0] * len(ns)
    h = [0.004] * len(ns)
    for _ in range(3):
        res = batch_attempt(a, t, h, 1e-6)
        for i in range(len(ns)):
            one = b[i].attempt(t[i], h[i], 1e-6)
            assert not isinstance(res[i], Exception), res[i]
            assert res[i]["accepted"] == one["accepted"]
            assert res[i]["te"] == one["te"] and res[i]["h_next"] == one["h_next"]
            np.testing.assert_array_equal(res[i]["J"], one["J"])
            if one["accepted"]:
                t[i] += one["h_next"]
            h[i] = one["h_next"]
    for i in range(len(ns)):
        np.testing.assert_array_equal(a[i].sigma(), b[i].sigma())
    for q in a + b:
        q.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c4_instance_attempt_vs_oracle(native, precision):
    from zandpack_amd.engine import Propagator
    p = builder.config_c4_instance(300, seed=7)
    rng = np.random.default_rng(3)
    NK, n, NL, NX, NOT = p.dims
    psi = 0.01 * (rng.standard_normal(p.xi.shape) + 1j * rng.standard_normal(p.xi.shape))
    y0 = (p.sigma0.astype(complex), psi, np.zeros((NK, NL, NX, NOT, NL, NX, NOT), complex))
    t0, h = 3.0, 0.004
    say("c4 instance vs oracle", precision, p.dims, p.M)
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, t0, h, p.H0, dH_of(p), p.bias, NL)
    prop = Propagator(p, precision, state=y0)
    r = prop.attempt(t0, h, 1.0, force_accept=True)
    s, ps, om = prop.get_state()
    prop.close()
    fp64 = precision == "fp64"
    assert abs(r["te"] - TE) <= (1e-10 if fp64 else 1e-4) * TE
    assert rel(r["J"], J0) < (1e-12 if fp64 else 1e-5)
    for got, want in zip((s, ps, om), y_new):
        assert rel(got, want) < (1e-12 if fp64 else 1e-5)


# ------------------------------------------------------------------------- C5
def test_c5_stream_request_vs_oracle(native):
    """A Poisson stream of six C5 requests (n ~ U[50, 1200], M ~ U[2n, 5n], step /
    Gaussian / sine bias) served together; the smallest request, run in FP64, takes
    the oracle's accepted steps with its currents (≤ 1e-10)."""
    from zandpack_amd.serving import Request, Server
    rng = np.random.default_rng(11)
    reqs, t = [], 0.0
    for i in range(6):
        t += rng.exponential(1.0 / 40.0)
        n = int(rng.integers(50, 1201))
        M = int(rng.integers(2 * n, 5 * n + 1))
        p = builder.config_sized(n, M, seed=900 + i, pulse=builder.random_pulse(rng))
        reqs.append(Request(i, p, t_end=0.05, eps=1e-6, h0=0.005, arrival=t, precision="mixed"))
    small = min(reqs, key=lambda r: r.problem.dims[1])
    small.precision = "fp64"
    small.t_end = 0.4
    say("c5 stream", [(r.problem.dims[1], r.problem.M) for r in reqs], "checked",
        small.rid)
    st = Server(hbm_bytes=200 << 30, max_live=8).run(reqs, realtime=True)
    assert st["failed"] == 0 and st["requests"] == 6
    q = small.problem
    ref = O.propagate(q.consts(), q.zero_state(), 0.0, small.t_end, small.h0, small.eps,
                      q.H0, dH_of(q), q.bias, q.dims[2])
    assert small.steps == len(ref["times"]) - 1 and small.rejected == ref["n_reject"]
    np.testing.assert_allclose(np.array(small.times), ref["times"], rtol=1e-10)
    assert rel(np.array(small.J), ref["J"]) < 1e-10
    assert np.abs(ref["J"]).max() > 0
    assert rel(small.sigma, ref["y"][0]) < 1e-10
