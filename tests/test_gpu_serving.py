"""Serving loop: continuous batching of device-resident propagations.

Each request's result is bitwise the same alone and in the mix (same kernels,
own buffers and stream, fixed reduction orders); admission respects the HBM
budget; a failing request is retired without disturbing the others."""
import sys

import numpy as np
import pytest

from zandpack_amd import builder

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _requests(precision="mixed"):
    from zandpack_amd.serving import Request
    rng = np.random.default_rng(5)
    out = []
    for i, (n, M) in enumerate([(40, 80), (70, 210), (100, 400)]):
        p = builder.config_sized(n, M, seed=10 + i, pulse=builder.random_pulse(rng))
        out.append(Request(i, p, t_end=0.15 + 0.05 * i, eps=1e-6, h0=0.005, arrival=0.01 * i,
                           precision=precision))
    return out


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_alone_equals_mix(native, precision):
    from zandpack_amd.serving import Server
    say("serving alone vs mix", precision)
    alone = []
    for r in _requests(precision):
        st = Server().run([r], realtime=False)
        assert st["failed"] == 0
        alone.append(st["done"][0])
    mixed = Server().run(_requests(precision), realtime=True)
    assert mixed["failed"] == 0 and mixed["requests"] == 3
    by_id = {r.rid: r for r in mixed["done"]}
    for a in alone:
        b = by_id[a.rid]
        assert a.steps == b.steps and a.rejected == b.rejected
        np.testing.assert_array_equal(np.array(a.J), np.array(b.J))
        np.testing.assert_array_equal(np.array(a.times), np.array(b.times))
        np.testing.assert_array_equal(a.sigma, b.sigma)
    assert mixed["p50_s"] is not None and mixed["p99_s"] >= mixed["p50_s"]


def test_admission_by_footprint(native):
    from zandpack_amd.serving import Server, footprint_bytes
    from zandpack_amd.engine import Propagator
    reqs = _requests()
    fps = [footprint_bytes(r.problem, r.precision) for r in reqs]
    # the estimate covers what the engine allocates
    prop = Propagator(reqs[2].problem, "mixed")
    assert prop.device_bytes() <= fps[2]
    prop.close()
    budget = max(fps) + min(fps) // 2  # never room for the two largest together
    srv = Server(hbm_bytes=budget)
    st = srv.run(reqs, realtime=False)
    assert st["failed"] == 0 and st["requests"] == 3
    assert srv.peak_bytes <= budget


def test_failing_request_is_retired(native):
    from zandpack_amd.serving import Server
    reqs = _requests()
    reqs[1].eps = 1e-30      # never accepted
    reqs[1].max_attempts = 5
    st = Server().run(reqs, realtime=False)
    errs = {r.rid: r.error for r in st["done"]}
    assert errs[1] == "attempt limit reached"
    assert errs[0] is None and errs[2] is None


def test_batch_attempt_matches_single(native):
    """zh_batch_attempt: the same results as one zh_attempt per problem, bitwise, and
    a problem whose attempt is refused (a non-finite step) leaves the others going."""
    from zandpack_amd.engine import Propagator, batch_attempt
    say("batch attempt vs single")
    reqs = _requests("mixed")
    a = [Propagator(r.problem, "mixed") for r in reqs]
    b = [Propagator(r.problem, "mixed") for r in reqs]
    t = [0.0] * 3
    h = [0.004, 0.005, 0.006]
    for _ in range(3):
        res = batch_attempt(a, t, h, 1e-6)
        for i in range(3):
            one = b[i].attempt(t[i], h[i], 1e-6)
            assert res[i]["accepted"] == one["accepted"]
            assert res[i]["te"] == one["te"] and res[i]["h_next"] == one["h_next"]
            np.testing.assert_array_equal(res[i]["J"], one["J"])
            if one["accepted"]:
                t[i] += one["h_next"]
            h[i] = one["h_next"]
    for i in range(3):
        np.testing.assert_array_equal(a[i].sigma(), b[i].sigma())
    res = batch_attempt(a, t, [h[0], float("inf"), h[2]], 1e-6)
    assert isinstance(res[1], Exception) and "step size" in str(res[1])
    for i in (0, 2):
        one = b[i].attempt(t[i], h[i], 1e-6)
        assert res[i]["te"] == one["te"]
    for p in a + b:
        p.close()


def test_same_shape_requests_grouped(native):
    """Requests of one shape are admitted together as the instances of one problem (one
    launch per kernel class for all of them); each result is bitwise the same as the
    request served alone, finished instances held while the others go on."""
    from zandpack_amd.serving import Request, Server
    say("serving grouped")

    def reqs():
        return [Request(i, builder.config_c4_instance(100, seed=20 + i), t_end=60.0, eps=1e-6,
                        h0=0.005, precision="mixed", max_steps=3 + 2 * i) for i in range(5)]

    alone = []
    for r in reqs():
        st = Server(max_group=1).run([r], realtime=False)
        assert st["failed"] == 0
        alone.append(st["done"][0])
    srv = Server(max_group=16)
    st = srv.run(reqs(), realtime=False)
    assert st["failed"] == 0 and st["requests"] == 5
    assert srv.groups == [5]
    by_id = {r.rid: r for r in st["done"]}
    for a in alone:
        b = by_id[a.rid]
        assert b.steps == a.steps == 3 + 2 * a.rid and b.rejected == a.rejected
        np.testing.assert_array_equal(np.array(a.J), np.array(b.J))
        np.testing.assert_array_equal(np.array(a.times), np.array(b.times))
        np.testing.assert_array_equal(a.sigma, b.sigma)
