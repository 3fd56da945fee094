"""Failure paths of the device propagation loop end in an error status (the
reference's loop would spin or run on with h = inf: Zandpack/mpi/zand:396-407,
Zandpack/mpi_RK4pars_dev.py:148-149).  Every case is an ordinary kernel run on
finite or NaN data; nothing here faults or hangs the device."""
import sys

import numpy as np
import pytest

from zandpack_amd import builder

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_nan_state_is_nonfinite(native):
    from zandpack_amd._native import ZHError
    from zandpack_amd.engine import Propagator
    p = builder.config_small(seed=1)
    sig = p.sigma0.astype(complex).copy()
    sig[0, 1, 1] = np.nan
    say("nan state")
    prop = Propagator(p, "fp64", state=(sig, None, None))
    with pytest.raises(ZHError) as e:
        prop.propagate(0.0, 1.0, 0.01, 1e-6)
    assert e.value.code == -5 and "non-finite" in str(e.value)
    prop.close()


def test_zero_rhs_next_step_not_finite(native):
    """A state the equations leave unchanged (σ = I/2 with a real symmetric H, Ψ = Ω
    = 0, Λ^{<+} = Λ^{>+} = 0 so that no lead term sources Ψ): every stage derivative
    is exactly zero, TE = 0 and
    h_next = 0.9 h (eps/0)^(1/5) is infinite.  The step is refused with an error and
    not committed; the resident state stays where it was."""
    from zandpack_amd._native import ZHError
    from zandpack_amd.engine import Propagator
    p = builder.config_small(seed=2)
    for a in ("GG_P", "GL_P"):
        setattr(p, a, np.zeros_like(getattr(p, a)))
    p.pulse = builder.Pulse("none")
    n = p.dims[1]
    sig = np.broadcast_to(0.5 * np.eye(n), p.sigma0.shape).astype(complex).copy()
    say("zero rhs")
    prop = Propagator(p, "fp64", state=(sig, None, None))
    r = prop.attempt(0.0, 0.01, 1e-6)  # the single attempt reports it as the reference does
    assert r["te"] == 0.0 and np.isinf(r["h_next"])
    prop.put_state(sig)
    with pytest.raises(ZHError) as e:
        prop.propagate(0.0, 1.0, 0.01, 1e-6)
    assert e.value.code == -5 and "not committed" in str(e.value)
    np.testing.assert_array_equal(prop.sigma(), sig)
    prop.close()


def test_endless_rejection_hits_the_limit(native):
    from zandpack_amd._native import ZHError
    from zandpack_amd.engine import Propagator
    p = builder.config_small(seed=3)
    rng = np.random.default_rng(0)
    psi = 0.05 * (rng.normal(size=p.xi.shape) + 1j * rng.normal(size=p.xi.shape))
    say("endless rejection")
    prop = Propagator(p, "fp64", state=(p.sigma0, psi, None))
    s0 = prop.sigma()
    # eps far below what any step reaches within three attempts (each rejection shrinks
    # h by 1e-40 or more; many more would drive TE^2 into underflow, TE = 0 and an
    # infinite next step, which the non-finite check reports instead)
    with pytest.raises(ZHError) as e:
        prop.propagate(0.0, 1.0, 0.01, 1e-300, max_attempts=3)
    assert e.value.code == -6 and "attempt limit" in str(e.value)
    np.testing.assert_array_equal(prop.sigma(), s0)
    prop.close()
