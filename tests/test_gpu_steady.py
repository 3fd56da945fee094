"""GPU steady state (cmdtools/psinought equivalent) against the oracle's dense
solve of psinought's linear system, and the fixed-point property the reference
checks after solving (dΨ = dΩ = 0)."""
import functools
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


CASES = {
    "small": lambda: builder.config_small(seed=1),
    "small_spin_hubbard": lambda: builder.config_small(NK=2, hubbard_U=2.0, seed=2),
    "shared_orbitals": lambda: builder.config_small(n_w=4, n_l=1, seed=13),
    "tile_aligned": lambda: builder.config_small(n_w=8, n_l=3, N_L=4, slots=4, NOT=8, seed=6),
    "chain20": lambda: builder.config_chain(),
}


def _bias(p, t0):
    return np.array([p.bias(t0, a) for a in range(p.dims[2])])


def _problem(case):
    p = CASES[case]()
    p.pulse = builder.Pulse("step", amp=0.2, p1=0.0, weights=(0.5, -0.5))
    return p


@functools.lru_cache(maxsize=None)
def _oracle_solution(case, t0):
    p = _problem(case)
    H = p.H0 + dH_of(p)(t0, p.sigma0)
    return O.steady_state_dense(p.consts(), H, p.sigma0, _bias(p, t0))


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_steady_state_matches_oracle(native, case, precision):
    from zandpack_amd.engine import Propagator
    p = _problem(case)
    t0 = 1.0
    sig0 = p.sigma0
    psi_ref, om_ref = _oracle_solution(case, t0)
    say("steady", case, precision, p.dims)
    prop = Propagator(p, precision)
    tol = 1e-12 if precision == "fp64" else 1e-5
    its, res = prop.steady_state(t0, tol=tol, maxiter=2000)
    say("  iterations", its, "residual", res)
    assert res < tol
    s, ps, om = prop.get_state()
    bar = 1e-8 if precision == "fp64" else 1e-3
    assert rel(ps, psi_ref) < bar, rel(ps, psi_ref)
    assert rel(om, om_ref) < bar, rel(om, om_ref)
    assert rel(s, sig0) == 0.0
    prop.close()


@pytest.mark.parametrize("case", ["small_spin_hubbard", "chain20"])
def test_steady_state_is_a_fixed_point(native, case):
    """psinought's own check (cmdtools/psinought:455-462): the equations of motion
    at (σ₀, Ψ₀, Ω₀) give dΨ = 0 and dΩ = 0 (here through the device RHS)."""
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    t0 = 0.0
    prop = Propagator(p, "fp64")
    its, res = prop.steady_state(t0, tol=1e-13, maxiter=2000)
    s, ps, om = prop.get_state()
    Ht = p.H0 + dH_of(p)(t0, s)
    ds, dp, do, J = prop.rhs(Ht, _bias(p, t0), s, ps, om)
    # compare with the size of the terms that cancel
    scale_p = np.abs(O.steady_state_rhs(p.consts(), s)).max() / O.HBAR
    assert np.abs(dp).max() < 1e-9 * scale_p
    assert np.abs(do).max() < 1e-9 * max(np.abs(om).max(), 1e-300) / O.HBAR
    assert its < 200
    prop.close()


def test_unpreconditioned_limit_is_reported(native):
    """Without eigenpairs the Krylov iteration runs on the bare operator, whose real
    spectrum surrounds zero (BiCGSTAB stalls, as scipy's does on the same matrix):
    the iteration limit comes back as ZH_ELIMIT with a finite residual and a usable
    state; the preconditioned solve from there converges."""
    from zandpack_amd import _native as N
    from zandpack_amd.engine import Propagator
    p = _problem("small")
    prop = Propagator(p, "fp64")
    with pytest.raises(N.ZHError) as ei:
        prop.steady_state(1.0, tol=1e-14, maxiter=20, precondition=False)
    assert ei.value.code in (-6, -5)  # ZH_ELIMIT, or ZH_ENONFINITE on a Krylov breakdown
    s, ps, om = prop.get_state()
    assert np.isfinite(ps).all() and np.isfinite(om).all()
    its, res = prop.steady_state(1.0, tol=1e-12, maxiter=2000)
    psi_ref, _ = _oracle_solution("small", 1.0)
    assert rel(prop.get_state()[1], psi_ref) < 1e-8
    prop.close()
