"""Device-evaluated stage Hamiltonians against the oracle's models: drive terms
H(t) = H0 + Σ f_i(t) M_i in the forms of the reference's examples (a Rabi coupling
switched on and off, switched site potentials, a Gaussian-enveloped dipole ramp),
alone and with the Hubbard term, through dense, real and row-compressed H products;
complex drive entries; and the Hubbard term and drive terms in the non-orthogonal
(nozand) form, where S⁻¹H(t, σ) is formed every stage.  FP64 bar: 1e-10 on propagated
currents and states, 1e-11 on one attempt; mixed: 1e-5 on one attempt's state."""
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


def _drives(p, complex_entries=False):
    """Three terms after the reference's examples on the problem's orbitals."""
    NK, n = p.dims[:2]
    rabi = np.zeros((NK, n, n), complex)
    rabi[:, 0, 1] = rabi[:, 1, 0] = 1.0
    sites = np.zeros((NK, n, n), complex)
    sites[0, 0, 0] = 1.0
    sites[0, 5, 5] = 0.7
    sites[NK - 1, 3, 9] = sites[NK - 1, 9, 3] = 0.5
    x = np.arange(n, dtype=float)
    dip = np.zeros((NK, n, n), complex)
    for k in range(NK):
        dip[k][np.diag_indices(n)] = -2 * (x - x.mean()) / (x.max() - x.min())
    if complex_entries:
        sites[0, 2, 7] = 0.3j
        sites[0, 7, 2] = -0.3j
    return [builder.Drive("cos_window", 0.3, (0.9, 0.0, 0.2, 1.5), rabi),
            builder.Drive("sin_fermi", 0.6, (2.0, 0.3, 1.2, 0.1), sites),
            builder.Drive("gauss_sin", 0.8, (0.8, 0.5, 0.0, 0.6), dip)]


CASES = {
    "drive": lambda: builder.config_small(n_w=4, n_l=3, seed=21),
    "drive_hubbard": lambda: builder.config_small(NK=2, hubbard_U=2.0, seed=22),
    "drive_chain": lambda: builder.config_chain(),
}


def _replay(prop, ref):
    """Device attempts forced onto the oracle's accepted (t, h) sequence: stage-0 currents
    of every step and the final state."""
    J = [prop.attempt(t, h, 1.0, force_accept=True)["J"]
         for t, h in zip(ref["times"][:-1], ref["h_accepted"])]
    return np.array(J), prop.sigma()


# A drive term's time factor enters the error estimate, whose embedded difference
# cancels to ~1e-8 of the stage derivatives: a one-ulp change of f(t) (the device's and
# numpy's cos/sin/exp may differ by that) moves the oracle's own accepted times by
# 1.8e-8 and J by 2.7e-8 over this run.  The free-running comparison is held to 1e-6;
# the replay of the oracle's steps to the FP64 bar.
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("hproducts", ["dense", "sparse"])
def test_drive_propagation_fp64(native, case, hproducts):
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    p.drive = _drives(p)
    y0 = p.zero_state()
    say("drive propagation", case, hproducts, p.dims)
    ref = O.propagate(p.consts(), y0, 0.0, 2.0, 0.01, 1e-8, p.H0, dH_of(p), p.bias, p.dims[2])
    prop = Propagator(p, "fp64", hproducts=hproducts)
    assert prop.hamiltonian_info()["sparse"] == (hproducts == "sparse")
    J, sig = _replay(prop, ref)
    assert rel(J, ref["J"]) < 1e-10
    assert rel(sig, ref["y"][0]) < 1e-10
    prop.put_state(*y0)
    out = prop.propagate(0.0, 2.0, 0.01, 1e-8)
    assert out["n_steps"] == len(ref["times"]) - 1
    np.testing.assert_allclose(out["times"], ref["times"], rtol=1e-6)
    assert rel(out["J"], ref["J"]) < 1e-6
    # the drive changes the answer (not a no-op)
    p.drive = []
    ref0 = O.propagate(p.consts(), y0, 0.0, 2.0, 0.01, 1e-8, p.H0, dH_of(p), p.bias, p.dims[2])
    assert rel(ref0["y"][0], ref["y"][0]) > 1e-4
    prop.close()


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("complex_entries", [False, True])
def test_drive_attempt(native, case, precision, complex_entries):
    from zandpack_amd.engine import Propagator
    p = CASES[case]()
    p.drive = _drives(p, complex_entries)
    rng = np.random.default_rng(3)
    psi = 0.02 * (rng.normal(size=p.xi.shape) + 1j * rng.normal(size=p.xi.shape))
    y0 = (p.sigma0.astype(complex), psi, p.zero_state()[2])
    t0, h = 1.1, 0.03  # the stages straddle the sin_fermi switch-off
    say("drive attempt", case, precision, complex_entries)
    ks, TE, J0, y_new = O.dop54_attempt(p.consts(), y0, t0, h, p.H0, dH_of(p), p.bias, p.dims[2])
    prop = Propagator(p, precision, state=y0)
    assert prop.hamiltonian_info()["real"] == (not complex_entries)
    r = prop.attempt(t0, h, 1.0, force_accept=True)
    s, ps, om = prop.get_state()
    fp64 = precision == "fp64"
    assert abs(r["te"] - TE) <= (1e-10 if fp64 else 2e-3) * TE
    for got, want in zip((s, ps, om), y_new):
        assert rel(got, want) < (1e-12 if fp64 else 1e-5)
    prop.close()


def _nonortho(p, seed=1, band=2):
    NK, n = p.dims[:2]
    Sm05 = builder.lowdin_factors(builder.random_overlap(NK, n, band=band, seed=seed))
    return builder.to_nonorthogonal(p, Sm05)


@pytest.mark.parametrize("with_drive", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_overlap_hubbard_attempt(native, with_drive, precision):
    """nozand with a Hubbard dH (mpi/nozand:427): S⁻¹(H0 + dH(σ)) in the commutator and
    on Ψ, formed on the device every stage."""
    from zandpack_amd.engine import Propagator
    p = builder.config_small(NK=2, hubbard_U=2.0, seed=23)
    if with_drive:
        p.drive = _drives(p)
    q = _nonortho(p, seed=3)
    q.pulse = builder.Pulse("step", amp=0.2, p1=0.0, weights=(0.5, -0.5))
    rng = np.random.default_rng(4)
    psi = 0.02 * (rng.normal(size=q.xi.shape) + 1j * rng.normal(size=q.xi.shape))
    y0 = (q.sigma0, psi, q.zero_state()[2])
    t0, h = 1.1, 0.02
    say("overlap hubbard attempt", with_drive, precision)
    ks, TE, J0, y_new = O.dop54_attempt(q.consts(), y0, t0, h, q.H0, dH_of(q), q.bias, 2,
                                        Sinv=q.Sinv)
    prop = Propagator(q, precision, state=y0)
    r = prop.attempt(t0, h, 1.0, force_accept=True)
    s, ps, om = prop.get_state()
    fp64 = precision == "fp64"
    assert abs(r["te"] - TE) <= (1e-10 if fp64 else 2e-3) * TE
    for got, want in zip((s, ps, om), y_new):
        assert rel(got, want) < (1e-11 if fp64 else 1e-5)
    # the Hubbard term enters: without it the step differs by far more
    _, _, _, y_nd = O.dop54_attempt(q.consts(), y0, t0, h, q.H0,
                                     O.model_dH([(O.waveform(d.kind, d.amp, d.par), d.M)
                                                 for d in q.drive]), q.bias, 2, Sinv=q.Sinv)
    assert rel(y_nd[0], y_new[0]) > 1e-6
    prop.close()


def test_overlap_hubbard_propagation(native):
    from zandpack_amd.engine import Propagator
    p = builder.config_small(NK=2, hubbard_U=2.0, seed=24)
    p.drive = _drives(p)
    q = _nonortho(p, seed=5)
    q.pulse = builder.Pulse("gauss", 0.3, 1.0, 0.8, (0.5, -0.5))
    y0 = q.zero_state()
    say("overlap hubbard propagation")
    ref = O.propagate(q.consts(), y0, 0.0, 1.5, 0.01, 1e-8, q.H0, dH_of(q), q.bias, 2,
                      Sinv=q.Sinv)
    prop = Propagator(q, "fp64")
    J, sig = _replay(prop, ref)
    assert rel(J, ref["J"]) < 1e-10
    assert rel(sig, ref["y"][0]) < 1e-10
    prop.put_state(*y0)
    out = prop.propagate(0.0, 1.5, 0.01, 1e-8)
    assert out["n_steps"] == len(ref["times"]) - 1
    assert rel(out["J"], ref["J"]) < 1e-6
    prop.close()
