"""The oracle's stage-Hamiltonian models, pinned by what the mathematics fixes.

* Mean-field Hubbard dH (examples/Hubbard/Calculation/Bias.py:13-20): on an isolated
  device the evolution dσ_k/dt = -(i/ħ)[H0 + dH_k(σ), σ_k] conserves the mean-field
  energy E = Σ_k Tr(H0 σ_k) + U Σ_i (σ↑_ii - n0↑_i)(σ↓_ii - n0↓_i), because
  ∂E/∂σ_k = H0 + dH_k exactly when dH couples each spin to the other's density.
  A spin-swapped, half-U, sign-flipped or n0-swapped form breaks the invariant
  (planted here in copies of the oracle's form, on the CPU).
* Drive terms H(t) = H0 + Σ f_i(t) M_i: with [H0, M] = 0 the propagator is
  exp(-i(H0 t + F(t) M)/ħ), F = ∫ f (scipy quadrature), which fixes every waveform's
  shape, the sign and the ħ of its application.
* Dissipator rate: the closed form of the relaxation is in test_oracle_rhs.py; here
  the step rate's limits.
"""
import numpy as np
import pytest
from scipy.integrate import quad

from oracle import zand_oracle as O
from zandpack_amd import builder


def _isolated(NK=2, seed=3):
    p = builder.config_small(NK=NK, hubbard_U=2.0, seed=seed)
    p.xi[:] = 0
    p.Ixi[:] = 0
    return p


def _start(p, seed=4):
    rng = np.random.default_rng(seed)
    NK, n = p.dims[0], p.dims[1]
    A = rng.normal(size=(NK, n, n)) + 1j * rng.normal(size=(NK, n, n))
    sig = p.sigma0 + 0.08 * (A + A.conj().transpose(0, 2, 1))
    return (sig,) + p.zero_state()[1:]


def _energy_drift(p, dH, n0, U, T=1.5, eps=1e-11):
    y0 = _start(p)
    out = O.propagate(p.consts(), y0, 0.0, T, 0.005, eps, p.H0, dH, lambda t, a: 0.0, 2,
                      time_update="correct")
    E = np.array([O.hubbard_energy(p.H0, s, U, n0) for s in out["sigma"]])
    return np.abs(E - E[0]).max() / abs(E[0]), len(out["times"])


def _asym_n0(p):
    n0 = np.array(p.hubbard_n0, float)
    n0[1] = n0[1][::-1] + 0.05  # spin-asymmetric reference densities
    return n0


def test_hubbard_energy_conserved():
    p = _isolated()
    U, n0 = 2.0, _asym_n0(p)
    drift, nsteps = _energy_drift(p, O.hubbard_dH(U, n0), n0, U)
    assert nsteps > 20
    assert drift < 1e-8, drift


def _planted(kind, U, n0):
    """Plausible mistakes in the Hubbard form (test copies; the oracle is untouched)."""
    def dH(t, sig):
        no = sig.shape[-1]
        i = np.arange(no)
        out = np.zeros(sig.shape, complex)
        if kind == "same_spin":
            out[0, i, i] = U * (sig[0, i, i] - n0[0])
            out[1, i, i] = U * (sig[1, i, i] - n0[1])
        elif kind == "half_U":
            out[0, i, i] = 0.5 * U * (sig[1, i, i] - n0[1])
            out[1, i, i] = 0.5 * U * (sig[0, i, i] - n0[0])
        elif kind == "sign":
            out[0, i, i] = -U * (sig[1, i, i] - n0[1])
            out[1, i, i] = -U * (sig[0, i, i] - n0[0])
        elif kind == "n0_swapped":
            out[0, i, i] = U * (sig[1, i, i] - n0[0])
            out[1, i, i] = U * (sig[0, i, i] - n0[1])
        return out
    return dH


@pytest.mark.parametrize("kind", ["same_spin", "half_U", "sign", "n0_swapped"])
def test_hubbard_planted_mistakes_break_the_invariant(kind):
    p = _isolated()
    U, n0 = 2.0, _asym_n0(p)
    drift, _ = _energy_drift(p, _planted(kind, U, n0), n0, U)
    assert drift > 1e-5, (kind, drift)


def test_hubbard_dH_structure():
    """Diagonal, real on a Hermitian σ, zero at the reference densities."""
    p = _isolated()
    U, n0 = 2.0, _asym_n0(p)
    sig = _start(p)[0]
    d = O.hubbard_dH(U, n0)(0.0, sig)
    off = d - np.einsum("kii->ki", d)[:, :, None] * np.eye(d.shape[-1])
    assert np.abs(off).max() == 0.0
    assert np.abs(d.imag).max() < 1e-14
    s0 = np.zeros_like(sig)
    for k in range(2):
        s0[k][np.diag_indices(sig.shape[-1])] = n0[k]
    assert np.abs(O.hubbard_dH(U, n0)(0.0, s0)).max() == 0.0


# ------------------------------------------------------------------ drive terms
WAVES = {
    "cos_window": (0.7, (2.1, 0.3, 0.2, 1.1)),
    "sin_fermi": (0.9, (3.0, 0.1, 0.8, 0.1)),
    "gauss_sin": (1.3, (0.5, 0.9, 0.2, 0.6)),
    "const": (0.4, ()),
}


@pytest.mark.parametrize("kind", list(WAVES))
def test_drive_commuting_closed_form(kind):
    """[H0, M] = 0: σ(T) = U σ0 U†, U = exp(-i(H0 T + F(T) M)/ħ), F = ∫_0^T f."""
    p = _isolated(NK=1, seed=6)
    n = p.dims[1]
    e, v = np.linalg.eigh(p.H0[0])
    rng = np.random.default_rng(2)
    m = rng.normal(size=n)
    M = ((v * m) @ v.conj().T)[None]           # commutes with H0
    amp, par = WAVES[kind]
    f = O.waveform(kind, amp, par)
    dH = O.model_dH([(f, M)])
    T = 1.4
    y0 = _start(p)
    out = O.propagate(p.consts(), y0, 0.0, T, 0.005, 1e-11, p.H0, dH, lambda t, a: 0.0, 2,
                      time_update="correct")
    t_end = out["times"][-1]
    pts = [x for x in (par[2], par[3]) if 0 < x < t_end] if kind in ("cos_window",) else None
    F = quad(f, 0.0, t_end, points=pts, epsabs=1e-14, epsrel=1e-13, limit=400)[0]
    U = (v * np.exp(-1j * (e * t_end + F * m) / O.HBAR)) @ v.conj().T
    exact = U @ y0[0][0] @ U.conj().T
    assert np.abs(out["sigma"][-1][0] - exact).max() < 1e-8
    # the drive really entered
    U0 = (v * np.exp(-1j * e * t_end / O.HBAR)) @ v.conj().T
    assert np.abs(U0 @ y0[0][0] @ U0.conj().T - exact).max() > 1e-4


def test_waveform_limits():
    cw = O.waveform("cos_window", 2.0, (1.5, 0.0, 0.0, 80.0))
    assert cw(0.0) == 0.0 and cw(80.5) == 0.0 and np.isclose(cw(1e-9), 2.0)
    sf = O.waveform("sin_fermi", 1.0, (1.0, np.pi / 2, 12.0, 0.1))
    assert np.isclose(sf(0.0), 1.0, atol=1e-12) and abs(sf(20.0)) < 1e-30
    assert np.isclose(sf(12.0), 0.5 * np.cos(12.0))
    gs = O.waveform("gauss_sin", 1.0, (5.0, 0.05, 0.0, 0.0))
    assert gs(0.0) == 0.0 and abs(gs(40.0)) < 1e-25


def test_dissipator_rate_limits():
    r = O.dissipator_rate("step", 2.0, t0=1.0, width=0.1)
    assert np.isclose(r(-5.0), 2.0) and np.isclose(r(1.0), 1.0) and r(5.0) < 1e-15
    assert O.dissipator_rate("const", 0.3)(7.0) == 0.3
