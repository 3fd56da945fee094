"""The propagation-constant builder (TimedependentTransport.py:858-1038) pinned by
equilibrium and algebra, for multi-channel (rank > 1) level widths and the Fermi B/C
slots:

* equilibrium is a fixed point: at zero bias the equilibrium density of the same
  pole-expanded Fermi function f̃ (PadeDecomp.py:36-42) and the leads' Lorentzian
  self-energy, closed by residues,
      σ_eq = ∫ dE/2π f̃(E) A(E) = I/2 + kT Σ_p R_p [G^r(w_p) + G^r(w_p)†],  w_p = μ + i q_p kT
  (the spectral sum ∫ A dE/2π = I gives the 1/2; each pole of f̃ picks G^r in the upper and
  G^a in the lower half-plane), with psinought's Ψ₀(σ_eq) and its closed-form Ω₀, makes
  dσ/dt vanish to rounding.  The Ψ source Λ^{<+} of the Fermi slots (GL_P = R_p λ_u / β
  on the B and C eigenvectors of Γ at the upper poles) is what balances the commutator;
  planted mistakes in the slot construction (the C slots' factor i dropped, the residues
  halved, the channel vectors swapped between B and C) leave it out of balance;
* the lower-pole eigenvalues: Γ(z̄) = Γ(z)† for Hermitian weights, so B_l = B_u and
  C_l = -C_u, the upper eigenvectors diagonalise both and λ_l = +λ_u (B) or -λ_u (C) —
  which the builder's Rayleigh quotient reproduces.
"""
import numpy as np
import pytest

from oracle import zand_oracle as O
from zandpack_amd import builder

def _problem(N_F=20, NOT=2, seed=3, kT=0.025, width=4):
    p = builder.config_honeycomb(width, 3, 3, 2 * N_F, NOT, NK=1, seed=seed, kT=kT,
                                 pulse=builder.Pulse("none"))
    return p, kT


def _imbalance(consts, H, sig_eq):
    delta = np.zeros(2)
    psi0, om0 = O.steady_state_dense(consts, H, sig_eq, delta)
    ds, _, _ = O.rhs(consts, sig_eq, psi0, om0, H, delta)
    comm = -1j * (H @ sig_eq - sig_eq @ H) / O.HBAR
    return np.abs(ds).max() / np.abs(comm).max()


def _sigma_eq(p, kT, N_F=20, mu=0.0):
    return O.equilibrium_density(p.consts(), p.H0, kT, N_F, mu)


# the balance holds for any pole count: both sides use the same f̃
@pytest.mark.parametrize("NOT,kT,width,N_F", [(2, 0.025, 4, 20), (3, 0.1, 3, 8)])
def test_equilibrium_is_a_fixed_point(NOT, kT, width, N_F):
    p, kT = _problem(N_F=N_F, NOT=NOT, kT=kT, width=width)
    sig_eq = _sigma_eq(p, kT, N_F=N_F)
    r = _imbalance(p.consts(), np.asarray(p.H0, complex), sig_eq)
    assert r < 1e-10, r
    # the equilibrium density is a genuine one: Hermitian, 0 <= occupations <= 1
    e = np.linalg.eigvalsh(sig_eq[0])
    assert e.min() > -1e-10 and e.max() < 1 + 1e-10


def _planted(p, kind):
    """Copies of the constants with a mistake in the Fermi slots (test only)."""
    N_L, N_F = 3, 20
    q = {k: np.array(v, copy=True) for k, v in p.consts().items()}
    B = slice(N_L, N_L + N_F)
    Cs = slice(N_L + N_F, N_L + 2 * N_F)
    if kind == "C_factor_i_dropped":
        q["GL_P"][:, :, Cs] *= -1j
    elif kind == "residues_halved":
        q["GL_P"][:, :, B] *= 0.5
        q["GL_P"][:, :, Cs] *= 0.5
    elif kind == "B_C_vectors_swapped":
        for key in ("xi", "Ixi"):
            b = q[key][:, :, B].copy()
            q[key][:, :, B] = q[key][:, :, Cs]
            q[key][:, :, Cs] = b
    return q


@pytest.mark.parametrize("kind", ["C_factor_i_dropped", "residues_halved", "B_C_vectors_swapped"])
def test_planted_slot_mistakes_break_the_balance(kind):
    p, kT = _problem()
    sig_eq = _sigma_eq(p, kT)
    r = _imbalance(_planted(p, kind), np.asarray(p.H0, complex), sig_eq)
    assert r > 1e-3, (kind, r)


def test_lower_pole_eigenvalues():
    """λ_l = +λ_u on the B slots and -λ_u on the C slots, through the builder."""
    N_L, N_F, NOT = 3, 20, 2
    p, kT = _problem(N_F=N_F, NOT=NOT)
    R = O.hu_poles(N_F)[1]
    beta = 1.0 / kT
    for x in range(N_L, N_L + 2 * N_F):
        p_idx = (x - N_L) % N_F
        fac = 1.0 if x < N_L + N_F else 1j
        lam_u = p.GL_P[0, :, x] * beta / (R[p_idx] * fac)          # (NL, NOT)
        lam_l = -p.GL_M[0, :, x] * beta / (R[p_idx] * fac)
        sign = 1.0 if x < N_L + N_F else -1.0
        np.testing.assert_allclose(lam_l, sign * lam_u, rtol=1e-12, atol=1e-14 * np.abs(lam_u).max())
        assert np.abs(lam_u.imag).max() < 1e-12 * max(1.0, np.abs(lam_u).max())
