"""The oracle's equilibrium-density pieces (cmdtools/SCF, FittedSelfEnergy) pinned
by closed forms: the Lorentzian pair's Kramers-Kronig relation, the level width
-2 Im Σ = Γ on the real axis, analyticity in the upper half plane, and, with no
self-energy, the contour sum against U f(e) U† with the contour's temperature."""
import os

import numpy as np
import pytest

from oracle import zand_oracle as O
from zandpack_amd import builder

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_lorentzian_pair_is_a_hilbert_pair():
    """KK(E) = (1/π) P∫ L(E')/(E - E') dE' on the real axis (numerical PV)."""
    eps, gam = 0.3, 0.7
    x = np.linspace(-400, 400, 800001)
    L, _ = O.lorentzian_basis(x, eps, gam)
    for E in (-1.1, 0.05, 2.4):
        # principal value by the symmetric subtraction on a grid centred at E
        xs = x + (E - x[np.argmin(np.abs(x - E))])
        Ls, _ = O.lorentzian_basis(xs, eps, gam)
        d = E - xs
        mask = np.abs(d) > 1e-9
        pv = np.trapezoid(np.where(mask, (Ls - O.lorentzian_basis(E, eps, gam)[0]) / np.where(mask, d, 1), 0), xs)
        pv += O.lorentzian_basis(E, eps, gam)[0] * np.log(abs((E - xs[0]) / (xs[-1] - E)))
        assert abs(pv / np.pi - O.lorentzian_basis(E, eps, gam)[1]) < 2e-3


def test_self_energy_width_and_analyticity():
    p = builder.config_small(NK=2, seed=3)
    c = p.consts()
    NK, n, NL, NX, NOT = p.dims
    # -2 Im Σ_α(E) = Γ_α(E) = Σ_x C_x L_x(E) on the real axis, Γ from the fit weights
    for E in (-0.7, 0.1, 1.3):
        S = O.lead_self_energy(c, E)
        G = 1j * (S - S.conj().transpose(0, 1, 3, 2))
        # Γ from the builder's own modes: λ ξ ⊗ Ixi weighted by the Lorentzian
        gam = np.zeros_like(S)
        for k in range(NK):
            for a in range(NL):
                for x in range(NX):
                    for cc in range(NOT):
                        d = c["dGP"][k, a, x, cc]
                        if d == 0:
                            continue
                        g = c["Xpp"][k, a, x, cc].imag
                        L = builder.lorentz(E, c["Xpp"][k, a, x, cc].real, g)
                        gam[k, a] += L * (2j * d / g) * np.outer(c["xi"][k, a, x, cc], c["Ixi"][k, a, x, cc])
        assert np.abs(G - gam).max() < 1e-12 * max(1, np.abs(gam).max())
        assert np.abs(gam).max() > 1e-3
    # analytic in the upper half plane: Cauchy-Riemann by finite differences
    z, h = 0.2 + 0.3j, 1e-6
    f = lambda e: O.lead_self_energy(c, e)
    dx = (f(z + h) - f(z - h)) / (2 * h)
    dy = (f(z + 1j * h) - f(z - 1j * h)) / (2j * h)
    assert np.abs(dx - dy).max() < 1e-6 * np.abs(dx).max()


@pytest.mark.parametrize("mu", [0.0, 0.3])
def test_contour_sum_is_the_fermi_function(mu):
    """No self-energy: ρ = U f(e) U† at the contour's kT (pole spacing 2πkT)."""
    C = np.load(os.path.join(GOLD, "scf_default_contour.npy"))
    kT = C[-1, 1] - C[-2, 1]
    kT /= 2 * np.pi
    assert abs(kT - 0.02585) < 1e-5
    rng = np.random.default_rng(0)
    n = 24
    A = rng.normal(size=(n, n)) + 0.3j * rng.normal(size=(n, n))
    H = ((A + A.conj().T) / 2 * 1.4)[None]
    e, V = np.linalg.eigh(H[0])
    Cs = C.copy()
    Cs[:, 0] += mu
    rho = O.contour_density(H, Cs, lambda E: np.zeros_like(H))
    f = 1 / (1 + np.exp((e - mu) / kT))
    ref = (V * f) @ V.conj().T
    assert np.abs(rho[0] - ref).max() < 5e-6
    assert np.abs(rho[0] - rho[0].conj().T).max() < 1e-12


def test_host_self_energy_matches_oracle():
    """The product's host evaluation of Σ on the coupled orbitals (scf.self_energy_cc)
    against the oracle's reference-form evaluation, with a bias shift."""
    from zandpack_amd import scf
    p = builder.config_small(NK=2, seed=5)
    corb = scf.coupled_orbitals(p)
    E = np.array([-0.4 + 0.05j, 0.2 + 1.1j, 1.5 + 0.3j])
    bias = np.array([0.1, -0.2])
    got = scf.self_energy_cc(p, E, corb, bias)
    for i, e in enumerate(E):
        ref = O.lead_self_energy(p.consts(), e, bias).sum(axis=1)
        assert np.abs(got[:, i] - ref[:, corb][:, :, corb]).max() < 1e-13 * np.abs(ref).max()
        # nothing outside corb
        mask = np.ones(ref.shape[-1], bool)
        mask[corb] = False
        assert np.abs(ref[:, mask]).max() == 0 and np.abs(ref[:, :, mask]).max() == 0


def test_pulay_mixer_finds_fixed_point():
    from zandpack_amd.scf import PulayMixer
    rng = np.random.default_rng(1)
    A = rng.normal(size=(20, 20))
    A = A + A.T
    A *= 0.9 / np.abs(np.linalg.eigvalsh(A)).max()
    b = rng.normal(size=20)
    xs = np.linalg.solve(np.eye(20) - A, b)
    x = np.zeros(20)
    mix = PulayMixer(weight=0.3, history=6)
    for it in range(60):
        r = A @ x + b - x
        if np.abs(r).max() < 1e-10:
            break
        x = mix(x, r)
    assert np.abs(x - xs).max() < 1e-9 and it < 80
    # plain linear mixing at the same weight needs far more iterations
    x = np.zeros(20)
    for lin in range(2000):
        r = A @ x + b - x
        if np.abs(r).max() < 1e-10:
            break
        x = x + 0.3 * r
    assert lin > 5 * it


# ---------------------------------------------------------------- bias window
def _wide_band(gl, gr):
    return lambda E: np.array([[[[-0.5j * gl]], [[-0.5j * gr]]]])


def test_window_single_level_closed_form():
    """One level ε between wide-band leads (Σ_α = -iΓ_α/2): the window density is
    (Γ_L/2π) ∫ (f_L - f_R) / ((E-ε)² + Γ²/4) dE, integrated here as a scalar."""
    from scipy.integrate import quad
    eps, gl, gr, kT = 0.1, 0.05, 0.08, 0.025
    mus = np.array([0.25, 0.0])
    H = np.array([[[eps]]], complex)
    e1, e2 = O.window_limits(mus, [kT, kT])
    x, w = O.gauss_legendre_panels(e1, e2, 200, 8)
    rho = O.window_density(H, x, w, _wide_band(gl, gr), mus, [kT, kT])
    g = gl + gr
    ref = quad(lambda E: gl / (2 * np.pi) * (O.fermi(E, mus[0], kT) - O.fermi(E, mus[1], kT))
               / ((E - eps) ** 2 + g * g / 4), e1, e2, epsabs=1e-14, epsrel=1e-12, limit=400)[0]
    assert abs(rho[0, 0, 0] - ref) < 1e-10 * abs(ref)
    # the sign of the bias sets which lead fills the window
    rho2 = O.window_density(H, x, w, _wide_band(gr, gl), mus[::-1], [kT, kT])
    assert rho2[0, 0, 0].real > 0 and abs(rho2[0, 0, 0].imag) < 1e-15


def _small_device(seed=2):
    p = builder.config_small(seed=seed)
    c = p.consts()
    H = np.asarray(p.H0, complex)
    return p, c, H


def test_window_panels_match_adaptive_quadrature():
    """The fixed Gauss-Legendre panels against scipy's adaptive quad_vec (the
    reference's integrator) on a small device with Lorentzian-fitted leads."""
    from scipy.integrate import quad_vec
    p, c, H = _small_device()
    mus, kTs = np.array([0.3, -0.1]), np.array([0.025, 0.025])
    sig = lambda E: O.lead_self_energy(c, E)
    e1, e2 = O.window_limits(mus, kTs)
    x, w = O.gauss_legendre_panels(e1, e2, 48, 8)
    rho = O.window_density(H, x, w, sig, mus, kTs)
    ref, err = quad_vec(lambda E: O.window_integrand(H, E, sig, mus, kTs), e1, e2,
                        epsabs=1e-11, epsrel=1e-10, quadrature="gk21")
    assert np.abs(rho - ref).max() < 1e-8 * np.abs(ref).max()
    assert np.abs(rho - rho.conj().transpose(0, 2, 1)).max() < 1e-12 * np.abs(rho).max()
    # the window adds charge for a positive bias on the lead at the higher potential
    assert np.trace(rho[0]).real > 0


def test_window_vanishes_in_equilibrium():
    p, c, H = _small_device()
    mus, kTs = np.array([0.1, 0.1]), np.array([0.025, 0.025])
    x, w = O.gauss_legendre_panels(-0.5, 0.5, 8, 8)
    rho = O.window_density(H, x, w, lambda E: O.lead_self_energy(c, E), mus, kTs)
    assert np.abs(rho).max() == 0.0


def test_window_closes_the_contour_with_one_lead():
    """With the low-potential lead decoupled (Σ_R = 0), the device is in equilibrium
    with the other lead: contour(μ_R) + window(f_L - f_R) = contour(μ_L)."""
    C = np.load(os.path.join(GOLD, "scf_default_contour.npy"))
    kT = (C[-1, 1] - C[-2, 1]) / (2 * np.pi)
    p, c, H = _small_device(seed=4)

    def sig_leads(E):
        s = O.lead_self_energy(c, E)
        s[:, 1] = 0.0
        return s

    sig_tot = lambda E: sig_leads(E).sum(axis=1)
    mus, kTs = np.array([0.2, 0.0]), np.array([kT, kT])
    lo, hi = C.copy(), C.copy()
    lo[:, 0] += mus[1]
    hi[:, 0] += mus[0]
    e1, e2 = O.window_limits(mus, kTs, fact_kT=20.0)
    x, w = O.gauss_legendre_panels(e1, e2, 96, 8)
    rho = O.contour_density(H, lo, sig_tot) + O.window_density(H, x, w, sig_leads, mus, kTs)
    ref = O.contour_density(H, hi, sig_tot)
    assert np.abs(rho - ref).max() < 2e-5 * np.abs(ref).max()


def test_bias_window_nodes():
    """The product's window nodes: the reference's limits (mu_min - 10 kT_max,
    mu_max + 10 kT_max), panels no wider than the smallest kT, and a rule exact for
    polynomials of degree 2*order - 1 on each panel."""
    from zandpack_amd import scf
    mus, kTs = np.array([0.15, -0.15]), np.array([0.025, 0.02])
    x, w = scf.bias_window(mus, kTs)
    e1, e2 = O.window_limits(mus, kTs)
    assert e1 < x.min() and x.max() < e2
    assert abs(w.sum() - (e2 - e1)) < 1e-13
    panels = x.size // 8
    assert (e2 - e1) / panels <= kTs.min() + 1e-15
    for deg in (1, 5, 15):
        exact = (e2 ** (deg + 1) - e1 ** (deg + 1)) / (deg + 1)
        assert abs(np.dot(w, x ** deg) - exact) < 1e-12 * max(1.0, abs(exact))
    xo, wo = O.gauss_legendre_panels(e1, e2, panels, 8)
    np.testing.assert_allclose(np.sort(x), np.sort(xo), rtol=0, atol=1e-14)
