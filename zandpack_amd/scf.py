"""Density matrix σ₀ on the GPU: the contour sum, the bias-window term and the
self-consistency loop of ``Zandpack/cmdtools/SCF``.

The lead self-energy is the one the propagation constants encode
(``FittedSelfEnergy.evaluate``: Σ = ½·KK[Γ] - ½i·Γ over the Lorentzian fit): for a
Lorentzian mode m with centre ε, width γ (χ⁺ = ε + iγ = Xpp) and
GG_P - GL_P = -½iλγ,

    Σ(E) = Σ_m  i (GG_P - GL_P)_m / (E - bias_α - Xpm_m)  ξ_m ⊗ Ixi_m ,

Xpm = ε - iγ; Fermi-pole modes have GG_P = GL_P and drop out.  It lives on the
orbitals the Lorentzian modes couple (``corb``), where it is evaluated on the host
once per contour.  Each SCF iteration then takes the eigenpairs of H(ρ) (host
LAPACK, as the steady-state preconditioner does) and one ``zh_contour_density`` call,
which forms Σ_i (i/2π)(W_i G_i - h.c.) with MFMA GEMMs through the Woodbury identity
on ``corb`` instead of one n x n inversion per contour point.

The contour is an (nE, 4) array (Re E, Im E, Re W, Im W) around μ = 0, the format
of the SCF tool's ``Contour`` option and its default contour file; it is shifted by
the lowest chemical potential (SCF:119).
"""
from __future__ import annotations

import ctypes as C
import warnings

import numpy as np

from . import _native as N


def coupled_orbitals(problem):
    """Orbitals touched by a Lorentzian mode (GG_P != GL_P) of any lead or k."""
    NK, n = problem.dims[:2]
    live = problem.dGP != 0
    nz = (np.abs(problem.xi) > 0) & live[..., None]
    nz |= (np.abs(problem.Ixi) > 0) & live[..., None]
    return np.nonzero(nz.reshape(-1, n).any(axis=0))[0].astype(np.int32)


def self_energy_cc(problem, E, corb, bias=None):
    """Σ(E_i) summed over leads on the orbitals corb: (NK, nE, nc, nc) complex128."""
    return self_energy_cc_leads(problem, E, corb, bias).sum(axis=0)


def self_energy_cc_leads(problem, E, corb, bias=None):
    """Σ_α(E_i) per lead on the orbitals corb: (NL, NK, nE, nc, nc) complex128."""
    NK, n, NL, NX, NOT = problem.dims
    E = np.asarray(E, complex).reshape(-1)
    b = np.zeros(NL) if bias is None else np.asarray(bias, float)
    out = np.zeros((NL, NK, len(E), len(corb), len(corb)), complex)
    for k in range(NK):
        for a in range(NL):
            d = problem.dGP[k, a].reshape(-1)
            m = np.nonzero(d != 0)[0]
            if m.size == 0:
                continue
            xm = problem.Xpm[k, a].reshape(-1)[m]
            X = problem.xi[k, a].reshape(-1, n)[m][:, corb]        # (Mm, nc)
            Y = problem.Ixi[k, a].reshape(-1, n)[m][:, corb]
            s = 1j * d[m][None, :] / (E[:, None] - b[a] - xm[None, :])   # (nE, Mm)
            out[a, k] = np.matmul(np.ascontiguousarray((s[:, :, None] * X[None]).transpose(0, 2, 1)), Y)
    return out


def eigenpairs(H):
    """H = V diag(e) V† per k (host LAPACK), in the layout the density entries take."""
    e, V = np.linalg.eigh(np.asarray(H, complex))
    return np.ascontiguousarray(e, np.float64), np.ascontiguousarray(V, np.complex128)


def contour_density(H, contour, corb, sigma_cc, device=0, eig=None, noalloc=False):
    """One device evaluation of Σ_i (i/2π)(W_i G_i - h.c.) for H (NK, n, n);
    ``eig`` = eigenpairs(H) when the caller has them already.  ``noalloc`` runs it
    without device allocation (``zh_contour_density_noalloc``: the library's fixed
    scratch array, sizes up to about n = 1000)."""
    lib = N.lib()
    H = np.asarray(H, complex)
    NK, n, _ = H.shape
    e, V = eigenpairs(H) if eig is None else eig
    cont = np.ascontiguousarray(contour, np.float64)
    assert cont.ndim == 2 and cont.shape[1] == 4
    corb = np.ascontiguousarray(corb, np.int32)
    nc = int(corb.size)
    sig = None
    if nc:
        sig = N.c128(sigma_cc)
        assert sig.shape == (NK, cont.shape[0], nc, nc)
    rho = np.empty((NK, n, n), complex)
    fn = lib.zh_contour_density_noalloc if noalloc else lib.zh_contour_density
    N.check(fn(
        NK, n, cont.shape[0], nc, N.ptr(cont), N.ptr(e), N.ptr(V),
        corb.ctypes.data_as(C.POINTER(C.c_int)) if nc else None,
        N.ptr(sig) if nc else None, N.ptr(rho), int(device)))
    return rho


def fermi(E, mu, kT):
    """1 / (1 + exp((E - mu)/kT))."""
    return 1.0 / (1.0 + np.exp((np.asarray(E, float) - mu) / kT))


def bias_window(mus, kTs, fact_kT=10.0, panels=None, order=8):
    """Real-axis nodes and weights over [mu_min - fact_kT kT_max, mu_max + fact_kT kT_max]
    (the integration range of cmdtools/SCF:273-276): composite Gauss-Legendre of
    ``order`` points on ``panels`` equal panels (default: panels no wider than the
    smallest kT).  The reference integrates adaptively (scipy quad_vec, gk21,
    epsabs 1e-5, epsrel 1e-4); a fixed rule keeps the nodes, and so the self-energy
    evaluations, the same in every SCF iteration."""
    mus, kTs = np.asarray(mus, float), np.asarray(kTs, float)
    e1 = mus.min() - fact_kT * kTs.max()
    e2 = mus.max() + fact_kT * kTs.max()
    if panels is None:
        panels = max(8, int(np.ceil((e2 - e1) / kTs.min())))
    x, w = np.polynomial.legendre.leggauss(int(order))
    edges = np.linspace(e1, e2, int(panels) + 1)
    half = 0.5 * np.diff(edges)
    mid = 0.5 * (edges[:-1] + edges[1:])
    nodes = (mid[:, None] + half[:, None] * x[None]).reshape(-1)
    weights = (half[:, None] * w[None]).reshape(-1)
    return nodes, weights


def window_weights(sigma_leads, nodes, mus, kTs):
    """M_i = Σ_α (f_α(E_i) - f_min(E_i)) Γ_α(E_i) with Γ_α = i(Σ_α - Σ_α†)
    (the SE_l of cmdtools/SCF:246-247 with the fdf of :301-302, times -i):
    (NK, nE, nc, nc) from ``sigma_leads`` (NL, NK, nE, nc, nc)."""
    mus, kTs = np.asarray(mus, float), np.asarray(kTs, float)
    imin = int(np.argmin(mus))
    f_min = fermi(nodes, mus[imin], kTs[imin])
    out = np.zeros(sigma_leads.shape[1:], complex)
    for a in range(sigma_leads.shape[0]):
        fd = fermi(nodes, mus[a], kTs[a]) - f_min
        S = sigma_leads[a]
        out += fd[None, :, None, None] * (1j * (S - S.conj().swapaxes(-1, -2)))
    return out


def window_density(H, nodes, weights, corb, sigma_cc, gamma_w, device=0, eig=None):
    """One device evaluation of Σ_i w_i (1/2π) G(E_i)[:, c] M_i G(E_i)[:, c]† for H
    (NK, n, n): the bias-window density of cmdtools/SCF:299-306; ``eig`` as for
    ``contour_density``."""
    lib = N.lib()
    H = np.asarray(H, complex)
    NK, n, _ = H.shape
    e, V = eigenpairs(H) if eig is None else eig
    nw = np.ascontiguousarray(np.stack([nodes, weights], axis=1), np.float64)
    corb = np.ascontiguousarray(corb, np.int32)
    nc = int(corb.size)
    sig = N.c128(sigma_cc)
    gam = N.c128(gamma_w)
    assert sig.shape == gam.shape == (NK, nw.shape[0], nc, nc)
    rho = np.empty((NK, n, n), complex)
    N.check(lib.zh_window_density(
        NK, n, nw.shape[0], nc, N.ptr(nw), N.ptr(e), N.ptr(V),
        corb.ctypes.data_as(C.POINTER(C.c_int)), N.ptr(sig), N.ptr(gam), N.ptr(rho),
        int(device)))
    return rho


class PulayMixer:
    """Pulay (DIIS) mixing of ρ_in with the residual ρ_out - ρ_in, the role of
    ``sisl.mixing.PulayMixer(weight, history)`` in cmdtools/SCF:127-130 (sisl is an
    external package, absent here; this is the published method): keep the last
    ``history`` pairs (x_j, r_j), take the coefficients c minimising |Σ c_j r_j| with
    Σ c_j = 1, and return Σ c_j (x_j + weight r_j)."""

    def __init__(self, weight=0.2, history=6):
        self.weight = float(weight)
        self.history = int(history)
        self.x, self.r = [], []

    def __call__(self, x, r):
        self.x.append(x.copy())
        self.r.append(r.copy())
        if len(self.x) > self.history:
            self.x.pop(0)
            self.r.pop(0)
        m = len(self.r)
        if m == 1:
            return x + self.weight * r
        # c = (γ, 1 - Σγ) with γ the least-squares solution of [r_j - r_m] γ = -r_m
        # (a QR solve on the residual differences, not the normal equations)
        rm = self.r[-1].reshape(-1)
        D = np.stack([(rj.reshape(-1) - rm) for rj in self.r[:-1]], axis=1)
        D = np.concatenate([D.real, D.imag]) if np.iscomplexobj(D) else D
        rr = np.concatenate([rm.real, rm.imag]) if np.iscomplexobj(rm) else rm
        g = np.linalg.lstsq(D, -rr, rcond=None)[0]
        c = np.append(g, 1.0 - g.sum())
        return sum(cj * (xj + self.weight * rj) for cj, xj, rj in zip(c, self.x, self.r))


def self_consistent_density(problem, contour, mu, rho0=None, t0=0.0, tol=1e-5, weight=0.2,
                            history=6, max_iter=500, device=0, mixer=None, callback=None,
                            kT=0.025, nonequilibrium=False, fact_kT=10.0, window_panels=None):
    """The SCF tool's loop (cmdtools/SCF:282-309): H = H0 + dH(t0, ρ_in), ρ_out from
    the contour sum at the lowest chemical potential (the leads' ``mu`` shifted by the
    bias at t0, SCF:93-99, 119), plus with ``nonequilibrium`` the bias-window term on
    the real axis (SCF:299-306; ``kT`` per lead or one value), stop when
    max|ρ_out - ρ_in| <= tol, else ρ_in <- mixer(ρ_in, δ).  Unequal chemical
    potentials without ``nonequilibrium`` give a warning and the equilibrium density
    at the lowest one, as the tool does (SCF:100-101).  Returns (ρ_out, iterations,
    converged)."""
    NK, n, NL = problem.dims[:3]
    bias = np.array([problem.bias(t0, a) for a in range(NL)])
    mus = np.broadcast_to(np.asarray(mu, float), (NL,)) + bias
    kTs = np.broadcast_to(np.asarray(kT, float), (NL,))
    window = mus.max() - mus.min() > 1e-6
    if window and not nonequilibrium:
        warnings.warn("unequal chemical potentials without nonequilibrium=True: "
                      "equilibrium density at the lowest one", stacklevel=2)
        window = False
    cont = np.array(contour, float)
    cont[:, 0] += mus.min()
    corb = coupled_orbitals(problem)
    sig = self_energy_cc(problem, cont[:, 0] + 1j * cont[:, 1], corb, bias)
    if window:
        if corb.size == 0:
            raise ValueError("nonequilibrium density needs lead-coupled orbitals")
        nodes, wts = bias_window(mus, kTs, fact_kT, window_panels)
        sig_l = self_energy_cc_leads(problem, nodes.astype(complex), corb, bias)
        sig_w = sig_l.sum(axis=0)
        gam_w = window_weights(sig_l, nodes, mus, kTs)
    rho_in = np.array(problem.sigma0 if rho0 is None else rho0, complex)
    mixer = mixer or PulayMixer(weight, history)
    it, ok = 0, False
    rho_out = rho_in
    while it < max_iter:
        H = np.asarray(problem.H0, complex) + problem.dH(t0, rho_in)
        eig = eigenpairs(H)
        rho_out = contour_density(H, cont, corb, sig, device, eig)
        if window:
            rho_out = rho_out + window_density(H, nodes, wts, corb, sig_w, gam_w, device, eig)
        delta = rho_out - rho_in
        it += 1
        if callback:
            callback(it, rho_out, delta)
        if np.abs(delta).max() <= tol:
            ok = True
            break
        rho_in = mixer(rho_in, delta)
    return rho_out, it, ok
