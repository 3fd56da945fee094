"""Propagation-constant builder and seeded synthetic problems (setup side).

Builds the per-mode arrays the propagator consumes -- ξ, Ixi, Λ^{<,>}_± (GL_P,
GG_P, GL_M, GG_M), χ± (Xpp, Xpm) -- from a Lorentzian expansion of each lead's
level-width matrix and a Hu (2011) pole expansion of the Fermi function,
following Zandpack/TimedependentTransport.py:932-1038 (``get_propagation_quantities``),
:858-930 (``diagonalise``: B/C split of Γ at the Fermi poles), :72-83 (pole
placement μ ± z_p kT) and Zandpack/PadeDecomp.py:84-152 (Hu poles/residues).

The mode index is the reference's: ``x < N_L`` Lorentzian terms, then ``N_F``
"B" slots and ``N_F`` "C" slots (NX = N_L + 2 N_F).  Arrays are complex128 in
the reference layout ``(NK, NL, NX, NOT[, n])``.

Runs once per problem on the host (numpy); it is not on the timed path.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from fractions import Fraction

import numpy as np

HBAR = 6.582119569e-1  # eV*fs (Zandpack/td_constants.py:3)


# ----------------------------------------------------------------- Fermi poles
def hu_poles(N: int):
    """N-pole Hu expansion of the Fermi function: (z_p = i q_p, R_p).

    q_p are 2/λ for the positive eigenvalues λ of the 2N x 2N tridiagonal
    matrix with off-diagonals 1/sqrt((2i-1)(2i+1)); the residues follow the
    closed product formula of Hu, Xu, Yan, J. Chem. Phys. 134, 244106 (2011)
    as used by PadeDecomp.py:114-147: float64 products up to N=40, exact
    rational products of the float64 roots above (where float64 products lose
    the last digits).
    """
    if N <= 0:
        return np.zeros(0, complex), np.zeros(0)
    m = 2 * N
    b = 2.0 * np.arange(1, m + 1) - 1.0
    off = 1.0 / np.sqrt(b[:-1] * b[1:])
    T = (np.diag(off, 1) + np.diag(off, -1)).astype(complex)
    lam = np.linalg.eigh(T)[0]
    q = 2.0 / np.sort(lam[lam > 1e-15])[::-1]
    lam_p = np.linalg.eigh(T[1:, 1:])[0]
    p = 2.0 / lam_p[lam_p > 1e-15]
    R = np.empty(N)
    if N <= 40:
        c0 = N * (2 * N + 1) / 2.0
        for i in range(N):
            dq = q ** 2 - q[i] ** 2
            dq[i] = 1.0
            R[i] = c0 * np.prod(p ** 2 - q[i] ** 2) / np.prod(dq)
    else:
        qf = [Fraction(float(v)) for v in q]
        pf = [Fraction(float(v)) for v in p]
        c0 = Fraction(N * (2 * N + 1), 2)
        for i in range(N):
            num = Fraction(1)
            for v in pf:
                num *= v * v - qf[i] * qf[i]
            den = Fraction(1)
            for j, v in enumerate(qf):
                if j != i:
                    den *= v * v - qf[i] * qf[i]
            R[i] = float(c0 * num / den)
    return 1j * q, R


def fermi_expanded(E, z, R, beta, mu=0.0):
    """Pole-expanded Fermi function (PadeDecomp.py:36-42)."""
    E = np.asarray(E, dtype=complex)
    up = 1.0 / (E[..., None] - (mu + z / beta))
    dn = 1.0 / (E[..., None] - (mu - z / beta))
    return 0.5 - ((up + dn) * R).sum(axis=-1) / beta


def lorentz(z, centre, width):
    """Lorentzian basis γ²/((z-ε)²+γ²) of the level-width fit."""
    return width ** 2 / ((z - centre) ** 2 + width ** 2)


# ------------------------------------------------------------------ devices
def chain_hamiltonian(n, hop=-1.0, onsite=0.0):
    H = np.diag(np.full(n, onsite, float))
    i = np.arange(n - 1)
    H[i, i + 1] = hop
    H[i + 1, i] = hop
    return H


def honeycomb_patch(width, length, hop=-2.7, disorder=0.0, rng=None):
    """Brick-wall (honeycomb-topology) patch of ``width*length`` orbitals.

    Site (r, c) has index c*width + r; bonds (r,c)-(r+1,c) always and
    (r,c)-(r,c+1) when r+c is even, giving coordination 3 in the bulk like a
    zigzag ribbon.  Returns (H, left_edge, right_edge).
    """
    n = width * length
    H = np.zeros((n, n))
    idx = lambda r, c: c * width + r
    for c in range(length):
        for r in range(width):
            if r + 1 < width:
                H[idx(r, c), idx(r + 1, c)] = H[idx(r + 1, c), idx(r, c)] = hop
            if c + 1 < length and (r + c) % 2 == 0:
                H[idx(r, c), idx(r, c + 1)] = H[idx(r, c + 1), idx(r, c)] = hop
    if disorder and rng is not None:
        H[np.arange(n), np.arange(n)] += rng.uniform(-disorder, disorder, n)
    left = np.array([idx(r, 0) for r in range(width)])
    right = np.array([idx(r, length - 1) for r in range(width)])
    return H, left, right


def fermi_dirac_dm(H, kT, mu=0.0):
    e, v = np.linalg.eigh(H)
    f = 0.5 * (1.0 - np.tanh((e - mu) / (2.0 * kT)))
    return (v * f) @ v.conj().T


# -------------------------------------------------------------- lead constants
def _sorted_eigh(Mat, n_keep):
    """Hermitian eigendecomposition sorted by |λ| descending, first n_keep."""
    e, v = np.linalg.eigh(Mat)
    order = np.argsort(-np.abs(e), kind="stable")[:n_keep]
    return e[order], v[:, order]


def lead_mode_constants(n, orbitals, centres, widths, G, N_F, NOT, mu, kT):
    """Constants of one lead at one k.

    ``G[x]`` is the (n_α x n_α) Hermitian PSD weight of Lorentzian x on the
    lead's orbitals.  Returns per-mode arrays of shape (NX, NOT) and vectors
    (NX, NOT, n), following TimedependentTransport.py:964-1032.
    """
    N_L = len(centres)
    NX = N_L + 2 * N_F
    beta = 1.0 / kT
    z, R = hu_poles(N_F)
    out = {k: np.zeros((NX, NOT), complex) for k in ("GG_P", "GG_M", "GL_P", "GL_M", "Xpp", "Xpm")}
    xi = np.zeros((NX, NOT, n), complex)
    for x in range(N_L):
        lam, vec = _sorted_eigh(G[x], NOT)
        chi_p = centres[x] + 1j * widths[x]
        chi_m = np.conj(chi_p)
        fp = fermi_expanded(np.array([chi_p]), z, R, beta, mu)[0]
        fm = fermi_expanded(np.array([chi_m]), z, R, beta, mu)[0]
        W = widths[x]
        out["GG_P"][x] = -0.5j * lam * W * (1 - fp)
        out["GG_M"][x] = -0.5j * lam * W * (1 - fm)
        out["GL_P"][x] = 0.5j * lam * W * fp
        out["GL_M"][x] = 0.5j * lam * W * fm
        out["Xpp"][x] = chi_p
        out["Xpm"][x] = chi_m
        xi[x][:, orbitals] = vec.T
    for p in range(N_F):
        zu = mu + z[p] * kT
        zl = mu - z[p] * kT
        Gam_u = sum(lorentz(zu, centres[x], widths[x]) * G[x] for x in range(N_L))
        Gam_l = sum(lorentz(zl, centres[x], widths[x]) * G[x] for x in range(N_L))
        Bu = (Gam_u + Gam_u.conj().T) / 2
        Cu = (Gam_u - Gam_u.conj().T) / 2j
        Bl = (Gam_l + Gam_l.conj().T) / 2
        Cl = (Gam_l - Gam_l.conj().T) / 2j
        for slot, (Mu, Ml, fac) in enumerate(((Bu, Bl, 1.0), (Cu, Cl, 1j))):
            lam_u, vec = _sorted_eigh(Mu, NOT)
            # eigenvalue at the lower pole on the same eigenvector
            lam_l = np.real(np.einsum("oc,op,pc->c", vec.conj(), Ml, vec))
            x = N_L + slot * N_F + p
            out["GG_P"][x] = +R[p] * fac * lam_u / beta
            out["GL_P"][x] = +R[p] * fac * lam_u / beta
            out["GG_M"][x] = -R[p] * fac * lam_l / beta
            out["GL_M"][x] = -R[p] * fac * lam_l / beta
            out["Xpp"][x] = zu
            out["Xpm"][x] = zl
            xi[x][:, orbitals] = vec.T
    return out, xi


# ---------------------------------------------------------------- problem type
@dataclass
class Pulse:
    """Bias waveform V(t); lead α sees Δ_α(t) = weights[α] * V(t).

    kinds: 'none', 'step' (amp for t>=t0), 'gauss' amp*exp(-((t-tc)/w)^2),
    'ramp' amp*(1-exp(-t/tau)), 'sin' amp*sin(omega*t/ħ) (omega in eV),
    'box' amp*(tanh(t/ts)-tanh((t-tp)/ts))/2 (Zandpack/Pulses.py:76-77),
    'generic' amp*exp(-((t-d)/s)^2)*sin(2πf(t-d)-phase) with (p1, p2, p3, p4) =
    (s, f, phase, d) (generic_pulse with env1, Pulses.py:52-66),
    'stairs' amp*Σ_{i=1}^{n-1}(1-1/(1+exp((t-i w)/soft))) with (p1, p2, p3) =
    (w, n, soft) (Pulses.py:79-88),
    'spline' amp*S(t)*(tanh((t-a)/ts)-tanh((t-b)/ts))/2 inside the knots with
    (p1, p2, p3) = (a, b, ts) and S the piecewise cubic (spline_x, spline_c) in
    scipy's CubicSpline layout (the data-driven pulses, Pulses.py:5-50).
    """
    kind: str = "none"
    amp: float = 0.0
    p1: float = 0.0
    p2: float = 0.0
    weights: tuple = (0.5, -0.5)
    p3: float = 0.0
    p4: float = 0.0
    spline_x: np.ndarray | None = None
    spline_c: np.ndarray | None = None

    KINDS = {"none": 0, "step": 1, "gauss": 2, "ramp": 3, "sin": 4, "box": 5, "generic": 6,
             "stairs": 7, "spline": 8}

    def value(self, t):
        k = self.kind
        if k == "none":
            return 0.0
        if k == "step":
            return self.amp if t >= self.p1 else 0.0
        if k == "gauss":
            return self.amp * np.exp(-((t - self.p1) / self.p2) ** 2)
        if k == "ramp":
            return self.amp * (1.0 - np.exp(-t / self.p1))
        if k == "sin":
            return self.amp * np.sin(self.p1 * t / HBAR)
        if k == "box":
            return self.amp * (np.tanh(t / self.p2) - np.tanh((t - self.p1) / self.p2)) / 2
        if k == "generic":
            x = (t - self.p4) / self.p1
            return self.amp * np.exp(-x * x) * np.sin(2 * np.pi * self.p2 * (t - self.p4) - self.p3)
        if k == "stairs":
            r = 0.0
            with np.errstate(over="ignore"):  # exp -> inf gives the exact limit 0
                for i in range(1, int(self.p2)):
                    r += 1.0 - 1.0 / (1.0 + np.exp((t - i * self.p1) / self.p3))
            return self.amp * r
        if k == "spline":
            x, c = self.spline_x, self.spline_c
            if t < x[0] or t > x[-1]:
                return 0.0
            i = min(int(np.searchsorted(x, t, side="right")) - 1, len(x) - 2)
            d = t - x[i]
            S = ((c[0, i] * d + c[1, i]) * d + c[2, i]) * d + c[3, i]
            return self.amp * S * (np.tanh((t - self.p1) / self.p3) - np.tanh((t - self.p2) / self.p3)) / 2
        raise ValueError(k)

    def bias(self, t, a):
        return float(self.weights[a] * self.value(t))


def spline_pulse(t, y, window, amp=1.0, weights=(0.5, -0.5)):
    """A 'spline' pulse through the samples (t, y): scipy's CubicSpline (its default
    not-a-knot ends, as Pulses.py:19-22 builds it) inside the tanh window
    (a, b, ts)."""
    from scipy.interpolate import CubicSpline
    cs = CubicSpline(np.asarray(t, float), np.asarray(y, float))
    a, b, ts = window
    return Pulse("spline", amp, a, b, tuple(weights), ts, 0.0,
                 np.ascontiguousarray(cs.x, float), np.ascontiguousarray(cs.c, float))


# The two measured waveforms of Zandpack/ExpPulses and how Pulses.py:5-50 prepares
# them: rows [lo:hi] of the (t [ps], V) table, t in fs from 0, V / max|V|, window.
EXPERIMENTAL_PULSES = {
    "air_photonics": dict(rows=(10, 60), window=(100.0, 550.0, 15.0)),
    "toptica": dict(rows=(90, 170), window=(300.0, 4000.0, 300.0)),
}


def experimental_pulse(name, table, amp=1.0, weights=(0.5, -0.5)):
    """The data-driven pulse ``name`` ('air_photonics' or 'toptica') from its
    ExpPulses table (the (N, 2) array of the reference's .npy file)."""
    spec = EXPERIMENTAL_PULSES[name]
    t, v = np.asarray(table, float).T
    lo, hi = spec["rows"]
    t = t[lo:hi] * 1000.0
    v = v[lo:hi]
    v = v / np.abs(v).max()
    t = t - t.min()
    return spline_pulse(t, v, spec["window"], amp, weights)


@dataclass
class Dissipator:
    """dissipator(t, σ) = η(t)(σ - σ_eq), subtracted in the σ equation as
    D -= ħ·dissipator (Zandpack/mpi/zand:343; the form of
    examples/Hubbard/Calculation/Bias.py:29-39).  kinds: 'none', 'const'
    (η = eta0), 'step' (η = eta0 / (1 + exp((t - t0)/width)))."""
    kind: str = "none"
    eta0: float = 0.0
    t0: float = 0.0
    width: float = 1.0
    sig_eq: np.ndarray | None = None  # (NK, n, n)

    KINDS = {"none": 0, "const": 1, "step": 2}

    def eta(self, t):
        if self.kind == "none":
            return 0.0
        if self.kind == "const":
            return self.eta0
        if self.kind == "step":
            return self.eta0 / (1.0 + np.exp((t - self.t0) / self.width))
        raise ValueError(self.kind)

    def __call__(self, t, sig):
        if self.kind == "none":
            return np.zeros_like(sig)
        return self.eta(t) * (sig - self.sig_eq)


@dataclass
class Drive:
    """One drive term f(t) M of a time-dependent Hamiltonian H(t) = H0 + Σ f_i(t) M_i,
    the dH(t, σ) of a Bias.py (Zandpack/mpi/zand:318-319).  kinds of f:
      'const'       amp
      'cos_window'  amp cos(ω t + φ) for t_on < t <= t_off, else 0; par = (ω, φ, t_on, t_off)
                    (the Rabi drive of examples/nb_tutorial_updated/Tight-binding/T1_Rabi/Bias.py:16-24)
      'sin_fermi'   amp sin(ω t + φ) / (1 + exp((t - t_off)/w)); par = (ω, φ, t_off, w)
                    (examples/Reproduce_and_checks/scipy_vs_ARK4_Zand/shared.py:4-19)
      'gauss_sin'   amp exp(-((t - d)/s)^2) sin(2π f (t - d) - φ); par = (s, f, φ, d)
                    (examples/nb_tutorial_updated/Tight-binding/T2_TB_Chain/Bias.py:9-10, 21-25)
    M is (NK, n, n); only its non-zero entries are stored on the device."""
    kind: str
    amp: float
    par: tuple
    M: np.ndarray

    def value(self, t):
        a, b, c, d = (tuple(float(x) for x in self.par) + (0.0,) * 4)[:4]
        if self.kind == "const":
            return self.amp
        if self.kind == "cos_window":
            return self.amp * np.cos(a * t + b) if c < t <= d else 0.0
        if self.kind == "sin_fermi":
            return self.amp * np.sin(a * t + b) / (1.0 + np.exp((t - c) / d))
        if self.kind == "gauss_sin":
            return self.amp * np.exp(-((t - d) / a) ** 2) * np.sin(2 * np.pi * b * (t - d) - c)
        raise ValueError(self.kind)


@dataclass
class AMEProblem:
    """Everything one propagation needs, in the reference's array layout."""
    H0: np.ndarray          # (NK, n, n)
    sigma0: np.ndarray      # (NK, n, n)
    xi: np.ndarray          # (NK, NL, NX, NOT, n)
    Ixi: np.ndarray         # (NK, NL, NX, NOT, n)
    GG_P: np.ndarray        # (NK, NL, NX, NOT)
    GL_P: np.ndarray
    GG_M: np.ndarray
    GL_M: np.ndarray
    Xpp: np.ndarray
    Xpm: np.ndarray
    pulse: Pulse = field(default_factory=Pulse)
    hubbard_U: float = 0.0
    hubbard_n0: np.ndarray | None = None  # (NK, n) reference densities
    dissipator: Dissipator = field(default_factory=Dissipator)
    Sinv: np.ndarray | None = None  # (NK, n, n) S⁻¹: the non-orthogonal (nozand) form
    drive: list = field(default_factory=list)  # Drive terms of H(t)
    hbar: float = HBAR
    name: str = "problem"
    t_end: float = 10.0
    eps: float = 1e-6
    h0: float = 0.01

    @property
    def dims(self):
        NK, NL, NX, NOT, n = self.xi.shape
        return NK, n, NL, NX, NOT

    @property
    def M(self):
        NK, NL, NX, NOT, n = self.xi.shape
        return NL * NX * NOT

    @property
    def dGP(self):
        return self.GG_P - self.GL_P

    @property
    def dGM(self):
        return self.GG_M - self.GL_M

    def consts(self):
        """Dict in the form the CPU oracle takes."""
        return dict(xi=self.xi, Ixi=self.Ixi, GL_P=self.GL_P, dGP=self.dGP, dGM=self.dGM,
                    Xpp=self.Xpp, Xpm=self.Xpm)

    def zero_state(self):
        NK, n, NL, NX, NOT = self.dims
        M = self.M
        return (self.sigma0.astype(complex).copy(),
                np.zeros((NK, NL, NX, NOT, n), complex),
                np.zeros((NK, NL, NX, NOT, NL, NX, NOT), complex))

    def dH(self, t, sig):
        """Host evaluation of the stage Hamiltonian increment dH(t, σ)."""
        out = np.zeros(sig.shape, complex)
        if self.hubbard_U != 0.0:
            n = sig.shape[-1]
            i = np.arange(n)
            out[0, i, i] = self.hubbard_U * (sig[1, i, i] - self.hubbard_n0[1])
            out[1, i, i] = self.hubbard_U * (sig[0, i, i] - self.hubbard_n0[0])
        for d in self.drive:
            out = out + d.value(t) * np.asarray(d.M, complex)
        return out

    def bias(self, t, a):
        return self.pulse.bias(t, a)


def lowdin_factors(S):
    """S^(-1/2) of a Hermitian positive-definite overlap (NK, n, n), the ``S^(-0.5)``
    array a zand run directory holds (Zandpack/Writer.py:83)."""
    e, v = np.linalg.eigh(np.asarray(S, complex))
    return (v * (1.0 / np.sqrt(e))[:, None, :]) @ v.conj().transpose(0, 2, 1)


def to_nonorthogonal(p, Sm05, zero_tol=1e-10):
    """The problem in the non-orthogonal basis that ``mpi/nozand`` propagates, from
    the orthogonalised arrays of ``p`` and S^(-1/2) (Zandpack/mpi/nozand:259-300,
    376, 808-812): H = S^½ H_o S^½ (entries below 1e-13 dropped, Hermitian part; drive
    matrices the same way, the Hubbard parameters as they are),
    σ₀ = S^-½ σ_o S^-½, ξ = S^½ ξ_o, Ixi = Ixi_o S^½, S⁻¹ = (S^-½)².  Per lead, the
    orbitals where every ξ and Ixi entry is at most ``zero_tol`` are dropped, as
    nozand's orbital index lists do (``dot_prod_zero_tol``, :242-245, 378-383, 917-928)."""
    Sm05 = np.asarray(Sm05, complex)
    S05 = np.linalg.inv(Sm05)
    H = S05 @ np.asarray(p.H0, complex) @ S05
    H[np.abs(H) < 1e-13] = 0.0
    H = 0.5 * (H + H.conj().transpose(0, 2, 1))
    sig = Sm05 @ np.asarray(p.sigma0, complex) @ Sm05
    NK, n, NL, NX, NOT = p.dims
    xi = (np.asarray(p.xi, complex).reshape(NK, -1, n) @ S05.transpose(0, 2, 1)).reshape(p.xi.shape)
    Ixi = (np.asarray(p.Ixi, complex).reshape(NK, -1, n) @ S05).reshape(p.Ixi.shape)
    for a in range(NL):
        live = (np.abs(xi[:, a]) > zero_tol).any(axis=(0, 1, 2)) | \
               (np.abs(Ixi[:, a]) > zero_tol).any(axis=(0, 1, 2))
        xi[:, a, ..., ~live] = 0.0
        Ixi[:, a, ..., ~live] = 0.0
    # drive matrices are Hamiltonian terms: carried like H (S^½ M S^½)
    drive = [Drive(d.kind, d.amp, d.par, S05 @ np.asarray(d.M, complex) @ S05) for d in p.drive]
    q = AMEProblem(H0=H, sigma0=sig, xi=xi, Ixi=Ixi, GG_P=p.GG_P, GL_P=p.GL_P, GG_M=p.GG_M,
                   GL_M=p.GL_M, Xpp=p.Xpp, Xpm=p.Xpm, pulse=p.pulse, dissipator=p.dissipator,
                   Sinv=Sm05 @ Sm05, hbar=p.hbar, name=p.name + "_nonortho", t_end=p.t_end,
                   eps=p.eps, h0=p.h0, hubbard_U=p.hubbard_U, hubbard_n0=p.hubbard_n0,
                   drive=drive)
    return q


def random_overlap(NK, n, strength=0.15, band=2, seed=0):
    """A banded Hermitian positive-definite overlap I + B (|B| small, |i-j| <= band)."""
    rng = np.random.default_rng(seed)
    S = np.zeros((NK, n, n), complex)
    for k in range(NK):
        B = np.zeros((n, n), complex)
        for d in range(1, band + 1):
            v = strength * (rng.normal(size=n - d) + 0.3j * rng.normal(size=n - d)) / d
            B += np.diag(v, d)
        S[k] = np.eye(n) + B + B.conj().T
    return S


def assemble(H0k, leads, N_F, NOT, mu, kT, sigma0=None, **kw):
    """Build an AMEProblem from per-k device Hamiltonians and lead fits.

    ``leads`` is a list (one per lead) of dicts with ``orbitals``, ``centres``,
    ``widths`` and ``G`` (list over k of (N_L, n_α, n_α) weights).
    """
    H0k = np.asarray(H0k, complex)
    NK, n, _ = H0k.shape
    NL = len(leads)
    N_L = len(leads[0]["centres"])
    NX = N_L + 2 * N_F
    shp = (NK, NL, NX, NOT)
    arr = {k: np.zeros(shp, complex) for k in ("GG_P", "GG_M", "GL_P", "GL_M", "Xpp", "Xpm")}
    xi = np.zeros(shp + (n,), complex)
    for k in range(NK):
        for a, L in enumerate(leads):
            o, v = lead_mode_constants(n, L["orbitals"], L["centres"], L["widths"], L["G"][k],
                                       N_F, NOT, mu[a], kT)
            for key in arr:
                arr[key][k, a] = o[key]
            xi[k, a] = v
    if sigma0 is None:
        sigma0 = np.array([fermi_dirac_dm(H0k[k], kT) for k in range(NK)])
    return AMEProblem(H0=H0k, sigma0=np.asarray(sigma0, complex), xi=xi, Ixi=xi.conj(),
                      **arr, **kw)


# ------------------------------------------------------------- synthetic configs
def _chain_gamma(E, hop):
    """Exact level width of a semi-infinite 1D chain coupled with ``hop``."""
    E = np.asarray(E, float)
    g = np.zeros_like(E)
    m = np.abs(E) < 2 * abs(hop)
    g[m] = np.sqrt(4 * hop ** 2 - E[m] ** 2)
    return g


def config_chain(n=20, N_L=8, N_F=20, kT=0.025, seed=0, hop=-1.0):
    """1D chain, two one-channel leads at sites 0 and n-1; Lorentzian weights by NNLS."""
    from scipy.optimize import nnls

    rng = np.random.Generator(np.random.PCG64(seed))
    H = chain_hamiltonian(n, hop)
    centres = np.sort(rng.uniform(-4.0, 4.0, N_L))
    widths = np.full(N_L, 0.5)
    E = np.linspace(-4.0, 4.0, 801)
    basis = lorentz(E[:, None], centres[None, :], widths[None, :]).real
    w, _ = nnls(basis, _chain_gamma(E, hop))
    leads = []
    for site in (0, n - 1):
        G = [np.array([[[wx]] for wx in w], complex)]
        leads.append(dict(orbitals=np.array([site]), centres=centres, widths=widths, G=G))
    p = assemble(H[None], leads, N_F, 1, mu=[0.0, 0.0], kT=kT,
                 pulse=Pulse("ramp", 0.1, 5.0, 0.0, (1.0, 0.0)),
                 name="chain20", t_end=50.0, eps=1e-8, h0=0.01)
    return p


def config_honeycomb(width, length, N_L, n_fermi_slots, NOT, NK=1, seed=0, kT=0.025,
                     hop=-2.7, disorder=0.1, pulse=None, hubbard_U=0.0, name="honeycomb",
                     t_end=10.0, eps=1e-6, h0=0.01, gamma_scale=1.0, k_shared=False):
    """Honeycomb patch with two leads on its zigzag edges, rank-NOT PSD Lorentzian weights.

    Each lead has one orthonormal channel basis V (n_α x NOT) and non-negative
    weights w_{x,c}: G_x = V diag(w_x) V†, so every Γ(z) has rank NOT and the NOT
    cut is exact.  ``n_fermi_slots`` = 2 N_F (B and C halves).  With ``k_shared`` every k
    block (e.g. the two spins of a Hubbard problem) has the same lead couplings.
    """
    assert n_fermi_slots % 2 == 0
    N_F = n_fermi_slots // 2
    rng = np.random.Generator(np.random.PCG64(seed))
    H, left, right = honeycomb_patch(width, length, hop, disorder, rng)
    H0k = np.array([H] * NK, complex)
    leads = []
    for orb in (left, right):
        na = len(orb)
        centres = np.sort(rng.uniform(-9.0, 9.0, N_L))
        widths = rng.uniform(0.6, 1.2, N_L)
        Gk = []
        for k in range(1 if k_shared else NK):
            Z = rng.normal(size=(na, NOT)) + 1j * rng.normal(size=(na, NOT))
            V, _ = np.linalg.qr(Z)
            w = gamma_scale * rng.uniform(0.05, 1.0, (N_L, NOT))
            Gk.append(np.array([(V * w[x]) @ V.conj().T for x in range(N_L)]))
        if k_shared:
            Gk = Gk * NK
        leads.append(dict(orbitals=orb, centres=centres, widths=widths, G=Gk))
    p = assemble(H0k, leads, N_F, NOT, mu=[0.0, 0.0], kT=kT,
                 pulse=pulse or Pulse("none"), hubbard_U=hubbard_U, name=name,
                 t_end=t_end, eps=eps, h0=h0)
    if hubbard_U:
        p.hubbard_n0 = np.real(np.einsum("kii->ki", p.sigma0))
    return p


def config_c2(seed=0):
    """n=400 honeycomb, 2 leads x rank 8, N_L=40 + 60 Fermi slots -> M=1600 per k."""
    return config_honeycomb(20, 20, 40, 60, 8, NK=1, seed=seed,
                            pulse=Pulse("gauss", 0.5, 20.0, 10.0, (0.5, -0.5)),
                            name="c2_n400_M1600", t_end=100.0, eps=1e-10, h0=0.005)


def config_c3(seed=0):
    """n=1000 honeycomb, NK=2 spin, Hubbard U=2, 2 leads x rank 20, N_L=50 + 50 slots -> M=4000;
    both spins see the same (spin-independent) lead couplings."""
    return config_honeycomb(40, 25, 50, 50, 20, NK=2, seed=seed,
                            pulse=Pulse("step", 0.3, 0.0, 0.0, (0.5, -0.5)),
                            hubbard_U=2.0, name="c3_n1000_M4000", t_end=10.0, eps=1e-8,
                            h0=0.005, k_shared=True)


def config_small(n_w=4, n_l=3, N_L=3, slots=4, NOT=2, NK=1, seed=0, hubbard_U=0.0,
                 pulse=None):
    """Tiny honeycomb case for parity tests (M = 2*(N_L+slots)*NOT)."""
    return config_honeycomb(n_w, n_l, N_L, slots, NOT, NK=NK, seed=seed,
                            pulse=pulse or Pulse("gauss", 0.3, 1.0, 0.8, (0.5, -0.5)),
                            hubbard_U=hubbard_U, name="small", t_end=1.0, eps=1e-7, h0=0.01)


def config_sized(n_target, M_target, seed=0, pulse=None, hubbard_U=0.0, NK=1, t_end=1.0,
                 eps=1e-6, h0=0.005, name=None):
    """Honeycomb device of about ``n_target`` orbitals with about ``M_target`` modes
    (two leads, edge-localised rank-NOT couplings, half Lorentzian / half Fermi
    slots), for the batched and serving workloads."""
    rng = np.random.Generator(np.random.PCG64(seed))
    width = int(np.clip(round(np.sqrt(n_target / 1.5)), 4, 40))
    length = max(2, int(round(n_target / width)))
    NOT = int(min(width, max(1, round(M_target / 100.0))))
    NX = max(4, int(round(M_target / (2.0 * NOT))))
    slots = 2 * max(1, (NX // 2) // 2)
    N_L = max(1, NX - slots)
    return config_honeycomb(width, length, N_L, slots, NOT, NK=NK, seed=int(rng.integers(1 << 30)),
                            pulse=pulse or Pulse("none"), hubbard_U=hubbard_U,
                            name=name or f"sized_n{width * length}", t_end=t_end, eps=eps, h0=h0)


def random_pulse(rng):
    """Step, Gaussian or sinusoidal bias at 0.1-1 eV (sine frequency 0.1-2 eV/ħ)."""
    kind = ["step", "gauss", "sin"][int(rng.integers(3))]
    amp = float(rng.uniform(0.1, 1.0))
    if kind == "step":
        return Pulse("step", amp, 0.0, 0.0, (0.5, -0.5))
    if kind == "gauss":
        return Pulse("gauss", amp, float(rng.uniform(0.0, 1.0)), float(rng.uniform(0.2, 2.0)),
                     (0.5, -0.5))
    return Pulse("sin", amp, float(rng.uniform(0.1, 2.0)), 0.0, (0.5, -0.5))


def config_c4_instance(n, seed=0, t_end=60.0):
    """One instance of the batched workload: honeycomb of n orbitals (edge 20),
    two leads with n/50 channels, N_L = 50 Lorentzians + 50 Fermi slots, so
    M = 4n; Gaussian pulse with amplitude U[0.05, 0.5] eV and width U[2, 20] fs."""
    rng = np.random.Generator(np.random.PCG64(seed))
    NOT = max(1, n // 50)
    amp = float(rng.uniform(0.05, 0.5))
    width = float(rng.uniform(2.0, 20.0))
    return config_honeycomb(20, max(2, n // 20), 50, 50, NOT, NK=1, seed=seed,
                            pulse=Pulse("gauss", amp, 2.0 * width, width, (0.5, -0.5)),
                            name=f"c4_n{n}", t_end=t_end, eps=1e-6, h0=0.005)
