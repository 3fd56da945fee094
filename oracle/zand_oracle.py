"""CPU FP64 oracle for the auxiliary-mode-expansion (AME) propagation hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``zandpack_amd`` imports this module; only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker / the timed CPU baseline.

It is a plain numpy (complex128) restatement of what the reference computes in
its orthogonal MPI propagator, written as one process:

* σ master path            -- Zandpack/mpi/zand:300-408 (``SIGMA``)
* Ψ/Ω worker path          -- Zandpack/mpi/zand:697-821 (``Worker``)
* Π/Q, currents            -- Zandpack/mpi/zand:91-110, 363-375
* RK helpers               -- Zandpack/mpi_funcs.py:23-25, 91-102, 247-259
* DOP54 tables, controller -- Zandpack/mpi_RK4pars_dev.py:118-149
* loop form of the RHS     -- Zandpack/TimedependentTransport.py:2798-2873
* Hu poles / Fermi expansion -- Zandpack/PadeDecomp.py:36-42, 84-152
* non-orthogonal form      -- Zandpack/mpi/nozand:456-497, 1003, 1075-1085, 1208-1210
* SCF contour and window   -- Zandpack/cmdtools/SCF:209-215, 236-253, 273-306

Array layout is the reference's (per k):
``sigma (NK,n,n)``, ``psi (NK,NL,NX,NOT,n)``, ``omega (NK,NL,NX,NOT,NL,NX,NOT)``,
mode constants ``(NK,NL,NX,NOT)``.

Pinning (what makes this oracle trustworthy; see tests/test_oracle_*.py):

* ``dop54_tableau``   -- exact Dormand-Prince rationals, order conditions, row sums.
* ``hu_poles``        -- reproduces the reference's golden accuracy table
                         ``Pade_data/JieHu2011_accuracy_table.npz``.
* ``rhs``             -- agrees with the independent loop form ``rhs_loop``;
                         invariants: dσ Hermitian, Tr dσ = Σ_α J_α, the Fermi x Fermi
                         block of dΩ vanishes, von-Neumann closed form when ξ=0,
                         closed-form stationary Ω makes dΩ=0.
* ``propagate``       -- von-Neumann closed form σ(t)=U σ0 U† to the DOP54 error level.
* ``rkf45_propagate`` -- Fehlberg rationals (order conditions, row sums) and the same
                         von-Neumann closed form.
* ``window_density``  -- a single level between wide-band leads against the scalar
                         Lorentzian integral; the fixed Gauss-Legendre panels against
                         scipy's adaptive quad_vec (the reference's integrator) on a
                         random device; zero for equal chemical potentials; with one
                         lead decoupled, contour(mu_min) + window = contour(mu_max).
* ``equilibrium_density`` -- with psinought's Ψ₀, Ω₀ a fixed point of ``rhs`` to 1e-14
                         (multi-channel widths, Fermi B/C slots; planted slot mistakes
                         break it); Hermitian with occupations in [0, 1].
* ``hubbard_dH``      -- conserves the mean-field energy ``hubbard_energy`` on an isolated
                         device; spin-swapped, half-U, sign or n0 mistakes do not.
* ``waveform`` / ``model_dH`` -- commuting drive: exp(-i(H0 t + ∫f M)/ħ) closed form.
* ``rhs(Sinv=...)``   -- the nozand form equals the orthogonal form carried to the
                         non-orthogonal basis (dσ_n = S^-½ dσ_o S^-½, dΨ_n = S^-½ dΨ_o,
                         dΩ and J unchanged); S = I reduces it to zand.
"""
from __future__ import annotations

from fractions import Fraction as Fr

import numpy as np

# Zandpack/td_constants.py:3
HBAR = 6.582119569e-1  # eV * fs


# ---------------------------------------------------------------------------
# DOP54 tableau and controller -- Zandpack/mpi_RK4pars_dev.py:118-149
# ---------------------------------------------------------------------------
def dop54_rationals():
    """Exact Dormand-Prince 5(4) coefficients (mpi_RK4pars_dev.py:123-132)."""
    c = [Fr(0), Fr(1, 5), Fr(3, 10), Fr(4, 5), Fr(8, 9), Fr(1), Fr(1)]
    a = [
        [],
        [Fr(1, 5)],
        [Fr(3, 40), Fr(9, 40)],
        [Fr(44, 45), Fr(-56, 15), Fr(32, 9)],
        [Fr(19372, 6561), Fr(-25360, 2187), Fr(64448, 6561), Fr(-212, 729)],
        [Fr(9017, 3168), Fr(-355, 33), Fr(46732, 5247), Fr(49, 176), Fr(-5103, 18656)],
        [Fr(35, 384), Fr(0), Fr(500, 1113), Fr(125, 192), Fr(-2187, 6784), Fr(11, 84)],
    ]
    b5 = [Fr(35, 384), Fr(0), Fr(500, 1113), Fr(125, 192), Fr(-2187, 6784), Fr(11, 84), Fr(0)]
    b4 = [Fr(5179, 57600), Fr(0), Fr(7571, 16695), Fr(393, 640), Fr(-92097, 339200),
          Fr(187, 2100), Fr(1, 40)]
    return c, a, b5, b4


def rkf45_rationals():
    """Exact Runge-Kutta-Fehlberg 4(5) coefficients of the in-process driver
    AdaptiveRK4 (TimedependentTransport.py:3191-3203; Optimized_RK45.py:34-46):
    nodes c, stage matrix a, the fifth-order weights CH the step is taken with and
    the fourth-order weights C (error weights CT = CH - C)."""
    c = [Fr(0), Fr(1, 4), Fr(3, 8), Fr(12, 13), Fr(1), Fr(1, 2)]
    a = [
        [],
        [Fr(1, 4)],
        [Fr(3, 32), Fr(9, 32)],
        [Fr(1932, 2197), Fr(-7200, 2197), Fr(7296, 2197)],
        [Fr(439, 216), Fr(-8), Fr(3680, 513), Fr(-845, 4104)],
        [Fr(-8, 27), Fr(2), Fr(-3544, 2565), Fr(1859, 4104), Fr(-11, 40)],
    ]
    ch = [Fr(16, 135), Fr(0), Fr(6656, 12825), Fr(28561, 56430), Fr(-9, 50), Fr(2, 55)]
    c4 = [Fr(25, 216), Fr(0), Fr(1408, 2565), Fr(2197, 4104), Fr(-1, 5), Fr(0)]
    return c, a, ch, c4


def rkf45_propagate(consts, y0, t0, t1, h, eps, H0, dH, bias, NL, hbar=HBAR,
                    fixed_h=False, max_attempts=100000):
    """The in-process adaptive RKF45 loop (AdaptiveRK4, TimedependentTransport.py:
    3179-3390; AdaptiveRK4_Opti, Optimized_RK45.py:19-228): six stages per attempt,
    TE = sqrt(Σ |Σ_j CT_j k_j|²) over σ, Ψ, Ω (:3215-3221), h <- 0.9 h (eps/TE)^(1/5)
    after every attempt, accept iff TE <= eps (the ``while TE > eps`` loop), step
    y += Σ CH_j k_j with the k of the accepted attempt, currents and σ recorded from
    the state before the step, ``while t0 <= t1`` and t += h with the new h.
    Returns dict(times (start time of every step), J, sigma, y, h_accepted, n_reject)."""
    c, a, ch, c4 = rkf45_rationals()
    c = [float(x) for x in c]
    a = [[float(x) for x in row] for row in a]
    ch = np.array([float(x) for x in ch])
    ct = ch - np.array([float(x) for x in c4])
    y = tuple(np.array(v, complex) for v in y0)
    times, Js, sigmas, hs = [], [], [], []
    n_rej, attempts, t = 0, 0, t0
    while t <= t1:
        while True:
            attempts += 1
            if attempts > max_attempts:
                raise RuntimeError("attempt limit reached")
            ks = []
            for i in range(6):
                ts = t + c[i] * h
                yi = [y[p] + sum(a[i][j] * ks[j][p] for j in range(i)) for p in range(3)]
                Ht = H0 + dH(ts, yi[0])
                delta = np.array([bias(ts, al) for al in range(NL)], float)
                ks.append(rhs(consts, yi[0], yi[1], yi[2], Ht, delta, h=h, hbar=hbar))
            h_used = h
            if fixed_h:
                break
            te = 0.0
            for p in range(3):
                e = sum(ct[j] * ks[j][p] for j in range(6))
                te += float(np.sum(e.real ** 2 + e.imag ** 2))
            TE = te ** 0.5
            if not np.isfinite(TE):
                raise FloatingPointError("non-finite error estimate")
            h = 0.9 * h * (eps / TE) ** 0.2
            if TE <= eps:
                break
            n_rej += 1
        sigmas.append(y[0].copy())
        Js.append(currents(consts, y[1], hbar=hbar))
        y = tuple(y[p] + sum(ch[j] * ks[j][p] for j in range(6)) for p in range(3))
        times.append(t)
        hs.append(h_used)
        t = t + h
    return dict(times=np.array(times), J=np.array(Js), sigma=np.array(sigmas), y=y,
                h_accepted=np.array(hs), n_reject=n_rej)


def dop54_tableau():
    """Float tableau in the reference's slot convention.

    Slots are ``[k1..k7, y0]``; ``B`` is 7x8 with ``B[i,7] = 1`` for i>=1
    (mpi_RK4pars_dev.py:134-137), ``CT = b5-b4`` (:138), ``CH = [b5, 1]`` (:139-140).
    """
    c, a, b5, b4 = dop54_rationals()
    A = np.array([float(x) for x in c])
    B = np.zeros((7, 8))
    for i in range(1, 7):
        for j, v in enumerate(a[i]):
            B[i, j] = float(v)
        B[i, 7] = 1.0
    b5f = np.array([float(x) for x in b5])
    b4f = np.array([float(x) for x in b4])
    CT = b5f - b4f
    CH = np.hstack((b5f, [1.0]))
    return dict(A=A, B=B, CT=CT, CH=CH, b5=b5f, b4=b4f, N=7)


def hnew(h, te, eps):
    """Step update, mpi_RK4pars_dev.py:148-149 (no clamp)."""
    return 0.9 * h * ((eps / te) ** 0.2)


# ---------------------------------------------------------------------------
# Fermi-function pole expansion -- Zandpack/PadeDecomp.py:36-42, 84-152
# ---------------------------------------------------------------------------
def _hu_tridiag(m):
    """Hu et al. 2011 matrix: off-diagonal 1/sqrt(b_i b_j), b_i = 2i-1 (PadeDecomp.py:90-99)."""
    b = 2.0 * np.arange(1, m + 1) - 1.0
    off = 1.0 / np.sqrt(b[:-1] * b[1:])
    return np.diag(off, 1) + np.diag(off, -1)


def hu_poles(N):
    """Poles z_p (i * positive real) and residues R_p of the N-pole Hu expansion.

    Restates PadeDecomp.py:101-127,151-152: roots q = 2/λ from the positive
    eigenvalues λ of the 2N tridiagonal matrix (complex Hermitian eigh, as the
    reference), the auxiliary roots from its trailing (2N-1) block, and the
    residues from the product formula.  For N <= 40 the products are float64
    (PadeDecomp.py:120-126); for N > 40 the reference switches to sympy
    (:115-116, 131-147), restated here as exact rational products of the same
    float64 roots.
    """
    if N == 0:
        return np.zeros(0, complex), np.zeros(0)
    m = 2 * N
    T = _hu_tridiag(m).astype(complex)
    eq = np.linalg.eigh(T)[0]
    Q = 2.0 / np.sort(eq[eq > 1e-15])[::-1]
    ep = np.linalg.eigh(T[1:, 1:])[0]
    P = 2.0 / ep[ep > 1e-15]
    const = N * (2 * (N + 1) - 1) / 2.0
    R = np.empty(N)
    if N <= 40:
        for i in range(N):
            p1 = Q ** 2 - Q[i] ** 2
            p1[np.abs(p1) < 1e-15] = 1.0
            R[i] = const * np.prod(P ** 2 - Q[i] ** 2) / np.prod(p1)
    else:
        Qf = [Fr(float(q)) for q in Q]
        Pf = [Fr(float(p)) for p in P]
        cf = Fr(N * (2 * (N + 1) - 1), 2)
        for i in range(N):
            p1 = Fr(1)
            for j in range(N):
                if j != i:
                    p1 *= Qf[j] ** 2 - Qf[i] ** 2
            p2 = Fr(1)
            for p in Pf:
                p2 *= p ** 2 - Qf[i] ** 2
            R[i] = float(cf * p2 / p1)
    return 1j * Q, R


def fd_expanded(E, zi, R, beta, mu=0.0):
    """f(E) ~ 1/2 - (1/β) Σ_p R_p [1/(E-μ-z_p/β) + 1/(E-μ+z_p/β)]  (PadeDecomp.py:36-42)."""
    E = np.asarray(E, dtype=complex)
    xpp = mu + zi / beta
    xpm = mu - zi / beta
    d = (1.0 / beta) * (1.0 / np.subtract.outer(E, xpp) + 1.0 / np.subtract.outer(E, xpm)) * R
    return 0.5 - d.sum(axis=-1)


# ---------------------------------------------------------------------------
# Right-hand side (zand master + worker) -- one RK stage, k = h * f
# ---------------------------------------------------------------------------
def _flat(consts):
    xi = consts["xi"]
    NK, NL, NX, NOT, n = xi.shape
    M = NL * NX * NOT
    return NK, NL, NX, NOT, n, M


def lead_of_mode(NL, NX, NOT):
    return np.repeat(np.arange(NL), NX * NOT)


def coupled_orbitals(consts):
    """Orbitals where any ξ or Ixi entry is non-zero (mpi/zand:274-280, 589-602)."""
    NK, NL, NX, NOT, n, M = _flat(consts)
    nz = (consts["xi"] != 0) | (consts["Ixi"] != 0)
    return np.where(nz.reshape(-1, n).any(axis=0))[0]


def rhs(consts, sig, psi, omg, Ht, delta, h=1.0, hbar=HBAR, return_pi=False, prune=False,
        diss=None, Sinv=None):
    """One stage of zand: returns (h*dσ/dt, h*dΨ/dt, h*dΩ/dt).

    σ: mpi/zand:340-387; Ψ: mpi/zand:746-778; Ω: mpi/zand:784-809.
    ``diss`` is the value of Bias.dissipator(ts, σ) at the stage, subtracted as
    ``D -= ħ·dissipator`` (mpi/zand:343); None means 0.
    ``Ht`` is the full Hamiltonian at the stage (H0 + dH(ts, σ)); ``delta`` is
    Δ_α(ts) per lead.  ``prune=False`` is dense in the orbital index;
    ``prune=True`` restricts every ξ/Ixi contraction to the coupled orbitals as
    zand does (mpi/zand:589-602, 641-658) -- the same sums, since ξ vanishes
    elsewhere, at the reference's cost.

    ``Sinv`` (NK, n, n) selects the non-orthogonal form of ``mpi/nozand``, where
    ξ, Ixi, H, σ are in the non-orthogonal basis and S⁻¹ enters four terms:
    S⁻¹Hσ in the commutator (nozand:456-461), Π_α S⁻¹ in Q (:489-497), S⁻¹H in
    the Ψ propagator term (:1075-1085), and S⁻¹ on the constant and Ωξ terms of
    dΨ (:1003, 1208-1210).  The ΔP σξ term, Ω and the currents carry no S⁻¹.
    """
    NK, NL, NX, NOT, n, M = _flat(consts)
    xi = consts["xi"].reshape(NK, M, n)
    Ixi = consts["Ixi"].reshape(NK, M, n)
    glp = consts["GL_P"].reshape(NK, M)
    dgp = consts["dGP"].reshape(NK, M)
    dgm = consts["dGM"].reshape(NK, M)
    lead = lead_of_mode(NL, NX, NOT)
    d_mode = np.asarray(delta, float)[lead]
    xpp = consts["Xpp"].reshape(NK, M) + d_mode[None, :]
    xpm = consts["Xpm"].reshape(NK, M) + d_mode[None, :]
    p = psi.reshape(NK, M, n)
    W = omg.reshape(NK, M, M)
    cidx = coupled_orbitals(consts) if prune else slice(None)
    xc, ixc, pc = xi[:, :, cidx], Ixi[:, :, cidx], p[:, :, cidx]

    # σ: D = Hσ; D -= D†; D *= -i  (mpi/zand:340-342); nozand: D = S⁻¹(Hσ)
    D = Ht @ sig
    if Sinv is not None:
        D = Sinv @ D
    D = -1j * (D - D.conj().transpose(0, 2, 1))
    if diss is not None:
        D = D - hbar * diss
    # Q_α = (Ψ_αᵀ Ixi_α + h.c.)/ħ  (mpi/zand:91-110)
    Q = np.zeros((NK, NL, n, n), complex)
    for a in range(NL):
        sl = slice(a * NX * NOT, (a + 1) * NX * NOT)
        r = np.zeros((NK, n, n), complex)
        r[:, :, cidx] = p[:, sl, :].transpose(0, 2, 1) @ ixc[:, sl, :]
        if Sinv is not None:
            r = r @ Sinv
        Q[:, a] = (r + r.conj().transpose(0, 2, 1)) / hbar
    dsig = (h / hbar) * (D + Q.sum(axis=1))

    # Ψ (mpi/zand:746-778)
    HS = Ht if Sinv is None else Sinv @ Ht
    dpsi = p @ (HS / hbar).transpose(0, 2, 1)
    dpsi -= (xpp / hbar)[..., None] * p
    if Sinv is None:
        dpsi += glp[..., None] * xi
        dpsi += dgp[..., None] * (xc @ sig[:, :, cidx].transpose(0, 2, 1))
        dpsi[:, :, cidx] += W @ (xc / hbar ** 2)
    else:
        dpsi += dgp[..., None] * (xc @ sig[:, :, cidx].transpose(0, 2, 1))
        v = glp[..., None] * xi
        v[:, :, cidx] += W @ (xc / hbar ** 2)
        dpsi += v @ Sinv.transpose(0, 2, 1)
    dpsi *= -1j * h

    # Ω (mpi/zand:784-809)
    domg = h * (pc @ (dgm[..., None] * ixc).transpose(0, 2, 1))            # ΔM_j (ψ_i·Ixi_j)
    domg += h * ((dgp[..., None] * xc) @ pc.conj().transpose(0, 2, 1))     # ΔP_i (ξ_i·ψ_j*)
    domg += (1j * h / hbar) * (xpp[:, :, None] - xpm[:, None, :]) * W     # outer subtraction
    out = (dsig, dpsi.reshape(psi.shape), domg.reshape(omg.shape))
    if return_pi:
        return out, Q
    return out


def rhs_loop(consts, sig, psi, omg, Ht, delta, hbar=HBAR):
    """Brute-force loop form (TimedependentTransport.py:2816-2873, with the Ixi
    convention of zand).  Returns the derivatives (h = 1).  Tiny shapes only."""
    NK, NL, NX, NOT, n, M = _flat(consts)
    xi = consts["xi"]
    Ixi = consts["Ixi"]
    dsig = np.zeros_like(sig)
    dpsi = np.zeros_like(psi)
    domg = np.zeros_like(omg)
    for k in range(NK):
        H = Ht[k]
        s = sig[k]
        for i in range(n):
            for j in range(n):
                acc = 0j
                for o in range(n):
                    acc += H[i, o] * s[o, j] - s[i, o] * H[o, j]
                dsig[k, i, j] = -1j * acc
        for a in range(NL):
            pi = np.zeros((n, n), complex)
            for x in range(NX):
                for c in range(NOT):
                    for i in range(n):
                        for j in range(n):
                            pi[i, j] += psi[k, a, x, c, i] * Ixi[k, a, x, c, j]
            pi /= hbar
            dsig[k] += pi + pi.conj().T
        dsig[k] /= hbar
        for a in range(NL):
            for x in range(NX):
                for c in range(NOT):
                    xp = consts["Xpp"][k, a, x, c] + delta[a]
                    v = np.zeros(n, complex)
                    for i in range(n):
                        acc = 0j
                        for o in range(n):
                            acc += H[i, o] * psi[k, a, x, c, o]
                        acc -= xp * psi[k, a, x, c, i]
                        v[i] = acc / hbar
                        v[i] += consts["GL_P"][k, a, x, c] * xi[k, a, x, c, i]
                        ss = 0j
                        for o in range(n):
                            ss += s[i, o] * xi[k, a, x, c, o]
                        v[i] += consts["dGP"][k, a, x, c] * ss
                        ow = 0j
                        for a2 in range(NL):
                            for x2 in range(NX):
                                for c2 in range(NOT):
                                    ow += omg[k, a, x, c, a2, x2, c2] * xi[k, a2, x2, c2, i]
                        v[i] += ow / hbar ** 2
                    dpsi[k, a, x, c] = -1j * v
                    for a2 in range(NL):
                        for x2 in range(NX):
                            for c2 in range(NOT):
                                xm = consts["Xpm"][k, a2, x2, c2] + delta[a2]
                                t1 = 0j
                                t2 = 0j
                                for o in range(n):
                                    t1 += Ixi[k, a2, x2, c2, o] * psi[k, a, x, c, o]
                                    t2 += np.conj(psi[k, a2, x2, c2, o]) * xi[k, a, x, c, o]
                                domg[k, a, x, c, a2, x2, c2] = (
                                    -1j * (xm - xp) * omg[k, a, x, c, a2, x2, c2] / hbar
                                    + consts["dGM"][k, a2, x2, c2] * t1
                                    + consts["dGP"][k, a, x, c] * t2)
    return dsig, dpsi, domg


def currents(consts, psi, hbar=HBAR):
    """J_α[k] = Tr(Q_α[k]).real/ħ = 2 Re Tr Π_α / ħ²  (mpi/zand:368-375)."""
    NK, NL, NX, NOT, n, M = _flat(consts)
    J = np.zeros((NL, NK))
    for a in range(NL):
        tr = np.einsum("kxco,kxco->k", psi[:, a], consts["Ixi"][:, a])
        J[a] = 2.0 * tr.real / hbar ** 2
    return J


# ---------------------------------------------------------------------------
# Hamiltonian / bias models evaluated per stage
# ---------------------------------------------------------------------------
def hubbard_dH(U, n0):
    """Mean-field Hubbard dH (examples/Hubbard/Calculation/Bias.py:13-20):
    dH[0,ii] = U(σ[1,ii]-n0[1,i]); dH[1,ii] = U(σ[0,ii]-n0[0,i])."""
    n0 = np.asarray(n0)

    def dH(t, sig):
        no = sig.shape[-1]
        out = np.zeros(sig.shape, complex)
        idx = np.arange(no)
        out[0, idx, idx] = U * (sig[1, idx, idx] - n0[1])
        out[1, idx, idx] = U * (sig[0, idx, idx] - n0[0])
        return out

    return dH


def zero_dH(t, sig):
    return np.zeros(sig.shape, complex)


def hubbard_energy(H0, sig, U, n0):
    """Mean-field Hubbard energy E[σ] = Σ_k Tr(H0_k σ_k) + U Σ_i (σ↑_ii - n0↑_i)(σ↓_ii - n0↓_i),
    whose derivative ∂E/∂σ_k is H0_k + dH_k with dH the cross-spin diagonal of
    ``hubbard_dH``; the isolated-device evolution dσ_k/dt = -(i/ħ)[H_k(σ), σ_k]
    conserves it."""
    n0 = np.asarray(n0)
    E = np.einsum("kij,kji->", H0, sig).real
    d0 = np.einsum("ii->i", sig[0]).real - n0[0]
    d1 = np.einsum("ii->i", sig[1]).real - n0[1]
    return float(E + U * np.dot(d0, d1))


def waveform(kind, amp, par):
    """Time factor f(t) of a drive term H(t) = H0 + Σ_i f_i(t) M_i, restating the
    forms the reference's examples write into dH(t, σ):
      'cos_window'  amp cos(ω t + φ) for t_on < t <= t_off, else 0; par = (ω, φ, t_on, t_off)
                    (examples/nb_tutorial_updated/Tight-binding/T1_Rabi/Bias.py:16-24)
      'sin_fermi'   amp sin(ω t + φ) / (1 + exp((t - t_off)/w)); par = (ω, φ, t_off, w)
                    (examples/Reproduce_and_checks/scipy_vs_ARK4_Zand/shared.py:4-19)
      'gauss_sin'   amp exp(-((t - d)/s)^2) sin(2π f (t - d) - φ); par = (s, f, φ, d)
                    (examples/nb_tutorial_updated/Tight-binding/T2_TB_Chain/Bias.py:9-10, 21-26;
                    Zandpack/Pulses.py:52-66)
      'const'       amp."""
    par = tuple(float(x) for x in par) + (0.0,) * (4 - len(par))
    a, b, c, d = par[:4]
    if kind == "cos_window":
        return lambda t: amp * np.cos(a * t + b) if c < t <= d else 0.0
    if kind == "sin_fermi":
        return lambda t: amp * np.sin(a * t + b) / (1.0 + np.exp((t - c) / d))
    if kind == "gauss_sin":
        return lambda t: amp * np.exp(-((t - d) / a) ** 2) * np.sin(2 * np.pi * b * (t - d) - c)
    if kind == "const":
        return lambda t: amp
    raise ValueError(kind)


def model_dH(terms=(), U=0.0, n0=None):
    """dH(t, σ) = Σ_i f_i(t) M_i (+ the Hubbard term when U != 0): the time-dependent
    Hamiltonian increment a Bias.py hands the propagator per stage (mpi/zand:319;
    mpi/nozand:427).  ``terms`` is a sequence of (f, M) with f a callable of t and M
    a (NK, n, n) matrix."""
    hub = hubbard_dH(U, n0) if U else None

    def dH(t, sig):
        out = np.zeros(sig.shape, complex) if hub is None else hub(t, sig)
        for f, M in terms:
            out = out + f(t) * np.asarray(M, complex)
        return out

    return dH


# ---------------------------------------------------------------------------
# DOP54 attempt / propagation with zand's control semantics
# ---------------------------------------------------------------------------
def relaxation_dissipator(eta, sig_eq):
    """dissipator(t, σ) = η(t)(σ - σ_eq), the form of
    examples/Hubbard/Calculation/Bias.py:29-39 (``eta`` a callable of t)."""
    def diss(t, sig):
        return eta(t) * (sig - sig_eq)
    return diss


def dissipator_rate(kind, eta0, t0=0.0, width=1.0):
    """η(t) of the relaxation dissipator: 'const' η0; 'step' η0 / (1 + exp((t - t0)/w)),
    the smooth switch-off of examples/Hubbard/Calculation/Bias.py:29-39."""
    if kind == "const":
        return lambda t: eta0
    if kind == "step":
        return lambda t: eta0 / (1.0 + np.exp((t - t0) / width))
    raise ValueError(kind)


def dop54_attempt(consts, y0, t0, h, H0, dH, bias, NL, hbar=HBAR, tab=None, dissipator=None,
                  Sinv=None, rtol=0.0, atol=0.0, prune=False):
    """One 7-stage attempt (mpi/zand:301-396).  Returns (ks, TE, J_stage0, y_new).

    ``y0`` = (σ, Ψ, Ω); ``ks[i]`` are the h-scaled stage derivatives.  J is taken
    from the stage-0 state (mpi/zand:368-375).  ``y_new`` = y0 + Σ b5_j k_j
    (step_fourth_*, mpi_funcs.py:254-259).  TE is zand's absolute norm
    (mpi_funcs.py:247-252), or with atol > 0 the scaled RMS norm
    sqrt(mean |err/(atol + rtol max(|y0|, |y_new|))|²) over all entries.  ``prune``
    restricts the ξ contractions to the coupled orbitals as zand does (same sums).
    """
    tab = tab or dop54_tableau()
    A, B, CT = tab["A"], tab["B"], tab["CT"]
    ks = []
    J0 = None
    for i in range(7):
        ts = t0 + A[i] * h
        yi = []
        for part in range(3):
            v = y0[part].copy()
            for j in range(i):
                v = v + B[i, j] * ks[j][part]
            yi.append(v)
        Ht = H0 + dH(ts, yi[0])
        delta = np.array([bias(ts, a) for a in range(NL)], float)
        dv = None if dissipator is None else dissipator(ts, yi[0])
        k = rhs(consts, yi[0], yi[1], yi[2], Ht, delta, h=h, hbar=hbar, diss=dv, Sinv=Sinv,
                prune=prune)
        if i == 0:
            J0 = currents(consts, yi[1], hbar=hbar)
        ks.append(k)
    b5 = tab["b5"]
    y_new = tuple(y0[p] + sum(b5[j] * ks[j][p] for j in range(7)) for p in range(3))
    te2, count = 0.0, 0
    for part in range(3):
        e = sum(CT[j] * ks[j][part] for j in range(7))
        w = 1.0
        if atol > 0:
            sc = atol + rtol * np.maximum(np.abs(y0[part]), np.abs(y_new[part]))
            w = 1.0 / sc ** 2
        te2 += float(np.sum((e.real ** 2 + e.imag ** 2) * w))
        count += e.size
    TE = (te2 / count) ** 0.5 if atol > 0 else te2 ** 0.5
    return ks, TE, J0, y_new


def propagate(consts, y0, t0, t1, h, eps, H0, dH, bias, NL, hbar=HBAR,
              time_update="reference", max_attempts=100000, fixed_h=False, dissipator=None,
              Sinv=None, rtol=0.0, atol=0.0):
    """zand's outer loop (mpi/zand:297-431).

    Accept iff TE < eps; h <- 0.9 h (eps/TE)^0.2 after every attempt; on accept
    t advances by the *new* h when ``time_update='reference'`` (mpi/zand:406-407,427),
    by the step just taken when ``'correct'``.  ``fixed_h`` accepts every attempt
    and keeps h (the reference's ``fixed_mode``).
    Returns dict(times, sigma, J (list of (NL,NK)), h_accepted, TE, n_reject).
    """
    tab = dop54_tableau()
    y = tuple(np.array(v, complex) for v in y0)
    times = [t0]
    sigmas = [y[0].copy()]
    Js, hs, tes = [], [], []
    n_rej = 0
    attempts = 0
    t = t0
    while t < t1:
        while True:
            attempts += 1
            if attempts > max_attempts:
                raise RuntimeError("attempt limit reached")
            ks, TE, J0, y_new = dop54_attempt(consts, y, t, h, H0, dH, bias, NL, hbar, tab,
                                              dissipator=dissipator, Sinv=Sinv, rtol=rtol,
                                              atol=atol)
            if not np.isfinite(TE):
                raise FloatingPointError("non-finite error estimate")
            accept = fixed_h or (TE < eps)
            h_old = h
            if not fixed_h:
                h = hnew(h, TE, eps)
            if accept:
                break
            n_rej += 1
        y = y_new
        Js.append(J0)
        hs.append(h_old)
        tes.append(TE)
        t = t + (h if time_update == "reference" else h_old)
        times.append(t)
        sigmas.append(y[0].copy())
    return dict(times=np.array(times), sigma=np.array(sigmas), J=np.array(Js),
                h_accepted=np.array(hs), TE=np.array(tes), n_reject=n_rej, y=y)


# ---------------------------------------------------------------------------
# Closed forms used to pin the RHS
# ---------------------------------------------------------------------------
def stationary_omega(consts, psi, delta, hbar=HBAR):
    """Ω that makes dΩ = 0 for a fixed Ψ (cmdtools/psinought:447-449 form):
    Ω_ij = -(ΔM_j(ψ_i·Ixi_j) + ΔP_i(ξ_i·ψ_j*)) * iħ/(χ⁻_j - χ⁺_i)  (with Δ shifts)."""
    NK, NL, NX, NOT, n, M = _flat(consts)
    lead = lead_of_mode(NL, NX, NOT)
    d_mode = np.asarray(delta, float)[lead]
    xpp = consts["Xpp"].reshape(NK, M) + d_mode
    xpm = consts["Xpm"].reshape(NK, M) + d_mode
    p = psi.reshape(NK, M, n)
    xi = consts["xi"].reshape(NK, M, n)
    Ixi = consts["Ixi"].reshape(NK, M, n)
    src = (p @ (consts["dGM"].reshape(NK, M)[..., None] * Ixi).transpose(0, 2, 1)
           + (consts["dGP"].reshape(NK, M)[..., None] * xi) @ p.conj().transpose(0, 2, 1))
    om = -src * 1j * hbar / (xpm[:, None, :] - xpp[:, :, None])
    return om.reshape(NK, NL, NX, NOT, NL, NX, NOT)


# ---------------------------------------------------------------------------
# Steady state (cmdtools/psinought)
# ---------------------------------------------------------------------------
def steady_state_operator(consts, psi, H, delta, hbar=HBAR):
    """A(ψ) of cmdtools/psinought:199-226 (``Av3``): with Ω₀(ψ) the closed form of
    ``stationary_omega`` (psinought:447-449),
        A(ψ)_i = (χ⁺_i + Δ_α(i) - H) ψ_i - (Ω₀(ψ) ξ)_i / ħ
    (psinought writes the Ω₀ term as Σ_j ξ_j i/(χ⁻_j-χ⁺_i) (ΔM_j Ixi_j·ψ_i + ΔP_i ψ_j*·ξ_i)).
    Real-linear in ψ (ψ* enters)."""
    NK, NL, NX, NOT, n, M = _flat(consts)
    lead = lead_of_mode(NL, NX, NOT)
    xpp = consts["Xpp"].reshape(NK, M) + np.asarray(delta, float)[lead]
    p = psi.reshape(NK, M, n)
    om = stationary_omega(consts, psi, delta, hbar).reshape(NK, M, M)
    W = om @ consts["xi"].reshape(NK, M, n)
    out = xpp[..., None] * p - p @ H.transpose(0, 2, 1) - W / hbar
    return out.reshape(psi.shape)


def steady_state_rhs(consts, sig0, hbar=HBAR):
    """b of cmdtools/psinought:228-239 (``B3``): ħ Λ^{<+}_i ξ_i + ħ ΔP_i σ₀ ξ_i."""
    NK, NL, NX, NOT, n, M = _flat(consts)
    xi = consts["xi"].reshape(NK, M, n)
    b = hbar * consts["GL_P"].reshape(NK, M)[..., None] * xi
    b = b + hbar * consts["dGP"].reshape(NK, M)[..., None] * (xi @ sig0.transpose(0, 2, 1))
    return b.reshape(consts["xi"].shape)


def steady_state_dense(consts, H, sig0, delta, hbar=HBAR):
    """Ψ₀ solving A(ψ) = b as the real 2N x 2N system psinought hands to its Krylov
    solver (psinought:203-209, 418-419), here by a dense solve (tiny systems only).
    Returns (Ψ₀, Ω₀)."""
    shape = consts["xi"].shape
    N = int(np.prod(shape))
    b = steady_state_rhs(consts, sig0, hbar).reshape(N)
    A = np.zeros((2 * N, 2 * N))
    e = np.zeros(N, complex)
    for k in range(2 * N):
        e[:] = 0
        e[k % N] = 1.0 if k < N else 1j
        col = steady_state_operator(consts, e.reshape(shape), H, delta, hbar).reshape(N)
        A[:N, k] = col.real
        A[N:, k] = col.imag
    x = np.linalg.solve(A, np.concatenate([b.real, b.imag]))
    psi0 = (x[:N] + 1j * x[N:]).reshape(shape)
    return psi0, stationary_omega(consts, psi0, delta, hbar)


# ---------------------------------------------------------------------------
# Equilibrium density (cmdtools/SCF)
# ---------------------------------------------------------------------------
def lorentzian_basis(E, eps, gam):
    """Γ-basis L(E) = γ²/((E-ε)²+γ²) and its Kramers-Kronig partner
    KK(E) = γ(E-ε)/((E-ε)²+γ²), both continued to complex E (the Lorentzian basis of
    the level-width fit that FittedSelfEnergy.py:35-59 evaluates through the
    external Block_matrices.Croy routines; restated from the Lorentzian form)."""
    d = E - eps
    den = d * d + gam * gam
    return gam * gam / den, gam * d / den


def lead_self_energy(consts, E, bias=0.0):
    """Σ_α(E) per lead, (NK, NL, n, n) at one complex E, as FittedSelfEnergy.evaluate
    builds it (:35-53): Σ = 0.5·KK - 0.5i·Γ with Γ(E) = Σ_x C_x L_x(E - bias) and
    C_x = Σ_c λ_xc ξ_xc ⊗ Ixi_xc (from_eigendecomp, :72-76).  The Lorentzian
    modes' eigenvalues come from the propagation constants, λ = 2i(GG_P - GL_P)/γ
    (TimedependentTransport.py:964-1000: GG_P - GL_P = -½iλγ); the Fermi-pole modes
    have GG_P = GL_P and drop out."""
    NK, NL, NX, NOT, n, M = _flat(consts)
    xi = consts["xi"]
    Ixi = consts["Ixi"]
    dgp = consts["dGP"]
    eps = consts["Xpp"].real
    gam = consts["Xpp"].imag
    b = np.broadcast_to(np.asarray(bias, float), (NL,))
    out = np.zeros((NK, NL, n, n), complex)
    for k in range(NK):
        for a in range(NL):
            for x in range(NX):
                for c in range(NOT):
                    if dgp[k, a, x, c] == 0:
                        continue
                    lam = 2j * dgp[k, a, x, c] / gam[k, a, x, c]
                    L, KK = lorentzian_basis(E - b[a], eps[k, a, x, c], gam[k, a, x, c])
                    out[k, a] += (0.5 * KK - 0.5j * L) * lam * np.outer(xi[k, a, x, c], Ixi[k, a, x, c])
    return out


def contour_density(H, contour, sigma_fn):
    """ρ = Σ_i (i/2π)(W_i G_i - conj(W_i) G_i†), G_i = (E_i - H - Σ(E_i))⁻¹
    (cmdtools/SCF:209-215, the non-memory-conserving funci).  ``contour`` (nE, 4):
    Re E, Im E, Re W, Im W (already shifted by μ); ``sigma_fn(E)`` -> (NK, n, n)."""
    NK, n, _ = H.shape
    I = np.eye(n)
    rho = np.zeros((NK, n, n), complex)
    for Er, Ei, Wr, Wi in contour:
        E, W = Er + 1j * Ei, Wr + 1j * Wi
        G = np.linalg.inv(E * I - H - sigma_fn(E))
        rho += (1j / (2 * np.pi)) * (W * G - np.conj(W) * G.conj().transpose(0, 2, 1))
    return rho


def equilibrium_density(consts, H, kT, N_F, mu=0.0, hbar=HBAR):
    """Equilibrium density of the model the propagation constants describe:
    σ_eq = ∫ dE/2π f̃(E) A(E) with the N_F-pole Hu expansion f̃ (PadeDecomp.py:36-42) and
    the leads' Lorentzian self-energy, closed by residues,
        σ_eq = I/2 + kT Σ_p R_p [G^r(w_p) + G^r(w_p)†],  w_p = μ + i q_p kT,
    (the spectral sum ∫ A dE/2π = I gives I/2; each pole of f̃ picks up G^r in the upper and
    G^a in the lower half-plane).  It is what the SCF tool's contour sum (cmdtools/SCF:
    209-215) converges to for this f̃, and a fixed point of the equations of motion
    together with psinought's Ψ₀ and Ω₀ (tests/test_builder_constants.py)."""
    H = np.asarray(H, complex)
    NK, n, _ = H.shape
    z, R = hu_poles(N_F)
    rho = np.broadcast_to(0.5 * np.eye(n), H.shape).astype(complex)
    for zp, Rp in zip(z, R):
        w = mu + zp * kT
        G = np.linalg.inv(w * np.eye(n) - H - lead_self_energy(consts, w).sum(axis=1))
        rho = rho + kT * Rp * (G + G.conj().transpose(0, 2, 1))
    return rho


def fermi(E, mu, kT):
    """1 / (1 + exp((E - mu)/kT)) (the exact Fermi function of cmdtools/SCF:228-229)."""
    return 1.0 / (1.0 + np.exp((np.asarray(E, float) - mu) / kT))


def window_limits(mus, kTs, fact_kT=10.0):
    """[mu_min - fact_kT kT_max, mu_max + fact_kT kT_max] (cmdtools/SCF:273-276)."""
    mus, kTs = np.asarray(mus, float), np.asarray(kTs, float)
    return mus.min() - fact_kT * kTs.max(), mus.max() + fact_kT * kTs.max()


def window_integrand(H, E, sigma_leads_fn, mus, kTs):
    """G_l(E) of cmdtools/SCF:236-253 with the nonequilibrium fdf of :301-302:
    (-i/2π) G SE_l G†, SE_l = Σ_α -(f_α(E) - f_min(E)) (Σ_α - Σ_α†), G = (E - H - ΣΣ_α)⁻¹;
    ``sigma_leads_fn(E)`` -> (NK, NL, n, n).  Full n x n here (the reference keeps
    the coupled columns only, which carry every non-zero of SE_l)."""
    NK, n, _ = H.shape
    mus, kTs = np.asarray(mus, float), np.asarray(kTs, float)
    imin = int(np.argmin(mus))
    S = sigma_leads_fn(E)
    G = np.linalg.inv(E * np.eye(n) - H - S.sum(axis=1))
    f_min = fermi(E, mus[imin], kTs[imin])
    SE_l = np.zeros((NK, n, n), complex)
    for a in range(S.shape[1]):
        fd = fermi(E, mus[a], kTs[a]) - f_min
        SE_l += -fd * (S[:, a] - S[:, a].conj().transpose(0, 2, 1))
    return (-1j / (2 * np.pi)) * (G @ SE_l @ G.conj().transpose(0, 2, 1))


def gauss_legendre_panels(e1, e2, panels, order=8):
    """Composite Gauss-Legendre nodes and weights on [e1, e2] (the fixed quadrature
    used in place of scipy's adaptive quad_vec of cmdtools/SCF:303-305)."""
    x, w = np.polynomial.legendre.leggauss(order)
    edges = np.linspace(e1, e2, panels + 1)
    nodes, weights = [], []
    for a, b in zip(edges[:-1], edges[1:]):
        nodes.append(0.5 * (b - a) * x + 0.5 * (a + b))
        weights.append(0.5 * (b - a) * w)
    return np.concatenate(nodes), np.concatenate(weights)


def window_density(H, nodes, weights, sigma_leads_fn, mus, kTs):
    """Σ_i w_i G_l(E_i): the bias-window density on given real nodes."""
    rho = np.zeros(H.shape, complex)
    for E, w in zip(nodes, weights):
        rho += w * window_integrand(H, E, sigma_leads_fn, mus, kTs)
    return rho
