"""The oracle's non-orthogonal (nozand) form pinned by the change of basis it is:
with ψ_n = S^-½ ψ_o, σ_n = S^-½ σ_o S^-½ and the arrays of ``to_nonorthogonal``,
the nozand right-hand side must be the orthogonal one carried to the same basis
(dσ_n = S^-½ dσ_o S^-½, dΨ_n = S^-½ dΨ_o, dΩ and the currents unchanged).  A
dropped or misplaced S⁻¹ in any of its four terms breaks this."""
import numpy as np
import pytest

from oracle import zand_oracle as O
from zandpack_amd import builder


def _pair(seed=3, NK=2, zero_tol=0.0, band=2):
    p = builder.config_small(NK=NK, seed=seed)
    p.pulse = builder.Pulse("step", amp=0.2, p1=0.0, weights=(0.5, -0.5))
    NK, n, NL, NX, NOT = p.dims
    Sm05 = builder.lowdin_factors(builder.random_overlap(NK, n, band=band, seed=seed + 1))
    return p, builder.to_nonorthogonal(p, Sm05, zero_tol=zero_tol), Sm05


def _random_state(p, rng):
    NK, n, NL, NX, NOT = p.dims
    c = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    s = c(NK, n, n)
    s = 0.5 * (s + s.conj().transpose(0, 2, 1))
    return s, 0.1 * c(NK, NL, NX, NOT, n), 0.1 * c(NK, NL, NX, NOT, NL, NX, NOT)


def _to_n(Sm05, sig, psi):
    return Sm05 @ sig @ Sm05, np.einsum("kij,kaxcj->kaxci", Sm05, psi)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("prune", [False, True])
def test_rhs_is_the_orthogonal_rhs_in_the_other_basis(prune):
    p, q, Sm05 = _pair()
    rng = np.random.default_rng(0)
    so, po, om = _random_state(p, rng)
    sn, pn = _to_n(Sm05, so, po)
    delta = np.array([0.1, -0.1])
    do = O.rhs(p.consts(), so, po, om, p.H0, delta, h=0.7)
    dn = O.rhs(q.consts(), sn, pn, om, q.H0, delta, h=0.7, Sinv=q.Sinv, prune=prune)
    ds_exp, dp_exp = _to_n(Sm05, do[0], do[1])
    assert rel(dn[0], ds_exp) < 1e-11
    assert rel(dn[1], dp_exp) < 1e-11
    assert rel(dn[2], do[2]) < 1e-11
    assert rel(O.currents(q.consts(), pn), O.currents(p.consts(), po)) < 1e-11


def test_identity_overlap_is_zand():
    p = builder.config_small(NK=1, seed=4)
    NK, n = p.dims[:2]
    rng = np.random.default_rng(1)
    s, ps, om = _random_state(p, rng)
    I = np.broadcast_to(np.eye(n), (NK, n, n)).astype(complex)
    a = O.rhs(p.consts(), s, ps, om, p.H0, np.zeros(p.dims[2]))
    b = O.rhs(p.consts(), s, ps, om, p.H0, np.zeros(p.dims[2]), Sinv=I)
    for x, y in zip(a, b):
        assert rel(y, x) < 1e-14


def test_each_overlap_term_matters():
    """Removing S⁻¹ from any one term changes the result (so the basis test pins all four)."""
    p, q, Sm05 = _pair()
    rng = np.random.default_rng(2)
    so, po, om = _random_state(p, rng)
    sn, pn = _to_n(Sm05, so, po)
    delta = np.array([0.05, -0.05])
    ref = O.rhs(q.consts(), sn, pn, om, q.H0, delta, Sinv=q.Sinv)
    # S⁻¹H replaced by H in both products: the commutator and the Ψ term change
    alt = O.rhs(q.consts(), sn, pn, om, q.H0, delta)
    assert rel(alt[0], ref[0]) > 1e-3 and rel(alt[1], ref[1]) > 1e-3


def test_attempt_stages_transform():
    p, q, Sm05 = _pair(seed=5)
    y0o = p.zero_state()
    rng = np.random.default_rng(3)
    _, po, om = _random_state(p, rng)
    y0o = (y0o[0], po, om)
    sn, pn = _to_n(Sm05, y0o[0], y0o[1])
    zero = lambda t, s: np.zeros(s.shape, complex)
    ko, TEo, Jo, yo = O.dop54_attempt(p.consts(), y0o, 0.0, 0.05, p.H0, zero, p.bias, 2)
    kn, TEn, Jn, yn = O.dop54_attempt(q.consts(), (sn, pn, om), 0.0, 0.05, q.H0, zero, q.bias, 2,
                                      Sinv=q.Sinv)
    for i in range(7):
        s_exp, p_exp = _to_n(Sm05, ko[i][0], ko[i][1])
        assert rel(kn[i][0], s_exp) < 1e-10
        assert rel(kn[i][1], p_exp) < 1e-10
        assert rel(kn[i][2], ko[i][2]) < 1e-10
    assert rel(Jn, Jo) < 1e-11
    assert TEn != TEo  # the error norm is taken in the non-orthogonal basis (nozand:503-509)


def test_zero_tol_drops_only_small_orbitals():
    p, q0, Sm05 = _pair(zero_tol=0.0, band=1)
    q = builder.to_nonorthogonal(p, Sm05, zero_tol=1e-3)
    kept = np.abs(q.xi).max(axis=(0, 2, 3)) > 0      # (NL, n)
    full = np.abs(q0.xi).max(axis=(0, 2, 3))
    assert np.all(full[~kept] <= 1e-3) and np.array_equal(q.xi[..., :][np.broadcast_to(
        kept[None, :, None, None, :], q.xi.shape)], q0.xi[np.broadcast_to(
            kept[None, :, None, None, :], q.xi.shape)])
