"""Host factors of the steady-state preconditioner: the Woodbury form used on the
device equals the dense inverse of the mode block (χ⁺_i + Δ - H + B_i) of
psinought's operator (cmdtools/psinought:256-301)."""
import numpy as np
import pytest

from zandpack_amd import builder
from zandpack_amd.engine import steady_preconditioner


@pytest.mark.parametrize("case", ["small", "spin_hubbard", "shared"])
def test_woodbury_equals_block_inverse(case):
    p = {"small": lambda: builder.config_small(seed=1),
         "spin_hubbard": lambda: builder.config_small(NK=2, hubbard_U=2.0, seed=2),
         "shared": lambda: builder.config_small(n_w=4, n_l=1, seed=13)}[case]()
    p.pulse = builder.Pulse("step", amp=0.2, p1=0.0, weights=(0.5, -0.5))
    t0 = 1.0
    NK, n, NL, NX, NOT = p.dims
    M = p.M
    H = p.H0 + p.dH(t0, p.sigma0)
    pre = steady_preconditioner(p, H, t0)
    v, e, corb, K, group = pre["evecs"], pre["evals"], pre["corb"], pre["K"], pre["group"]
    lead = np.repeat(np.arange(NL), NX * NOT)
    dl = np.array([p.bias(t0, a) for a in range(NL)])[lead]
    xpp = p.Xpp.reshape(NK, M) + dl
    xpm = p.Xpm.reshape(NK, M) + dl
    xi = p.xi.reshape(NK, M, n)
    Ixi = p.Ixi.reshape(NK, M, n)
    rng = np.random.default_rng(0)
    for k in range(NK):
        for i in rng.choice(M, size=5, replace=False):
            f = 1j * p.dGM.reshape(NK, M)[k] / (xpm[k] - xpp[k, i])
            B = (xi[k].T * f) @ Ixi[k]          # Σ_j ξ_j Ixi_jᵀ f_j
            A = xpp[k, i] * np.eye(n) - H[k] + B
            r = rng.normal(size=n) + 1j * rng.normal(size=n)
            want = np.linalg.solve(A, r)
            D = 1.0 / (xpp[k, i] - e[k])
            y = D * (v[k].conj().T @ r)
            zc = (v[k] @ y)[corb]
            w = K[group[k, i]] @ zc
            t = np.zeros(n, complex)
            t[corb] = w
            y = y - D * (v[k].conj().T @ t)
            got = v[k] @ y
            assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()
