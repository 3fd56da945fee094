"""Writes c1_chain20.npz: the oracle's propagation of config 1 (20-site chain,
two one-channel leads, 8 Lorentzians + 20 Hu poles, M = 96, ramp bias
V_L = 0.1 (1 - exp(-t/5 fs)) eV) from the equilibrium steady state (the SCF and
psinought equivalents: the equilibrium density of the model's own Fermi expansion,
oracle.equilibrium_density, then Ψ₀ and the closed-form Ω₀ of psinought's linear
system at zero bias), DOP54 with zand's control (absolute eps = 1e-8, t += h_next) to
t = 50 fs.  Run from the repository root: python tests/golden/make_c1_fixture.py"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import zand_oracle as O  # noqa: E402
from zandpack_amd import builder  # noqa: E402

T_END, H0, EPS = 50.0, 0.01, 1e-8
KT, N_F = 0.025, 20  # config_chain's Fermi expansion


def initial_state(p):
    """(σ_eq, Ψ₀, Ω₀) at t = 0 (zero bias)."""
    H = np.asarray(p.H0, complex)
    sig = O.equilibrium_density(p.consts(), H, KT, N_F)
    delta = np.array([p.bias(0.0, a) for a in range(p.dims[2])])
    psi, omg = O.steady_state_dense(p.consts(), H, sig, delta)
    return sig, psi, omg


def run():
    p = builder.config_chain()
    y0 = initial_state(p)
    out = O.propagate(p.consts(), y0, 0.0, T_END, H0, EPS, p.H0, O.zero_dH, p.bias, 2)
    return p, y0, out


if __name__ == "__main__":
    p, y0, out = run()
    every = 50
    np.savez_compressed(
        os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_chain20.npz"),
        times=out["times"], J=out["J"], h=out["h_accepted"], TE=out["TE"],
        sigma_every=out["sigma"][::every], every=every, n_reject=out["n_reject"],
        psi_final=out["y"][1], omega_final=out["y"][2], t_end=T_END, h0=H0, eps=EPS,
        sigma0=y0[0], psi0=y0[1], omega0=y0[2], kT=KT, N_F=N_F)
    print(len(out["times"]), "steps written")
