"""zh_contour_density_noalloc (no device allocation: slices of the library's static
scratch array) against zh_contour_density, through the C ABI: two calls in a row, the
second smaller, to the FP64 bar (1e-10 relative to max|ρ|); a size beyond the
scratch array is refused before any device work."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from zandpack_amd import builder, scf
from zandpack_amd import _native as N

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _inputs(p):
    cont = np.load(os.path.join(GOLD, "scf_default_contour.npy"))
    corb = scf.coupled_orbitals(p)
    sig = scf.self_energy_cc(p, cont[:, 0] + 1j * cont[:, 1], corb)
    H = np.asarray(p.H0, complex) + p.dH(0.0, p.sigma0)
    return H, cont, corb, sig


def test_noalloc_matches_allocating(native):
    # C3 size (n = 1000, 80 coupled orbitals, 102 points: several chunks), then C2 (n = 400)
    for name, p in (("c3", builder.config_c3(seed=0)), ("c2", builder.config_c2(seed=1))):
        H, cont, corb, sig = _inputs(p)
        eig = scf.eigenpairs(H)
        say("noalloc contour density", name, H.shape, corb.size)
        ref = scf.contour_density(H, cont, corb, sig, eig=eig)
        got = scf.contour_density(H, cont, corb, sig, eig=eig, noalloc=True)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        say("rel err", err)
        assert err < 1e-10
        assert np.isfinite(got).all()


def test_noalloc_refuses_oversize(native):
    n, nE = 3000, 4
    e = np.zeros((1, n))
    V = np.zeros((1, n, n), complex)
    cont = np.zeros((nE, 4))
    rho = np.zeros((1, n, n), complex)
    rc = N.lib().zh_contour_density_noalloc(1, n, nE, 0, N.ptr(cont), N.ptr(e), N.ptr(V), None,
                                            None, N.ptr(rho), 0)
    assert rc == -1
    assert b"scratch" in N.lib().zh_last_error()
