"""zandpack_amd: MI355X-native engine for the auxiliary-mode-expansion (AME)
equations of motion of Zandpack and their DOP54 adaptive Runge-Kutta driver.

Product path: ``libzandhip.so`` (HIP, gfx950) through the C ABI in
``include/zandhip.h``; ``Propagator`` (device-resident propagation),
``make_f`` / ``AdaptiveDOP`` (the reference's in-process call surface) and
``builder`` (propagation constants and seeded synthetic problems).
"""
from .builder import AMEProblem, Pulse, HBAR  # noqa: F401


def __getattr__(name):
    # lazy: importing the package must not require the built library
    if name in ("Propagator", "selftest_gemm"):
        from . import engine
        return getattr(engine, name)
    if name in ("make_f", "AdaptiveDOP"):
        from . import drop_in
        return getattr(drop_in, name)
    raise AttributeError(name)
