"""The reference side of the parity tests: the oracle's own stage-Hamiltonian and
dissipator models for a problem, built from the problem's parameters (U, n0, drive
terms, dissipator rate), so that no product code computes what the device result
is checked against."""
from oracle import zand_oracle as O


def dH_of(p):
    """dH(t, σ) of problem p from the oracle: Hubbard term and drive terms."""
    terms = [(O.waveform(d.kind, d.amp, d.par), d.M) for d in getattr(p, "drive", ())]
    return O.model_dH(terms, p.hubbard_U, p.hubbard_n0)


def diss_of(p):
    """dissipator(t, σ) of problem p from the oracle, or None."""
    d = p.dissipator
    if d.kind == "none":
        return None
    return O.relaxation_dissipator(O.dissipator_rate(d.kind, d.eta0, d.t0, d.width), d.sig_eq)
