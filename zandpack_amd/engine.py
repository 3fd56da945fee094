"""Host-side handle over the native propagation engine (libzandhip.so).

``Propagator`` owns one ``zh_problem``: constants, device-evaluated H(t, σ)
model and bias, and the device-resident state (σ, Ψ, Ω).  It replaces one run of
the reference's ``mpirun -np P zand`` (Zandpack/mpi/zand:200-495, 498-853) on a
single GPU, and exposes the explicit one-evaluation RHS that the reference's
in-process ``f`` closures compute (Zandpack/TimedependentTransport.py:2605-2660).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .builder import AMEProblem

PRECISIONS = {"fp64": N.ZH_FP64, "mixed": N.ZH_MIXED}
LAYOUTS = {"auto": 0, "reference": 1, "pruned": 2}
SPARSE = {"auto": 0, "dense": 1, "sparse": 2}


class Propagator:
    def __init__(self, problem: AMEProblem, precision: str = "fp64", device: int = 0,
                 graph: bool = True, state=None, layout: str = "auto", autotune: bool = False,
                 hproducts: str = "auto", antihermitian: str = "auto", _instances: int = 1):
        self.lib = N.lib()
        self.problem = problem
        self.precision = precision
        NK, n, NL, NX, NOT = problem.dims
        self.NK, self.n, self.NL, self.NX, self.NOT = NK, n, NL, NX, NOT
        self.M = NL * NX * NOT
        G = _instances
        d = N.zh_dims(NK // G, n, NL, NX, NOT, PRECISIONS[precision], float(problem.hbar), device)
        h = C.c_void_p()
        if G == 1:
            N.check(self.lib.zh_problem_create(C.byref(d), C.byref(h)))
        else:
            N.check(self.lib.zh_problem_create_instances(C.byref(d), G, C.byref(h)))
        self.h = h
        N.check(self.lib.zh_set_layout(self.h, LAYOUTS[layout]))
        N.check(self.lib.zh_set_sparse(self.h, SPARSE[hproducts]))
        p = problem
        self._keep = [N.c128(a) for a in (p.H0, p.xi, p.Ixi, p.GL_P, p.dGP, p.dGM, p.Xpp, p.Xpm)]
        N.check(self.lib.zh_set_constants(self.h, *[N.ptr(a) for a in self._keep]))
        if p.hubbard_U:
            n0 = np.ascontiguousarray(p.hubbard_n0, dtype=np.float64)
            N.check(self.lib.zh_set_hubbard(self.h, float(p.hubbard_U), N.ptr(n0)))
        self.set_pulse(p.pulse)
        if p.drive:
            self.set_drive(p.drive)
        if p.dissipator.kind != "none":
            self.set_dissipator(p.dissipator)
        if p.Sinv is not None:
            self.set_overlap(p.Sinv)
        N.check(self.lib.zh_set_graph(self.h, 1 if graph else 0))
        if antihermitian != "on":  # "on" is checked against the state, which comes next
            self.set_antihermitian(antihermitian)
        if state is not None:
            self.put_state(*state)
        else:
            self.put_state(p.sigma0)
        if antihermitian == "on":
            self.set_antihermitian("on")
        self.tiles = self.autotune() if autotune else None

    # --------------------------------------------------------------- set-up
    def set_pulse(self, pulse):
        w = np.zeros(max(self.NL, 1))
        w[: len(pulse.weights)] = pulse.weights[: self.NL]
        w = np.ascontiguousarray(w[: self.NL])
        if pulse.kind == "spline":
            x = np.ascontiguousarray(pulse.spline_x, np.float64)
            c = np.ascontiguousarray(pulse.spline_c, np.float64)
            assert c.shape == (4, len(x) - 1)
            N.check(self.lib.zh_set_pulse_spline(self.h, len(x), N.ptr(x), N.ptr(c)))
        par = np.array([pulse.p1, pulse.p2, pulse.p3, pulse.p4], np.float64)
        N.check(self.lib.zh_set_pulse_ex(self.h, N.PULSE_KINDS[pulse.kind], float(pulse.amp),
                                         N.ptr(par), N.ptr(w)))

    def set_drive(self, terms):
        """Drive terms H(t) = H0 + Σ f_i(t) M_i (builder.Drive list) evaluated on the
        device every stage; an empty list removes them."""
        kinds = (C.c_int * max(1, len(terms)))(*[N.WAVE_KINDS[d.kind] for d in terms])
        amps = np.array([float(d.amp) for d in terms] or [0.0])
        par = np.zeros((max(1, len(terms)), 4))
        ks, rows, cols, vals, nnz = [], [], [], [], []
        for i, d in enumerate(terms):
            pv = tuple(float(x) for x in d.par)[:4]
            par[i, :len(pv)] = pv
            M = np.asarray(d.M, complex)
            assert M.shape == (self.NK, self.n, self.n)
            k, r, c = np.nonzero(M)
            ks.append(k); rows.append(r); cols.append(c); vals.append(M[k, r, c])
            nnz.append(len(k))
        cat = lambda xs, dt: np.ascontiguousarray(np.concatenate(xs) if xs else np.zeros(0), dt)
        ks, rows, cols = (cat(x, np.int32) for x in (ks, rows, cols))
        vals = cat(vals, np.complex128)
        nnz_a = (C.c_int * max(1, len(terms)))(*nnz)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        self._drive_keep = (amps, par, ks, rows, cols, vals)
        N.check(self.lib.zh_set_drive(self.h, len(terms), kinds, N.ptr(amps), N.ptr(par), nnz_a,
                                      ip(ks), ip(rows), ip(cols), N.ptr(vals)))

    def set_tolerance(self, rtol=0.0, atol=0.0):
        """Scaled RMS error norm with these tolerances (attempts then take eps = 1);
        (0, 0) restores zand's absolute norm."""
        N.check(self.lib.zh_set_tolerance(self.h, float(rtol), float(atol)))

    def set_overlap(self, Sinv):
        """Non-orthogonal (nozand) form with S⁻¹ (NK, n, n); None: orthogonal form."""
        if Sinv is None:
            N.check(self.lib.zh_set_overlap(self.h, None))
            return
        self._sinv = N.c128(Sinv)
        assert self._sinv.shape == (self.NK, self.n, self.n)
        N.check(self.lib.zh_set_overlap(self.h, N.ptr(self._sinv)))

    def steady_state(self, t0=0.0, tol=1e-10, maxiter=500, precondition=True,
                     self_energy=True, raise_on_limit=True):
        """Ψ₀, Ω₀ of cmdtools/psinought for the resident σ (as σ₀), the bias at t0 and
        H(t0, σ₀); they become the resident state.  Returns (iterations, relative
        residual).  The preconditioner (χ⁺_i + Δ - H + B_i)^{-1} is built on the host:
        eigenpairs of H(t0, σ₀) = H0 + dH(t0, σ₀) (LAPACK) and, with self_energy, the
        Woodbury factors of the lead self-energy B_i (``steady_preconditioner``).  With
        raise_on_limit=False an unconverged solve returns its counts instead of raising."""
        keep = []
        pc = None
        if precondition:
            s0 = self.sigma()
            Hs = np.asarray(self.problem.H0 + self.problem.dH(t0, s0), complex)
            pre = steady_preconditioner(self.problem, Hs, t0, self_energy)
            keep = list(pre.values())
            pc = N.zh_precond(N.ptr(pre["evecs"]), N.ptr(pre["evals"]), 0, None, 0, None, None)
            if self_energy and pre["corb"].size:
                pc.ncr = int(pre["corb"].size)
                pc.corb = pre["corb"].ctypes.data_as(C.POINTER(C.c_int))
                pc.ngroups = int(pre["K"].shape[0])
                pc.group = pre["group"].ctypes.data_as(C.POINTER(C.c_int))
                pc.K = N.ptr(pre["K"])
        its, res = C.c_int(0), C.c_double(0.0)
        code = self.lib.zh_steady_state(self.h, float(t0), float(tol), int(maxiter),
                                        C.byref(pc) if pc is not None else None,
                                        C.byref(its), C.byref(res))
        del keep
        if code != N.ZH_ELIMIT or raise_on_limit:
            N.check(code)
        return its.value, res.value

    def set_dissipator(self, d):
        """dissipator(t, σ) = η(t)(σ - σ_eq) of the attempts (builder.Dissipator)."""
        kind = d.KINDS[d.kind]
        seq = None
        if kind:
            seq = N.c128(d.sig_eq)
            assert seq.shape == (self.NK, self.n, self.n)
        N.check(self.lib.zh_set_dissipator(self.h, kind, float(d.eta0), float(d.t0),
                                           float(d.width), N.ptr(seq) if kind else None))
        self._diss = seq

    def layout_info(self):
        out = (C.c_int * 5)()
        N.check(self.lib.zh_layout_info(self.h, out))
        return dict(active_first=bool(out[0]), skip_passive_block=bool(out[1]),
                    device_modes=out[2], active_block=out[3], coupled_orbitals=out[4])

    def set_hproducts(self, mode: str):
        """H products as dense MFMA GEMMs ("dense"), gathers over a row-compressed H
        ("sparse") or chosen from H's non-zeros ("auto")."""
        N.check(self.lib.zh_set_sparse(self.h, SPARSE[mode]))

    def hamiltonian_info(self):
        out = (C.c_int * 3)()
        N.check(self.lib.zh_hamiltonian_info(self.h, out))
        return dict(real=bool(out[0]), sparse=bool(out[1]), row_entries=out[2])

    def set_fusion(self, mode="auto"):
        """Fused Ω stage kernel of mixed precision (the Ωξ product inside the Ω kernel):
        "auto" (the autotuner's timed choice), "on" or "off"; same results to rounding.
        True / False mean "on" / "off"."""
        if mode is True or mode is False:
            mode = "on" if mode else "off"
        N.check(self.lib.zh_set_fusion(self.h, {"auto": 0, "off": 1, "on": 2}[mode]))

    def set_omgxi(self, mode="auto"):
        """Kernel of the Ωξ product of the Ψ equation: "auto" (the autotuner's timed
        choice), "gemm" (the tiled GEMM) or "stream" (the streaming kernel: Ω and ξ_c
        staged through LDS by direct global-to-LDS loads; falls back to the GEMM where
        the layout does not allow it); same results to rounding."""
        N.check(self.lib.zh_set_omgxi(self.h, {"auto": 0, "gemm": 1, "stream": 2}[mode]))

    def set_psi_kernel(self, mode="auto"):
        """How the Ψ right-hand side runs with a row-compressed H: "auto" (one pass when
        the rows of H stay near their tile), "split" (matrix-core ΔPξ_cσ_cᵀ product, then
        the gather kernel) or "fused" (one pass); same results to rounding."""
        N.check(self.lib.zh_set_psi_kernel(self.h, {"auto": 0, "split": 1, "fused": 2}[mode]))

    def psi_kernel_info(self):
        v = C.c_int()
        N.check(self.lib.zh_psi_kernel_info(self.h, C.byref(v)))
        return "fused" if v.value else "split"

    def set_fsal(self, mode="on"):
        """First same as last: "on" (an attempt after an accepted one takes that one's k_7
        as its k_1 where f is the same there) or "off" (every attempt evaluates all
        seven stages); same results to rounding."""
        N.check(self.lib.zh_set_fsal(self.h, {"on": 0, "off": 1}[mode]))

    def fsal_info(self):
        """{"taken": instance attempts that reused k_7, "on": whether attempts may}."""
        t, o = C.c_long(), C.c_int()
        N.check(self.lib.zh_fsal_info(self.h, C.byref(t), C.byref(o)))
        return {"taken": t.value, "on": bool(o.value)}

    def omgxi_info(self):
        """{"stream", "possible"}: whether the attempts run the streaming Ωξ kernel,
        whether the problem's layout allows it."""
        out = (C.c_int * 2)()
        N.check(self.lib.zh_omgxi_info(self.h, out))
        return {"stream": bool(out[0]), "possible": bool(out[1])}

    def set_antihermitian(self, mode="auto"):
        """Ω held as its block upper triangle (half the Ω stage kernel's work) when the
        constants keep Ω anti-Hermitian (ΔM_j Ixi_j = -conj(ΔP_j ξ_j), χ⁻ = conj χ⁺) and
        the resident Ω is anti-Hermitian: "auto" (when both hold), "on" (refused
        otherwise) or "off"; same results to rounding."""
        if mode is True or mode is False:
            mode = "on" if mode else "off"
        N.check(self.lib.zh_set_antihermitian(self.h, {"auto": 0, "off": 1, "on": 2}[mode]))

    def antihermitian_info(self):
        """{"active", "constants", "state"}: whether Ω is held as its block upper triangle
        now, whether the constants have the symmetry, whether the resident Ω is
        anti-Hermitian."""
        v = C.c_int()
        N.check(self.lib.zh_antihermitian_info(self.h, C.byref(v)))
        return {"active": bool(v.value & 1), "constants": bool(v.value & 2),
                "state": bool(v.value & 4)}

    def set_graph(self, on: bool):
        N.check(self.lib.zh_set_graph(self.h, 1 if on else 0))

    def set_profile(self, on: bool):
        N.check(self.lib.zh_set_profile(self.h, 1 if on else 0))

    # --------------------------------------------------------------- state
    def put_state(self, sig, psi=None, omg=None):
        """Resident state; a None Ψ or Ω is zero (cleared on the device)."""
        s = N.c128(sig)
        assert s.shape == (self.NK, self.n, self.n)
        p = o = None
        if psi is not None:
            p = N.c128(psi)
            assert p.shape == (self.NK, self.NL, self.NX, self.NOT, self.n)
        if omg is not None:
            o = N.c128(omg)
            assert o.shape == (self.NK, self.NL, self.NX, self.NOT, self.NL, self.NX, self.NOT)
        N.check(self.lib.zh_state_put(self.h, N.ptr(s), None if p is None else N.ptr(p),
                                      None if o is None else N.ptr(o)))

    def get_state(self):
        s = np.empty((self.NK, self.n, self.n), complex)
        p = np.empty((self.NK, self.NL, self.NX, self.NOT, self.n), complex)
        o = np.empty((self.NK, self.NL, self.NX, self.NOT, self.NL, self.NX, self.NOT), complex)
        N.check(self.lib.zh_state_get(self.h, N.ptr(s), N.ptr(p), N.ptr(o)))
        return s, p, o

    def pi(self, hermitian=True):
        """Per-lead Π of the resident state (NK, NL, n, n): (Π_α + Π_α†)/ħ as zand's
        Q_np, or Π_α/ħ as nozand's PI_np with ``hermitian=False``."""
        Q = np.empty((self.NK, self.NL, self.n, self.n), complex)
        N.check(self.lib.zh_pi_get(self.h, 1 if hermitian else 0, N.ptr(Q)))
        return Q

    def sigma(self):
        s = np.empty((self.NK, self.n, self.n), complex)
        N.check(self.lib.zh_sigma_get(self.h, N.ptr(s)))
        return s

    # --------------------------------------------------------------- compute
    def rhs(self, Ht, delta, sig, psi, omg, want_omega=True):
        """Time derivatives (dσ, dΨ, dΩ) and currents J (NL, NK) of one state."""
        Ht, s, p, o = (N.c128(a) for a in (Ht, sig, psi, omg))
        d = np.ascontiguousarray(delta, dtype=np.float64)
        assert d.shape == (self.NL,)
        ds = np.empty_like(s)
        dp = np.empty_like(p)
        do = np.empty_like(o) if want_omega else None
        J = np.empty((self.NL, self.NK))
        N.check(self.lib.zh_rhs(self.h, N.ptr(Ht), N.ptr(d), N.ptr(s), N.ptr(p), N.ptr(o),
                                N.ptr(ds), N.ptr(dp), N.ptr(do), N.ptr(J)))
        return ds, dp, do, J

    def attempt(self, t0, h, eps, force_accept=False):
        """One DOP54 attempt from the resident state; commits when accepted."""
        info = N.zh_attempt_info(float(t0), float(h), float(eps), 1 if force_accept else 0,
                                 0.0, 0.0, 0)
        J = np.empty((self.NL, self.NK))
        N.check(self.lib.zh_attempt(self.h, C.byref(info), N.ptr(J)))
        return dict(te=info.te, h_next=info.h_next, accepted=bool(info.accepted), J=J)

    def launch(self, t0, h, eps, force_accept=False):
        """Enqueue one attempt and return at once (see ``collect``)."""
        self._pending = N.zh_attempt_info(float(t0), float(h), float(eps),
                                          1 if force_accept else 0, 0.0, 0.0, 0)
        N.check(self.lib.zh_attempt_launch(self.h, C.byref(self._pending)))

    def ready(self):
        """True when the launched attempt has finished (``collect`` will not wait)."""
        r = C.c_int(0)
        N.check(self.lib.zh_attempt_ready(self.h, C.byref(r)))
        return bool(r.value)

    def collect(self):
        """Wait for the launched attempt; commits the state when accepted."""
        info = self._pending
        J = np.empty((self.NL, self.NK))
        N.check(self.lib.zh_attempt_collect(self.h, C.byref(info), N.ptr(J)))
        return dict(te=info.te, h_next=info.h_next, accepted=bool(info.accepted), J=J)

    def autotune(self, force=False):
        """Pick the fastest tile shape of each GEMM on this problem's resident data."""
        out = (C.c_int * 5)()
        N.check(self.lib.zh_autotune(self.h, 1 if force else 0, out))
        return dict(zip(("psi", "hsig", "omega", "omgxi", "pi"), list(out)))

    def set_tiles(self, choice):
        arr = (C.c_int * 5)(*[int(choice[k]) for k in ("psi", "hsig", "omega", "omgxi", "pi")])
        N.check(self.lib.zh_set_tiles(self.h, arr))

    def device_bytes(self):
        b = C.c_long(0)
        N.check(self.lib.zh_problem_bytes(self.h, C.byref(b)))
        return b.value

    def propagate(self, t0, t1, h0, eps, time_update="reference", max_attempts=10 ** 7,
                  cap=100000):
        """zand's adaptive loop on the device; returns times, J, h of accepted steps."""
        times = np.empty(cap)
        J = np.empty((cap, self.NL, self.NK))
        hs = np.empty(cap)
        ns, nr = C.c_long(0), C.c_long(0)
        hf = C.c_double(0.0)
        tu = N.ZH_TIME_REFERENCE if time_update == "reference" else N.ZH_TIME_CORRECT
        N.check(self.lib.zh_propagate(self.h, float(t0), float(t1), float(h0), float(eps), tu,
                                      int(max_attempts), int(cap), N.ptr(times), N.ptr(J),
                                      N.ptr(hs), C.byref(ns), C.byref(nr), C.byref(hf)))
        k = min(ns.value, cap)
        return dict(times=np.concatenate(([t0], times[:k])), J=J[:k], h=hs[:k],
                    n_steps=ns.value, n_rejected=nr.value, h_final=hf.value)

    def timing(self, reset=False, cap=16):
        names = C.create_string_buffer(32 * cap)
        ms = np.zeros(cap)
        cnt = np.zeros(cap, dtype=np.int64)
        k = self.lib.zh_timing(self.h, 1 if reset else 0, cap, names,
                               N.ptr(ms), cnt.ctypes.data_as(C.POINTER(C.c_long)))
        raw = names.raw
        out = {}
        for i in range(k):
            nm = raw[32 * i: 32 * (i + 1)].split(b"\0")[0].decode()
            out[nm] = (ms[i], int(cnt[i]))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.zh_problem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def steady_preconditioner(problem, H, t0, self_energy=True):
    """Host factors of the steady-state preconditioner (χ⁺_i + Δ - H + B_i)^{-1}
    (psinought's block preconditioner, Zandpack/cmdtools/psinought:256-301):
    eigenpairs of H per k, and for each group g of modes with one χ⁺ the Woodbury
    matrix K_g = C_g (I + G_g C_g)^{-1}, where on the coupled orbitals corb
      C_g = Σ_j ξ_j Ixi_jᵀ · i ΔM_j / (χ⁻_j - χ⁺_g)   (the T1 part of the operator)
      G_g = [(χ⁺_g - H)^{-1}]_corb,corb.
    χ± include the bias at t0."""
    NK, n, NL, NX, NOT = problem.dims
    M = problem.M
    e, v = np.linalg.eigh(np.asarray(H, complex))
    out = dict(evecs=np.ascontiguousarray(v, np.complex128),
               evals=np.ascontiguousarray(e, np.float64),
               corb=np.zeros(0, np.int32))
    if not self_energy:
        return out
    xi = problem.xi.reshape(NK, M, n)
    Ixi = problem.Ixi.reshape(NK, M, n)
    corb = np.nonzero(np.any(xi.reshape(-1, n) != 0, axis=0) |
                      np.any(Ixi.reshape(-1, n) != 0, axis=0))[0].astype(np.int32)
    if corb.size == 0:
        return out
    lead = np.repeat(np.arange(NL), NX * NOT)
    dl = np.array([problem.bias(t0, a) for a in range(NL)])[lead]
    xpp = problem.Xpp.reshape(NK, M) + dl
    xpm = problem.Xpm.reshape(NK, M) + dl
    dgm = problem.dGM.reshape(NK, M)
    group = np.zeros((NK, M), np.int32)
    Ks = []
    for k in range(NK):
        vals, inv = np.unique(xpp[k], return_inverse=True)
        group[k] = inv + len(Ks)
        Xc = xi[k][:, corb].T            # (ncr, M)
        Ic = Ixi[k][:, corb].T
        Vc = v[k][corb]                  # (ncr, n)
        for chi in vals:
            f = 1j * dgm[k] / (xpm[k] - chi)
            Cg = (Xc * f) @ Ic.T
            Gg = (Vc / (chi - e[k])) @ Vc.conj().T
            Ks.append(Cg @ np.linalg.inv(np.eye(corb.size) + Gg @ Cg))
    out.update(corb=np.ascontiguousarray(corb), group=np.ascontiguousarray(group),
               K=np.ascontiguousarray(np.array(Ks), np.complex128))
    return out


def selftest_gemm(A, B, transA=False, transB=False, precision="fp64"):
    """C = op(A) op(B) on the device with the engine's MFMA GEMM."""
    A = N.c128(A)
    B = N.c128(B)
    M, K = (A.shape[1], A.shape[0]) if transA else A.shape
    K2, Nn = (B.shape[1], B.shape[0]) if transB else B.shape
    assert K == K2
    Cm = np.empty((M, Nn), complex)
    N.check(N.lib().zh_selftest_gemm(PRECISIONS[precision], M, Nn, K, int(transA), int(transB),
                                     N.ptr(A), N.ptr(B), N.ptr(Cm)))
    return Cm


def batch_attempt(props, t0s, hs, eps, force_accept=False):
    """One attempt of each Propagator in ``props`` from times ``t0s`` with steps
    ``hs`` (C: ``zh_batch_attempt``: every launch, then every collect).  Returns one
    entry per problem: the ``collect`` dict, or a ``ZHError`` for a problem whose
    attempt failed (the others go on)."""
    n = len(props)
    if n == 0:
        return []
    lib = N.lib()
    eps = np.broadcast_to(np.asarray(eps, dtype=np.float64), (n,))
    infos = (N.zh_attempt_info * n)(*[
        N.zh_attempt_info(float(t), float(h), float(e), 1 if force_accept else 0, 0.0, 0.0, 0)
        for t, h, e in zip(t0s, hs, eps)])
    Js = [np.empty((p.NL, p.NK)) for p in props]
    jp = (N._dp * n)(*[N.ptr(J) for J in Js])
    hs_ = (C.c_void_p * n)(*[p.h.value for p in props])
    status = (C.c_int * n)()
    rc = lib.zh_batch_attempt(hs_, n, infos, jp, status)
    msg = lib.zh_last_error().decode() if rc != 0 else ""
    out = []
    for i in range(n):
        if status[i] != 0:
            out.append(N.ZHError(status[i], msg if status[i] == rc else "attempt failed"))
            continue
        info = infos[i]
        out.append(dict(te=info.te, h_next=info.h_next, accepted=bool(info.accepted), J=Js[i]))
    return out


class _Stacked:
    """Problems of one shape stacked on the k axis (the constants of instance g are its
    k blocks); pulses stay per instance."""

    def __init__(self, problems):
        p0 = problems[0]
        for p in problems:
            assert p.dims == p0.dims and p.hbar == p0.hbar, "instances must share one shape"
            assert bool(p.hubbard_U) == bool(p0.hubbard_U) and p.hubbard_U == p0.hubbard_U
            assert not p.drive and p.dissipator.kind == "none" and p.Sinv is None, \
                "instances carry a pulse each; drive, dissipator and overlap stay single"
        NK, n, NL, NX, NOT = p0.dims
        G = len(problems)
        self.dims = (NK * G, n, NL, NX, NOT)
        self.hbar = p0.hbar
        cat = lambda name: np.concatenate([np.asarray(getattr(p, name)) for p in problems])
        for name in ("H0", "xi", "Ixi", "GL_P", "dGP", "dGM", "Xpp", "Xpm", "sigma0"):
            setattr(self, name, cat(name))
        self.hubbard_U = p0.hubbard_U
        self.hubbard_n0 = cat("hubbard_n0") if p0.hubbard_U else None
        self.pulse = p0.pulse
        self.drive = None
        self.dissipator = p0.dissipator
        self.Sinv = None


class Instances(Propagator):
    """Several independent propagations of one shape in one native problem: every
    kernel of an attempt runs once for all of them (zh_problem_create_instances).  Each
    instance has its own pulse, time, step, tolerance, error norm and acceptance;
    ``attempt`` takes one (t0, h, eps) per instance.  Results of an instance are bitwise
    those of the same problem alone on the same tiles."""

    def __init__(self, problems, precision: str = "fp64", device: int = 0, graph: bool = True,
                 states=None, layout: str = "auto", autotune: bool = False,
                 hproducts: str = "auto", antihermitian: str = "auto"):
        self.problems = list(problems)
        self.G = len(self.problems)
        st = _Stacked(self.problems)
        self.KI = self.problems[0].dims[0]
        state = None
        if states is not None:
            state = tuple(np.concatenate([np.asarray(s[i]) for s in states]) for i in range(3))
        super().__init__(st, precision, device, graph, state, layout, False, hproducts,
                         antihermitian, _instances=self.G)
        for g, p in enumerate(self.problems):
            self.set_instance_pulse(g, p.pulse)
        self.tiles = self.autotune() if autotune else None

    def set_instance_pulse(self, g, pulse):
        w = np.zeros(max(self.NL, 1))
        w[: len(pulse.weights)] = pulse.weights[: self.NL]
        w = np.ascontiguousarray(w[: self.NL])
        assert pulse.kind != "spline", "spline pulses are shared: use set_pulse"
        par = np.array([pulse.p1, pulse.p2, pulse.p3, pulse.p4], np.float64)
        N.check(self.lib.zh_instance_pulse(self.h, int(g), N.PULSE_KINDS[pulse.kind],
                                           float(pulse.amp), N.ptr(par), N.ptr(w)))

    def instance_state(self, g, state=None):
        """Instance g's (σ, Ψ, Ω) out of the stacked state (the resident one by default)."""
        s, p, o = self.get_state() if state is None else state
        k = slice(g * self.KI, (g + 1) * self.KI)
        return s[k], p[k], o[k]

    def _infos(self, t0, h, eps, force_accept):
        G = self.G
        vec = lambda v: np.broadcast_to(np.asarray(v, dtype=float), (G,))
        t0, h, eps = vec(t0), vec(h), vec(eps)
        force = np.broadcast_to(np.asarray(force_accept, dtype=bool), (G,))
        arr = (N.zh_attempt_info * G)()
        for g in range(G):
            arr[g] = N.zh_attempt_info(float(t0[g]), float(h[g]), float(eps[g]),
                                       1 if force[g] else 0, 0.0, 0.0, 0)
        return arr

    def _results(self, arr, J):
        return [dict(te=arr[g].te, h_next=arr[g].h_next, accepted=bool(arr[g].accepted),
                     J=J[:, g * self.KI:(g + 1) * self.KI])
                for g in range(self.G)]

    def attempt(self, t0, h, eps, force_accept=False):
        """One attempt of every instance (per-instance or broadcast t0, h, eps); h = 0
        with force_accept holds a finished instance.  Returns one dict per instance."""
        arr = self._infos(t0, h, eps, force_accept)
        J = np.empty((self.NL, self.NK))
        N.check(self.lib.zh_attempt_instances(self.h, arr, N.ptr(J)))
        return self._results(arr, J)

    def launch(self, t0, h, eps, force_accept=False):
        self._pending = self._infos(t0, h, eps, force_accept)
        N.check(self.lib.zh_launch_instances(self.h, self._pending))

    def collect(self):
        arr = self._pending
        J = np.empty((self.NL, self.NK))
        N.check(self.lib.zh_collect_instances(self.h, arr, N.ptr(J)))
        return self._results(arr, J)
