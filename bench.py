#!/usr/bin/env python3
"""Headline benchmark: accepted DOP54 steps per second of the AME propagation
at n=1000 orbitals, M=4000 auxiliary modes per spin block (NK=2, mean-field
Hubbard U=2 eV, two leads x 20 channels, 50 Lorentzians + 50 Fermi slots),
FP64 drop-in vs mixed precision on one MI355X.

One "step" is one accepted DOP54 step (seven right-hand-side evaluations, the
embedded error norm and the controller) of the whole spin-polarised system,
advanced from device-resident state.  ``value`` is the mixed-precision rate;
the FP64 rate, the mixed/FP64 speedup and the mixed-vs-FP64 current error in
replay mode (mixed forced onto the FP64 run's accepted (t, h) sequence) are
reported beside it, with the roofline of the dominant kernel and the CPU
oracle timed on the host.

    python bench.py [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RK45 steps/s @n=1000,M=4000: FP64 vs mixed at ≤1e-6 J err; p50/p99 latency"
PEAK = {"fp64": 78.6, "mixed": 157.3}  # TFLOP/s, MI355X dense FP64 / FP32 matrix
HBM_PEAK = 8000.0  # GB/s, MI355X HBM3E


def tri_fraction(p):
    """Share of the Ω entries outside the passive x passive block that are independent
    when Ω is anti-Hermitian: the pairs (i, j), (j, i) once and the active diagonal."""
    act = ((p.dGP != 0) | (p.dGM != 0)).any(axis=0)
    M, ML = p.M, int(act.sum())
    nonff = M * M - (M - ML) ** 2
    return (nonff + ML) / (2.0 * nonff)


def algorithmic_flops(p, tri=False):
    """Real flops per launch of each GEMM class for the work the path performs.
    Dense complex count per k: n³ (Hσ) + M n² + M n n_c (ψHᵀ, ξσᵀ) + Σ_α n n_α M_α
    (Π) + M² n_c (Ωξ) + Σ_α M M_α n_α + M² n_c (Ω sources), 8 flops per complex MAC
    and 4 per complex x real MAC.  What the kernels execute is less: a real
    Hamiltonian halves ψHᵀ and Hσ and a sparse one reduces them to its non-zeros;
    the coupling vectors of a mode live on its own lead's orbitals (n_α of them), so
    ξσᵀ, Ωξ, Π and both Ω sources contract over those only; Fermi-block pruning drops
    the Ω sources on passive rows/columns and the passive x passive block of Ωξ.
    With Ω held as its block upper triangle (tri) the Ω sources are formed for its
    independent entries only; Ωξ stays the full product."""
    NK, n, NL, NX, NOT = p.dims
    M = p.M
    Ma = NX * NOT
    xi = p.xi.reshape(NK, NL, Ma, n)
    na = [int(np.count_nonzero(np.any(xi[:, a] != 0, axis=(0, 1)))) for a in range(NL)]
    nc = int(np.count_nonzero(np.any(p.xi.reshape(-1, n) != 0, axis=0)))
    h_real = not np.any(p.H0.imag)
    hmac = 4.0 if h_real else 8.0
    # H products as gathers over the row-compressed H (the engine's automatic choice
    # for at most 64 and at most n/8 non-zeros per row): nnz MACs per column of σ / row of ψ
    nnz_row = (p.H0 != 0).sum(axis=2)
    sparse = nnz_row.max() <= 64 and 8 * nnz_row.max() <= n
    nnz = float(nnz_row.sum(axis=1).max())
    act = ((p.dGP != 0) | (p.dGM != 0)).any(axis=0).reshape(NL, Ma)  # per lead, per mode
    colM = (p.dGM != 0).any(axis=0).reshape(NL, Ma).sum(axis=1)     # term1 columns per lead
    rowP = (p.dGP != 0).any(axis=0).reshape(NL, Ma).sum(axis=1)     # term2 rows per lead
    ML = int(act.sum())
    act_l = act.sum(axis=1)
    c = {
        "gemm_hsig": hmac * (nnz * n if sparse else n ** 3),
        "gemm_psi": hmac * M * (nnz if sparse else n * n)
        + 8.0 * sum(int(rowP[a]) * n * na[a] for a in range(NL)),
        "gemm_pi": 8.0 * sum(n * na[a] * Ma for a in range(NL)),
        "gemm_omgxi": 8.0 * sum(na[a] * (ML * Ma + (M - ML) * int(act_l[a])) for a in range(NL)),
        "gemm_omega": 8.0 * sum(M * int(colM[a]) * na[a] + int(rowP[a]) * M * na[a]
                                for a in range(NL)),
    }
    if tri:
        c["gemm_omega"] *= tri_fraction(p)
    return {k: NK * v for k, v in c.items()}, na, nc


def algorithmic_bytes(p, precision, tri=False):
    """HBM bytes per launch of the Ω kernel (average over the seven stages): stage s
    reads y0 and the s-1 earlier stage derivatives with a non-zero coefficient and
    writes k_s (the last stage writes the new state instead), over the Ω elements
    outside the passive x passive block (with Ω anti-Hermitian, tri, its independent
    entries); plus its two coupled-orbital operands.  Ωξ: one Ω slot outside the passive
    x passive block, ξ_c, and its four split-K partials of Ω ξ_c."""
    NK, n, NL, NX, NOT = p.dims
    M = p.M
    Ma = NX * NOT
    act = ((p.dGP != 0) | (p.dGM != 0)).any(axis=0)
    ML = int(act.sum())
    nc = int(np.count_nonzero(np.any(p.xi.reshape(-1, n) != 0, axis=0)))
    elem = 16 if precision == "fp64" else 8
    full = NK * (M * M - (M - ML) ** 2) * elem
    slot = full * (tri_fraction(p) if tri else 1.0)
    reads = [1, 2, 3, 4, 5, 6, 6]   # y0 + k_j with a_sj != 0 (a_72 = 0)
    writes = [1] * 7
    per = slot * (sum(reads) + sum(writes)) / 7.0
    return {"gemm_omega": per + 2 * NK * M * 2 * nc * elem,
            "gemm_omgxi": full + NK * M * nc * elem * (1 + 4)}


def measured_traffic(kernel, precision):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_c3.json")
    try:
        with open(path) as f:
            d = json.load(f)["per_launch"][precision][kernel]
        return d["fetch_bytes"] + d["write_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def run_leg(p, precision, K, W, eps, h0):
    from zandpack_amd.engine import Propagator
    prop = Propagator(p, precision, autotune=True)
    t, h = 0.0, h0

    def one():
        nonlocal t, h
        r = prop.attempt(t, h, eps)
        if not np.isfinite(r["te"]):
            raise FloatingPointError("non-finite error estimate")
        h_old, h = h, r["h_next"]
        if r["accepted"]:
            t += h  # reference bookkeeping (t += h_next)
        return r, h_old

    for _ in range(W):
        one()
    prop.timing(reset=True)
    acc = rej = 0
    taken0 = prop.fsal_info()["taken"]
    t_start = time.perf_counter()
    for _ in range(K):
        r, _ = one()
        acc += r["accepted"]
        rej += not r["accepted"]
    el = time.perf_counter() - t_start
    # attempts that took the previous accepted attempt's k_7 as their k_1 (first same as
    # last) evaluate six right-hand sides
    taken = prop.fsal_info()["taken"] - taken0
    dev = prop.timing()
    lay = prop.layout_info()
    out = dict(layout=lay, tiles=prop.tiles,
               omega_antihermitian_triangle=prop.antihermitian_info()["active"],
               steps_per_s=acc / el, attempts_per_s=K / el,
               rhs_evals_per_s=(7 * K - taken) / el, fsal_attempts=taken,
               reject_ratio=rej / K, ms_per_attempt=1e3 * el / K,
               device_ms_per_attempt=dev["attempt"][0], accepted=acc, rejected=rej,
               h_last=h, t_reached=t)
    prop.close()
    return out


def profile_leg(p, precision, attempts, h0):
    from zandpack_amd.engine import Propagator
    prop = Propagator(p, precision, autotune=True)
    # every stage of every attempt evaluated, so that each class mean is over full
    # launches (with first same as last the first stage's launches of an attempt after an
    # accepted one exit at once)
    prop.set_fsal("off")
    prop.set_profile(True)
    prop.attempt(0.0, h0, 1.0, force_accept=True)
    prop.timing(reset=True)
    t = 0.0
    for _ in range(attempts):
        prop.attempt(t, h0, 1.0, force_accept=True)
        t += h0
    tm = prop.timing()
    tri = prop.antihermitian_info()["active"]
    prop.close()
    return {k: v[0] for k, v in tm.items() if v[1] > 0}, tri


def replay_error(p, eps, h0, n_steps):
    """Mixed precision forced onto the FP64 run's accepted (t, h) sequence.  J error at
    every accepted step, max over leads and spins of |J_mixed - J_fp64| / max_t |J_fp64|,
    and σ's relative Frobenius error every 50 steps and at the end."""
    from zandpack_amd.engine import Propagator
    every = 50
    a = Propagator(p, "fp64", autotune=True)
    seq, J64, s64 = [], [], []
    t, h = 0.0, h0
    while len(seq) < n_steps:
        r = a.attempt(t, h, eps)
        h_old, h = h, r["h_next"]
        if r["accepted"]:
            seq.append((t, h_old))
            J64.append(r["J"])
            t += h
            if len(seq) % every == 0 or len(seq) == n_steps:
                s64.append(a.sigma())
    a.close()
    b = Propagator(p, "mixed", autotune=True)
    Jm, sm = [], []
    for i, (ts, hs) in enumerate(seq):
        Jm.append(b.attempt(ts, hs, 1.0, force_accept=True)["J"])
        if (i + 1) % every == 0 or i + 1 == n_steps:
            sm.append(b.sigma())
    b.close()
    J64, Jm = np.array(J64), np.array(Jm)
    scale = max(np.abs(J64).max(), 1e-300)
    per_step = np.abs(Jm - J64).max(axis=(1, 2)) / scale
    serr = [float(np.linalg.norm(x - y) / np.linalg.norm(y)) for x, y in zip(sm, s64)]
    return dict(steps=n_steps, t_reached_fs=float(t), J_rel_err_max=float(per_step.max()),
                J_rel_err_every_50=[float(x) for x in per_step[every - 1::every]],
                sigma_rel_err_every_50=serr[:-1] if n_steps % every == 0 else serr,
                sigma_rel_err_final=serr[-1], J_max=float(scale))


def freerun_h_diff(p, eps, h0, n_steps):
    """Free-running mode: both precisions with their own step control for n_steps
    accepted steps; the relative difference of the accepted-h sequences."""
    from zandpack_amd.engine import Propagator
    seqs = {}
    for prec in ("fp64", "mixed"):
        a = Propagator(p, prec, autotune=True)
        t, h, hs, rej = 0.0, h0, [], 0
        while len(hs) < n_steps:
            r = a.attempt(t, h, eps)
            h_old, h = h, r["h_next"]
            if r["accepted"]:
                hs.append(h_old)
                t += h
            else:
                rej += 1
        a.close()
        seqs[prec] = (np.array(hs), rej)
    h64, hm = seqs["fp64"][0], seqs["mixed"][0]
    return dict(steps=n_steps, h_rel_diff=float(np.abs(hm - h64).max() / np.abs(h64).max()),
                rejected_fp64=seqs["fp64"][1], rejected_mixed=seqs["mixed"][1])


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, attempts=3):
    """The CPU oracle (numpy/BLAS, zand-style orbital pruning in every right-hand side) on
    ``attempts`` fixed-h DOP54 attempts of the same workload with all BLAS threads, and
    one attempt with one thread: seconds per attempt (= per accepted step) and per RHS."""
    from oracle import zand_oracle as O
    try:
        from threadpoolctl import threadpool_info, threadpool_limits
        threads = max([d.get("num_threads", 1) for d in threadpool_info()] or [1])
    except Exception:
        threadpool_limits = None
        threads = os.cpu_count() or 1
    rng = np.random.default_rng(0)
    sig, psi, omg = p.zero_state()
    psi = 1e-3 * (rng.normal(size=psi.shape) + 1j * rng.normal(size=psi.shape))
    y = (sig, psi, omg)
    NL = p.dims[2]

    dH = O.hubbard_dH(p.hubbard_U, p.hubbard_n0)

    def attempt(y, t):
        return O.dop54_attempt(p.consts(), y, t, 0.002, p.H0, dH, p.bias, NL, prune=True)[3]

    t0 = time.perf_counter()
    for i in range(attempts):
        y = attempt(y, 0.002 * i)
    el = (time.perf_counter() - t0) / attempts
    el1 = None
    if threadpool_limits is not None:
        with threadpool_limits(limits=1):
            t0 = time.perf_counter()
            attempt(y, 0.002 * attempts)
            el1 = time.perf_counter() - t0
    return {"value": 1.0 / el, "unit": "steps/s", "cores": int(threads), "kind": "port",
            "sample": f"{attempts} fixed-h DOP54 attempts (7 RHS each) of the same n=1000, "
                      f"M=4000, NK=2 workload, {el:.2f} s per attempt",
            "s_per_attempt": el, "s_per_rhs": el / 7, "s_per_attempt_1_thread": el1,
            "steps_per_s_1_thread": (1.0 / el1) if el1 else None,
            "cpu_model": _cpu_model(), "cpu_count": os.cpu_count()}


def scf_contour():
    """The SCF tool's default contour (102 points, kT = 0.02585 eV), kept as test data."""
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                "scf_default_contour.npy"))


def scf_leg(p, reps=3):
    """Equilibrium density (cmdtools/SCF equivalent) of the same workload: one SCF
    iteration's density from H(σ₀) — host eigenpairs of H plus the device contour sum
    over the 102-point default contour — and the once-per-run self-energy set-up."""
    from zandpack_amd import scf
    C = scf_contour()
    corb = scf.coupled_orbitals(p)
    t0 = time.perf_counter()
    sig = scf.self_energy_cc(p, C[:, 0] + 1j * C[:, 1], corb)
    t_setup = time.perf_counter() - t0
    H = np.asarray(p.H0, complex) + p.dH(0.0, p.sigma0)
    scf.contour_density(H, C, corb, sig)  # warm-up
    t_eigh, t_all = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        np.linalg.eigh(H)
        t_eigh.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        rho = scf.contour_density(H, C, corb, sig)
        t_all.append(time.perf_counter() - t0)
    tr = np.trace(rho, axis1=1, axis2=2).real
    out = {"contour_points": int(C.shape[0]), "n_c": int(corb.size),
           "s_per_iteration": float(np.median(t_all)),
           "host_eigh_s": float(np.median(t_eigh)),
           "device_s": float(np.median(t_all) - np.median(t_eigh)),
           "self_energy_setup_s": t_setup, "charge_per_k": [float(x) for x in tr]}
    # under the workload's bias (cmdtools/SCF Nonequilibrium=True): the bias-window
    # term on fixed Gauss-Legendre nodes, kT = 0.025 eV for both leads
    NL = p.dims[2]
    bias = np.array([p.bias(1.0, a) for a in range(NL)])
    if bias.max() - bias.min() > 1e-6:
        kTs = np.full(NL, 0.025)
        nodes, w = scf.bias_window(bias, kTs)
        t0 = time.perf_counter()
        sig_l = scf.self_energy_cc_leads(p, nodes.astype(complex), corb, bias)
        gam = scf.window_weights(sig_l, nodes, bias, kTs)
        sig_w = sig_l.sum(axis=0)
        t_wsetup = time.perf_counter() - t0
        scf.window_density(H, nodes, w, corb, sig_w, gam)  # warm-up
        t_w = []
        for _ in range(reps):
            t0 = time.perf_counter()
            rw = scf.window_density(H, nodes, w, corb, sig_w, gam)
            t_w.append(time.perf_counter() - t0)
        out["window"] = {"nodes": int(nodes.size), "bias_eV": [float(b) for b in bias],
                         "s_per_iteration": float(np.median(t_w)),
                         "setup_s": t_wsetup,
                         "charge_per_k": [float(x) for x in np.trace(rw, axis1=1, axis2=2).real]}
    return out


def scf_cpu_sample(p, points=4):
    """The reference form of one SCF iteration (one n x n inversion per contour point,
    cmdtools/SCF:209-215) on the host, timed over ``points`` contour points and
    scaled to the full contour."""
    from oracle import zand_oracle as O
    from zandpack_amd import scf
    C = scf_contour()
    H = np.asarray(p.H0, complex) + p.dH(0.0, p.sigma0)
    sub = C[:: max(1, C.shape[0] // points)][:points]
    # the tool evaluates the self-energies once before its loop (SCF:181-189)
    corb = scf.coupled_orbitals(p)
    Ec = sub[:, 0] + 1j * sub[:, 1]
    scc = scf.self_energy_cc(p, Ec, corb)
    full = {}
    for i, E in enumerate(Ec):
        S = np.zeros(H.shape, complex)
        S[:, corb[:, None], corb[None, :]] = scc[:, i]
        full[complex(E)] = S
    t0 = time.perf_counter()
    O.contour_density(H, sub, lambda E: full[complex(E)])
    el = time.perf_counter() - t0
    return {"s_per_iteration": el * C.shape[0] / len(sub),
            "sample": f"{len(sub)} of {C.shape[0]} contour points ({el:.2f} s), scaled"}


def replay_fp64_error(req):
    """FP64 run of one served request forced onto its accepted (t, h) sequence:
    max |J_mixed - J_fp64| / max |J_fp64|."""
    from zandpack_amd.engine import Propagator
    prop = Propagator(req.problem, "fp64")
    J = [prop.attempt(t, h, 1.0, force_accept=True)["J"] for t, h in zip(req.times[:-1], req.h)]
    prop.close()
    J, Jm = np.array(J), np.array(req.J)
    return float(np.abs(Jm - J).max() / max(np.abs(J).max(), 1e-300))


def c2_leg(t_end=100.0, time_cap=40.0, warmup=3):
    """Config 2, the FP64 drop-in (n = 400, M = 1600, Gaussian pulse) with adaptive
    steps under rtol = 1e-10 (the scaled norm also needs an atol: 1e-12), from the
    zero state, to t_end = 100 fs or ``time_cap`` seconds of stepping: accepted
    steps/s after ``warmup`` steps and the time reached."""
    from zandpack_amd import builder
    from zandpack_amd.engine import Propagator
    rtol, atol = 1e-10, 1e-12
    p = builder.config_c2()
    prop = Propagator(p, "fp64", autotune=True)
    prop.set_tolerance(rtol, atol)
    t, h = 0.0, 0.005
    acc = rej = 0
    t_start = None
    while t < t_end:
        if acc == warmup and t_start is None:
            t_start = time.perf_counter()
        if t_start is not None and time.perf_counter() - t_start > time_cap:
            break
        r = prop.attempt(t, h, 1.0)
        if r["accepted"]:
            t += r["h_next"]
            acc += 1
        else:
            rej += 1
        h = r["h_next"]
    dt = time.perf_counter() - t_start
    steps = acc - warmup
    prop.close()
    NK, n, NL, NX, NOT = p.dims
    return {"workload": "honeycomb 20x20 n=400, 2 leads x 8 ch, N_L=40 + 60 Fermi slots, "
                        "M=1600, Gaussian pulse 0.5 eV", "precision": "fp64", "rtol": rtol,
            "atol": atol, "steps": steps, "steps_per_s": steps / dt, "t_reached_fs": t,
            "t_end_fs": t_end,
            "rejected": rej, "M": NL * NX * NOT, "n": n}


def steady_leg(p, tol=1e-10, maxiter=200):
    """Steady state (cmdtools/psinought equivalent) of the same workload in FP64:
    wall time including the host eigendecomposition of H(σ₀), iterations, relative
    residual, and the fixed-point check max|dΨ| / max|dΨ(Ψ=0, Ω=0)| from the device
    RHS (the denominator is the σ-driven source, b/ħ of the linear system)."""
    from zandpack_amd.engine import Propagator
    prop = Propagator(p, "fp64")
    t0 = time.perf_counter()
    its, res = prop.steady_state(0.0, tol=tol, maxiter=maxiter)
    el = time.perf_counter() - t0
    s, ps, om = prop.get_state()
    Ht = p.H0 + p.dH(0.0, s)
    delta = np.array([p.bias(0.0, a) for a in range(p.dims[2])])
    _, dp, _, _ = prop.rhs(Ht, delta, s, ps, om, want_omega=False)
    _, d0, _, _ = prop.rhs(Ht, delta, s, np.zeros_like(ps), np.zeros_like(om), want_omega=False)
    scale = np.abs(d0).max()
    prop.close()
    return {"precision": "fp64", "seconds": el, "iterations": its, "rel_residual": res,
            "tol": tol, "max_dpsi_rel": float(np.abs(dp).max() / scale)}


def batched_leg(n_inst, steps, seed=0, workers=1):
    """Config 4: n_inst uneven propagations (n in {100..800}, M = 4n) advanced
    together, each for ``steps`` accepted steps; aggregate steps/s.  Requests of one n
    share a shape and run as the instances of one problem (one launch per kernel class
    for the group); the same workload one problem per request beside it.  The figures
    with admission include the set-up and upload of every instance (on worker threads,
    beside the stepping); the stepping figures admit everything first."""
    from zandpack_amd import builder
    from zandpack_amd.serving import Request, Server
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.arange(100, 801, 100), size=n_inst)

    def reqs():
        return [Request(i, builder.config_c4_instance(int(n), seed=seed * 1000 + i),
                        t_end=60.0, eps=1e-6, h0=0.005, precision="mixed", max_steps=steps)
                for i, n in enumerate(ns)]

    out = {"instances": n_inst, "n_values": sorted(int(x) for x in ns), "steps_each": steps,
           "precision": "mixed"}
    for name, g in (("grouped", 16), ("one_problem_each", 1)):
        srv = Server(hbm_bytes=200 << 30, max_live=n_inst, max_group=g, admit_workers=workers)
        st = srv.run(reqs(), realtime=False)
        sp = Server(hbm_bytes=200 << 30, max_live=n_inst, max_group=g,
                    admit_workers=workers).run(reqs(), realtime=False, preadmit=True)
        out[name] = {"steps_per_s": st["steps_per_s"], "wall_s": st["wall_s"],
                     "failed": st["failed"] + sp["failed"], "admit_s_on_workers": st["admit_s"],
                     "stepping_steps_per_s": sp["steps_per_s"], "stepping_wall_s": sp["wall_s"],
                     "groups": srv.groups}
    out["steps_per_s"] = out["grouped"]["steps_per_s"]
    out["stepping_steps_per_s"] = out["grouped"]["stepping_steps_per_s"]
    return out


def serving_leg(n_req, rate, t_scale, seed=0):
    """Config 5: Poisson arrivals at ``rate``/s; n ~ U[50,1200], M ~ U[2n,5n],
    step/Gaussian/sine bias, t_end ~ U[20,200] fs x t_scale; mixed precision,
    continuous batching.  p50/p99 time-to-completion, aggregate steps/s, and the
    FP64 current error of a 5% sample."""
    from zandpack_amd import builder
    from zandpack_amd.serving import Request, Server
    rng = np.random.default_rng(seed)
    reqs, t = [], 0.0
    for i in range(n_req):
        t += rng.exponential(1.0 / rate)
        n = int(rng.integers(50, 1201))
        M = int(rng.integers(2 * n, 5 * n + 1))
        p = builder.config_sized(n, M, seed=seed * 1000 + i, pulse=builder.random_pulse(rng))
        reqs.append(Request(i, p, t_end=float(rng.uniform(20, 200)) * t_scale, eps=1e-6,
                            h0=0.005, arrival=t, precision="mixed"))
    st = Server(hbm_bytes=200 << 30, max_live=64).run(reqs, realtime=True)
    k = max(1, n_req // 20)
    sample = [r for r in st["done"] if r.error is None][:k]
    errs = [replay_fp64_error(r) for r in sample]
    return {"requests": n_req, "rate_per_s": rate, "t_end_scale": t_scale,
            "p50_s": st["p50_s"], "p99_s": st["p99_s"], "steps_per_s": st["steps_per_s"],
            "wall_s": st["wall_s"], "failed": st["failed"], "steps": st["steps"],
            "admit_s_on_worker": st["admit_s"], "sampled_J_rel_err_vs_fp64": errs,
            "t_end_fs": "U[20, 200] x t_end_scale"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--eps", type=float, default=1e-6)
    ap.add_argument("--h0", type=float, default=0.002)
    ap.add_argument("--replay-steps", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--batched-instances", type=int, default=64)
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--serving-requests", type=int, default=64)
    ap.add_argument("--serving-rate", type=float, default=20.0)
    ap.add_argument("--serving-t-scale", type=float, default=0.01)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "mode sharding is a design, not a build", file=sys.stderr)
        sys.exit(2)

    from zandpack_amd import builder
    p = builder.config_c3(seed=0)
    NK, n, NL, NX, NOT = p.dims
    print(f"# workload n={n} M={p.M} NK={NK}", file=sys.stderr, flush=True)

    legs = {}
    for prec in ("fp64", "mixed"):
        print(f"# timing {prec}", file=sys.stderr, flush=True)
        legs[prec] = run_leg(p, prec, args.steps, args.warmup, args.eps, args.h0)
    print("# replay error", file=sys.stderr, flush=True)
    replay = replay_error(p, args.eps, args.h0, args.replay_steps)
    free = freerun_h_diff(p, args.eps, args.h0, 20)

    roof = {}
    for prec in ("fp64", "mixed"):
        print(f"# kernel profile {prec}", file=sys.stderr, flush=True)
        tm, tri = profile_leg(p, prec, 3, args.h0)
        flops, na, nc = algorithmic_flops(p, tri)
        gemms = {k: v for k, v in tm.items() if k in flops}
        dom = max(gemms, key=lambda k: gemms[k])
        # kernel-class time is summed over NK-batched launches of one stage; the bound
        # is the roof (matrix-core flops or HBM bytes) the kernel's work takes longest on
        sec = gemms[dom] * 1e-3
        peak_f = PEAK[prec] if dom not in ("gemm_hsig", "gemm_pi") else PEAK["fp64"]
        nbytes = algorithmic_bytes(p, prec, tri).get(dom)
        t_f = flops[dom] / (peak_f * 1e12)
        t_b = nbytes / (HBM_PEAK * 1e9) if nbytes else 0.0
        if t_b > t_f:
            ach, peak, unit, bound = nbytes / sec / 1e9, HBM_PEAK, "GB/s", "hbm"
        else:
            ach, peak, unit, bound = flops[dom] / sec / 1e12, peak_f, "TFLOP/s", "mfma"
        roof[prec] = {"bound": bound, "kernel": dom, "achieved": ach, "peak": peak,
                      "unit": unit, "frac": ach / peak,
                      "traffic": measured_traffic(dom, prec),
                      "algorithmic_bytes": nbytes, "algorithmic_flops": flops[dom],
                      "mfma_tflops": flops[dom] / sec / 1e12,
                      "omega_antihermitian_triangle": tri,
                      "avg_ms": gemms[dom], "kernel_ms": tm}

    value = legs["mixed"]["steps_per_s"]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 / value if value else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "mixed: c64 Psi/Omega on f32 MFMA, c128 sigma/Hsigma/Pi/error/controller",
        "data": "synthetic",
        "config": {"workload": "honeycomb 40x25 n=1000, NK=2 spin, Hubbard U=2 eV, 2 leads x "
                               "20 ch, N_L=50 + 50 Fermi slots, M=4000 per k, step bias 0.3 eV",
                   "n": n, "M": p.M, "NK": NK, "n_c": nc, "n_alpha": na, "eps": args.eps,
                   "h0": args.h0},
        "fp64": legs["fp64"],
        "mixed": legs["mixed"],
        "speedup_mixed_vs_fp64": value / legs["fp64"]["steps_per_s"],
        "replay": dict(replay, free_running=free),
        "roofline": {k: v for k, v in roof["mixed"].items() if k != "kernel_ms"},
        "roofline_fp64": {k: v for k, v in roof["fp64"].items() if k != "kernel_ms"},
        "kernel_ms": {"fp64": roof["fp64"]["kernel_ms"], "mixed": roof["mixed"]["kernel_ms"]},
    }
    if not args.no_serving:
        print("# batched (config 4)", file=sys.stderr, flush=True)
        print("# steady state", file=sys.stderr, flush=True)
        line["steady_state"] = steady_leg(p)
        print("# config 2 (FP64, rtol 1e-10)", file=sys.stderr, flush=True)
        line["config2"] = c2_leg()
        print("# equilibrium density", file=sys.stderr, flush=True)
        line["scf"] = scf_leg(p)
        line["batched"] = batched_leg(args.batched_instances, args.batched_steps)
        print("# serving (config 5)", file=sys.stderr, flush=True)
        line["serving"] = serving_leg(args.serving_requests, args.serving_rate,
                                      args.serving_t_scale)
    if not args.no_cpu:
        print("# cpu baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(p)
        if "scf" in line:
            line["cpu_baseline"]["scf"] = scf_cpu_sample(p)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
