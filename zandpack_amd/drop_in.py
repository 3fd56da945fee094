"""The reference's in-process call surface, backed by the GPU engine.

* ``make_f(propagator)`` returns ``f(t, old_sig, old_psi, old_omega, dH,
  delta_variant, dH_given=True) -> (D_sig, D_psi, D_omega)``, the signature of
  ``TD_Transport.make_f`` / ``make_f_purenp`` (Zandpack/TimedependentTransport.py:2767-2777,
  2605-2660): H = H0 + dH(t, σ) when ``dH_given`` else dH(t, σ); Δ_α = delta_variant(t, α).
  The derivatives follow the zand formulation (Zandpack/mpi/zand:340-387, 746-809).
* ``AdaptiveDOP(f, sig0, psi0, omega0, eps, t0, t1, dH, delta_func, Ixi, ...)``
  mirrors the driver signature of Zandpack/TimedependentTransport.py:3392 with
  the step semantics of the production propagator (Zandpack/mpi/zand:297-431):
  seven stages (k7 from stage 7, not the ``k7 = h*k6`` slip at :3537), absolute
  error norm over σ, Ψ and Ω, h <- 0.9 h (eps/TE)^(1/5) after every attempt,
  currents from the stage-0 state, and ``t += h_next`` on acceptance unless
  ``time_update='correct'``.

* ``AdaptiveRK4(...)`` / ``AdaptiveRK4_Opti(...)``: the in-process Runge-Kutta-
  Fehlberg 4(5) drivers (Zandpack/TimedependentTransport.py:3179-3390,
  Zandpack/Optimized_RK45.py:19-228) with their own semantics: six stages, step with
  the fifth-order weights, accept iff TE <= eps, h <- 0.9 h (eps/TE)^(1/5) after every
  attempt, ``while t0 <= t1`` and t += the new h, h_guess default (t1 - t0)/1000.

* ``scipy_ode(f, sig0, psi0, omg0, t0, t_eval, dH, delta_variant, Ixi, ...)``: the
  reference's hand-off of ``f`` to ``scipy.integrate.solve_ivp``
  (Zandpack/TimedependentTransport.py:3591-3673): σ, Ψ, Ω flattened into one complex
  vector, one solve_ivp call per output interval; returns (t_eval, σ[k=0] at each
  output time, the first lead's current (NK,) at each output time).

The three adaptive drivers write the reference's run files into the working
directory (``<name>.txt``, ``<name>_last_{sig,psi,omega,time}.npy``, ``_times.npy``,
the current files and ``_#electrons_device.npy``) every ``print_step`` accepted steps,
resume from ``<name>_last_*`` when the state is None, and take the reference's
``atol``/``rtol`` error form; ``print_to_file=False`` writes nothing.

All run every evaluation on the GPU (host arrays in and out per call); the
device-resident path with no per-stage transfers is ``Propagator.propagate``.
"""
from __future__ import annotations

import time

import numpy as np

from .engine import Propagator

_TAB_C = np.array([0.0, 1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0])
_TAB_A = [
    [],
    [1 / 5],
    [3 / 40, 9 / 40],
    [44 / 45, -56 / 15, 32 / 9],
    [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
    [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
    [35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84],
]
_B5 = np.array([35 / 384, 0.0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0.0])
_B4 = np.array([5179 / 57600, 0.0, 7571 / 16695, 393 / 640, -92097 / 339200, 187 / 2100, 1 / 40])
_CT = _B5 - _B4


class _RunFiles:
    """The run files the in-process drivers write into the working directory
    (Zandpack/TimedependentTransport.py:3279-3281, 3343-3375, 3384-3388; the same in
    Zandpack/Optimized_RK45.py): ``<name>.txt`` (start line, then every ``print_step``
    accepted steps the progress in %, the new step and the elapsed time, and
    ``write_func(file, t, σ, Ψ, Ω)`` for the step just taken; at the end ``100%`` and
    the runtime), and with each progress entry the snapshots ``<name>_last_sig``,
    ``_last_psi``, ``_last_omega``, ``_last_time`` (the start time of the last step),
    ``_times``, the currents, and ``_#electrons_device`` (Tr σ per step and k).
    ``enabled=False`` writes nothing (the reference writes regardless of its
    ``print_to_file`` argument)."""

    def __init__(self, enabled, name, print_step, write_func, t_start, t_end, current_files):
        self.on = bool(enabled)
        self.name = str(name)
        self.every = max(1, int(print_step))
        self.write_func = write_func
        self.T0, self.t1 = float(t_start), float(t_end)
        self.current_files = current_files  # [(file stem, data key)]
        self.step = 0
        self.t_start = time.time()
        if self.on:
            with open(self.name + ".txt", "w") as fh:
                fh.write("\n\n\n\n\nStart (Wait for compilation)\n")

    def accepted(self, t_new, h, state, times, data, dms):
        if self.on and self.step % self.every == 0:
            with open(self.name + ".txt", "a") as fh:
                fh.write(str(((t_new - self.T0) / (self.t1 - self.T0)) * 100) + " %\n")
                fh.write("current timestep: " + str(h) + "fs\n")
                fh.write("delta t: " + str(time.time() - self.t_start) + "seconds\n")
                if self.write_func is not None:
                    self.write_func(fh, t_new - h, state[0].copy(), state[1].copy(),
                                    state[2].copy())
            np.save(self.name + "_last_sig", state[0])
            np.save(self.name + "_last_psi", state[1])
            np.save(self.name + "_last_omega", state[2])
            np.save(self.name + "_last_time", np.array(times[-1]))
            np.save("_times", np.array(times))
            for stem, key in self.current_files:
                np.save(stem, np.asarray(data.get(key, [])))
            np.save("_#electrons_device", np.trace(np.asarray(dms), axis1=2, axis2=3))
        self.step += 1

    def finish(self):
        if self.on:
            with open(self.name + ".txt", "a") as fh:
                fh.write("100%\n")
                fh.write("Runtime: " + str(time.time() - self.t_start) + "seconds")


def _resume(name, sig0, psi0, omega0, t0):
    """A missing state is read from ``<name>_last_*.npy`` and a missing t0 from
    ``<name>_last_time.npy`` (Zandpack/TimedependentTransport.py:3258-3264)."""
    if sig0 is None or psi0 is None or omega0 is None:
        sig0 = np.load(name + "_last_sig.npy")
        psi0 = np.load(name + "_last_psi.npy")
        omega0 = np.load(name + "_last_omega.npy")
        if t0 is None:
            t0 = float(np.load(name + "_last_time.npy"))
    return sig0, psi0, omega0, t0


def _tolerance_error(err, state, atol, rtol):
    """The drivers' error with ``atol``/``rtol`` given
    (Zandpack/TimedependentTransport.py:3226-3235): f_x = max((|e_x| - atol_x)/rtol_x -
    |x|) over σ, Ψ, Ω and error = 1 ** max(f); this is 1.0 for every finite or NaN
    f, so a step is accepted iff eps >= 1 and h <- 0.9 h eps^(1/5)."""
    f = [float(np.max((np.abs(e) - atol[i]) / rtol[i] - np.abs(x)))
         for i, (e, x) in enumerate(zip(err, state))]
    return 1.0 ** max(f)


def make_f(prop: Propagator):
    """Reference-style RHS closure over the constants held by ``prop``."""
    H0 = prop.problem.H0
    NL = prop.NL

    def f(t, old_sig, old_psi, old_omega, dH, delta_variant, dH_given=True):
        Ht = H0 + dH(t, old_sig) if dH_given else dH(t, old_sig)
        delta = np.array([delta_variant(t, a) for a in range(NL)], dtype=np.float64)
        ds, dp, do, J = prop.rhs(Ht, delta, old_sig, old_psi, old_omega)
        f.last_J = J
        return ds, dp, do

    f.last_J = None
    f.propagator = prop
    return f


def AdaptiveDOP(f, sig0, psi0, omega0, eps, t0, t1, dH, delta_func, Ixi=None, h_guess=None,
                dH_given=True, print_to_file=True, fixed_mode=False, name="Runge-Kutta",
                write_func=None, print_step=50, atol=None, rtol=None,
                elec_names=("left", "right"), time_update="reference", max_attempts=10 ** 6,
                **_unused):
    """Adaptive DOP54 over an ``f`` made by ``make_f``.  Returns (times, data) with
    ``data['current_<elec>']`` (list of (NK,) arrays) and ``data['density matrix']``.
    Run files as the reference driver writes them (see ``_RunFiles``); its ``_JL`` /
    ``_JR`` hold the first two leads' currents here (the reference saves two lists it
    never fills).  ``plot`` and ``use_GPU`` are accepted and ignored."""
    sig0, psi0, omega0, t0 = _resume(name, sig0, psi0, omega0, t0)
    h = 0.01 if h_guess is None else float(h_guess)
    y = [np.array(sig0, complex), np.array(psi0, complex), np.array(omega0, complex)]
    data = {f"current_{e}": [] for e in elec_names}
    dms, times = [], []
    t = float(t0)
    cur = [(stem, f"current_{e}") for stem, e in zip(("_JL", "_JR"), elec_names)]
    files = _RunFiles(print_to_file, name, print_step, write_func, t0, t1, cur)
    attempts = 0
    while t < t1:
        while True:
            attempts += 1
            if attempts > max_attempts:
                raise RuntimeError("attempt limit reached")
            ks = []
            J0 = None
            for i in range(7):
                yi = [y[p] + sum(_TAB_A[i][j] * ks[j][p] for j in range(i)) for p in range(3)]
                d = f(t + _TAB_C[i] * h, yi[0], yi[1], yi[2], dH, delta_func, dH_given=dH_given)
                if i == 0:
                    J0 = f.last_J
                ks.append(tuple(h * v for v in d))
            err = [sum(_CT[j] * ks[j][p] for j in range(7)) for p in range(3)]
            if atol is None:
                te = np.sqrt(sum(float(np.sum(np.abs(e) ** 2)) for e in err))
            else:
                te = _tolerance_error(err, y, atol, rtol)
            if not np.isfinite(te):
                raise FloatingPointError("non-finite error estimate")
            h_old = h
            if fixed_mode:
                break
            h = 0.9 * h * (eps / te) ** 0.2
            if te < eps:
                break
        dms.append(y[0].copy())
        for a, e in enumerate(elec_names):
            data[f"current_{e}"].append(J0[a].copy())
        y = [y[p] + sum(_B5[j] * ks[j][p] for j in range(7)) for p in range(3)]
        times.append(t)
        t += h if (time_update == "reference" and not fixed_mode) else h_old
        files.accepted(t, h, y, times, data, dms)
    files.finish()
    data["density matrix"] = np.array(dms)
    data["state"] = tuple(y)
    return np.array(times), data


_RKF_C = np.array([0.0, 1 / 4, 3 / 8, 12 / 13, 1.0, 1 / 2])
_RKF_A = [
    [],
    [1 / 4],
    [3 / 32, 9 / 32],
    [1932 / 2197, -7200 / 2197, 7296 / 2197],
    [439 / 216, -8.0, 3680 / 513, -845 / 4104],
    [-8 / 27, 2.0, -3544 / 2565, 1859 / 4104, -11 / 40],
]
_RKF_CH = np.array([16 / 135, 0.0, 6656 / 12825, 28561 / 56430, -9 / 50, 2 / 55])
_RKF_CT = np.array([1 / 360, 0.0, -128 / 4275, -2197 / 75240, 1 / 50, 2 / 55])


def AdaptiveRK4(f, sig0, psi0, omega0, eps, t0, t1, dH, delta_func, Ixi=None, h_guess=None,
                dH_given=True, print_to_file=True, fixed_mode=False, name="Runge-Kutta",
                write_func=None, print_step=50, atol=None, rtol=None,
                elec_names=("left", "right"), max_attempts=10 ** 6, **_unused):
    """Adaptive RKF45 over an ``f`` made by ``make_f`` (the in-process AdaptiveRK4's
    loop).  Returns (times, data): the start time of every accepted step, and
    ``data['current_<elec>']`` (list of (NK,) arrays, from the state before the
    step), ``data['density matrix']`` (σ before each step), ``data['state']``.
    Run files as the reference driver writes them (see ``_RunFiles``), with the
    currents in ``_current_<elec>``; ``plot`` and ``use_GPU`` are accepted and
    ignored."""
    sig0, psi0, omega0, t0 = _resume(name, sig0, psi0, omega0, t0)
    h = (t1 - t0) / 1000 if h_guess is None else float(h_guess)
    y = [np.array(sig0, complex), np.array(psi0, complex), np.array(omega0, complex)]
    data = {f"current_{e}": [] for e in elec_names}
    dms, times = [], []
    t = float(t0)
    cur = [(f"_current_{e}", f"current_{e}") for e in elec_names]
    files = _RunFiles(print_to_file, name, print_step, write_func, t0, t1, cur)
    attempts = 0
    while t <= t1:
        while True:
            attempts += 1
            if attempts > max_attempts:
                raise RuntimeError("attempt limit reached")
            ks = []
            J0 = None
            for i in range(6):
                yi = [y[p] + sum(_RKF_A[i][j] * ks[j][p] for j in range(i)) for p in range(3)]
                d = f(t + _RKF_C[i] * h, yi[0], yi[1], yi[2], dH, delta_func, dH_given=dH_given)
                if i == 0:
                    J0 = f.last_J
                ks.append(tuple(h * v for v in d))
            if fixed_mode:
                break
            err = [sum(_RKF_CT[j] * ks[j][p] for j in range(6)) for p in range(3)]
            if atol is None:
                te = np.sqrt(sum(float(np.sum(np.abs(e) ** 2)) for e in err))
            else:
                te = _tolerance_error(err, y, atol, rtol)
            if not np.isfinite(te):
                raise FloatingPointError("non-finite error estimate")
            h = 0.9 * h * (eps / te) ** 0.2
            if te <= eps:
                break
        dms.append(y[0].copy())
        for a, e in enumerate(elec_names):
            data[f"current_{e}"].append(J0[a].copy())
        y = [y[p] + sum(_RKF_CH[j] * ks[j][p] for j in range(6)) for p in range(3)]
        times.append(t)
        t += h
        files.accepted(t, h, y, times, data, dms)
    files.finish()
    data["density matrix"] = np.array(dms)
    data["state"] = tuple(y)
    return np.array(times), data


# Optimized_RK45.AdaptiveRK4_Opti is the same method with vectorised stage sums.
AdaptiveRK4_Opti = AdaptiveRK4


def scipy_ode(f, sig0, psi0, omg0, t0, t_eval, dH, delta_variant, Ixi=None, dH_given=True,
              method="RK45", dt_guess=None, atol=1e-6, rtol=1e-4):
    """solve_ivp over an ``f`` made by ``make_f`` (TimedependentTransport.py:3625-3673).

    One ``solve_ivp`` call per interval (t_prev, t_eval[i]) from the state reached at
    the previous output time; after each, σ[k=0] and the first lead's k-resolved
    current are recorded.  The current is the one the GPU forms with the problem's
    Ixi (the reference recomputes it on the host from the ``Ixi`` argument, which must
    be the same array; it is accepted here for the call signature).  ``method`` is
    passed to solve_ivp (the reference accepts it but always runs solve_ivp's default
    RK45); ``dt_guess`` is accepted and unused, as in the reference."""
    from scipy.integrate import solve_ivp
    shapes = [np.shape(sig0), np.shape(psi0), np.shape(omg0)]
    sizes = [int(np.prod(s)) for s in shapes]

    def split(y):
        out, o = [], 0
        for sh, sz in zip(shapes, sizes):
            out.append(y[o:o + sz].reshape(sh))
            o += sz
        return out

    def F(t, y):
        s, p, o = split(y)
        d = f(t, s, p, o, dH, delta_variant, dH_given=dH_given)
        return np.concatenate([np.ravel(v) for v in d])

    t_eval = np.asarray(t_eval, float)
    assert (t0 < t_eval).all()
    y = np.concatenate([np.ravel(np.asarray(v, complex)) for v in (sig0, psi0, omg0)])
    rho, cl = [], []
    t_prev = float(t0)
    for te in t_eval:
        sol = solve_ivp(F, (t_prev, te), y, method=method, t_eval=[te], rtol=rtol, atol=atol)
        y = sol.y[:, -1]
        s, p, o = split(y)
        rho.append(s[0].copy())
        f(te, s, p, o, dH, delta_variant, dH_given=dH_given)  # currents of the new state
        cl.append(np.array(f.last_J[0]))
        t_prev = float(te)
    return t_eval, np.array(rho), np.array(cl)
