"""Zand run directories on the GPU engine.

Reads the ``<name>/Arrays/`` inputs of ``Zandpack/mpi/zand`` and writes its
``<name>_save/`` outputs with the same file names, dtypes and bookkeeping.  A
directory prepared for ``mpirun -np P zand Dir=<dir>`` runs here with its
``Arrays/`` and the settings of its ``Initial.py`` (read from the syntax tree, not
imported: ``read_initial``); its ``Bias.py`` (arbitrary Python importing numba and
Zandpack) is not run, and the device models it would describe are given as keywords
instead.  The reference's post-processing (``mpi_tools.combine_currents``,
``td_info``) reads the results:

    python -m zandpack_amd.rundir Dir=<dir> name=<name> t1=50 eps=1e-8 stepsize=0.01 \\
        [t0=0 saveevery=100 n_dm_compress=1 usesave=False checkpoints=10,20 \\
         precision=fp64 hbar=0.6582119569 pulse=step amp=0.1 p1=0 p2=0 p3=0 p4=0 \\
         weights=0.5,-0.5 [pulse=air_photonics|toptica pulse_table=<ExpPulses .npy>] \\
         hubbard_U=0 hubbard_n0=<file.npy> steady_state=False nonorthogonal=False save_PI=False]

Keywords follow ``Initial.py`` (Zandpack/mpi/zand:57-71).  ``nonorthogonal=True``
runs the directory as ``mpi/nozand`` does: ``Arrays/S^(-0.5).npy`` carries the
problem to the non-orthogonal basis (``builder.to_nonorthogonal``, with
``dot_prod_zero_tol`` from the environment as nozand reads it, default 1e-10).  The ``Bias.py`` hooks
(arbitrary Python) are replaced by the device models (``pulse*``, ``hubbard_*``);
a run that needs an arbitrary ``dH``/``bias`` uses the in-process drivers of
``zandpack_amd.drop_in`` instead.
"""
from __future__ import annotations

import os
import re
import sys

import numpy as np

from .builder import (AMEProblem, EXPERIMENTAL_PULSES, HBAR, Pulse, experimental_pulse,
                      to_nonorthogonal)

ARRAY_KEYS = ("H_Ortho", "DM_Ortho", "psi_shape", "GG_P", "GL_P", "GG_M", "GL_M", "Xpp", "Xpm",
              "xi", "Ixi", "zero_tol")


def flexload(path):
    """Array at ``path`` (.npy, an .npz's first array, or the .npz twin of a missing
    .npy), as Zandpack/Loader.py:12-32 resolves it.  Loads never unpickle."""
    for p in (path, path.replace(".npy", ".npz")):
        if not os.path.exists(p):
            continue
        f = np.load(p, allow_pickle=False)
        if isinstance(f, np.ndarray):
            return f
        with f:
            return f[f.files[0]]
    raise FileNotFoundError(path)


def not_per_lead(GG_P, GL_P, zero_tol):
    """Channels kept per lead: one past the last channel index whose GG_P or GL_P
    exceeds zero_tol for some (k, pole) (Zandpack/mpi_splitter.py:555-580)."""
    nk, nl, nx, nc = GG_P.shape
    live = (np.abs(GG_P).max(axis=0) >= zero_tol) | (np.abs(GL_P).max(axis=0) >= zero_tol)
    out = []
    for a in range(nl):
        idx = np.nonzero(live[a].any(axis=0))[0]
        out.append(int(idx.max()) + 1 if idx.size else 0)
    return np.array(out)


def load_arrays(arrays_dir, hbar=HBAR, name="problem", nonorthogonal=False):
    """AMEProblem from a zand ``Arrays/`` directory, channels cut to NOT = max over
    leads of ``not_per_lead`` (Zandpack/mpi/zand:113-124, 218, 568-576).  With
    ``nonorthogonal`` the problem nozand propagates (Zandpack/mpi/nozand:259-311):
    the arrays carried to the basis of ``S^(-0.5).npy``."""
    d = {k: flexload(os.path.join(arrays_dir, k + ".npy")) for k in ARRAY_KEYS}
    zt = float(np.asarray(d["zero_tol"]))
    NOT = int(not_per_lead(d["GG_P"], d["GL_P"], zt).max())
    cut = lambda a: np.ascontiguousarray(np.asarray(a)[:, :, :, :NOT], dtype=complex)
    p = AMEProblem(H0=np.asarray(d["H_Ortho"], complex), sigma0=np.asarray(d["DM_Ortho"], complex),
                   xi=cut(d["xi"]), Ixi=cut(d["Ixi"]), GG_P=cut(d["GG_P"]), GL_P=cut(d["GL_P"]),
                   GG_M=cut(d["GG_M"]), GL_M=cut(d["GL_M"]), Xpp=cut(d["Xpp"]),
                   Xpm=cut(d["Xpm"]), hbar=hbar, name=name)
    if not nonorthogonal:
        return p
    Sm05 = flexload(os.path.join(arrays_dir, "S^(-0.5).npy"))
    q = to_nonorthogonal(p, Sm05, float(os.environ.get("dot_prod_zero_tol", 1e-10)))
    q.name = name
    q.Sm05 = np.asarray(Sm05, complex)
    return q


def write_arrays(problem, arrays_dir, zero_tol=1e-12, Sm05=None):
    """The inputs zand reads, under the names Zandpack/Writer.py:61-99 writes
    (xi and Ixi compressed, as ``arr_0``), for a problem built here; ``Sm05`` adds
    the ``S^(-0.5)`` array nozand reads (Writer.py:83)."""
    os.makedirs(arrays_dir, exist_ok=True)
    j = lambda k: os.path.join(arrays_dir, k)
    NK, n, NL, NX, NOT = problem.dims
    np.save(j("H_Ortho"), problem.H0)
    np.save(j("DM_Ortho"), problem.sigma0)
    np.save(j("psi_shape"), np.array([NK, NL, NX, NOT, n]))
    np.save(j("omg_shape"), np.array([NK, NL, NX, NOT, NL, NX, NOT]))
    for k in ("Xpp", "Xpm", "GG_P", "GL_P", "GG_M", "GL_M"):
        np.save(j(k), getattr(problem, k))
    np.save(j("diff_GGP_GLP"), problem.dGP)
    np.save(j("diff_GGM_GLM"), problem.dGM)
    np.save(j("num_leads"), NL)
    np.save(j("noT"), NOT)
    np.save(j("zero_tol"), zero_tol)
    np.savez_compressed(j("xi"), problem.xi)
    np.savez_compressed(j("Ixi"), problem.Ixi)
    if Sm05 is not None:
        np.save(j("S^(-0.5)"), np.asarray(Sm05, complex))


class SaveWriter:
    """``<name>_save/`` as zand writes it (Zandpack/mpi/zand:256-265, 412-460):
    every ``saveevery`` accepted steps DM<step> (complex64, every n_dm_compress-th),
    DMt<step>, times<step>, current_<a>_<step> ((steps, NK) per lead); checkpoints
    last_sig, last_time, last_stepsize and the full last_psi / last_omg that
    ``LoadFromFull`` resumes from (:535-542).  ``S05`` (non-orthogonal runs) writes
    last_psi in the orthogonal basis, S^½ Ψ, which nozand's full-state loader
    carries back with S^-½ (Zandpack/mpi/nozand:720-723); σ stays in its own basis."""

    def __init__(self, save_dir, NL, saveevery=100, n_dm_compress=1, t0=0.0, sig0=None,
                 S05=None, save_pi=False):
        self.S05 = S05
        self.save_pi = bool(save_pi)
        self.PI = []
        self.dir = save_dir
        self.NL = NL
        self.saveevery = int(saveevery)
        self.n = int(n_dm_compress)
        os.makedirs(save_dir, exist_ok=True)
        for f in os.listdir(save_dir):  # a new run drops earlier DM files (:262-265)
            if f.startswith("DM"):
                os.remove(os.path.join(save_dir, f))
        self.steps = 0
        self.currents = [[] for _ in range(NL)]
        self.DM = [] if sig0 is None else [np.array(sig0)]
        self.times = [float(t0)]

    def _p(self, f):
        return os.path.join(self.dir, f)

    def step(self, t, sig, J, pi=None):
        """One accepted step: the new time, the new σ, the stage-0 currents J (NL, NK)
        and, with save_pi, the stage-0 per-lead Π (NK, NL, n, n)."""
        if self.save_pi:
            self.PI.append(np.array(pi))
        for a in range(self.NL):
            self.currents[a].append(np.array(J[a], float))
        self.steps += 1
        self.DM.append(np.array(sig))
        self.times.append(float(t))
        if self.steps % self.saveevery == 0 and self.steps >= self.saveevery:
            self.flush()

    def flush(self):
        s, L = self.steps, len(self.DM)
        np.save(self._p(f"DM{s}"), np.array(self.DM[::self.n]).astype(np.complex64))
        dm_t = np.array(self.times[-L:])
        if len(dm_t) > L:
            dm_t = dm_t[1:]
        np.save(self._p(f"DMt{s}"), dm_t[::self.n])
        if self.save_pi:  # PIa<step> (Zandpack/mpi/zand:444-448)
            np.save(self._p(f"PIa{s}"), np.array(self.PI[::self.n]).astype(np.complex64))
            self.PI = []
        np.save(self._p(f"times{s}"), self.times[-L:])
        for a, c in enumerate(self.currents):
            np.save(self._p(f"current_{a}_{s}"), np.array(c[-L:]))
        self.DM = []

    def checkpoint(self, t, h, sig, psi, omg):
        np.save(self._p("last_sig"), sig)
        np.save(self._p("last_time"), float(t))
        np.save(self._p("last_stepsize"), float(h))
        if self.S05 is not None:
            psi = np.einsum("kij,kaxcj->kaxci", self.S05, psi)
        np.save(self._p("last_psi"), psi)
        np.save(self._p("last_omg"), omg)


def load_checkpoint(save_dir, Sm05=None):
    """(t, h, σ, Ψ, Ω) written by ``SaveWriter.checkpoint`` (or zand's full-state
    files); with ``Sm05`` Ψ is carried to the non-orthogonal basis as nozand does."""
    p = lambda f: os.path.join(save_dir, f)
    psi = np.load(p("last_psi.npy"))
    if Sm05 is not None:
        psi = np.einsum("kij,kaxcj->kaxci", Sm05, psi)
    return (float(np.load(p("last_time.npy"))), float(np.load(p("last_stepsize.npy"))),
            np.load(p("last_sig.npy")), psi, np.load(p("last_omg.npy")))


def propagate_to_save(prop, writer, t0, t1, h, eps, checkpoints=(), time_update="reference",
                      max_attempts=10 ** 7):
    """zand's outer loop (Zandpack/mpi/zand:297-431) on the device engine with zand's
    outputs: currents from stage 0 of each accepted attempt, t += h_next after
    acceptance (``time_update='correct'``: t += h).  Once t has passed a checkpoint
    time, the state at the start of the next accepted step, that t and the step size
    taken from it are saved (:412-420), so resuming repeats that step.
    Returns (t, h, accepted, rejected)."""
    ck = np.array(sorted(c for c in checkpoints if t0 < c < t1), float)
    done = np.zeros(len(ck), bool)
    t, acc, rej, attempts = float(t0), 0, 0, 0
    while t < t1:
        due = (t > ck) & ~done
        snap = prop.get_state() if due.any() else None
        # the stage-0 Π of the step (every attempt of the step starts from the same state):
        # zand saves Q_np = (Π + Π†)/ħ, nozand PI_np = Π/ħ
        pi = prop.pi(hermitian=prop.problem.Sinv is None) if writer.save_pi else None
        while True:
            attempts += 1
            if attempts > max_attempts:
                raise RuntimeError("attempt limit reached")
            r = prop.attempt(t, h, eps)
            if not np.isfinite(r["te"]):
                raise FloatingPointError("non-finite error estimate")
            h_old, h = h, r["h_next"]
            if r["accepted"]:
                break
            rej += 1
        acc += 1
        if snap is not None:
            writer.checkpoint(t, h_old, *snap)
            done[np.nonzero(due)[0][0]] = True
        t = t + (h if time_update == "reference" else h_old)
        writer.step(t, prop.sigma(), r["J"], pi)
    return t, h, acc, rej


def _parse(argv):
    kw = {}
    for a in argv:
        k, _, v = a.replace(" ", "").partition("=")
        kw[k] = v
    return kw


# the settings zand imports from a run's Initial.py (Zandpack/mpi/zand:57-71)
INITIAL_KEYS = ("name", "eps", "t0", "t1", "stepsize", "usesave", "saveevery", "checkpoints",
                "save_checkpoints", "n_dm_compress", "save_PI", "print_timings", "LoadFromFull")


def read_initial(run_dir):
    """The settings of ``<run_dir>/Initial.py`` without importing or running it.

    zand imports the file (Zandpack/mpi/zand:57-71); its settings are plain top-level
    assignments, which are evaluated here from the syntax tree with a closed set of
    forms: literals, names assigned earlier in the file, tuple unpacking, unary and
    binary arithmetic and string concatenation, ``list``/``tuple``/``float``/``int``/
    ``bool``, ``np.linspace``/``np.arange``/``np.array`` on such values and
    ``np.load`` of an ``.npy`` file in the run directory (no pickles).  Returns
    ``(values, skipped)``: the evaluated settings, and the names of assignments in
    another form (the loaded arrays dictionary among them), which are left out.
    """
    import ast
    path = os.path.join(run_dir, "Initial.py")
    with open(path) as fh:
        tree = ast.parse(fh.read(), filename=path)
    env, skipped = {}, []

    class Unsupported(Exception):
        pass

    def npy(fname):
        if not isinstance(fname, str) or not fname.endswith(".npy"):
            raise Unsupported
        full = os.path.realpath(os.path.join(run_dir, fname))
        if os.path.dirname(full) != os.path.realpath(run_dir):
            raise Unsupported
        return np.load(full, allow_pickle=False)

    calls = {"list": list, "tuple": tuple, "float": float, "int": int, "bool": bool}
    npcalls = {"linspace": np.linspace, "arange": np.arange, "array": np.array, "load": npy}
    binops = {ast.Add: lambda a, b: a + b, ast.Sub: lambda a, b: a - b,
              ast.Mult: lambda a, b: a * b, ast.Div: lambda a, b: a / b}

    def ev(node):
        if isinstance(node, ast.Constant):
            return node.value
        if isinstance(node, (ast.Tuple, ast.List)):
            return [ev(e) for e in node.elts]
        if isinstance(node, ast.Name):
            if node.id in env:
                return env[node.id]
            raise Unsupported
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            v = ev(node.operand)
            return -v if isinstance(node.op, ast.USub) else +v
        if isinstance(node, ast.BinOp) and type(node.op) in binops:
            return binops[type(node.op)](ev(node.left), ev(node.right))
        if isinstance(node, ast.Call) and not node.keywords:
            f = node.func
            args = [ev(a) for a in node.args]
            if isinstance(f, ast.Name) and f.id in calls:
                return calls[f.id](*args)
            if (isinstance(f, ast.Attribute) and isinstance(f.value, ast.Name)
                    and f.value.id in ("np", "numpy") and f.attr in npcalls):
                return npcalls[f.attr](*args)
        raise Unsupported

    for st in tree.body:
        if not isinstance(st, ast.Assign) or len(st.targets) != 1:
            continue
        tgt = st.targets[0]
        names = ([tgt.id] if isinstance(tgt, ast.Name) else
                 [e.id for e in tgt.elts if isinstance(e, ast.Name)]
                 if isinstance(tgt, ast.Tuple) else [])
        try:
            v = ev(st.value)
            if isinstance(tgt, ast.Tuple):
                v = list(v)
                if len(v) != len(tgt.elts) or len(names) != len(v):
                    raise Unsupported
                env.update(zip(names, v))
            elif names:
                env[names[0]] = v
        except (Unsupported, TypeError, ValueError, ZeroDivisionError, OSError):
            skipped.extend(names)
    return {k: env[k] for k in INITIAL_KEYS if k in env}, skipped


def main(argv=None):
    from .engine import Propagator
    kw = _parse(sys.argv[1:] if argv is None else argv)
    run_dir = kw.pop("Dir", ".")
    # Initial.py's settings, which keywords on the command line override
    init = {}
    if os.path.exists(os.path.join(run_dir, "Initial.py")):
        init, _ = read_initial(run_dir)
        if not init.get("save_checkpoints", True):
            init["checkpoints"] = []
        if "checkpoints" in init:
            init["checkpoints"] = ",".join(repr(float(c)) for c in np.ravel(init["checkpoints"]))
    for k, v in init.items():
        kw.setdefault(k, v)
    if "name" not in kw:
        raise ValueError("no run name: give name=<name> or an Initial.py that sets it")
    name = str(kw.pop("name"))
    f = lambda k, d: float(kw.get(k, d))
    b = lambda k, d: str(kw.get(k, d)).lower() in ("1", "true", "yes")
    hbar = f("hbar", HBAR)
    nonortho = b("nonorthogonal", False)
    p = load_arrays(os.path.join(run_dir, name, "Arrays"), hbar=hbar, name=name,
                    nonorthogonal=nonortho)
    Sm05 = p.Sm05 if nonortho else None
    weights = tuple(float(x) for x in kw.get("weights", "0.5,-0.5").split(","))
    kind = kw.get("pulse", "none")
    if kind in EXPERIMENTAL_PULSES:  # pulse=air_photonics pulse_table=<ExpPulses .npy>
        p.pulse = experimental_pulse(kind, np.load(kw["pulse_table"], allow_pickle=False),
                                     f("amp", 1.0), weights)
    else:
        p.pulse = Pulse(kind, f("amp", 0.0), f("p1", 0.0), f("p2", 0.0), weights,
                        f("p3", 0.0), f("p4", 0.0))
    if f("hubbard_U", 0.0):
        p.hubbard_U = f("hubbard_U", 0.0)
        p.hubbard_n0 = np.load(kw["hubbard_n0"], allow_pickle=False)
    save_dir = os.path.join(run_dir, name + "_save")
    t0, t1, h = f("t0", 0.0), f("t1", 10.0), f("stepsize", 0.01)
    state = None
    if b("usesave", False):
        t0, h, s, ps, om = load_checkpoint(save_dir, Sm05)
        state = (s, ps, om)
    prop = Propagator(p, kw.get("precision", "fp64"), state=state, autotune=True)
    if state is None and b("steady_state", False):
        its, res = prop.steady_state(t0)
        print(f"steady state: {its} iterations, relative residual {res:.2e}", flush=True)
    ck = [float(x) for x in kw.get("checkpoints", "").split(",") if x]
    w = SaveWriter(save_dir, p.dims[2], int(f("saveevery", 100)), int(f("n_dm_compress", 1)),
                   t0, prop.sigma(), S05=None if Sm05 is None else np.linalg.inv(Sm05),
                   save_pi=b("save_PI", False))
    t, h, acc, rej = propagate_to_save(prop, w, t0, t1, h, f("eps", 1e-8), ck,
                                       kw.get("time_update", "reference"))
    print(f"t={t:.6g} accepted={acc} rejected={rej} h={h:.6g}", flush=True)
    prop.close()
    return 0


def combine_currents(save_dirs, NL=2):
    """The reading side of the format, as Zandpack/mpi_tools.py:19-49 reads it: the
    current files of each directory in natural step order, stacked; the times files
    concatenated with one trailing time dropped per directory."""
    nat = lambda s: [int(c) if c.isdigit() else c for c in re.split(r"(\d+)", s)]
    C = [[] for _ in range(NL)]
    T = []
    for d in save_dirs:
        fs = sorted((v for v in os.listdir(d) if v.startswith("current_")), key=nat)
        for ff in fs:
            C[int(ff.split("_")[1])].append(np.load(os.path.join(d, ff)))
        for ff in sorted((v for v in os.listdir(d) if v.startswith("times")), key=nat):
            T.append(np.load(os.path.join(d, ff)))
    return np.hstack(T)[:-len(save_dirs)], [np.vstack(c) for c in C]


if __name__ == "__main__":
    sys.exit(main())
