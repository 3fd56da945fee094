"""Continuous-batching serving loop over device-resident propagations.

Requests (one AME propagation each: a device, its lead constants, a bias pulse
and an end time) arrive over time.  The server admits them in arrival order as
HBM frees up (by the footprint of σ, Ψ, Ω, the seven RK stages and the GEMM
operands at the request's precision), keeps every admitted state resident on
the device and keeps one DOP54 attempt of every live request in flight on its
own HIP streams, so the attempts of different requests run concurrently on the
GPU.  Each request is collected as soon as its attempt has finished and its next
attempt launched at once: a small request does not wait for the largest one.
Admission (problem set-up, upload and tile tuning) runs on a host worker thread
while the live requests keep stepping.  Finished requests are retired at once and
their memory admits the next ones.

Requests of one shape (dimensions, precision and coupling structure) that have
arrived when one of them is admitted are admitted together as the instances of one
native problem (engine.Instances): every kernel of their attempts runs once for all
of them, each with its own time, step, pulse and acceptance; an instance that has
finished is held (h = 0) until the group's last one is done.

Every request's arithmetic is the same kernels on its own buffers in its own
order, so its result is bitwise the same alone and in any mix (grouped, on the
group's tiles).

There is no counterpart in the reference (one `mpirun zand` runs one problem);
the step semantics per request are zand's (Zandpack/mpi/zand:297-431).
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from .builder import AMEProblem
from .engine import Instances


def footprint_bytes(problem: AMEProblem, precision: str = "mixed") -> int:
    """Upper estimate of the device memory one request holds (bytes)."""
    NK, n, NL, NX, NOT = problem.dims
    M = problem.M
    Md = M + 64  # active-block padding bound
    cs = 16 if precision == "fp64" else 8
    nc = int(np.count_nonzero(np.any(problem.xi.reshape(-1, n) != 0, axis=0)))
    ncp = nc + 16 * NL  # lead groups padded to 16 columns
    slots = NK * (9 * Md * Md * cs + 10 * n * n * 16 + 9 * Md * n * cs)
    ops = NK * (Md * (n + ncp) * cs + (n + ncp) * n * cs       # Ap, Bp
                + 4 * Md * ncp * cs                              # Aomg, Bomg
                + (1 + 4 + 7 + 1) * Md * ncp * cs                # ξ_c, Ωξ parts, products, W
                + Md * ncp * 16 + 8 * n * ncp * 16               # Ixi_c, Π parts
                + 5 * n * n * 16                                 # H, its real copies, G, σ_eq
                + 64 * n * 20                                    # row-compressed H
                + 8 * Md * 16)                                   # per-mode constants
    if precision == "mixed" and 0 < ncp <= 128:  # fused Ω kernel: chunk partials, y0/y_new ξ_c
        ops += NK * (2 * ((Md + 255) // 256 + 1) + 2) * Md * ncp * cs
    te = 8 * (2 * 1024 + ((Md + 15) // 16) ** 2 * NK)
    return int(slots + ops + te + (1 << 20))


@dataclass
class Request:
    rid: int
    problem: AMEProblem
    t_end: float
    eps: float
    h0: float
    arrival: float = 0.0          # seconds after the server starts
    precision: str = "mixed"
    time_update: str = "reference"
    max_attempts: int = 10 ** 6
    max_steps: int | None = None  # finish after this many accepted steps (bounded runs)
    # filled in by the server
    times: list = field(default_factory=list)
    h: list = field(default_factory=list)
    J: list = field(default_factory=list)
    steps: int = 0
    rejected: int = 0
    admitted: float = float("nan")
    finished: float = float("nan")
    sigma: np.ndarray | None = None
    error: str | None = None

    @property
    def latency(self):
        return self.finished - self.arrival


def shape_key(req: Request):
    """Requests with equal keys can be instances of one native problem: same
    dimensions, precision and coupling structure (active modes, coupled orbitals), no
    per-problem drive, dissipator, overlap or spline pulse."""
    p = req.problem
    act = ((p.dGP != 0) | (p.dGM != 0)).reshape(p.dims[0], -1).any(axis=0)
    cpl = (p.xi != 0).reshape(-1, p.dims[1]).any(axis=0)
    single = bool(p.drive) or p.dissipator.kind != "none" or p.Sinv is not None \
        or p.pulse.kind == "spline"
    return (tuple(p.dims), req.precision, float(p.hbar), float(p.hubbard_U or 0.0),
            act.tobytes(), cpl.tobytes(), req.rid if single else -1)


class _Live:
    """Requests of one shape stepped as the instances of one native problem."""

    def __init__(self, reqs, graph: bool, autotune: bool):
        self.reqs = list(reqs)
        # tile choices are cached per (precision, size bucket): only the first
        # group of a bucket pays for the timing attempts
        t0 = time.perf_counter()
        self.prop = Instances([r.problem for r in self.reqs], self.reqs[0].precision,
                              graph=graph, autotune=autotune)
        self.setup_s = time.perf_counter() - t0
        self.bytes = sum(footprint_bytes(r.problem, r.precision) for r in self.reqs)
        G = len(self.reqs)
        self.t = np.zeros(G)
        self.h = np.array([r.h0 for r in self.reqs], dtype=float)
        self.eps = np.array([r.eps for r in self.reqs], dtype=float)
        self.attempts = np.zeros(G, dtype=np.int64)
        self.done = np.zeros(G, dtype=bool)
        self.inflight = False
        for r in self.reqs:
            r.times.append(0.0)

    def launch(self):
        # finished instances are held: h = 0, accepted by force, state unchanged
        self.prop.launch(self.t, np.where(self.done, 0.0, self.h), self.eps,
                         force_accept=self.done)
        self.inflight = True

    def step_done(self, results):
        """Bookkeeping of a collected attempt; returns the requests that finished."""
        out = []
        for g, (req, res) in enumerate(zip(self.reqs, results)):
            if self.done[g]:
                continue
            self.attempts[g] += 1
            if not (np.isfinite(res["te"]) and np.isfinite(res["h_next"])):
                req.error = f"non-finite error estimate at t={self.t[g]}"
                self.done[g] = True
                out.append(req)
                continue
            h_old, self.h[g] = self.h[g], res["h_next"]
            if res["accepted"]:
                req.J.append(res["J"])
                req.h.append(h_old)
                self.t[g] += self.h[g] if req.time_update == "reference" else h_old
                req.times.append(self.t[g])
                req.steps += 1
            else:
                req.rejected += 1
            if self.t[g] >= req.t_end or (req.max_steps is not None
                                          and req.steps >= req.max_steps):
                self.done[g] = True
                out.append(req)
            elif self.attempts[g] >= req.max_attempts:
                req.error = "attempt limit reached"
                self.done[g] = True
                out.append(req)
        return out


class Server:
    """Admission by HBM footprint (on a host worker thread), continuous stepping with
    one attempt of every live group in flight; max_group > 1 admits arrived requests
    of the oldest one's shape together as one group (one launch per kernel class)."""

    def __init__(self, hbm_bytes: int = 64 << 30, max_live: int = 64, graph: bool = True,
                 keep_sigma: bool = True, autotune: bool = True, admit_workers: int = 1,
                 max_group: int = 16):
        self.budget = int(hbm_bytes)
        self.max_live = int(max_live)
        self.graph = graph
        self.keep_sigma = keep_sigma
        self.autotune = autotune
        self.admit_workers = int(admit_workers)
        self.max_group = max(1, int(max_group))

    def run(self, requests, realtime: bool = True, preadmit: bool = False):
        """Serve ``requests`` (arrival times in seconds).  With ``realtime`` the
        arrivals follow the wall clock; otherwise all are present at start.  With
        ``preadmit`` (not realtime) every request is admitted before the first attempt
        and the clock starts then: the figures are of the stepping alone."""
        assert not (preadmit and realtime)
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in pending:
            if footprint_bytes(r.problem, r.precision) > self.budget:
                raise ValueError(f"request {r.rid} alone exceeds the HBM budget")
        live: list[_Live] = []
        admitting = []            # (requests, reserved bytes, future)
        done: list[Request] = []
        used = 0
        n_live = 0                # requests admitted and not finished
        self.peak_bytes = 0
        self.groups = []          # sizes of the admitted groups
        attempts = 0
        t_admit = 0.0
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: float("inf"))
        keys = {r.rid: shape_key(r) for r in pending}

        def finish(lv, reqs):
            nonlocal n_live
            now = time.perf_counter() - t0
            sig = lv.prop.sigma() if self.keep_sigma and reqs else None
            KI = lv.prop.KI
            for r in reqs:
                r.finished = now
                if self.keep_sigma and r.error is None:
                    g = lv.reqs.index(r)
                    r.sigma = sig[g * KI:(g + 1) * KI].copy()
                done.append(r)
                n_live -= 1

        def retire(lv, error=None):
            nonlocal used
            live.remove(lv)
            rest = [r for g, r in enumerate(lv.reqs) if not lv.done[g]]
            for r in rest:
                r.error = r.error or error
            finish(lv, rest)
            lv.prop.close()
            used -= lv.bytes

        with ThreadPoolExecutor(max_workers=self.admit_workers) as pool:
            while pending or live or admitting:
                now = clock()
                # admission: the oldest arrived request and the arrived ones of its shape
                # (up to max_group) reserve their footprint; set-up on the worker
                while (pending and pending[0].arrival <= now and n_live < self.max_live
                       and used + footprint_bytes(pending[0].problem, pending[0].precision)
                       <= self.budget):
                    k0 = keys[pending[0].rid]
                    grp, fp = [], 0
                    for r in list(pending):
                        if r.arrival > now or len(grp) >= self.max_group \
                                or n_live + len(grp) >= self.max_live:
                            break
                        if keys[r.rid] != k0:
                            continue
                        f = footprint_bytes(r.problem, r.precision)
                        if used + fp + f > self.budget:
                            break
                        grp.append(r)
                        fp += f
                    for r in grp:
                        pending.remove(r)
                        r.admitted = time.perf_counter() - t0
                    used += fp
                    n_live += len(grp)
                    self.groups.append(len(grp))
                    self.peak_bytes = max(self.peak_bytes, used)
                    admitting.append((grp, fp, pool.submit(_Live, grp, self.graph,
                                                           self.autotune)))
                for item in [a for a in admitting if a[2].done()]:
                    admitting.remove(item)
                    grp, fp, fut = item
                    try:
                        lv = fut.result()
                    except Exception as e:  # set-up failed: retired with its error
                        for r in grp:
                            r.error = f"admission failed: {e}"
                            r.finished = time.perf_counter() - t0
                            done.append(r)
                        n_live -= len(grp)
                        used -= fp
                        continue
                    t_admit += lv.setup_s
                    lv.bytes = fp
                    live.append(lv)
                    if preadmit:
                        continue
                    try:
                        lv.launch()
                    except Exception as e:
                        retire(lv, str(e))
                if preadmit:
                    if pending or admitting:
                        time.sleep(1e-3)
                        continue
                    # everything admitted: the stepping starts now
                    preadmit = False
                    t0 = time.perf_counter()
                    for lv in list(live):
                        try:
                            lv.launch()
                        except Exception as e:
                            retire(lv, str(e))
                # collect every finished attempt and launch the next one at once
                progressed = False
                for lv in list(live):
                    if not lv.inflight or not lv.prop.ready():
                        continue
                    progressed = True
                    lv.inflight = False
                    try:
                        res = lv.prop.collect()
                    except Exception as e:
                        retire(lv, str(e))
                        continue
                    attempts += 1
                    finished = lv.step_done(res)
                    if finished:
                        finish(lv, finished)
                    if lv.done.all():
                        retire(lv)
                        continue
                    try:
                        lv.launch()
                    except Exception as e:
                        retire(lv, str(e))
                if not progressed:
                    if not live and not admitting and pending:
                        time.sleep(max(0.0, min(pending[0].arrival - clock(), 0.01)))
                    else:
                        time.sleep(2e-5)
        wall = time.perf_counter() - t0
        return self.stats(done, wall, attempts, t_admit)

    @staticmethod
    def stats(done, wall, attempts, t_admit=0.0):
        """Aggregate figures over the whole wall time: steps/s, p50/p99 time to
        completion (arrival to result), and the host time spent setting requests up
        (on the admission worker, beside the stepping)."""
        ok = [r for r in done if r.error is None]
        lat = np.array([r.latency for r in ok]) if ok else np.zeros(0)
        steps = int(sum(r.steps for r in done))
        return dict(requests=len(done), failed=len(done) - len(ok), wall_s=wall,
                    attempts=attempts, steps=steps,
                    steps_per_s=steps / wall if wall > 0 else 0.0, admit_s=t_admit,
                    p50_s=float(np.percentile(lat, 50)) if lat.size else None,
                    p99_s=float(np.percentile(lat, 99)) if lat.size else None,
                    done=done)
