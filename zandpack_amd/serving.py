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

Every request's arithmetic is the same kernels on its own buffers in its own
order, so its result is bitwise the same alone and in any mix.

There is no counterpart in the reference (one `mpirun zand` runs one problem);
the step semantics per request are zand's (Zandpack/mpi/zand:297-431).
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from .builder import AMEProblem
from .engine import Propagator


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


class _Live:
    def __init__(self, req: Request, graph: bool, autotune: bool):
        self.req = req
        # tile choices are cached per (precision, size bucket): only the first
        # request of a bucket pays for the timing attempts
        t0 = time.perf_counter()
        self.prop = Propagator(req.problem, req.precision, graph=graph, autotune=autotune)
        self.setup_s = time.perf_counter() - t0
        self.bytes = footprint_bytes(req.problem, req.precision)
        self.t = 0.0
        self.h = req.h0
        self.attempts = 0
        self.inflight = False
        req.times.append(0.0)

    def launch(self):
        self.prop.launch(self.t, self.h, self.req.eps)
        self.inflight = True

    def step_done(self, res):
        """Bookkeeping of a collected attempt; True when the request is finished."""
        req = self.req
        self.attempts += 1
        h_old, self.h = self.h, res["h_next"]
        if res["accepted"]:
            req.J.append(res["J"])
            req.h.append(h_old)
            self.t += self.h if req.time_update == "reference" else h_old
            req.times.append(self.t)
            req.steps += 1
        else:
            req.rejected += 1
        if self.t >= req.t_end or (req.max_steps is not None and req.steps >= req.max_steps):
            return True
        if self.attempts >= req.max_attempts:
            req.error = "attempt limit reached"
            return True
        return False


class Server:
    """Admission by HBM footprint (on a host worker thread), per-request continuous
    stepping with one attempt of every live request in flight."""

    def __init__(self, hbm_bytes: int = 64 << 30, max_live: int = 64, graph: bool = True,
                 keep_sigma: bool = True, autotune: bool = True, admit_workers: int = 1):
        self.budget = int(hbm_bytes)
        self.max_live = int(max_live)
        self.graph = graph
        self.keep_sigma = keep_sigma
        self.autotune = autotune
        self.admit_workers = int(admit_workers)

    def run(self, requests, realtime: bool = True):
        """Serve ``requests`` (arrival times in seconds).  With ``realtime`` the
        arrivals follow the wall clock; otherwise all are present at start."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in pending:
            if footprint_bytes(r.problem, r.precision) > self.budget:
                raise ValueError(f"request {r.rid} alone exceeds the HBM budget")
        live: list[_Live] = []
        admitting = []            # (request, reserved bytes, future)
        done: list[Request] = []
        used = 0
        self.peak_bytes = 0
        attempts = 0
        t_admit = 0.0
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: float("inf"))

        def retire(lv):
            nonlocal used
            live.remove(lv)
            lv.req.finished = time.perf_counter() - t0
            if self.keep_sigma and lv.req.error is None:
                lv.req.sigma = lv.prop.sigma()
            lv.prop.close()
            used -= lv.bytes
            done.append(lv.req)

        with ThreadPoolExecutor(max_workers=self.admit_workers) as pool:
            while pending or live or admitting:
                now = clock()
                # admission: reserve the footprint, set the problem up on the worker
                while (pending and pending[0].arrival <= now
                       and len(live) + len(admitting) < self.max_live
                       and used + footprint_bytes(pending[0].problem, pending[0].precision)
                       <= self.budget):
                    r = pending.pop(0)
                    r.admitted = time.perf_counter() - t0
                    fp = footprint_bytes(r.problem, r.precision)
                    used += fp
                    self.peak_bytes = max(self.peak_bytes, used)
                    admitting.append((r, fp, pool.submit(_Live, r, self.graph, self.autotune)))
                for item in [a for a in admitting if a[2].done()]:
                    admitting.remove(item)
                    r, fp, fut = item
                    try:
                        lv = fut.result()
                    except Exception as e:  # set-up failed: retired with its error
                        r.error = f"admission failed: {e}"
                        r.finished = time.perf_counter() - t0
                        used -= fp
                        done.append(r)
                        continue
                    t_admit += lv.setup_s
                    lv.bytes = fp
                    live.append(lv)
                    try:
                        lv.launch()
                    except Exception as e:
                        lv.req.error = str(e)
                        retire(lv)
                # collect every finished attempt and launch the next one at once
                progressed = False
                for lv in list(live):
                    if not lv.inflight or not lv.prop.ready():
                        continue
                    progressed = True
                    lv.inflight = False
                    try:
                        res = lv.prop.collect()
                    except Exception as e:  # e.g. a step size driven to zero or infinity
                        lv.req.error = str(e)
                        retire(lv)
                        continue
                    attempts += 1
                    if lv.step_done(res):
                        retire(lv)
                        continue
                    try:
                        lv.launch()
                    except Exception as e:
                        lv.req.error = str(e)
                        retire(lv)
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
