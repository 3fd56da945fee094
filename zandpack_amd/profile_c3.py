"""Profiling target: a few DOP54 attempts of the n=1000, M=4000, NK=2 workload
(the headline configuration) in one precision, for rocprofv3 runs, e.g.

    rocprofv3 --kernel-trace --stats -d out -o c3 -- python3 -m zandpack_amd.profile_c3 --precision fp64
    rocprofv3 --pmc FETCH_SIZE -d out -o fetch -- python3 -m zandpack_amd.profile_c3 --copy-mib 256

``--copy-mib`` also runs one device copy of that many MiB (torch), whose counters
calibrate FETCH_SIZE / WRITE_SIZE against a known byte count; ``--copy-only`` runs
just that copy (one calibration for both precisions).  The tiles are
autotuned first (those timing attempts show up in a trace as well); the last
``--attempts`` attempts run with the chosen tiles.  ``--serial`` runs them the way
bench.py's per-class profile does (one stream, HIP events around each class).
``--fsal off`` evaluates every stage of every attempt (with first same as last on, the
first stage's dispatches of an attempt after an accepted one exit at once, which a
per-dispatch average would count).
"""
import argparse
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--attempts", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=0)
    ap.add_argument("--copy-only", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--serial", action="store_true",
                    help="per-class timing mode: one stream, no graph (as bench.py's kernel profile)")
    ap.add_argument("--fsal", default="on", choices=["on", "off"],
                    help="first same as last (off: every stage of every attempt evaluated)")
    a = ap.parse_args()
    from zandpack_amd import builder
    from zandpack_amd.engine import Propagator
    if a.copy_mib:
        import torch
        x = torch.ones(a.copy_mib << 20, dtype=torch.uint8, device="cuda")
        y = x.clone()
        torch.cuda.synchronize()
        del x, y
        if a.copy_only:
            print(f"profile_c3: copy of {a.copy_mib} MiB", flush=True)
            return 0
    p = builder.config_c3(seed=0)
    prop = Propagator(p, a.precision, graph=not a.no_graph, autotune=True)
    prop.set_fsal(a.fsal)
    if a.serial:
        prop.set_profile(True)
    t, h = 0.0, 0.002
    for _ in range(a.attempts):
        r = prop.attempt(t, h, 1.0, force_accept=True)
        t += h
    print(f"profile_c3 {a.precision}: {a.attempts} attempts, tiles {prop.tiles}, "
          f"TE {r['te']:.3e}", flush=True)
    prop.close()


if __name__ == "__main__":
    sys.exit(main())
