"""Batching workload: 64 independent Holstein chains (lengths U{16..48}, levels {8,12,16},
J ~ U[0.5,1.5], w ~ U[0.3,1.5], g ~ U[0.5,2.0], bond caps U{64..256}, seed 3), two-site
TDVP 50 steps at dt = 0.05, advanced concurrently by worker threads with one context
(HIP stream) each.  Prints one JSON line with chain-steps/s.

    python kylin-v_amd/run_batch.py [--chains 64] [--steps 50] [--workers 8] [--precision fp64]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--chains", type=int, default=64)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--first", type=int, default=0, help="run chains [first, first+count) of the set")
    ap.add_argument("--count", type=int, default=None)
    ap.add_argument("--pick", default=None, help="comma-separated chain indices (instead of first/count)")
    args = ap.parse_args()
    import __graft_entry__ as GE
    GE.load_package()
    from kylin_v_amd import batch

    specs = batch.seeded_chains(args.chains, args.seed)
    if args.pick:
        idx = [int(x) for x in args.pick.split(",")]
    else:
        count = len(specs) - args.first if args.count is None else args.count
        idx = list(range(args.first, args.first + count))
    specs = [specs[i] for i in idx]
    t0 = time.perf_counter()

    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")

    def free_gib():
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
            return float("nan")
        return free.value / 2**30

    def progress(i, r):
        print(f"batch: chain {idx[i]} N={specs[i]['N']} L={specs[i]['L']} D={specs[i]['D']} done in "
              f"{r['seconds']:.1f} s, max bond {max(r['bonds'])}, device free {free_gib():.1f} GiB",
              file=sys.stderr, flush=True)

    res, wall = batch.run_batch(specs, args.steps, dt=0.05, workers=args.workers,
                                precision=args.precision, on_progress=progress)
    out = {"workload": "64 Holstein chains, 2TDVP 50 steps dt=0.05 (seeded set)",
           "chains": args.chains, "picked": idx, "count": len(specs), "steps": args.steps,
           "workers": min(args.workers, batch.MAX_WORKERS),
           "precision": args.precision, "wall_s": wall,
           "chain_steps_per_s": len(specs) * args.steps / wall,
           "sum_chain_seconds": sum(r["seconds"] for r in res),
           "max_bond": max(max(r["bonds"]) for r in res)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
