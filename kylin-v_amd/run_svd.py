"""Bond-SVD measurement at the benchmark's middle bond: the two-site tensor A[15] A[16] of the
32-site merged Holstein chain (16 phonon levels, d = 32) in the seeded synthetic D = 256
start, truncated `reps` times at maxdim 256, cutoff 1e-10, in each sweep direction.  Its two
2048 x 4096 charge sectors take the block subspace route, so under rocprofv3 every FP64
grouped-GEMM dispatch of this run is a block-route product; the sector gathers give the
counter calibration (zero_kernel writes, nrm2sq_partial_kernel reads, each 2048 x 4096 x 16
bytes per sector).

    python kylin-v_amd/run_svd.py [--reps 3] [--D 256]
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
    ap.add_argument("--N", type=int, default=32)
    ap.add_argument("--L", type=int, default=16)
    ap.add_argument("--D", type=int, default=256)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    import __graft_entry__ as GE
    kv = GE.load_package()
    from kylin_v_amd import model as KM

    def log(m):
        print(f"svd: {m}", file=sys.stderr, flush=True)

    sites, terms, init, ex = KM.chain(args.N, [0.5], 1.0, [args.L], 1.0, "merged")
    blocks = KM.synthetic_state_blocks(sites, args.D, seed=7)
    mid = args.N // 2 - 1
    a = kv.Tensor.from_blocks((1, 1, -1), blocks[mid])
    b = kv.Tensor.from_blocks((1, 1, -1), blocks[mid + 1])
    x = kv.product(a, b, 1)
    log(f"two-site tensor: {x.structure()[4]} complex")
    out = {"workload": f"bond SVD of A[{mid}] A[{mid + 1}], merged N={args.N} L={args.L} D={args.D}, "
                       f"maxdim {args.D}, cutoff 1e-10", "ms_per_call": {}}
    for l2r in (True, False):
        kv.svd_impl(x, 2, 1e-10, args.D, l2r)  # first call: plans, pools, rocSOLVER workspace
        ctx = x.ctx
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.reps):
            kv.svd_impl(x, 2, 1e-10, args.D, l2r)
        ctx.sync()
        ms = 1e3 * (time.perf_counter() - t0) / args.reps
        out["ms_per_call"]["l2r" if l2r else "r2l"] = ms
        log(f"{'left-to-right' if l2r else 'right-to-left'}: {ms:.2f} ms per call")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
