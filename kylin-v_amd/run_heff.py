"""Kernel-level H_eff measurement: one two-site effective Hamiltonian of a merged-layout
Holstein chain at bond dimension D, applied `reps` times per precision (fp64, fp32,
bf16x3 = split-BF16 products with FP32 accumulation), HIP-event timed, with the error of
each precision against the FP64 result.

The default is the scale + split-precision shape: each site an exciton times two phonon
modes (w1 = 0.5, w2 = 1.5, 8 levels each, so d = 128), g1 = g2 = 0.7, J = 1.  The state is
a seeded synthetic MPS; only the contraction shapes matter for the timing.

    python kylin-v_amd/run_heff.py [--D 256] [--modes 2] [--L 8] [--reps 5]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0 / 3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=6, help="sites (the H_eff sits on the middle bond)")
    ap.add_argument("--modes", type=int, default=2)
    ap.add_argument("--L", type=int, default=8)
    ap.add_argument("--D", type=int, default=256)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--precisions", default="fp64,fp32,bf16x3")
    args = ap.parse_args()
    import numpy as np
    import __graft_entry__ as GE
    kv = GE.load_package()
    from kylin_v_amd import model as KM

    def log(m):
        print(f"heff: {m}", file=sys.stderr, flush=True)

    omegas = [0.5, 1.5][:args.modes] if args.modes <= 2 else [1.0] * args.modes
    ch = KM.HolsteinChain(args.N, omegas, 0.7, [args.L] * args.modes, 1.0, "merged")
    blocks = KM.synthetic_state_blocks(ch.sites, args.D, seed=9)
    A = [kv.Tensor.from_blocks((1, 1, -1), b) for b in blocks]
    mid = args.N // 2 - 1
    envL = kv.Tensor.from_blocks((1, -1, -1), {(0, 0, 0): np.ones((1, 1, 1))})
    for i in range(mid):
        envL = kv.env_sweep(envL, A[i], ch.mpo[i], True)
    envR = kv.Tensor.from_blocks((-1, 1, 1), {(1, 0, 1): np.ones((1, 1, 1))})
    for i in range(args.N - 1, mid + 1, -1):
        envR = kv.env_sweep(envR, A[i], ch.mpo[i], False)
    x = kv.product(A[mid], A[mid + 1], 1)
    d = ch.sites[0].d
    nx = x.structure()[4]
    log(f"d={d} D={args.D} x={nx} complex ({nx * 16 / 2**30:.2f} GiB FP64)")
    ref = kv.apply_double(envL, x, ch.mpo[mid], ch.mpo[mid + 1], envR, "fp64")
    yr = ref.download()
    ymax = float(np.abs(yr).max())
    del ref
    out = {"workload": f"two-site H_eff, merged d={d} ({args.modes} modes x {args.L} levels), "
                       f"D={args.D}", "x_complex": nx, "precisions": {}}
    for prec in args.precisions.split(","):
        t0 = time.perf_counter()
        p = kv.heff_profile(2, envL, x, ch.mpo[mid], ch.mpo[mid + 1], envR, prec, args.reps)
        y = kv.apply_double(envL, x, ch.mpo[mid], ch.mpo[mid + 1], envR, prec).download()
        err = float(np.abs(y - yr).max()) / ymax
        gemm_tf = p["gemm_flop"] / p["gemm_s"] / 1e12 if p["gemm_s"] > 0 else 0.0
        out["precisions"][prec] = dict(
            apply_ms=1e3 * p["apply_s"], heff_tflops=p["flop"] / p["apply_s"] / 1e12,
            gemm_tflops=gemm_tf, gemm_frac_of_peak=gemm_tf / PEAK[prec],
            gemm_ms_per_launch=1e3 * p["gemm_s"] / max(1, p["gemm_launches"]),
            gemm_gflop_per_launch=p["gemm_flop"] / max(1, p["gemm_launches"]) / 1e9,
            wfold_ms=1e3 * p["wfold_s"], flop_per_apply=p["flop"],
            max_rel_err_vs_fp64=err)
        log(f"{prec}: {out['precisions'][prec]['apply_ms']:.2f} ms/apply, "
            f"{out['precisions'][prec]['heff_tflops']:.1f} TFLOP/s H_eff, GEMM {gemm_tf:.1f} TFLOP/s, "
            f"err {err:.2e} ({time.perf_counter() - t0:.1f} s)")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
