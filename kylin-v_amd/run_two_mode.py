"""Scale + split-precision workload at the TDVP level: a 64-site chain, each site an exciton
times two phonon modes (w1 = 0.5, w2 = 1.5, 8 levels each, so d = 128), g1 = g2 = 0.7,
J = 1, one exciton, two-site TDVP at dt = 0.02 with bond cap 512, cutoff 1e-10, from the
physical start (exciton on site 0, phonon vacuum).  The same steps run with the H_eff
contractions in FP64, FP32 and split-BF16 (FP64 Lanczos, expm, SVD and environments);
prints one JSON line with C(t) of the FP64 run, max|C_p(t) - C_fp64(t)| per precision,
the bond dimensions reached and the wall time per step.  With --budget-s the first
precision stops once it has used that much wall time, and the others run the same number
of steps.

    python kylin-v_amd/run_two_mode.py [--steps 10] [--N 64] [--budget-s 300]
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
    ap.add_argument("--N", type=int, default=64)
    ap.add_argument("--L", type=int, default=8)
    ap.add_argument("--D", type=int, default=512)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--dt", type=float, default=0.02)
    ap.add_argument("--precisions", default="fp64,fp32,bf16x3")
    ap.add_argument("--budget-s", type=float, default=0.0)
    args = ap.parse_args()
    import numpy as np
    import __graft_entry__ as GE
    kv = GE.load_package()
    from kylin_v_amd import model as KM

    def log(msg):
        print(f"two-mode: {msg}", file=sys.stderr, flush=True)

    ch = KM.HolsteinChain(args.N, [0.5, 1.5], 0.7, [args.L, args.L], 1.0, "merged")
    out = {"workload": f"Holstein chain N={args.N}, exciton x 2 modes (w 0.5, 1.5; {args.L} levels "
                       f"each, d={2 * args.L * args.L}), g=0.7, J=1, two-site TDVP dt={args.dt}, "
                       f"D<={args.D}, cutoff 1e-10, physical start",
           "steps": args.steps, "precisions": {}}
    ct = {}
    steps = args.steps
    for pi, prec in enumerate(args.precisions.split(",")):
        eng = kv.TDVP(ch.psi0, ch.mpo, cutoff=1e-10, maxdim=args.D, dt=args.dt, precision=prec)
        cs = []
        t0 = time.perf_counter()
        for k in range(steps):
            eng.two_site()
            cs.append(kv.overlap_states(ch.psi0, eng.mps))
            log(f"{prec} step {k + 1}: C = {cs[-1]:.10f}, max bond {max(eng.bond_dims())}")
            if pi == 0 and args.budget_s > 0 and time.perf_counter() - t0 > args.budget_s:
                break
        steps = len(cs)
        out["steps"] = steps
        wall = time.perf_counter() - t0
        ct[prec] = np.array(cs)
        st = eng.stats()
        out["precisions"][prec] = {"s_per_step": wall / steps, "max_bond": max(eng.bond_dims()),
                                   "heff_tflops": st["heff_flop"] / max(st["heff_gpu_s"], 1e-30) / 1e12,
                                   "modules_s": {k: st[k] for k in ("Evl-2", "Evl-1", "SVD", "Update")}}
        del eng
    if "fp64" in ct:
        out["C_t_fp64"] = [[float(c.real), float(c.imag)] for c in ct["fp64"]]
        for prec in ct:
            if prec != "fp64":
                out["precisions"][prec]["max_dC_vs_fp64"] = float(np.max(np.abs(ct[prec] - ct["fp64"])))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
