"""Times the CPU finite-difference baseline of the differentiation workload on a reduced
chain: the oracle's FP64 one-site TDVP forward (the numpy restatement of the reference's
algorithm, Krylov tolerance 1e-12) under central differences of every parameter, i.e.
2 P + 1 forwards for P = N + 2 couplings.  Prints one JSON line.

    python tests/time_oracle_fd.py [--N 8] [--L 6] [--D 32] [--steps 10] [--threads T]

Reference-algorithm test infrastructure (it runs the oracle); not collected by pytest.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=8)
    ap.add_argument("--L", type=int, default=6)
    ap.add_argument("--D", type=int, default=32)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--dt", type=float, default=0.05)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--h", type=float, default=1e-4)
    args = ap.parse_args()
    import numpy as np
    import adjoint_refs as R
    import importlib.util  # the package's model builder, loaded without the HIP library
    spec = importlib.util.spec_from_file_location(
        "kv_model_cpu", os.path.join(os.path.dirname(HERE), "kylin-v_amd", "model.py"))
    KM = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(KM)
    N = args.N
    rng = np.random.default_rng(args.seed)
    theta = np.concatenate([rng.uniform(0.8, 1.2, N), [1.0, 1.0]])
    sites, terms, init, ex = KM.chain(N, [1.0], np.ones((N, 1)), [args.L], 1.0, "merged")
    psi0 = KM.expanded_state_blocks(sites, init, args.D, seed=args.seed + 1, eps=0.05)
    t0 = time.perf_counter()
    f0 = R.oracle_autocorr(theta, N, args.L, psi0, args.dt, args.steps)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    grad = R.oracle_fd_gradient(theta, N, args.L, psi0, args.dt, args.steps, h=args.h)
    t_fd = time.perf_counter() - t0
    print(json.dumps({"workload": f"one-site TDVP N={N} L={args.L} D={args.D} {args.steps} steps "
                                  f"dt={args.dt}", "params": int(theta.size), "value": f0,
                      "forward_s": t_fwd, "fd_gradient_s": t_fd, "forwards": 2 * int(theta.size) + 1,
                      "threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                      "grad_max": float(np.abs(grad).max())}), flush=True)


if __name__ == "__main__":
    main()
