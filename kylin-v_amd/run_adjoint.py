"""Differentiation workload: one-site TDVP at fixed bond dimension 128 on a 24-site chain
with 12 phonon levels (J = 1, w = 1, g_i ~ U[0.8, 1.2] from numpy default_rng(seed)),
30 steps at dt = 0.05.  Objective |C(T)|^2, gradient with respect to g_1..g_24, J and w by
the hand-written adjoint with checkpointed environments, checked against central
differences (h = 1e-4) of the same device forward on the parameters listed in --fd and
along --dirs seeded random unit directions of all N + 2 parameters (each direction checks the
whole gradient: (f(theta + h u) - f(theta - h u)) / 2h against grad . u).
The start is the exciton-on-site-0 product state expanded to D = 128 (direct sum with a
seeded random MPS of weight eps, right-canonical, full-rank Schmidt spectra).

The forward's Lanczos runs to 1e-12 with up to 30 vectors (--krylov tight) because the
adjoint differentiates the exact exponential: the reference's setting (tolerance 1e-8 on the
product of residual norms, 15 iterations; --krylov reference) makes each local step's
truncation error, and the integer iteration count behind it, a discontinuous function of
the couplings, and central differences at h = 1e-4 divide that error by 2h.  --krylov
reference runs the forward that way as well, to measure the difference.

    python kylin-v_amd/run_adjoint.py [--steps 30] [--fd 0,11,23,24,25] [--dirs 3]
                                      [--nodes 4] [--krylov tight|reference]
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
    ap.add_argument("--N", type=int, default=24)
    ap.add_argument("--L", type=int, default=12)
    ap.add_argument("--D", type=int, default=128)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--dt", type=float, default=0.05)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--eps", type=float, default=0.05)
    ap.add_argument("--nodes", type=int, default=4)
    ap.add_argument("--fd", default="0,11,23,24,25")
    ap.add_argument("--h", type=float, default=1e-4)
    ap.add_argument("--dirs", type=int, default=0, help="random directions for central differences")
    ap.add_argument("--krylov", default="tight", choices=["tight", "reference"])
    args = ap.parse_args()
    import numpy as np
    import __graft_entry__ as GE
    kv = GE.load_package()
    from kylin_v_amd import adjoint as A
    from kylin_v_amd import model as KM

    N = args.N
    rng = np.random.default_rng(args.seed)
    theta = np.concatenate([rng.uniform(0.8, 1.2, N), [1.0, 1.0]])
    sites, terms, init, ex = KM.chain(N, [1.0], np.ones((N, 1)), [args.L], 1.0, "merged")
    psi0 = KM.expanded_state_blocks(sites, init, args.D, seed=args.seed + 1, eps=args.eps)
    bonds = [sum(KM._bond_dims(b, 2).values()) for b in psi0[:-1]]
    kry = dict(krylov_tol=1e-12, krylov_maxiter=30) if args.krylov == "tight" else \
        dict(krylov_tol=1e-8, krylov_maxiter=15)
    eng = A.OneSiteAdjoint(N, args.L, args.dt, args.steps, psi0,
                           A.DeviceOps(kv.default_context(), **kry), nodes=args.nodes)

    def log(msg):
        print(f"adjoint: {msg}", file=sys.stderr, flush=True)

    def prog(phase, k):
        log(f"{phase} step {k} done")

    t0 = time.perf_counter()
    c = eng.forward(theta, prog)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    grad = eng.backward(prog)
    t_bwd = time.perf_counter() - t0
    log(f"forward {t_fwd:.1f} s, backward {t_bwd:.1f} s")
    fd = {}
    for p in [int(x) for x in args.fd.split(",") if x != ""]:
        tp, tm = theta.copy(), theta.copy()
        tp[p] += args.h
        tm[p] -= args.h
        fp = abs(eng.forward(tp)) ** 2
        fm = abs(eng.forward(tm)) ** 2
        fd[p] = (fp - fm) / (2 * args.h)
        log(f"fd param {p}: {fd[p]:.10e} adjoint {grad[p]:.10e}")
    scale = float(np.abs(grad).max())
    rel = {p: abs(fd[p] - grad[p]) / scale for p in fd}
    drng = np.random.default_rng(args.seed + 7)
    dirs = []
    for k in range(args.dirs):
        u = drng.standard_normal(theta.size)
        u /= np.linalg.norm(u)
        fp = abs(eng.forward(theta + args.h * u)) ** 2
        fm = abs(eng.forward(theta - args.h * u)) ** 2
        d_fd, d_ad = (fp - fm) / (2 * args.h), float(grad @ u)
        dirs.append(dict(fd=d_fd, adjoint=d_ad, rel_err_vs_grad_norm=abs(d_fd - d_ad) / float(np.linalg.norm(grad))))
        log(f"fd direction {k}: {d_fd:.10e} adjoint {d_ad:.10e}")
    print(json.dumps({"workload": f"one-site TDVP N={N} L={args.L} D={args.D} "
                                  f"{args.steps} steps dt={args.dt}",
                      "max_bond": max(bonds), "value": abs(c) ** 2, "C_T": [c.real, c.imag],
                      "forward_s": t_fwd, "backward_s": t_bwd, "grad": list(map(float, grad)),
                      "fd_h": args.h, "fd": {str(k): v for k, v in fd.items()},
                      "fd_rel_err_vs_max_grad": {str(k): v for k, v in rel.items()},
                      "max_fd_rel_err": max(rel.values()) if rel else None,
                      "krylov": kry, "directions": dirs,
                      "max_dir_rel_err": max(d["rel_err_vs_grad_norm"] for d in dirs) if dirs else None}),
          flush=True)


if __name__ == "__main__":
    main()
