"""Writes tests/golden/chain8_<layout>.json: the oracle's C(t), exciton populations and bond
dimensions for the 8-site Holstein chain (exciton x one mode at 8 levels, J = w = g = 1,
two-site TDVP dt = 0.05, 40 steps, D <= 32, cutoff 1e-10, Krylov tol 1e-8, 15 iterations).
"pp_nk" is the projected-purification run with the stage list [3, 40, 0, 0]: three global
Krylov steps (holstein/krylov.h:206-284 as driven by holstein/tdvp.h:318-327) open the
exciton-hop sectors that two-site TDVP alone never creates in that layout, then 40 two-site
steps move the exciton along the chain.

The oracle (oracle/) is the numpy restatement of the reference's algorithm, pinned by
tests/test_oracle.py and tests/test_oracle_krylov.py; these files let the GPU test compare
the whole run without re-running the oracle trajectory inside the GPU test.

    python tests/golden/make_chain8_fixture.py [merged|pp|pp_nk ...]
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import model as M  # noqa: E402
from oracle import tdvp as T  # noqa: E402


def run(name, steps=40):
    from oracle import krylov as K
    layout = "pp" if name == "pp_nk" else name
    nk = 3 if name == "pp_nk" else 0
    p = M.chain8_params()
    lat = M.holstein(p["N"], p["omegas"], p["gs"], p["Ls"], p["J"], layout)
    mpo = M.build_mpo(lat)
    psi0 = M.product_state(lat)
    out = {"layout": layout, "N": p["N"], "L": p["Ls"][0], "J": p["J"], "omega": p["omegas"][0],
           "g": 1.0, "dt": 0.05, "D": 32, "cutoff": 1e-10, "steps": steps, "global_krylov_steps": nk,
           "C_global": [], "C": [], "pop_steps": [], "pops": [], "bonds": []}
    start = psi0
    for k in range(nk):  # exp(-i dt H) on the whole MPS, truncated with the run's cutoff and cap
        start, _ = K.global_main(K.scale_mpo(mpo, 0.05), start, 1e-10, 32)
        c = T.overlap_states(psi0, start)
        out["C_global"].append([c.real, c.imag])
        print(name, "global", k, c, flush=True)
    eng = T.TDVP(start, mpo, cutoff=1e-10, maxdim=32, dt=0.05, tol_krylov=1e-8, max_iter=15)
    dens = [M.density_mpo(lat, i) for i in range(p["N"])]
    for k in range(1, steps + 1):
        eng.two_site()
        c = T.overlap_states(psi0, eng.mps)
        out["C"].append([c.real, c.imag])
        if k % 10 == 0:
            out["pop_steps"].append(k)
            out["pops"].append([T.join(d, eng.mps, eng.mps).real for d in dens])
            out["bonds"].append(eng.bond_dims())
        print(name, k, c, flush=True)
    return out


if __name__ == "__main__":
    for name in (sys.argv[1:] or ["merged", "pp", "pp_nk"]):
        d = run(name)
        with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), f"chain8_{name}.json"), "w") as f:
            json.dump(d, f)
