"""The 8-site Holstein chain (one mode at 8 levels, J = w = g = 1, 2TDVP dt = 0.05, 40
steps, D <= 32, cutoff 1e-10) on the GPU against the oracle's fixtures
(tests/golden/chain8_<name>.json, written by tests/golden/make_chain8_fixture.py).  "pp_nk"
runs the projected-purification layout with three global Krylov steps before the 40
two-site steps (stage list [3, 40, 0, 0]), so the exciton travels: its population on the
second site must exceed 1e-3 by the last step."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import __graft_entry__ as GE  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["merged", "pp", "pp_nk"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_chain8_run_matches_oracle_fixture(name, precision):
    print(f"case chain8 {name} {precision}", flush=True)
    kv = GE.load_package()
    from kylin_v_amd import model as KM
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", f"chain8_{name}.json")))
    N, layout = ref["N"], ref["layout"]
    ch = KM.HolsteinChain(N, [ref["omega"]], ref["g"], [ref["L"]], ref["J"], layout)
    eng = kv.TDVP(ch.psi0, ch.mpo, cutoff=ref["cutoff"], maxdim=ref["D"], dt=ref["dt"],
                  precision=precision)
    for k, (re, im) in enumerate(ref.get("C_global", [])):
        eng.global_krylov()
        c = kv.overlap_states(ch.psi0, eng.mps)
        assert abs(c - complex(re, im)) < 1e-9, f"global Krylov step {k + 1}"
    C = []
    pops = []
    for k in range(1, ref["steps"] + 1):
        eng.two_site()
        C.append(kv.overlap_states(ch.psi0, eng.mps))
        if k in ref["pop_steps"]:
            mps = eng.mps
            pops.append([kv.join(ch.density(i), mps, mps).real for i in range(N)])
    C = np.array(C)
    Cref = np.array([complex(a, b) for a, b in ref["C"]])
    err_c = np.abs(C - Cref).max()
    err_p = np.abs(np.array(pops) - np.array(ref["pops"])).max()
    print(f"chain8 {name} {precision}: max|dC| = {err_c:.3e}, max|dP| = {err_p:.3e}", flush=True)
    tol = 1e-8 if precision == "fp64" else 1e-6
    assert err_c < tol and err_p < tol
    assert abs(np.array(pops[-1]).sum() - 1.0) < 1e-8
    if name == "pp_nk":
        assert ref["pops"][-1][1] > 1e-3 and pops[-1][1] > 1e-3  # the exciton left site 0
