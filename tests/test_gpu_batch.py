"""Concurrent chains (kylin-v_amd/batch.py): results equal the chains run one at a time,
and one chain equals the oracle."""
import os
import sys

import numpy as np
import pytest

from oracle import model as M
from oracle import tdvp as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import __graft_entry__ as GE  # noqa: E402

pytestmark = pytest.mark.gpu


def test_concurrent_chains_match_serial_and_oracle():
    print("case batch", flush=True)
    GE.load_package()
    from kylin_v_amd import batch
    specs = [dict(N=5, L=4, J=1.0, omega=1.0, g=0.8, D=16), dict(N=6, L=3, J=0.7, omega=0.5, g=1.5, D=12),
             dict(N=4, L=5, J=1.3, omega=1.2, g=1.0, D=20), dict(N=7, L=3, J=0.9, omega=0.8, g=0.6, D=10)]
    par, _ = batch.run_batch(specs, 3, workers=3)
    ser, _ = batch.run_batch(specs, 3, workers=1)
    for a, b in zip(par, ser):
        assert np.abs(np.array(a["C"]) - np.array(b["C"])).max() < 1e-12
        assert a["bonds"] == b["bonds"]
    s = specs[1]
    lat = M.holstein(s["N"], [s["omega"]], s["g"], [s["L"]], s["J"], "merged")
    psi0 = M.product_state(lat)
    ref = T.TDVP(psi0, M.build_mpo(lat), cutoff=1e-10, maxdim=s["D"], dt=0.05)
    for k in range(3):
        ref.two_site()
        assert abs(T.overlap_states(psi0, ref.mps) - par[1]["C"][k]) < 1e-9


def test_seeded_workload_chains_match_oracle():
    """Chains 36, 37 and 45 of the batching workload (seeded_chains(64, seed=3): N=16, L=8,
    the smallest of the set), run together for 2 two-site steps, each against the oracle:
    C(t) to 1e-9 and the same bond dimensions."""
    print("case batch seeded chains", flush=True)
    GE.load_package()
    from kylin_v_amd import batch
    every = batch.seeded_chains(64, seed=3)
    pick = [36, 37, 45]
    specs = [every[i] for i in pick]
    assert all(s["N"] == 16 and s["L"] == 8 for s in specs)
    res, _ = batch.run_batch(specs, 2, workers=3)
    for s, r in zip(specs, res):
        lat = M.holstein(s["N"], [s["omega"]], s["g"], [s["L"]], s["J"], "merged")
        psi0 = M.product_state(lat)
        ref = T.TDVP(psi0, M.build_mpo(lat), cutoff=1e-10, maxdim=s["D"], dt=0.05)
        for k in range(2):
            ref.two_site()
            assert abs(T.overlap_states(psi0, ref.mps) - r["C"][k]) < 1e-9, (s, k)
        assert list(r["bonds"]) == [int(b) for b in ref.bond_dims()], s
