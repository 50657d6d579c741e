"""CPU checks of the batching workload's host side (kylin-v_amd/batch.py): the seeded set of
64 Holstein chains and the order in which workers take them.  The GPU run of the same set
is tests/test_gpu_batch.py."""
import importlib.util
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _batch():
    spec = importlib.util.spec_from_file_location("kv_batch", os.path.join(ROOT, "kylin-v_amd", "batch.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_seeded_set_is_reproducible_and_in_range():
    """Lengths U{16..48}, levels {8, 12, 16}, J ~ U[0.5, 1.5], w ~ U[0.3, 1.5],
    g ~ U[0.5, 2.0], bond caps U{64..256}, drawn in a fixed order from default_rng(3)."""
    B = _batch()
    a, b = B.seeded_chains(64, 3), B.seeded_chains(64, 3)
    assert a == b and len(a) == 64
    assert a != B.seeded_chains(64, 4)
    for s in a:
        assert 16 <= s["N"] <= 48 and s["L"] in (8, 12, 16) and 64 <= s["D"] <= 256
        assert 0.5 <= s["J"] <= 1.5 and 0.3 <= s["omega"] <= 1.5 and 0.5 <= s["g"] <= 2.0
    # the first draw, restated from the generator directly
    rng = np.random.default_rng(3)
    first = dict(N=int(rng.integers(16, 49)), L=int(rng.choice([8, 12, 16])), J=float(rng.uniform(0.5, 1.5)),
                 omega=float(rng.uniform(0.3, 1.5)), g=float(rng.uniform(0.5, 2.0)), D=int(rng.integers(64, 257)))
    assert a[0] == first
    # the set spans its ranges (64 draws)
    assert len({s["L"] for s in a}) == 3 and min(s["N"] for s in a) < 24 and max(s["N"] for s in a) > 40


def test_workers_take_the_costliest_chains_first():
    """Longest-processing-time-first: the cost N * L^2 * D orders the hand-out, so the batch
    does not end on one long chain."""
    B = _batch()
    specs = B.seeded_chains(64, 3)
    order = sorted(range(len(specs)), key=lambda j: -B.chain_cost(specs[j]))
    costs = [B.chain_cost(specs[j]) for j in order]
    assert costs == sorted(costs, reverse=True)
    big = max(specs, key=lambda s: s["N"] * s["L"] ** 2 * s["D"])
    assert specs[order[0]] == big
    assert B.MAX_WORKERS == 8
