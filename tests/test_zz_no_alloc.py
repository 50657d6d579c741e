"""The allocation-free overlap entry point (kv_tensor_overlap_noalloc) against the
allocating one (kv_tensor_overlap) through the C ABI: the same kernels in the same order, so
the results must be bit-identical; both against numpy's vdot to rounding.  Two calls in a
row, the second on smaller tensors (the scratch keeps the first call's contents), and the
limit on matching block pairs."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import __graft_entry__ as GE  # noqa: E402

pytestmark = pytest.mark.gpu


def rand_tensor(kv, rng, a_dims, s_dims, b_dims):
    blocks = {}
    for qa, da in a_dims.items():
        for qs, ds in s_dims.items():
            qb = qa + qs
            if qb in b_dims:
                sh = (da, ds, b_dims[qb])
                blocks[(qa, qs, qb)] = rng.standard_normal(sh) + 1j * rng.standard_normal(sh)
    return kv.Tensor.from_blocks((1, 1, -1), blocks), blocks


def test_noalloc_overlap_matches_allocating_overlap():
    kv = GE.load_package()
    rng = np.random.default_rng(7)
    cases = [({0: 200, 1: 150, 2: 90}, {0: 16, 1: 16}, {0: 180, 1: 220, 2: 64, 3: 10}),
             ({0: 7, 1: 3}, {0: 2, 1: 5}, {0: 4, 1: 9, 2: 1})]
    for a_dims, s_dims, b_dims in cases:  # the second call is much smaller than the first
        print(f"case noalloc overlap a={a_dims}", flush=True)
        ta, ba = rand_tensor(kv, rng, a_dims, s_dims, b_dims)
        tb, bb = rand_tensor(kv, rng, a_dims, s_dims, b_dims)
        got = kv.overlap_noalloc(ta, tb)
        ref = kv.overlap(ta, tb)
        assert got == ref  # bit-identical
        exact = sum(np.vdot(bb[q], ba[q]) for q in ba)
        scale = sum(np.abs(bb[q]).ravel() @ np.abs(ba[q]).ravel() for q in ba)
        assert abs(got - exact) <= 1e-13 * scale
    # tensors with no common blocks: zero, with the scratch still holding the last result
    ta, _ = rand_tensor(kv, rng, {0: 3}, {0: 2}, {0: 3})
    tb, _ = rand_tensor(kv, rng, {1: 3}, {0: 2}, {1: 3})
    assert kv.overlap_noalloc(ta, tb) == 0j


def test_noalloc_overlap_rejects_more_pairs_than_its_scratch_holds():
    kv = GE.load_package()
    print("case noalloc overlap limit", flush=True)
    rng = np.random.default_rng(8)
    n = 2100  # matching blocks, above the 2048 the scratch holds
    ta, _ = rand_tensor(kv, rng, {q: 1 for q in range(n)}, {0: 1}, {q: 1 for q in range(n)})
    with pytest.raises(kv.KVError) as e:
        kv.overlap_noalloc(ta, ta)
    assert e.value.code == -1 and "2048" in str(e.value)
    ref = kv.overlap(ta, ta)  # the allocating entry point has no such limit
    assert abs(ref.real - ta.norm() ** 2) < 1e-12 * n
