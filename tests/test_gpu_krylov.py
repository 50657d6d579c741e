"""Global Krylov stage on the device (kv_tdvp_global_krylov, kln-tdvp nK stage) against the
oracle's restatement of holstein/krylov.h and against exact dynamics."""
import os
import sys

import numpy as np
import pytest
import scipy.linalg

from oracle import blocksparse as B
from oracle import krylov as K
from oracle import model as M
from oracle import tdvp as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import __graft_entry__ as GE  # noqa: E402

pytestmark = pytest.mark.gpu


def _down(t):
    dirs, blocks = t.to_blocks()
    return B.Tensor(dirs, blocks)


def _up(kv, t):
    return kv.Tensor.from_blocks(t.dirs, t.blocks)


def test_global_krylov_matches_oracle_and_exact_dynamics():
    print("case global krylov merged", flush=True)
    kv = GE.load_package()
    N, L, dt = 4, 3, 0.1
    g = np.array([1.0, 0.8, 1.2, 0.9]).reshape(N, 1)
    lat = M.holstein(N, [1.0], g, [L], 1.0, "merged")
    mpo, psi = M.build_mpo(lat), M.product_state(lat)
    eng = kv.TDVP([_up(kv, t) for t in psi], [_up(kv, w) for w in mpo], cutoff=1e-12,
                  maxdim=1000, dt=dt)
    ref = psi
    v0 = M.mps_to_dense(lat, psi)
    Hd = M.ed_hamiltonian(N, [1.0], g, [L], 1.0)
    for step in range(1, 3):
        n = eng.global_krylov()
        ref, nref = K.global_main(K.scale_mpo(mpo, dt), ref, 1e-12, 1000)
        assert n == nref
        got = M.mps_to_dense(lat, [_down(t) for t in eng.mps])
        np.testing.assert_allclose(got, M.mps_to_dense(lat, ref), atol=1e-10)
        np.testing.assert_allclose(got, scipy.linalg.expm(-1j * Hd * dt * step) @ v0, atol=1e-10)
    # the two-site sweep after the stage starts from rebuilt environments
    eng.two_site()
    oeng = T.TDVP(ref, mpo, cutoff=1e-12, maxdim=1000, dt=dt)
    oeng.two_site()
    assert abs(kv.overlap_states([_up(kv, t) for t in psi], eng.mps)
               - T.overlap_states(psi, oeng.mps)) < 1e-10


def test_global_krylov_opens_pp_sectors_like_oracle():
    print("case global krylov pp", flush=True)
    kv = GE.load_package()
    lat = M.holstein(2, [1.0], 1.0, [3], 1.0, "pp")
    mpo, psi = M.build_mpo(lat), M.product_state(lat)
    eng = kv.TDVP([_up(kv, t) for t in psi], [_up(kv, w) for w in mpo], cutoff=1e-12,
                  maxdim=1000, dt=0.1)
    eng.global_krylov()
    ref, _ = K.global_main(K.scale_mpo(mpo, 0.1), psi, 1e-12, 1000)
    got = [_down(t) for t in eng.mps]
    assert [sum(t.legs()[2].values()) for t in got[:-1]] == \
        [sum(t.legs()[2].values()) for t in ref[:-1]]
    assert abs(T.overlap_states(ref, got) - 1.0) < 1e-10
