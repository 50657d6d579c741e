"""CPU pins of the gradient references: the exact-diagonalisation closed form against
central differences of the dense propagator, and the oracle's one-site TDVP (no
truncation) against the exact dynamics on a full-rank state."""
import os
import sys

import numpy as np
import scipy.linalg

from oracle import blocksparse as B
from oracle import model as M

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import adjoint_refs as AR  # noqa: E402

ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
import __graft_entry__ as GE  # noqa: E402


def _exact_value(theta, N, L, v0, T):
    g, J, w = AR.theta_split(theta, N)
    H = M.ed_hamiltonian(N, [w], g.reshape(N, 1), [L], J)
    return abs(np.vdot(v0, scipy.linalg.expm(-1j * H * T) @ v0)) ** 2


def _full_rank_state(N, L, seed):
    GE.load_package()
    from kylin_v_amd import model as KM
    sites, terms, init, ex = KM.chain(N, [1.0], np.ones((N, 1)), [L], 1.0, "merged")
    return KM.expanded_state_blocks(sites, init, 10 ** 4, seed=seed, eps=0.3)


def _dense(N, L, blocks):
    lat = M.holstein(N, [1.0], np.ones((N, 1)), [L], 1.0, "merged")
    return M.mps_to_dense(lat, [B.Tensor((1, 1, -1), b) for b in blocks])


def test_ed_gradient_matches_central_differences():
    N, L, T = 3, 2, 0.7
    rng = np.random.default_rng(0)
    theta = np.concatenate([rng.uniform(0.8, 1.2, N), [0.9, 1.1]])
    v0 = _dense(N, L, _full_rank_state(N, L, 1))
    grad, _ = AR.ed_gradient(theta, N, L, v0, T)
    h = 1e-6
    fd = np.zeros_like(grad)
    for p in range(len(theta)):
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (_exact_value(tp, N, L, v0, T) - _exact_value(tm, N, L, v0, T)) / (2 * h)
    assert np.abs(grad - fd).max() < 1e-7 * max(1.0, np.abs(fd).max())


def test_oracle_one_site_is_exact_on_a_full_rank_state():
    """At full bond dimension the one-site projector is the identity: the oracle's
    one-site TDVP reproduces exp(-iHT)."""
    N, L, dt, steps = 3, 2, 0.1, 4
    theta = np.array([1.0, 0.9, 1.1, 0.8, 1.2])
    blocks = _full_rank_state(N, L, 2)
    v0 = _dense(N, L, blocks)
    got = AR.oracle_autocorr(theta, N, L, blocks, dt, steps)
    assert abs(got - _exact_value(theta, N, L, v0, dt * steps)) < 1e-9
