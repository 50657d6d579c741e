"""One-site TDVP adjoint on the device (kylin-v_amd/adjoint.py) against the oracle's FP64
forward under central differences, the device forward under central differences, and the
exact-dynamics closed form on a full-rank state (where one-site TDVP is exact)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import adjoint_refs as AR  # noqa: E402

from oracle import blocksparse as B  # noqa: E402
from oracle import model as M  # noqa: E402

ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
import __graft_entry__ as GE  # noqa: E402

pytestmark = pytest.mark.gpu


def _setup(N, L, D, seed, eps):
    kv = GE.load_package()
    from kylin_v_amd import model as KM
    sites, terms, init, ex = KM.chain(N, [1.0], np.ones((N, 1)), [L], 1.0, "merged")
    return kv, KM.expanded_state_blocks(sites, init, D, seed=seed, eps=eps)


def _engine(kv, N, L, dt, steps, psi0):
    from kylin_v_amd import adjoint as A
    return A.OneSiteAdjoint(N, L, dt, steps, psi0, A.DeviceOps(kv.default_context()), nodes=4)


def test_adjoint_matches_oracle_and_device_differences():
    print("case adjoint N=4 L=3 D=10 steps=3", flush=True)
    N, L, D, steps, dt = 4, 3, 10, 3, 0.1
    rng = np.random.default_rng(7)
    theta = np.concatenate([rng.uniform(0.8, 1.2, N), [1.0, 0.9]])
    kv, psi0 = _setup(N, L, D, 5, 0.1)
    eng = _engine(kv, N, L, dt, steps, psi0)
    f, grad = eng.value_and_grad(theta)
    assert abs(f - AR.oracle_autocorr(theta, N, L, psi0, dt, steps)) < 1e-10
    ref = AR.oracle_fd_gradient(theta, N, L, psi0, dt, steps)
    scale = np.abs(ref).max()
    print("grad", grad, "\noracle fd", ref, flush=True)
    assert np.abs(grad - ref).max() < 1e-6 * scale
    h = 1e-5
    for p in (0, N, N + 1):
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd = (abs(eng.forward(tp)) ** 2 - abs(eng.forward(tm)) ** 2) / (2 * h)
        assert abs(fd - grad[p]) < 1e-6 * scale


def test_adjoint_on_full_rank_state_matches_exact_dynamics():
    print("case adjoint exact N=3 L=2", flush=True)
    N, L, dt, steps = 3, 2, 0.1, 4
    theta = np.array([1.0, 0.9, 1.1, 0.8, 1.2])
    kv, psi0 = _setup(N, L, 10 ** 4, 2, 0.3)
    lat = M.holstein(N, [1.0], np.ones((N, 1)), [L], 1.0, "merged")
    v0 = M.mps_to_dense(lat, [B.Tensor((1, 1, -1), b) for b in psi0])
    ref, C = AR.ed_gradient(theta, N, L, v0, dt * steps)
    eng = _engine(kv, N, L, dt, steps, psi0)
    f, grad = eng.value_and_grad(theta)
    assert abs(f - abs(C) ** 2) < 1e-9
    assert np.abs(grad - ref).max() < 1e-7 * np.abs(ref).max()


def test_autograd_function_returns_adjoint_gradient():
    print("case adjoint autograd", flush=True)
    import torch
    N, L, D, steps, dt = 3, 3, 8, 2, 0.1
    theta0 = np.array([1.1, 0.9, 1.0, 1.0, 0.8])
    kv, psi0 = _setup(N, L, D, 3, 0.1)
    from kylin_v_amd import adjoint as A
    eng = _engine(kv, N, L, dt, steps, psi0)
    th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    loss = A.TdvpAutocorr.apply(th, eng)
    loss.backward()
    f, grad = _engine(kv, N, L, dt, steps, psi0).value_and_grad(theta0)
    assert abs(loss.item() - f) < 1e-13
    assert np.abs(th.grad.numpy() - grad).max() < 1e-12
