"""SA-1TDVP on the GPU against the oracle: the enrichment SVDs of whole sweeps (bond
dimensions equal at every step, C(t) and the sites' content against the oracle's state),
the SVS log of the two-site sweeps, and the state invariants an enrichment must keep."""
import os
import sys

import numpy as np
import pytest

from oracle import model as M
from oracle import sa
from oracle import tdvp as T
from oracle.blocksparse import Tensor as OT

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import __graft_entry__ as GE  # noqa: E402

pytestmark = pytest.mark.gpu


def to_gpu(kv, sites):
    return [kv.Tensor.from_blocks(t.dirs, dict(t.blocks)) for t in sites]


def to_oracle(sites):
    return [OT(*s.to_blocks()) for s in sites]


@pytest.mark.parametrize("layout", ["merged", "pp"])
def test_sa_sweeps_match_oracle(tmp_path, layout):
    """Two-site sweeps capped at D = 6 (with the SVS log), then SA-1TDVP sweeps at D = 40
    under fixed widths and seeds: the enrichment adds states and the GPU follows the
    oracle.  In the merged layout the enrichment counts are stable (the oracle's own run
    keeps every bond dimension under 1e-13 noise on its start), so the bond dimensions must
    be equal and C(t) agree to 1e-10.  In the pp layout the counts hang on singular values
    near 1e-8 that 1e-13 start noise moves by O(1) in the oracle itself (its C(t) then
    moves by 1.5e-10, the populations by 4e-9): there the state must agree (infidelity
    1e-8, C(t) 1e-8, populations 1e-7) and the bond dimensions may differ."""
    print(f"case sa sweeps {layout}", flush=True)
    kv = GE.load_package()
    N, Ls, om, J, dt = 4, [4], [1.0], 1.0, 0.05
    g = np.array([0.9, 1.1, 0.8, 1.0]).reshape(4, 1)
    lat = M.holstein(N, om, g, Ls, J, layout)
    mpo = M.build_mpo(lat)
    psi0 = M.product_state(lat)
    if layout == "pp":  # open the exciton-hop sectors first (a global Krylov prelude)
        from oracle import krylov as K
        psi0, _ = K.global_main(K.scale_mpo(mpo, dt), psi0, 1e-10, 24)
    ref = T.TDVP(psi0, mpo, cutoff=1e-10, maxdim=6, dt=dt)
    gpu = kv.TDVP(to_gpu(kv, psi0), to_gpu(kv, mpo), cutoff=1e-10, maxdim=6, dt=dt)
    svs = tmp_path / "SVS"
    gpu.svs_open(str(svs))
    for _ in range(6):
        ref.two_site()
        gpu.two_site()
    assert gpu.bond_dims() == ref.bond_dims()
    # the SVS log: the oracle's records.  Capped at D = 6 the sweeps truncate hard, and the
    # two runs' kept spectra agree to ~1e-11 absolute (s_max = 1): each entry, a log-ratio
    # of three kept values, to 1e-10 + 2e-10 (1/s_i + 1/s_{i-2} + 2/s_{i-1})
    got = sa.parse_svs(svs.read_text())
    ns = len(psi0)  # MPS sites (the pp layout has phonon sites too)
    want = sa.parse_svs(sa.svs_text(ns, ref.svs))

    def per_entry(f):
        return sa.parse_svs(sa.svs_text(ns, [(len(v), [f(v[i], v[i - 1], v[i - 2]) for i in range(2, len(v))])
                                             for v in ref.svs_values]))

    tol = per_entry(lambda a, b, c: 1e-10 + 2e-10 * (1 / a + 1 / c + 2 / b))
    assert [len(x) for x in got] == [len(x) for x in want]
    for a, b, t in zip(got, want, tol):
        assert np.all(np.abs(np.array(a) - np.array(b)) <= np.array(t))
    widths = sa.fit_widths(svs.read_text())
    betas = [1.0 / w for w in widths]
    start = ref.mps
    ref_sa = T.TDVP(start, mpo, cutoff=1e-10, maxdim=40, dt=dt)
    gpu_sa = kv.TDVP(to_gpu(kv, start), to_gpu(kv, mpo), cutoff=1e-10, maxdim=40, dt=dt)
    psi_ref = [t.copy() for t in psi0]
    dims = []
    exact = layout == "merged"
    for k in range(4):
        ref_sa.one_site(betas=betas, seed=77 + k)
        gpu_sa.sa_one_site(betas, 77 + k)
        dims.append((ref_sa.bond_dims(), gpu_sa.bond_dims()))
        if exact:
            assert gpu_sa.bond_dims() == ref_sa.bond_dims(), k
        got_sites = to_oracle(gpu_sa.mps)
        c_gpu = T.overlap_states(psi_ref, got_sites)
        c_ref = T.overlap_states(psi_ref, ref_sa.mps)
        assert abs(c_gpu - c_ref) < (1e-10 if exact else 1e-8), k
        # same state: |<ref|gpu>| = 1 (the sites agree up to gauge)
        assert abs(abs(T.overlap_states(ref_sa.mps, got_sites)) - 1.0) < (1e-10 if exact else 1e-8)
        if not exact:
            for i in range(N):
                n_i = M.density_mpo(lat, i)
                assert abs(T.join(n_i, got_sites, got_sites) - T.join(n_i, ref_sa.mps, ref_sa.mps)) < 1e-7
    assert max(dims[-1][0]) > 6 and max(dims[-1][1]) > 6
    assert gpu_sa.stats()["sa_added"] > 0
    print("bond dims", dims, flush=True)


def test_sa_sweep_without_room_is_the_one_site_step():
    """With the bond cap at the current spectrum sizes nothing can be added: the SA step
    is then the untruncated one-site step."""
    print("case sa no room", flush=True)
    kv = GE.load_package()
    N, Ls, om, J, dt = 3, [3], [1.0], 1.0, 0.05
    g = np.array([1.0, 0.8, 1.2]).reshape(3, 1)
    lat = M.holstein(N, om, g, Ls, J, "merged")
    mpo = M.build_mpo(lat)
    psi0 = M.product_state(lat)
    ref = T.TDVP(psi0, mpo, cutoff=1e-10, maxdim=5, dt=dt)
    for _ in range(4):
        ref.two_site()
    start = ref.mps
    a = kv.TDVP(to_gpu(kv, start), to_gpu(kv, mpo), cutoff=-1.0, maxdim=5, dt=dt)
    b = kv.TDVP(to_gpu(kv, start), to_gpu(kv, mpo), cutoff=-1.0, maxdim=5, dt=dt)
    for k in range(2):
        a.one_site()
        b.sa_one_site([2.0] * (N - 1), k)
    assert a.bond_dims() == b.bond_dims()
    assert b.stats()["sa_added"] == 0
    assert abs(abs(T.overlap_states(to_oracle(a.mps), to_oracle(b.mps))) - 1.0) < 1e-12


def test_sa_rejects_missing_widths():
    kv = GE.load_package()
    lat = M.holstein(3, [1.0], 1.0, [3], 1.0, "merged")
    psi0 = M.product_state(lat)
    d = kv.TDVP(to_gpu(kv, psi0), to_gpu(kv, M.build_mpo(lat)), cutoff=1e-10, maxdim=8, dt=0.05)
    with pytest.raises(kv.KVError):
        d.sa_one_site([1.0], 0)
    with pytest.raises(kv.KVError):
        d.sa_one_site([1.0, -2.0], 0)


def test_sa_sweeps_in_mixed_precision_follow_fp64():
    """The SA step with FP32 H_eff operands (FP64 enrichment SVD, Lanczos scalars and
    environments) against the FP64 SA step from the same state, widths and seeds: the
    states agree to the FP32 contraction error."""
    print("case sa mixed precision", flush=True)
    kv = GE.load_package()
    N, Ls, om, J, dt = 4, [3], [1.0], 1.0, 0.05
    g = np.array([0.9, 1.1, 0.8, 1.0]).reshape(4, 1)
    lat = M.holstein(N, om, g, Ls, J, "merged")
    mpo = M.build_mpo(lat)
    psi0 = M.product_state(lat)
    ref = T.TDVP(psi0, mpo, cutoff=1e-10, maxdim=6, dt=dt)
    for _ in range(6):
        ref.two_site()
    betas = [1.0 / 0.3] * (N - 1)
    runs = {}
    for prec in ("fp64", "fp32"):
        d = kv.TDVP(to_gpu(kv, ref.mps), to_gpu(kv, mpo), cutoff=1e-10, maxdim=40, dt=dt, precision=prec)
        for k in range(3):
            d.sa_one_site(betas, 11 + k)
        runs[prec] = (d.bond_dims(), to_oracle(d.mps), d.stats()["sa_added"])
    (d64, m64, a64), (d32, m32, a32) = runs["fp64"], runs["fp32"]
    assert a64 > 0 and a32 > 0
    assert abs(T.overlap_states(psi0, m64) - T.overlap_states(psi0, m32)) < 1e-5
    assert 1 - abs(T.overlap_states(m64, m32)) < 1e-5
