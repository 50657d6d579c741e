"""GPU parity on the 32-site chain the benchmark runs (N=32, 16 phonon levels, w=0.5, J=1,
g_i ~ U[0.8, 1.2] from numpy default_rng(2024), merged layout, d=32), against the oracle.

(a) Two-site TDVP from the physical start (exciton on site 0, phonon vacuum), dt=0.02,
    D<=256, cutoff 1e-10: C(t) every step, bond dimensions, populations and the bond
    energies E_{i,i+1} = <J(a+_i a_{i+1} + h.c.)> at the last step.
(b) The benchmark's middle bond at full size: the seeded synthetic right-canonical D=256
    MPS (seed 7) centred on sites (15, 16), where the two-site vector has four
    128 x 16 x 16 x 128 blocks and the bond SVD two sector matrices of 2048 x 4096.
    apply_double, local_krylov and svd_impl against the oracle on the same inputs.

Bars: FP64 C(t) 1e-9 (the oracle's own Krylov tolerance is 1e-8 on a product of
residuals; both run the same recurrence); FP32 H_eff C(t) 1e-7; FP64 contractions 1e-12 and
FP32 2e-5 of the same contraction over absolute values; the truncated SVD product 1e-9 of
its scale.
"""
import os
import sys

import numpy as np
import pytest

from oracle import model as M
from oracle import tdvp as T
from oracle.blocksparse import Tensor as OT
from oracle.blocksparse import norm as onorm
from oracle.blocksparse import product as oproduct

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import __graft_entry__ as GE  # noqa: E402

pytestmark = pytest.mark.gpu

N, L, OMEGA, J, DT, D, CUTOFF, MID = 32, 16, 0.5, 1.0, 0.02, 256, 1e-10, 15


@pytest.fixture(scope="module")
def kv():
    return GE.load_package()


@pytest.fixture(scope="module")
def chain(kv):
    from kylin_v_amd import model as KM
    p = M.chain32_params()
    ch = KM.HolsteinChain(N, [OMEGA], p["gs"].ravel(), [L], J, "merged")
    lat = M.holstein(N, p["omegas"], p["gs"], p["Ls"], p["J"], "merged")
    return ch, lat, M.build_mpo(lat)


def up(kv, t: OT):
    return kv.Tensor.from_blocks(t.dirs, t.blocks)


def down(t) -> OT:
    dirs, blocks = t.to_blocks()
    return OT(dirs, blocks)


def max_diff(got: OT, ref: OT):
    assert tuple(got.dirs) == tuple(ref.dirs)
    d = 0.0
    for q in set(got.blocks) | set(ref.blocks):
        a, b = got.blocks.get(q), ref.blocks.get(q)
        if a is None:
            d = max(d, np.abs(b).max())
        elif b is None:
            d = max(d, np.abs(a).max())
        else:
            d = max(d, np.abs(a - b).max())
    return d


def scale(t: OT):
    return max(np.abs(b).max() for b in t.blocks.values())


def absolute(t: OT) -> OT:
    return OT(t.dirs, {q: np.abs(b).astype(complex) for q, b in t.blocks.items()})


@pytest.mark.timeout(300)
def test_chain32_physical_start_matches_oracle(kv, chain):
    ch, lat, mpo = chain
    steps = 20
    print(f"case chain32 physical start, {steps} two-site steps", flush=True)
    psi0 = M.product_state(lat)
    ref = T.TDVP(M.product_state(lat), mpo, cutoff=CUTOFF, maxdim=D, dt=DT)
    e64 = kv.TDVP(ch.psi0, ch.mpo, cutoff=CUTOFF, maxdim=D, dt=DT, precision="fp64")
    e32 = kv.TDVP(ch.psi0, ch.mpo, cutoff=CUTOFF, maxdim=D, dt=DT, precision="fp32")
    d64 = d32 = 0.0
    for k in range(steps):
        ref.two_site()
        e64.two_site()
        e32.two_site()
        c_ref = T.overlap_states(psi0, ref.mps)
        d64 = max(d64, abs(kv.overlap_states(ch.psi0, e64.mps) - c_ref))
        d32 = max(d32, abs(kv.overlap_states(ch.psi0, e32.mps) - c_ref))
        assert e64.bond_dims() == ref.bond_dims(), f"step {k + 1}"
    print(f"max|dC| fp64 {d64:.2e}, fp32 {d32:.2e}; max bond {max(ref.bond_dims())}", flush=True)
    assert d64 < 1e-9
    assert d32 < 1e-7
    mps_g, mps_r = e64.mps, ref.mps
    pops_g = np.array([kv.join(ch.density(i), mps_g, mps_g).real for i in range(N)])
    pops_r = np.array([T.join(M.density_mpo(lat, i), mps_r, mps_r).real for i in range(N)])
    assert np.abs(pops_g - pops_r).max() < 1e-9
    assert pops_r[1] > 1e-3  # the exciton has left site 0
    for i in (0, 1, 2, N // 2):
        e_g = kv.join(ch.hopping(i), mps_g, mps_g)
        e_r = T.join(M.hopping_mpo(lat, i), mps_r, mps_r)
        assert abs(e_g - e_r) < 1e-9, f"bond energy {i}"


@pytest.fixture(scope="module")
def middle_bond(kv, chain):
    """The benchmark's D=256 middle bond on the device and as oracle tensors."""
    from kylin_v_amd import model as KM
    ch, lat, mpo = chain
    blocks = KM.synthetic_state_blocks(ch.sites, D, seed=7)
    psi = [kv.Tensor.from_blocks((1, 1, -1), b) for b in blocks]
    envL, a, b, envR2, envR1 = KM.centre_bond(psi, ch.mpo, MID)
    x = kv.product(a, b, 1)
    return dict(envL=envL, envR2=envR2, x=x, oL=down(envL), oR=down(envR2), ox=down(x),
                W1=mpo[MID], W2=mpo[MID + 1], dW1=ch.mpo[MID], dW2=ch.mpo[MID + 1])


@pytest.mark.timeout(300)
def test_chain32_middle_bond_heff_matches_oracle(kv, middle_bond):
    m = middle_bond
    shapes = sorted(set(b.shape for b in m["ox"].blocks.values()))
    print(f"case chain32 D=256 apply_double, x blocks {shapes}", flush=True)
    assert (128, 16, 16, 128) in shapes
    ref = T.apply_double(m["oL"], m["ox"], m["W1"], m["W2"], m["oR"])
    rs = scale(T.apply_double(absolute(m["oL"]), absolute(m["ox"]), absolute(m["W1"]),
                              absolute(m["W2"]), absolute(m["oR"])))
    for prec, bar in (("fp64", 1e-12), ("fp32", 2e-5)):
        got = down(kv.apply_double(m["envL"], m["x"], m["dW1"], m["dW2"], m["envR2"], prec))
        err = max_diff(got, ref) / rs
        print(f"apply_double {prec}: max|dy| / scale = {err:.2e}", flush=True)
        assert err <= bar


@pytest.mark.timeout(400)
def test_chain32_middle_bond_krylov_and_svd_match_oracle(kv, middle_bond):
    m = middle_bond
    print("case chain32 D=256 local_krylov + svd_impl", flush=True)
    xr, ex_r = T.local_krylov(m["oL"], m["ox"], m["W1"], m["W2"], m["oR"], DT, 1e-8, 15, False, 2)
    xg = kv.Tensor.from_blocks(m["ox"].dirs, m["ox"].blocks)  # local_krylov overwrites it
    xg, ex_g = kv.local_krylov(m["envL"], xg, m["dW1"], m["dW2"], m["envR2"], DT, 1e-8, 15, False, 2)
    print(f"krylov iterations gpu {ex_g['iters']} oracle {ex_r.iters}", flush=True)
    assert ex_g["iters"] == ex_r.iters
    xgo = down(xg)
    assert max_diff(xgo, xr) <= 1e-10 * scale(xr)
    # the bond SVD of the evolved vector: same kept dimension per sector, same product
    for l2r in (True, False):
        lef, rig, sv = T.svd_impl(xr, 2, CUTOFF, D, l2r, return_svals=True)
        gl, gr = kv.svd_impl(kv.Tensor.from_blocks(xr.dirs, xr.blocks), 2, CUTOFF, D, l2r)
        gl, gr = down(gl), down(gr)
        assert gl.legs()[2] == lef.legs()[2], f"kept dims l2r={l2r}"
        rec_o = oproduct(lef, rig, 1)
        rec_g = oproduct(gl, gr, 1)
        err = max_diff(rec_g, rec_o) / scale(rec_o)
        print(f"svd l2r={l2r}: kept {gl.legs()[2]}, product err {err:.2e}", flush=True)
        assert err <= 1e-9
        iso = gl if l2r else gr
        gram = oproduct(iso, iso, 2, "C", "N") if l2r else oproduct(iso, iso, 2, "N", "C")
        for q, blk in gram.blocks.items():
            np.testing.assert_allclose(blk, np.eye(blk.shape[0]), atol=1e-10)
        assert abs(onorm(rec_o) - onorm(rec_g)) < 1e-10
