"""GPU parity: the HIP path (through the C ABI) against the oracle on the same inputs.

Bars: FP64 kernels agree with the oracle to ~1e-12 (relative to the operand scale);
complex-FP32 contractions to 2e-5 relative; TDVP observables in FP64 to 1e-9.
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


@pytest.fixture(scope="module")
def kv():
    return GE.load_package()


def up(kv, t: OT):
    return kv.Tensor.from_blocks(t.dirs, t.blocks)


def down(t) -> OT:
    dirs, blocks = t.to_blocks()
    return OT(dirs, blocks)


def max_diff(got: OT, ref: OT):
    """max |got - ref| over the union of blocks (absent blocks count as zero)."""
    assert tuple(got.dirs) == tuple(ref.dirs)
    d = 0.0
    for q in set(got.blocks) | set(ref.blocks):
        a = got.blocks.get(q)
        b = ref.blocks.get(q)
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


def rounding_scale(fn, *tensors):
    """Magnitude the rounding error of a contraction scales with: the same contraction
    over elementwise absolute values (outputs that cancel to ~0 still carry it)."""
    return scale(fn(*[absolute(t) if t is not None else None for t in tensors]))


def evolved(layout, N=3, L=3, steps=2, D=12, seed=0):
    rng = np.random.default_rng(seed)
    g = rng.uniform(0.6, 1.4, size=(N, 1))
    lat = M.holstein(N, [1.0], g, [L], 0.8, layout)
    mpo = M.build_mpo(lat)
    eng = T.TDVP(M.product_state(lat), mpo, cutoff=1e-12, maxdim=D, dt=0.3)
    for _ in range(steps):
        eng.two_site()
    return lat, mpo, eng


def rand_block_tensor(rng, dirs, legs):
    import itertools
    t = OT(dirs)
    for qs in itertools.product(*[sorted(l) for l in legs]):
        if sum(d * q for d, q in zip(dirs, qs)) == 0:
            sh = tuple(legs[a][q] for a, q in enumerate(qs))
            t.push_back_ch(qs, rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
    return t


@pytest.mark.parametrize("spec", ["1NN", "2CN", "2NT"])
def test_grouped_gemm_product_large_ragged_blocks(kv, spec):
    """Sector dims straddling the 64x64 tile and 8-deep K step (edges, padding, segments)."""
    print(f"case product {spec}", flush=True)
    rng = np.random.default_rng(11)
    a_l = {0: 67, 1: 5, 2: 130}
    s_l = {0: 3, 1: 9}
    b_l = {0: 70, 1: 13, 2: 1, 3: 65}
    A = rand_block_tensor(rng, (1, 1, -1), [a_l, s_l, b_l])
    nc, o1, o2 = int(spec[0]), spec[1], spec[2]
    if spec == "1NN":
        B = rand_block_tensor(rng, (1, 1, -1), [b_l, s_l, {0: 33, 1: 77, 2: 2, 3: 64, 4: 100}])
    elif spec == "2CN":
        B = rand_block_tensor(rng, (1, 1, -1), [a_l, s_l, {0: 3, 1: 71, 2: 20, 3: 8}])
    else:
        B = rand_block_tensor(rng, (-1, -1, 1), [{-2: 2, -1: 80, 0: 31, 1: 66}, s_l, b_l])
    ref = oproduct(A, B, nc, o1, o2)
    got = down(kv.product(up(kv, A), up(kv, B), nc, o1, o2))
    assert max_diff(got, ref) <= 1e-12 * scale(ref)


# Bars relative to the contraction of absolute values: FP64 rounding; FP32 rounding over the
# four chained stages; split-BF16 drops the lo*lo term and rounds lo to 8 bits, ~3*2^-18 of
# |a||b| per product, over the same stages.
HEFF_TOL = {"fp64": 1e-12, "fp32": 2e-5, "bf16x3": 2e-4}


@pytest.mark.parametrize("layout", ["merged", "pp", "split"])
@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
def test_heff_matches_oracle(kv, layout, precision):
    print(f"case heff {layout} {precision}", flush=True)
    lat, mpo, eng = evolved(layout)
    tol = HEFF_TOL[precision]
    envl, envr = eng.envL[-1], eng.envR[-1]
    gL, gR = up(kv, envl), up(kv, envr)
    gW = [up(kv, w) for w in mpo]
    y = T.apply_single(envl, eng.mps[0], mpo[0], envr)
    got = down(kv.apply_single(gL, up(kv, eng.mps[0]), gW[0], gR, precision))
    assert max_diff(got, y) <= tol * rounding_scale(T.apply_single, envl, eng.mps[0], mpo[0], envr)
    envr2 = eng.envR[-2]
    x2 = oproduct(eng.mps[0], eng.mps[1], 1)
    y2 = T.apply_double(envl, x2, mpo[0], mpo[1], envr2)
    got2 = down(kv.apply_double(gL, up(kv, x2), gW[0], gW[1], up(kv, envr2), precision))
    assert max_diff(got2, y2) <= tol * rounding_scale(T.apply_double, envl, x2, mpo[0], mpo[1], envr2)
    cur, nex = T.svd_impl(eng.mps[0], 2, -1.0, 10 ** 6, True)
    envl1 = T.env_sweep(envl, cur, mpo[0], True)
    y0 = T.apply_zero(envl1, nex, envr)
    got0 = down(kv.apply_zero(up(kv, envl1), up(kv, nex), gR, precision))
    assert max_diff(got0, y0) <= tol * rounding_scale(T.apply_zero, envl1, nex, envr)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3", "bf16x3-128"])
def test_heff_large_ragged_sectors(kv, precision, monkeypatch):
    """One- and two-site H_eff on random block tensors whose sector dims straddle the 64x64
    output tile and the 32-deep split-BF16 K step (edge tiles, K tails, several K segments);
    bf16x3-128 forces the split-BF16 kernel's 128x128 tile."""
    print(f"case heff ragged {precision}", flush=True)
    if precision == "bf16x3-128":
        monkeypatch.setenv("KV_SPLIT_TILE", "128")
        precision = "bf16x3"
    rng = np.random.default_rng(5)
    a_l = {0: 70, 1: 33}
    s_l = {0: 5, 1: 3}
    w_l = {-1: 1, 0: 2, 1: 1}
    b_l = {0: 45, 1: 97}
    L = rand_block_tensor(rng, (1, -1, -1), [a_l, w_l, a_l])
    W = rand_block_tensor(rng, (1, 1, -1, -1), [s_l, w_l, s_l, w_l])
    R = rand_block_tensor(rng, (-1, 1, 1), [b_l, w_l, b_l])
    x = rand_block_tensor(rng, (1, 1, -1), [a_l, s_l, b_l])
    tol = HEFF_TOL[precision]
    y = T.apply_single(L, x, W, R)
    got = down(kv.apply_single(up(kv, L), up(kv, x), up(kv, W), up(kv, R), precision))
    assert max_diff(got, y) <= tol * rounding_scale(T.apply_single, L, x, W, R)
    x2 = rand_block_tensor(rng, (1, 1, 1, -1), [a_l, s_l, s_l, b_l])
    y2 = T.apply_double(L, x2, W, W, R)
    got2 = down(kv.apply_double(up(kv, L), up(kv, x2), up(kv, W), up(kv, W), up(kv, R), precision))
    assert max_diff(got2, y2) <= tol * rounding_scale(T.apply_double, L, x2, W, W, R)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_heff_plan_reused_like_lanczos(kv, precision):
    """An H_eff plan built once (kv_heff_plan_create) and applied repeatedly: the first apply
    equals the one-shot kv_heff_apply, scale is applied, and feeding each output back in
    (as a Lanczos loop does) tracks the oracle's repeated apply_double / apply_single /
    apply_zero."""
    print(f"case heff plan {precision}", flush=True)
    rng = np.random.default_rng(8)
    a_l = {0: 70, 1: 33}
    s_l = {0: 5, 1: 3}
    w_l = {-1: 1, 0: 2, 1: 1}
    b_l = {0: 45, 1: 97}
    L = rand_block_tensor(rng, (1, -1, -1), [a_l, w_l, a_l])
    W = rand_block_tensor(rng, (1, 1, -1, -1), [s_l, w_l, s_l, w_l])
    R = rand_block_tensor(rng, (-1, 1, 1), [b_l, w_l, b_l])
    R0 = rand_block_tensor(rng, (-1, 1, 1), [a_l, w_l, a_l])
    cases = [
        (2, rand_block_tensor(rng, (1, 1, 1, -1), [a_l, s_l, s_l, b_l]), W, W, R,
         lambda x: T.apply_double(L, x, W, W, R)),
        (1, rand_block_tensor(rng, (1, 1, -1), [a_l, s_l, b_l]), W, None, R,
         lambda x: T.apply_single(L, x, W, R)),
        (0, rand_block_tensor(rng, (1, -1), [a_l, a_l]), None, None, R0,
         lambda x: T.apply_zero(L, x, R0)),
    ]
    tol = {"fp64": 1e-12, "fp32": 1e-4}[precision]  # of max|y| on these random tensors
    gL = up(kv, L)
    for nsite, x, w1, w2, r, ref_apply in cases:
        gx, gr = up(kv, x), up(kv, r)
        gw1 = up(kv, w1) if w1 is not None else None
        gw2 = up(kv, w2) if w2 is not None else None
        plan = kv.HeffPlan(nsite, gL, gx, gw1, gw2, gr, precision)
        one_shot = down(kv._heff(nsite, gL, gw1, gw2, gr, gx, precision))
        y1 = plan.apply(gx, 1.0)
        assert max_diff(down(y1), one_shot) <= 1e-14 * scale(one_shot)
        y = plan.apply(gx, 0.5)
        r1 = ref_apply(x)
        ref = OT(r1.dirs, {q: 0.5 * b for q, b in r1.blocks.items()})
        assert max_diff(down(y), ref) <= tol * scale(ref)
        # two more applies on the plan's own outputs
        yg = y
        for _ in range(2):
            r1 = ref_apply(ref)
            ref = OT(r1.dirs, {q: 0.5 * b for q, b in r1.blocks.items()})
            yg = plan.apply(yg, 0.5)
            assert max_diff(down(yg), ref) <= 10 * tol * scale(ref)
        plan.close()


@pytest.mark.parametrize("layout", ["merged", "pp"])
def test_env_update_matches_oracle(kv, layout):
    print(f"case env {layout}", flush=True)
    lat, mpo, eng = evolved(layout)
    envl = eng.envL[-1]
    ref = T.env_sweep(envl, eng.mps[0], mpo[0], True)
    got = down(kv.env_sweep(up(kv, envl), up(kv, eng.mps[0]), up(kv, mpo[0]), True))
    assert max_diff(got, ref) <= 1e-12 * scale(ref)
    envr = T.boundary_envs(eng.mps)[1]
    ref_r = T.env_sweep(envr, eng.mps[-1], mpo[-1], False)
    got_r = down(kv.env_sweep(up(kv, envr), up(kv, eng.mps[-1]), up(kv, mpo[-1]), False))
    assert max_diff(got_r, ref_r) <= 1e-12 * scale(ref_r)
    # bra != ket form (join)
    ref_b = T.env_sweep_braket(envl, eng.mps[0], mpo[0], M.product_state(lat)[0], True)
    got_b = down(kv.env_sweep(up(kv, envl), up(kv, M.product_state(lat)[0]), up(kv, mpo[0]),
                              True, bra=up(kv, eng.mps[0])))
    assert max_diff(got_b, ref_b) <= 1e-12 * max(scale(ref_b), 1e-300)


def test_svd_truncation_matches_oracle(kv):
    print("case svd", flush=True)
    lat, mpo, eng = evolved("merged", N=4, L=3, steps=3, D=30)
    x2 = oproduct(eng.mps[1], eng.mps[2], 1)
    for cutoff, maxdim in [(1e-10, 1000), (1e-10, 7), (-1.0, 1000)]:
        for l2r in (True, False):
            lef, rig, sv = T.svd_impl(x2, 2, cutoff, maxdim, l2r, return_svals=True)
            gl, gr = kv.svd_impl(up(kv, x2), 2, cutoff, maxdim, l2r)
            gl, gr = down(gl), down(gr)
            # same kept dimension per sector
            assert gl.legs()[2] == lef.legs()[2]
            # gauge-free: the truncated product and the isometry property
            rec_o = oproduct(lef, rig, 1)
            rec_g = oproduct(gl, gr, 1)
            # Jacobi SVD (zgesvdj) and LAPACK's gesdd agree to ~1e-11 on truncated products
            assert max_diff(rec_g, rec_o) <= 1e-10 * scale(rec_o)
            iso = gl if l2r else gr
            gram = oproduct(iso, iso, 2, "C", "N") if l2r else oproduct(iso, iso, 2, "N", "C")
            for q, b in gram.blocks.items():
                np.testing.assert_allclose(b, np.eye(b.shape[0]), atol=1e-11)


@pytest.mark.parametrize("nsite", [2, 1, 0])
def test_local_krylov_matches_oracle(kv, nsite):
    print(f"case krylov nsite={nsite}", flush=True)
    lat, mpo, eng = evolved("merged", N=4, L=3)
    envl = eng.envL[-1]
    if nsite == 2:
        envr = eng.envR[-2]
        x = oproduct(eng.mps[0], eng.mps[1], 1)
        args = (mpo[0], mpo[1])
    elif nsite == 1:
        envr = eng.envR[-1]
        x = eng.mps[0]
        args = (mpo[0], None)
    else:
        cur, x = T.svd_impl(eng.mps[0], 2, -1.0, 10 ** 6, True)
        envl = T.env_sweep(envl, cur, mpo[0], True)
        envr = eng.envR[-1]
        args = (None, None)
    for rev in (False, True):
        ref, ex = T.local_krylov(envl, x, args[0], args[1], envr, 0.05, 1e-8, 15, rev, nsite)
        gx = up(kv, x)
        gw = [up(kv, a) if a is not None else None for a in args]
        _, gex = kv.local_krylov(up(kv, envl), gx, gw[0], gw[1], up(kv, envr), 0.05, 1e-8, 15,
                                 rev, nsite)
        assert gex["iters"] == ex.iters
        assert max_diff(down(gx), ref) <= 1e-11


def _ct_pair(kv, layout, N, L, g, J, dt, steps, D, cutoff, one_site=False, precision="fp64"):
    lat = M.holstein(N, [1.0], g, [L], J, layout)
    from kylin_v_amd import model as KM
    ch = KM.HolsteinChain(N, [1.0], g, [L], J, layout)
    psi0 = M.product_state(lat)
    ref = T.TDVP(psi0, M.build_mpo(lat), cutoff=cutoff, maxdim=D, dt=dt)
    eng = kv.TDVP(ch.psi0, ch.mpo, cutoff=cutoff, maxdim=D, dt=dt, precision=precision)
    c_ref, c_gpu, pops_ref, pops_gpu = [], [], [], []
    for s in range(steps):
        if one_site:
            ref.one_site()
            eng.one_site()
        else:
            ref.two_site()
            eng.two_site()
        c_ref.append(T.overlap_states(psi0, ref.mps))
        c_gpu.append(kv.overlap_states(ch.psi0, eng.mps))
    mps = eng.mps
    pops_gpu = [kv.join(ch.density(i), mps, mps).real for i in range(N)]
    pops_ref = [T.join(M.density_mpo(lat, i), ref.mps, ref.mps).real for i in range(N)]
    return (np.array(c_gpu), np.array(c_ref), np.array(pops_gpu), np.array(pops_ref),
            eng.bond_dims(), ref.bond_dims())


def test_product_side_mpo_equals_oracle_mpo(kv):
    from kylin_v_amd import model as KM
    for layout in ("merged", "pp", "split"):
        g = np.array([0.7, 1.2, 0.9])
        lat = M.holstein(3, [1.1], g, [3], 0.8, layout)
        ref = M.build_mpo(lat)
        sites, terms, init, ex = KM.chain(3, [1.1], g, [3], 0.8, layout)
        got = KM.mpo_blocks(sites, terms)
        for w_ref, w_got in zip(ref, got):
            assert set(w_ref.blocks) == set(w_got)
            for q in w_got:
                np.testing.assert_array_equal(w_got[q], w_ref.blocks[q])


@pytest.mark.parametrize("layout", ["merged", "pp"])
def test_two_site_tdvp_matches_oracle(kv, layout):
    print(f"case tdvp2 {layout}", flush=True)
    g = np.array([1.0, 0.8, 1.2])
    c_gpu, c_ref, p_gpu, p_ref, bd_g, bd_r = _ct_pair(kv, layout, 3, 3, g, 1.0, 0.05, 6, 12, 1e-10)
    assert bd_g == bd_r
    assert np.abs(c_gpu - c_ref).max() < 1e-9
    assert np.abs(p_gpu - p_ref).max() < 1e-9
    assert abs(p_gpu.sum() - 1.0) < 1e-9


def test_one_site_tdvp_matches_oracle(kv):
    print("case tdvp1 merged", flush=True)
    g = np.array([1.0, 0.8, 1.2])
    lat = M.holstein(3, [1.0], g, [3], 1.0, "merged")
    from kylin_v_amd import model as KM
    ch = KM.HolsteinChain(3, [1.0], g, [3], 1.0, "merged")
    ref = T.TDVP(M.product_state(lat), M.build_mpo(lat), cutoff=-1.0, maxdim=100, dt=0.1)
    eng = kv.TDVP(ch.psi0, ch.mpo, cutoff=-1.0, maxdim=100, dt=0.1)
    for _ in range(3):  # grow bonds with 2TDVP, then 1TDVP at fixed bond dimension
        ref.two_site()
        eng.two_site()
    psi0 = M.product_state(lat)
    for _ in range(3):
        ref.one_site()
        eng.one_site()
        c_ref = T.overlap_states(psi0, ref.mps)
        c_gpu = kv.overlap_states(ch.psi0, eng.mps)
        assert abs(c_ref - c_gpu) < 1e-9


def test_two_site_tdvp_matches_exact_diagonalisation(kv):
    print("case tdvp ED", flush=True)
    N, L, J, dt = 3, 3, 1.0, 0.05
    g = np.array([0.7, 1.1, 0.9])
    from kylin_v_amd import model as KM
    ch = KM.HolsteinChain(N, [1.0], g, [L], J, "merged")
    eng = kv.TDVP(ch.psi0, ch.mpo, cutoff=-1.0, maxdim=10 ** 5, dt=dt)
    H = M.ed_hamiltonian(N, [1.0], g.reshape(N, 1), [L], J)
    v0 = M.ed_initial_state(N, [L])
    w, V = np.linalg.eigh(H)
    for s in range(1, 9):
        eng.two_site()
        ex = np.vdot(v0, V @ (np.exp(-1j * w * dt * s) * (V.conj().T @ v0)))
        assert abs(kv.overlap_states(ch.psi0, eng.mps) - ex) < 1e-10


@pytest.mark.parametrize("precision,bar", [("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_mixed_precision_tdvp_tracks_fp64(kv, precision, bar):
    print(f"case tdvp {precision}", flush=True)
    g = np.array([1.0, 0.8, 1.2, 0.9])
    from kylin_v_amd import model as KM
    ch = KM.HolsteinChain(4, [1.0], g, [4], 1.0, "merged")
    e64 = kv.TDVP(ch.psi0, ch.mpo, cutoff=1e-10, maxdim=24, dt=0.05, precision="fp64")
    e32 = kv.TDVP(ch.psi0, ch.mpo, cutoff=1e-10, maxdim=24, dt=0.05, precision=precision)
    dmax = 0.0
    for _ in range(8):
        e64.two_site()
        e32.two_site()
        dmax = max(dmax, abs(kv.overlap_states(ch.psi0, e64.mps) - kv.overlap_states(ch.psi0, e32.mps)))
    print(f"{precision} vs fp64 max|dC| = {dmax:.3e}", flush=True)
    assert dmax < bar


def test_mps_mpo_files_round_trip(kv, tmp_path):
    """Binary formats of holstein/state.h:107-157 and operator.h:94-143, written here by an
    independent numpy writer, read by the engine, written back and compared byte-wise."""
    import struct
    lat, mpo, eng = evolved("merged")

    def write(path, tensors, magic, rank):
        with open(path, "wb") as f:
            f.write(struct.pack("ii", magic, len(tensors)))
            for t in tensors:
                items = t.sorted_items()
                f.write(struct.pack("i", len(items)))
                for q, b in items:
                    f.write(struct.pack(f"{rank}i", *b.shape))
                    f.write(struct.pack(f"{rank}i", *q))
                    f.write(np.ascontiguousarray(b).tobytes())
            f.write(b"\n")

    p_mps, p_mpo = tmp_path / "a.mps", tmp_path / "H"
    write(p_mps, eng.mps, 111, 3)
    write(p_mpo, mpo, 112, 4)
    sites = kv.load_mps(str(p_mps))
    ops = kv.load_mpo(str(p_mpo))
    for g, r in zip(sites, eng.mps):
        assert max_diff(down(g), r) == 0.0
    kv.save_mps(str(tmp_path / "b.mps"), sites)
    kv.save_mpo(str(tmp_path / "H2"), ops)
    assert open(p_mps, "rb").read() == open(tmp_path / "b.mps", "rb").read()
    assert open(p_mpo, "rb").read() == open(tmp_path / "H2", "rb").read()
    bad = tmp_path / "bad.mps"
    bad.write_bytes(struct.pack("ii", 112, 0))
    with pytest.raises(kv.KVError) as e:
        kv.load_mps(str(bad))
    assert e.value.code == -5


def _set_eig(monkeypatch, eig):
    """split: own tridiagonalisation; split-roc: rocSOLVER zhetrd; heevd: zheevd."""
    monkeypatch.setenv("KV_EIG", "split" if eig.startswith("split") else eig)
    monkeypatch.setenv("KV_TRD", "rocsolver" if eig == "split-roc" else "own")


@pytest.mark.parametrize("eig", ["split", "split-roc", "heevd", "block"])
def test_svd_gram_route_matches_oracle(kv, eig, monkeypatch):
    """Sectors whose smaller side is 512, at least twice the block of maxdim + 32, take the
    block subspace route by default ("block", its Rayleigh-Ritz on the engine's eigensolver);
    with the block route off they take the full Gram route, with the split Hermitian
    eigensolver (tridiagonalisation, dstedc, back-transformation of the kept vectors) or
    zheevd.  The kept singular values, the truncated product and the isometry match the
    oracle's dense SVD."""
    print(f"case svd gram eig={eig}", flush=True)
    if eig == "block":
        monkeypatch.delenv("KV_SVD", raising=False)
    else:
        _set_eig(monkeypatch, eig)
        monkeypatch.setenv("KV_SVD", "full")
    from kylin_v_amd import model as KM
    sites, terms, init, ex = KM.chain(6, [1.0], 1.0, [16], 1.0, "merged")
    blocks = KM.synthetic_state_blocks(sites, 64, seed=3, decay=0.3)
    a = OT((1, 1, -1), blocks[2])
    b = OT((1, 1, -1), blocks[3])
    x2 = oproduct(a, b, 1)
    # give x2 a generic spectrum (the product of two isometries has flat values)
    rng = np.random.default_rng(5)
    for q in x2.blocks:
        x2.blocks[q] = x2.blocks[q] + 1e-2 * (rng.standard_normal(x2.blocks[q].shape)
                                              + 1j * rng.standard_normal(x2.blocks[q].shape))
    # the sector matrices' sides, from x2's own blocks (rows (a, s1), columns (s2, b))
    rows, cols = T._sector_layout(x2, 2)
    sides = [(rows[q][1], cols[-q][1]) for q in rows if -q in cols]
    assert min(min(r, c) for r, c in sides) >= 512, sides  # Gram routes from 64 (iso_min)
    for maxdim, l2r in [(64, True), (100, False)]:
        lef, rig, sv = T.svd_impl(x2, 2, 1e-10, maxdim, l2r, return_svals=True)
        gl, gr = kv.svd_impl(up(kv, x2), 2, 1e-10, maxdim, l2r)
        gl, gr = down(gl), down(gr)
        assert gl.legs()[2] == lef.legs()[2]
        rec_o = oproduct(lef, rig, 1)
        rec_g = oproduct(gl, gr, 1)
        assert max_diff(rec_g, rec_o) <= 1e-9 * scale(rec_o)
        iso = gl if l2r else gr
        gram = oproduct(iso, iso, 2, "C", "N") if l2r else oproduct(iso, iso, 2, "N", "C")
        for q, blk in gram.blocks.items():
            np.testing.assert_allclose(blk, np.eye(blk.shape[0]), atol=1e-9)


def _spectrum_tensor(rng, svals_by_q):
    """Rank-3 tensor (a, s, b), dirs (+1,+1,-1), whose row sector q (rows (a, s) with
    qa + qs = q, columns b of charge q) is a dense matrix with singular values svals_by_q[q]."""
    a_l, s_l, b_l = {0: 40, 1: 30}, {0: 3, 1: 2}, {0: 100, 1: 120}
    t = OT((1, 1, -1))
    for q, sv in svals_by_q.items():
        keys = [(qa, qs) for qa in a_l for qs in s_l if qa + qs == q]
        nrow = sum(a_l[qa] * s_l[qs] for qa, qs in keys)
        ncol = b_l[q]
        u, _ = np.linalg.qr(rng.standard_normal((nrow, nrow)) + 1j * rng.standard_normal((nrow, nrow)))
        v, _ = np.linalg.qr(rng.standard_normal((ncol, ncol)) + 1j * rng.standard_normal((ncol, ncol)))
        k = len(sv)
        M = (u[:, :k] * np.asarray(sv)[None, :]) @ v[:, :k].conj().T
        pos = 0
        for qa, qs in keys:
            n = a_l[qa] * s_l[qs]
            t.push_back_ch((qa, qs, q), M[pos:pos + n].reshape(a_l[qa], s_l[qs], ncol))
            pos += n
    return t


@pytest.mark.parametrize("eig", ["split", "split-roc", "heevd"])
@pytest.mark.parametrize("l2r", [True, False])
def test_svd_isometric_route_tail_refinement(kv, l2r, eig, monkeypatch):
    """Sectors of 100 and 120 singular values spanning 1 .. 1e-14 with none in
    [0.5e-10, 2e-10] around the 1e-10 cutoff: the isometric-side Gram route must keep
    exactly the values above the cutoff (second pass over the tail), give the oracle's
    truncated product and an isometry orthonormal to machine precision; a maxdim-limited
    cut must match too."""
    print(f"case svd iso l2r={l2r} eig={eig}", flush=True)
    _set_eig(monkeypatch, eig)
    rng = np.random.default_rng(12)

    def spec(n, c):
        v = c * 10.0 ** (-14.0 * np.arange(n) / (n - 1))
        band = (v > 0.5e-10) & (v < 2e-10)
        return np.where(band & (v >= 1e-10), 2.5e-10, np.where(band, 0.4e-10, v))

    x = _spectrum_tensor(rng, {0: spec(100, 1.0), 1: spec(120, 0.5)})
    for cutoff, maxdim in [(1e-10, 1000), (1e-10, 90)]:
        lef, rig, sv = T.svd_impl(x, 2, cutoff, maxdim, l2r, return_svals=True)
        gl, gr = kv.svd_impl(up(kv, x), 2, cutoff, maxdim, l2r)
        gl, gr = down(gl), down(gr)
        assert gl.legs()[2] == lef.legs()[2]
        rec_o, rec_g = oproduct(lef, rig, 1), oproduct(gl, gr, 1)
        # one Gram pass alone leaves ~1e-9 here (the tail vectors are mixed at eps/s^2)
        assert max_diff(rec_g, rec_o) <= 1e-10 * scale(rec_o)
        iso = gl if l2r else gr
        gram = oproduct(iso, iso, 2, "C", "N") if l2r else oproduct(iso, iso, 1, "N", "C")
        for q, blk in gram.blocks.items():
            np.testing.assert_allclose(blk, np.eye(blk.shape[0]), atol=1e-12)


def test_two_mode_chain_rank_deficient_gram_sectors(kv):
    """Exciton x two phonon modes (d = 128) from the product state: the 64-wide sectors take
    the Gram routes on rank-deficient matrices whose tridiagonal form splits into decoupled
    blocks, where rocSOLVER's divide and conquer fails; those sectors are redone by QR
    iteration, and two steps match the oracle."""
    print("case svd rank-deficient gram", flush=True)
    from kylin_v_amd import model as KM
    N, om, g, Ls, J = 4, [0.5, 1.5], 0.7, [8, 8], 1.0
    ch = KM.HolsteinChain(N, om, g, Ls, J, "merged")
    eng = kv.TDVP(ch.psi0, ch.mpo, cutoff=1e-10, maxdim=512, dt=0.02)
    lat = M.holstein(N, om, g, Ls, J, "merged")
    psi0 = M.product_state(lat)
    ref = T.TDVP(psi0, M.build_mpo(lat), cutoff=1e-10, maxdim=512, dt=0.02)
    for k in range(2):
        eng.two_site()
        ref.two_site()
        c_ref = T.overlap_states(psi0, ref.mps)
        c_gpu = kv.overlap_states(ch.psi0, eng.mps)
        assert abs(c_gpu - c_ref) < 1e-10, (k, c_gpu, c_ref)
    assert list(eng.bond_dims()) == list(ref.bond_dims())


@pytest.mark.parametrize("eig", ["split", "split-roc"])
def test_svd_gram_side_beyond_register_rows(kv, eig, monkeypatch):
    """A Gram sector of side 2112: the tridiagonalisation's column kernel holds 2048 rows of
    the trailing column in registers (256 threads x 8) and passes the rows beyond them back
    through the matrix for its first 64 columns.  The full eigensolver route (the block
    subspace route switched off) on the engine's tridiagonalisation and on rocSOLVER's must
    both give the oracle's kept values, truncated product and an orthonormal isometry."""
    print(f"case svd gram side 2112 eig={eig}", flush=True)
    _set_eig(monkeypatch, eig)
    monkeypatch.setenv("KV_SVD", "full")
    rng = np.random.default_rng(21)
    nr, nc = 2112, 2240
    sv = np.exp(-np.arange(nr) / 150.0)
    u, _ = np.linalg.qr(rng.standard_normal((nr, nr)) + 1j * rng.standard_normal((nr, nr)))
    v, _ = np.linalg.qr(rng.standard_normal((nc, nr)) + 1j * rng.standard_normal((nc, nr)))
    x = OT((1, -1), {(0, 0): (u * sv[None, :]) @ v.conj().T})
    for l2r, maxdim in [(True, 200), (False, 260)]:
        lef, rig = T.svd_impl(x, 1, 1e-10, maxdim, l2r)
        gl, gr = kv.svd_impl(up(kv, x), 1, 1e-10, maxdim, l2r)
        gl, gr = down(gl), down(gr)
        assert gl.legs()[1] == lef.legs()[1]
        rec_o, rec_g = oproduct(lef, rig, 1), oproduct(gl, gr, 1)
        assert max_diff(rec_g, rec_o) <= 1e-10 * scale(rec_o)
        iso = gl if l2r else gr
        gram = oproduct(iso, iso, 1, "C", "N") if l2r else oproduct(iso, iso, 1, "N", "C")
        for q, blk in gram.blocks.items():
            np.testing.assert_allclose(blk, np.eye(blk.shape[0]), atol=1e-12)


@pytest.mark.parametrize("l2r", [True, False])
def test_svd_block_route_cliff_spectrum(kv, l2r):
    """Block subspace route on sectors of 700 x 900 and 640 x 800 whose values fall from 1
    to 1e-2 over 60 (50) values and then drop to 1e-7 and below.  Cut at maxdim 100 (maxdim
    bound, cutoff 1e-10): every sector keeps values above its cliff, the certified values and
    the ceiling on the rest settle the cut, and the kept values, truncated product and
    isometry match the oracle.  Cut at maxdim 120 the cut reaches the 1e-7 values, which the
    block cannot certify: the sectors are redone by the full eigensolver and still match."""
    print(f"case svd block cliff l2r={l2r}", flush=True)
    rng = np.random.default_rng(31)
    t = OT((1, -1))

    def sector(q, nr, nc, top):
        s = np.concatenate([np.logspace(0, -2, top), 1e-7 * np.logspace(0, -6, 150)])
        u, _ = np.linalg.qr(rng.standard_normal((nr, len(s))) + 1j * rng.standard_normal((nr, len(s))))
        v, _ = np.linalg.qr(rng.standard_normal((nc, len(s))) + 1j * rng.standard_normal((nc, len(s))))
        t.push_back_ch((q, q), (u * s[None, :]) @ v.conj().T)

    sector(0, 700, 900, 60)
    sector(1, 640, 800, 50)
    for maxdim in (100, 120):
        lef, rig = T.svd_impl(t, 1, 1e-10, maxdim, l2r)
        gl, gr = kv.svd_impl(up(kv, t), 1, 1e-10, maxdim, l2r)
        gl, gr = down(gl), down(gr)
        assert gl.legs()[1] == lef.legs()[1], maxdim
        rec_o, rec_g = oproduct(lef, rig, 1), oproduct(gl, gr, 1)
        assert max_diff(rec_g, rec_o) <= 1e-10 * scale(rec_o)
        iso = gl if l2r else gr
        gram = oproduct(iso, iso, 1, "C", "N") if l2r else oproduct(iso, iso, 1, "N", "C")
        for q, blk in gram.blocks.items():
            np.testing.assert_allclose(blk, np.eye(blk.shape[0]), atol=1e-12)


@pytest.mark.parametrize("l2r", [True, False])
def test_svd_householder_reduced_isometric_route(kv, l2r, monkeypatch):
    """Sectors of 160 x 1400 and 1100 x 240 (the shapes of the edge bonds of a D = 256 chain,
    scaled down) whose values fall from 1 to 1e-9, cut at maxdim 300 deep into the tail.  The
    kept values lie far below the Gram guard, so the sectors whose isometric side is the
    longer one (the 1400 columns right-to-left, the 1100 rows left-to-right) are reduced by a
    Householder QR (LQ) to their shorter side before the isometric Gram route; the kept
    dimension, the truncated product and the isometry match the oracle's dense SVD, with the
    reduction on and off."""
    print(f"case svd householder-reduced iso l2r={l2r}", flush=True)
    rng = np.random.default_rng(41)
    t = OT((1, -1))

    def sector(q, nr, nc):
        k = min(nr, nc)
        s = np.logspace(0, -9, k)
        u, _ = np.linalg.qr(rng.standard_normal((nr, k)) + 1j * rng.standard_normal((nr, k)))
        v, _ = np.linalg.qr(rng.standard_normal((nc, k)) + 1j * rng.standard_normal((nc, k)))
        t.push_back_ch((q, q), (u * s[None, :]) @ v.conj().T)

    sector(0, 160, 1400)
    sector(1, 1100, 240)
    lef, rig = T.svd_impl(t, 1, 1e-10, 300, l2r)
    rec_o = oproduct(lef, rig, 1)
    for reduce in ("1", "0"):
        monkeypatch.setenv("KV_QR_REDUCE", reduce)
        gl, gr = kv.svd_impl(up(kv, t), 1, 1e-10, 300, l2r)
        gl, gr = down(gl), down(gr)
        assert gl.legs()[1] == lef.legs()[1], reduce
        rec_g = oproduct(gl, gr, 1)
        assert max_diff(rec_g, rec_o) <= 1e-10 * scale(rec_o), reduce
        iso = gl if l2r else gr
        gram = oproduct(iso, iso, 1, "C", "N") if l2r else oproduct(iso, iso, 1, "N", "C")
        for q, blk in gram.blocks.items():
            np.testing.assert_allclose(blk, np.eye(blk.shape[0]), atol=1e-12)
