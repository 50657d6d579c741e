"""Oracle global Krylov stage (oracle/krylov.py, after holstein/krylov.h) pinned against
dense linear algebra: zip-up MPO x MPS equals H v, the MPS direct sum equals the vector
sum, canonicalisation keeps the state and leaves right-isometries, the local
orthogonalisation removes the La/Lb components, and global_main with H scaled by dt
equals exp(-i H dt) v when nothing is truncated.  The PP layout check shows what the
stage is for: from the product state it creates the exciton-hop sectors that two-site
TDVP in that layout cannot."""
import numpy as np
import scipy.linalg

from oracle import blocksparse as B
from oracle import krylov as K
from oracle import model as M

N, L = 4, 3
G = np.array([1.0, 0.8, 1.2, 0.9]).reshape(N, 1)


def _setup():
    lat = M.holstein(N, [1.0], G, [L], 1.0, "merged")
    return lat, M.build_mpo(lat), M.product_state(lat), M.ed_hamiltonian(N, [1.0], G, [L], 1.0)


def test_apply_operator_and_state_sum_match_dense():
    lat, H, psi, Hd = _setup()
    v0 = M.mps_to_dense(lat, psi)
    hp = K.apply_operator(H, psi, 0.0, 10 ** 6)
    np.testing.assert_allclose(M.mps_to_dense(lat, hp), Hd @ v0, atol=1e-13)
    hp2 = K.apply_operator(H, hp, 0.0, 10 ** 6)  # a state with nontrivial bonds
    np.testing.assert_allclose(M.mps_to_dense(lat, hp2), Hd @ Hd @ v0, atol=1e-12)
    s = K.state_add(hp, K.state_scale(hp2, -0.25 + 0.5j))
    np.testing.assert_allclose(M.mps_to_dense(lat, s), Hd @ v0 + (-0.25 + 0.5j) * Hd @ Hd @ v0,
                               atol=1e-12)


def test_stack_places_second_operand_after_first():
    a = B.Tensor((1, 1, -1))
    a.push_back_ch((0, 0, 0), np.ones((2, 1, 3)))
    b = B.Tensor((1, 1, -1))
    b.push_back_ch((0, 0, 0), 2 * np.ones((1, 1, 2)))
    b.push_back_ch((1, 0, 1), 3 * np.ones((4, 1, 5)))
    c = K.stack(a, b, [0, 2])
    blk = c.blocks[(0, 0, 0)]
    assert blk.shape == (3, 1, 5)
    np.testing.assert_array_equal(blk[:2, :, :3], 1)
    np.testing.assert_array_equal(blk[2:, :, 3:], 2)
    assert np.all(blk[:2, :, 3:] == 0) and np.all(blk[2:, :, :3] == 0)
    assert c.blocks[(1, 0, 1)].shape == (4, 1, 5)


def test_canonicalize_keeps_state_and_isometries():
    lat, H, psi, Hd = _setup()
    v0 = M.mps_to_dense(lat, psi)
    s = K.state_add(psi, K.state_scale(K.apply_operator(H, psi, 0.0, 10 ** 6), 0.3))
    c = K.canonicalize(s, 0.0, 10 ** 6)
    np.testing.assert_allclose(M.mps_to_dense(lat, c), v0 + 0.3 * Hd @ v0, atol=1e-13)
    for t in c[1:]:
        g = B.product(t, t, 2, "N", "C")
        for q, blk in g.blocks.items():
            np.testing.assert_allclose(blk, np.eye(blk.shape[0]), atol=1e-13)


def test_apply_ortho_site_removes_components():
    """krylov.h:12-45 solves the 2x2 Gram system with overlap(La, Nh) = <Nh|La> where the
    exact projection needs its conjugate; the two agree when the overlaps are real, as the
    Lanczos coefficients are.  Pinned on real data."""
    rng = np.random.default_rng(0)

    def rnd():
        t = B.Tensor((1, 1, 1, -1))
        t.push_back_ch((0, 0, 0, 0), rng.standard_normal((2, 3, 3, 2)).astype(complex))
        return t

    Nh, La, Lb = rnd(), rnd(), rnd()
    r = K.apply_ortho_site(Nh, La, Lb)
    assert abs(B.overlap(r, La)) < 1e-12 and abs(B.overlap(r, Lb)) < 1e-12


def test_global_krylov_step_is_exact_without_truncation():
    lat, H, psi, Hd = _setup()
    v0 = M.mps_to_dense(lat, psi)
    dt = 0.1
    Hc = K.scale_mpo(H, dt)
    s1, n1 = K.global_main(Hc, psi, 0.0, 10 ** 6)
    s2, n2 = K.global_main(Hc, s1, 0.0, 10 ** 6)
    np.testing.assert_allclose(M.mps_to_dense(lat, s1), scipy.linalg.expm(-1j * Hd * dt) @ v0, atol=1e-12)
    np.testing.assert_allclose(M.mps_to_dense(lat, s2), scipy.linalg.expm(-2j * Hd * dt) @ v0, atol=1e-12)
    assert 2 <= n1 <= 14


def test_global_krylov_opens_pp_hop_sectors():
    """Projected purification: from the product state the global step moves weight onto
    the next exciton site (bond dims grow past [1, d, 1, 1, ...])."""
    lat = M.holstein(2, [1.0], 1.0, [3], 1.0, "pp")
    H = M.build_mpo(lat)
    psi = M.product_state(lat)
    out, _ = K.global_main(K.scale_mpo(H, 0.1), psi, 1e-12, 1000)
    dims = [sum(t.legs()[2].values()) for t in out[:-1]]
    assert dims[2] > 1  # the bond left of the second exciton site now carries both charges
    nd = M.density_mpo(lat, 1)
    from oracle import tdvp as T
    assert T.join(nd, out, out).real > 1e-4
