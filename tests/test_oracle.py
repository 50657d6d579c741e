"""Pins for the oracle (CPU, numpy).

Every check compares the oracle with something it does not compute itself: dense
einsum on expanded tensors, the dense Hamiltonian and exact diagonalisation, closed
forms, scipy's expm, or hand-made singular-value spectra for the truncation rule.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import model as M
from oracle import tdvp as T
from oracle.blocksparse import Tensor, norm, overlap, product, to_dense_legs, transpose


def rand_tensor(rng, dirs, leg_sectors, fill=1.0):
    """Random charge-conserving tensor; leg_sectors[ax] = {qn: dim}."""
    import itertools
    t = Tensor(dirs)
    for qs in itertools.product(*[sorted(ls) for ls in leg_sectors]):
        if sum(d * q for d, q in zip(dirs, qs)) != 0:
            continue
        if rng.random() > fill:
            continue
        shape = tuple(leg_sectors[ax][q] for ax, q in enumerate(qs))
        t.push_back_ch(qs, rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    return t


def dense_of(t, legs):
    return to_dense_legs(t, legs)[0]


@pytest.mark.parametrize("t1,t2", [("N", "N"), ("N", "T"), ("C", "N"), ("T", "N"), ("N", "C")])
def test_product_matches_dense_einsum(t1, t2):
    rng = np.random.default_rng(1)
    a_s = {0: 3, 1: 2}
    s_s = {0: 2, 1: 3}  # physical sectors of dim > 1
    b_s = {0: 4, 1: 2, 2: 1}
    # m1: (a, s, b) dirs (+1,+1,-1);  m2 built so the contracted legs have opposite dirs
    m1 = rand_tensor(rng, (1, 1, -1), [a_s, s_s, b_s])
    if (t1, t2) == ("N", "N"):
        m2 = rand_tensor(rng, (1, 1, -1), [b_s, s_s, {0: 2, 1: 3, 2: 2, 3: 1}])
        out = product(m1, m2, 1, "N", "N")
        legs1, legs2 = m1.legs(), m2.legs()
        legs2[0] = legs1[2]
        ref = np.einsum("asb,btc->astc", dense_of(m1, legs1), dense_of(m2, legs2))
    elif (t1, t2) == ("N", "T"):
        m2 = rand_tensor(rng, (1, -1, 1), [{-1: 1, 0: 2, 1: 3, 2: 2}, s_s, b_s])
        out = product(m1, m2, 2, "N", "T")
        legs1, legs2 = m1.legs(), m2.legs()
        legs2[1], legs2[2] = legs1[1], legs1[2]
        ref = np.einsum("asb,csb->ac", dense_of(m1, legs1), dense_of(m2, legs2))
    elif (t1, t2) == ("C", "N"):
        m2 = rand_tensor(rng, (1, 1, -1), [a_s, s_s, {0: 2, 1: 1, 2: 3}])
        out = product(m1, m2, 2, "C", "N")
        legs1, legs2 = m1.legs(), m2.legs()
        legs2[0], legs2[1] = legs1[0], legs1[1]
        ref = np.einsum("asb,asc->bc", np.conj(dense_of(m1, legs1)), dense_of(m2, legs2))
    elif (t1, t2) == ("T", "N"):
        m2 = rand_tensor(rng, (-1, 1), [a_s, {0: 2, 1: 2}])
        out = product(m1, m2, 1, "T", "N")
        legs1, legs2 = m1.legs(), m2.legs()
        legs2[0] = legs1[0]
        ref = np.einsum("asb,ac->sbc", dense_of(m1, legs1), dense_of(m2, legs2))
    else:
        m2 = rand_tensor(rng, (1, 1, -1), [{-1: 2, 0: 2, 1: 2, 2: 1}, s_s, b_s])
        out = product(m1, m2, 2, "N", "C")
        legs1, legs2 = m1.legs(), m2.legs()
        legs2[1], legs2[2] = legs1[1], legs1[2]
        ref = np.einsum("asb,csb->ac", dense_of(m1, legs1), np.conj(dense_of(m2, legs2)))
    assert out.check_sym()
    if (t1, t2) == ("N", "N"):
        full = [legs1[0], legs1[1], legs2[1], legs2[2]]
    elif (t1, t2) in (("N", "T"), ("N", "C")):
        full = [legs1[0], legs2[0]]
    elif (t1, t2) == ("C", "N"):
        full = [legs1[2], legs2[2]]
    else:
        full = [legs1[1], legs1[2], legs2[1]]
    dense_out, _ = to_dense_legs(out, full)
    np.testing.assert_allclose(dense_out, ref, atol=1e-12)


def test_norm_overlap_transpose_identities():
    rng = np.random.default_rng(2)
    t = rand_tensor(rng, (1, 1, -1), [{0: 2, 1: 3}, {0: 1, 1: 2}, {0: 2, 1: 3, 2: 2}])
    u = rand_tensor(rng, (1, 1, -1), [{0: 2, 1: 3}, {0: 1, 1: 2}, {0: 2, 1: 3, 2: 2}])
    assert abs(overlap(t, t) - norm(t) ** 2) < 1e-12
    assert abs(overlap(t, u) - np.conj(overlap(u, t))) < 1e-12
    tt = transpose(transpose(t, (2, 0, 1)), (1, 2, 0))
    assert all(np.array_equal(tt.blocks[q], t.blocks[q]) for q in t.blocks)
    # linearity: <u | a t + b u> = a<u|t> + b<u|u>
    s = t.scaled(0.3 - 0.2j).add_(u, 1.7j)
    assert abs(overlap(s, u) - ((0.3 - 0.2j) * overlap(t, u) + 1.7j * overlap(u, u))) < 1e-10


def _small_chain(layout, N=3, L=3, seed=0):
    rng = np.random.default_rng(seed)
    g = rng.uniform(0.5, 1.5, size=(N, 1))
    lat = M.holstein(N, [1.1], g, [L], 0.9, layout)
    return lat, g


@pytest.mark.parametrize("layout", ["merged", "split"])
def test_mpo_dense_equals_ed_hamiltonian(layout):
    lat, g = _small_chain(layout)
    H = M.ed_hamiltonian(3, [1.1], g, [3], 0.9)
    assert all(t.check_sym() for t in M.build_mpo(lat))
    np.testing.assert_allclose(M.mpo_to_dense(lat, M.build_mpo(lat)), H, atol=1e-13)


def test_pp_mpo_restricted_to_physical_subspace_equals_ed():
    N, L = 2, 3
    lat, g = _small_chain("pp", N=N, L=L)
    Hpp = M.mpo_to_dense(lat, M.build_mpo(lat))
    H = M.ed_hamiltonian(N, [1.1], g, [L], 0.9)
    # merged basis index (n, k) per site -> pp basis (n, k_phys = k, k_aux = L-1-k)
    d_pp = 2 * L * L
    idx = []
    for flat in range((2 * L) ** N):
        digits, rem = [], flat
        for _ in range(N):
            digits.append(rem % (2 * L))
            rem //= 2 * L
        digits = digits[::-1]
        p = 0
        for dgt in digits:
            n, k = divmod(dgt, L)
            p = p * d_pp + (n * L + k) * L + (L - 1 - k)
        idx.append(p)
    np.testing.assert_allclose(Hpp[np.ix_(idx, idx)], H, atol=1e-13)
    # H maps the physical subspace into itself
    rest = np.setdiff1d(np.arange(Hpp.shape[0]), idx)
    assert np.abs(Hpp[np.ix_(rest, idx)]).max() < 1e-14


def _canonical_random_state(lat, D, rng):
    """Random state of bond dim <= D: evolve the product state with a random-time
    2TDVP so every bond gets populated, then use it as a generic MPS."""
    mpo = M.build_mpo(lat)
    eng = T.TDVP(M.product_state(lat), mpo, cutoff=1e-14, maxdim=D, dt=0.4)
    for _ in range(3):
        eng.two_site()
    return eng


def test_heff_matches_projected_dense_hamiltonian():
    lat, g = _small_chain("merged", N=3, L=2)
    eng = _canonical_random_state(lat, 8, np.random.default_rng(3))
    H = M.mpo_to_dense(lat, M.build_mpo(lat))
    mpo = M.build_mpo(lat)
    # after two_site the envR stack holds the right env of site 0 (sites 1..2), envL the boundary
    envl, envr = eng.envL[-1], eng.envR[-1]
    x = eng.mps[0]
    y = T.apply_single(envl, x, mpo[0], envr)
    # dense check: <psi'|H|psi> where psi = x . B1 . B2 embeds the one-site tensor
    def embed(t):
        return M.mps_to_dense(lat, [t] + eng.mps[1:])
    rng = np.random.default_rng(4)
    u = x.copy()
    for q in u.blocks:
        u.blocks[q] = rng.standard_normal(u.blocks[q].shape) + 1j * rng.standard_normal(u.blocks[q].shape)
    lhs = overlap(y, u)  # <u|H_eff x>
    rhs = np.vdot(embed(u), H @ embed(x))
    assert abs(lhs - rhs) < 1e-10
    # Hermiticity of H_eff
    yu = T.apply_single(envl, u, mpo[0], envr)
    assert abs(overlap(y, u) - np.conj(overlap(yu, x))) < 1e-10
    # two-site: <u|H_eff x> on sites (0,1) equals the dense value
    x2 = product(eng.mps[0], eng.mps[1], 1)
    envr2 = T.env_sweep(T.boundary_envs(eng.mps)[1], eng.mps[2], mpo[2], False)
    y2 = T.apply_double(envl, x2, mpo[0], mpo[1], envr2)
    u2 = x2.copy()
    for q in u2.blocks:
        u2.blocks[q] = rng.standard_normal(u2.blocks[q].shape) + 1j * rng.standard_normal(u2.blocks[q].shape)
    def embed2(t):
        # split t back into (site0 . site1) is not needed: contract with B2 directly
        lm = product(t, eng.mps[2], 1)
        v = np.zeros(H.shape[0], dtype=complex)
        d = lat.sites[0].dim
        full = np.zeros((d, d, d), dtype=complex)
        for q, b in lm.blocks.items():
            s0 = lat.sites[0].sectors[q[1]]
            s1 = lat.sites[1].sectors[q[2]]
            s2 = lat.sites[2].sectors[q[3]]
            full[np.ix_(s0, s1, s2)] += b[0, :, :, :, 0]
        return full.reshape(-1)
    lhs2 = overlap(y2, u2)
    rhs2 = np.vdot(embed2(u2), H @ embed2(x2))
    assert abs(lhs2 - rhs2) < 1e-10
    # zero-site: H_eff on a bond matrix at bond (0|1)
    cur, nex = T.svd_impl(eng.mps[0], 2, -1.0, 10 ** 6, True)
    envl1 = T.env_sweep(envl, cur, mpo[0], True)
    z = T.apply_zero(envl1, nex, envr)
    zu = T.apply_zero(envl1, nex.scaled(1j), envr)
    assert abs(overlap(zu, nex) - 1j * overlap(z, nex)) < 1e-10  # linearity
    assert abs(overlap(z, nex).imag) < 1e-10  # <c|H|c> real


def test_local_krylov_matches_expm_on_dense_heff():
    lat, g = _small_chain("merged", N=3, L=2)
    eng = _canonical_random_state(lat, 8, np.random.default_rng(5))
    mpo = M.build_mpo(lat)
    envl, envr = eng.envL[-1], eng.envR[-1]
    x = eng.mps[0]
    keys = sorted(x.blocks)
    sizes = [x.blocks[q].size for q in keys]
    n = sum(sizes)

    def vec(t):
        return np.concatenate([t.blocks[q].ravel() for q in keys])

    def unvec(v):
        t = Tensor(x.dirs)
        p = 0
        for q, s in zip(keys, sizes):
            t.blocks[q] = v[p:p + s].reshape(x.blocks[q].shape).copy()
            p += s
        return t

    Hd = np.zeros((n, n), dtype=complex)
    for j in range(n):
        e = np.zeros(n, dtype=complex)
        e[j] = 1
        y = T.apply_single(envl, unvec(e), mpo[0], envr)
        Hd[:, j] = np.concatenate([y.blocks.get(q, np.zeros_like(x.blocks[q])).ravel() for q in keys])
    dt = 0.3
    for rev in (False, True):
        out, ex = T.local_krylov(envl, x, mpo[0], None, envr, dt, tol=1e-14, max_iter=40,
                                 reverse=rev, nsite=1)
        v0 = vec(x) / np.linalg.norm(vec(x))
        sign = 1j if rev else -1j
        ref = sla.expm(sign * 0.5 * dt * Hd) @ v0
        assert np.abs(vec(out) - ref).max() < 1e-11
    # eigenvector input: exp(-i lambda dt/2) x
    w, V = np.linalg.eigh(Hd)
    xv = unvec(V[:, 2])
    out, _ = T.local_krylov(envl, xv, mpo[0], None, envr, dt, nsite=1)
    np.testing.assert_allclose(vec(out), np.exp(-0.5j * dt * w[2]) * V[:, 2], atol=1e-12)


def test_krylov_stop_rule_and_iteration_cap():
    lat, g = _small_chain("merged", N=3, L=3)
    eng = _canonical_random_state(lat, 16, np.random.default_rng(6))
    mpo = M.build_mpo(lat)
    x = eng.mps[0]
    _, ex = T.local_krylov(eng.envL[-1], x, mpo[0], None, eng.envR[-1], 5.0, tol=1e-30,
                           max_iter=15, nsite=1)
    assert ex.iters == 14 and not ex.converged  # at most MaxIter-1 vectors
    _, ex2 = T.local_krylov(eng.envL[-1], x, mpo[0], None, eng.envR[-1], 0.05, tol=1e-8,
                            max_iter=15, nsite=1)
    assert ex2.converged and ex2.err < 1e-8 and ex2.iters < 14


def test_krylov_stops_on_the_product_of_residual_norms():
    """holstein/exponential.h:139-147 multiplies the error estimate by every residual norm
    |h_k| and stops once the product falls below the tolerance.  On a Hermitian operator whose
    residual norms stay near 0.1 the product crosses 1e-8 after a few vectors while no single
    residual norm does before the 14-vector cap, so a last-residual rule would run to the cap.
    The residual norms come from a textbook three-term Lanczos on the dense operator."""
    rng = np.random.default_rng(11)
    n, dt, tol = 40, 0.05, 1e-8
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    H = (q * np.linspace(-8.0, 8.0, n)[None, :]) @ q.conj().T
    H = 0.5 * (H + H.conj().T)
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    # apply_zero with a one-dimensional right leg is y = H c (holstein/product.h:540-545)
    envl = Tensor((1, -1, -1))
    envl.push_back_ch((0, 0, 0), H.reshape(n, 1, n))
    envr = Tensor((-1, 1, 1))
    envr.push_back_ch((0, 0, 0), np.ones((1, 1, 1), dtype=complex))
    c = Tensor((1, -1))
    c.push_back_ch((0, 0), v.reshape(n, 1))
    A = 0.5 * dt * H
    vk, vprev, bprev, betas = v / np.linalg.norm(v), np.zeros(n, complex), 0.0, []
    for _ in range(14):
        w = A @ vk - bprev * vprev
        w -= np.vdot(vk, w) * vk
        betas.append(np.linalg.norm(w))
        vprev, vk, bprev = vk, w / betas[-1], betas[-1]
    prod_iters = next(k + 1 for k in range(14) if np.prod(betas[:k + 1]) < tol)
    last_iters = next((k + 1 for k in range(14) if betas[k] < tol), 14)
    assert prod_iters < last_iters == 14
    _, ex = T.local_krylov(envl, c, None, None, envr, dt, tol=tol, max_iter=15, nsite=0)
    assert ex.converged and ex.iters == prod_iters
    assert ex.err == pytest.approx(np.prod(betas[:prod_iters]), rel=1e-6)


def _svd_fixture(svals_by_sector):
    """Two-leg tensor whose sector q has the given singular values."""
    rng = np.random.default_rng(7)
    t = Tensor((1, -1))
    for q, s in svals_by_sector.items():
        n = len(s)
        u, _ = np.linalg.qr(rng.standard_normal((n + 1, n)) + 1j * rng.standard_normal((n + 1, n)))
        v, _ = np.linalg.qr(rng.standard_normal((n + 2, n)) + 1j * rng.standard_normal((n + 2, n)))
        t.push_back_ch((q, q), (u * np.asarray(s)) @ v.conj().T)
    return t


def test_svd_truncation_rule_on_synthetic_spectra():
    A = _svd_fixture({0: [0.9, 0.3, 1e-11], 1: [0.5, 0.3, 0.1, 2e-10]})
    # cutoff 1e-10: 1e-11 dropped (s <= cutoff); 2e-10 kept
    lef, rig, sv = T.svd_impl(A, 1, 1e-10, 100, True, return_svals=True)
    assert len(sv[0]) == 2 and len(sv[1]) == 4
    # equality with the cutoff drops
    lef, rig, sv = T.svd_impl(A, 1, 2e-10, 100, True, return_svals=True)
    assert len(sv[1]) == 3
    # maxdim = 3 keeps the three largest across sectors: 0.9, 0.5, and one of the 0.3 ties;
    # the tie goes to the lower sector charge
    lef, rig, sv = T.svd_impl(A, 1, 0.0, 3, True, return_svals=True)
    assert sorted(len(v) for v in sv.values()) == [1, 2] and len(sv[0]) == 2
    # index 0 is never dropped even above every cutoff
    lef, rig, sv = T.svd_impl(A, 1, 10.0, 100, True, return_svals=True)
    assert sum(len(v) for v in sv.values()) == 1 and abs(sv[0][0] - 0.9) < 1e-12
    # reconstruction without truncation, S on either side
    for l2r in (True, False):
        lef, rig = T.svd_impl(A, 1, -1.0, 100, l2r)
        rec = product(lef, rig, 1)
        for q, b in A.blocks.items():
            np.testing.assert_allclose(rec.blocks[q], b, atol=1e-12)
        iso = lef if l2r else rig
        g = product(iso, iso, 1, "C", "N") if l2r else product(iso, iso, 1, "N", "C")
        for q, b in g.blocks.items():
            np.testing.assert_allclose(b, np.eye(b.shape[0]), atol=1e-12)


def _ed_ct(N, om, g, Ls, J, dt, nsteps):
    H = M.ed_hamiltonian(N, om, g, Ls, J)
    v0 = M.ed_initial_state(N, Ls)
    w, V = np.linalg.eigh(H)
    c0 = V.conj().T @ v0
    return np.array([np.vdot(v0, V @ (np.exp(-1j * w * dt * k) * c0)) for k in range(1, nsteps + 1)])


def test_two_site_and_one_site_tdvp_match_exact_diagonalisation():
    N, Ls, om, J, dt = 3, [3], [1.0], 1.0, 0.05
    g = np.array([0.7, 1.1, 0.9]).reshape(3, 1)
    ed = _ed_ct(N, om, g, Ls, J, dt, 15)
    lat = M.holstein(N, om, g, Ls, J, "merged")
    mpo = M.build_mpo(lat)
    psi0 = M.product_state(lat)
    eng = T.TDVP(psi0, mpo, cutoff=-1.0, maxdim=10 ** 6, dt=dt)
    cs = []
    for _ in range(10):
        eng.two_site()
        cs.append(T.overlap_states(psi0, eng.mps))
    assert np.abs(np.array(cs) - ed[:10]).max() < 1e-10
    eng1 = T.TDVP(eng.mps, mpo, cutoff=-1.0, maxdim=10 ** 6, dt=dt)
    cs1 = []
    for _ in range(5):
        eng1.one_site()
        cs1.append(T.overlap_states(psi0, eng1.mps))
    assert np.abs(np.array(cs1) - ed[10:15]).max() < 1e-10


def test_free_exciton_closed_form():
    """g = 0: C(t) = sum_k phi_k(0)^2 exp(-i eps_k t); 2TDVP's splitting error is third
    order in dt here, so the bound scales with dt."""
    N, J = 6, 1.0
    k = np.arange(1, N + 1)
    eps = 2 * J * np.cos(k * np.pi / (N + 1))
    phi0 = np.sqrt(2 / (N + 1)) * np.sin(k * np.pi / (N + 1))
    errs = []
    for dt in (0.05, 0.025):
        lat = M.holstein(N, [1.0], 0.0, [4], J, "merged")
        psi0 = M.product_state(lat)
        eng = T.TDVP(psi0, M.build_mpo(lat), cutoff=1e-10, maxdim=64, dt=dt, tol_krylov=1e-12)
        for _ in range(int(round(0.5 / dt))):
            eng.two_site()
        c = T.overlap_states(psi0, eng.mps)
        errs.append(abs(c - np.sum(phi0 ** 2 * np.exp(-1j * eps * 0.5))))
        # populations sum to one, norm one
        pops = [T.join(M.density_mpo(lat, i), eng.mps, eng.mps).real for i in range(N)]
        assert abs(sum(pops) - 1) < 1e-10
    assert errs[0] < 5e-7 and errs[1] < errs[0] / 5


def test_displaced_oscillator_closed_form():
    L, g, w, dt = 10, 0.8, 1.3, 0.05
    lat = M.holstein(2, [w], g, [L], 0.0, "merged")
    psi0 = M.product_state(lat)
    eng = T.TDVP(psi0, M.build_mpo(lat), cutoff=1e-12, maxdim=64, dt=dt)
    b = np.diag(np.sqrt(np.arange(1, L)), 1)
    ev, U = np.linalg.eigh(w * b.T @ b + g * (b + b.T))
    for s in range(1, 21):
        eng.two_site()
        ex = (U @ (np.exp(-1j * ev * dt * s) * U[0].conj()))[0]
        assert abs(T.overlap_states(psi0, eng.mps) - ex) < 1e-10
    # truncation at L=10 is already converged against the untruncated Glauber form at t=1
    t = 1.0
    glauber = np.exp(-(g / w) ** 2 * (1 - np.exp(-1j * w * t)) + 1j * g * g / w * t)
    assert abs(ex - glauber) < 1e-9


def test_forward_backward_returns_initial_state_and_conjugate_symmetry():
    N, Ls, om, J, dt = 3, [3], [1.0], 1.0, 0.1
    g = 0.8
    lat = M.holstein(N, om, g, Ls, J, "merged")
    mpo = M.build_mpo(lat)
    psi0 = M.product_state(lat)
    fw = T.TDVP(psi0, mpo, cutoff=-1.0, maxdim=10 ** 6, dt=dt)
    for _ in range(4):
        fw.two_site()
    c_plus = T.overlap_states(psi0, fw.mps)
    bw = T.TDVP(fw.mps, mpo, cutoff=-1.0, maxdim=10 ** 6, dt=-dt)
    for _ in range(4):
        bw.two_site()
    assert abs(T.overlap_states(psi0, bw.mps) - 1.0) < 1e-10
    neg = T.TDVP(psi0, mpo, cutoff=-1.0, maxdim=10 ** 6, dt=-dt)
    for _ in range(4):
        neg.two_site()
    assert abs(T.overlap_states(psi0, neg.mps) - np.conj(c_plus)) < 1e-10


def test_join_energy_equals_dense_expectation():
    lat, g = _small_chain("merged", N=3, L=2)
    eng = _canonical_random_state(lat, 8, np.random.default_rng(8))
    H = M.mpo_to_dense(lat, M.build_mpo(lat))
    v = M.mps_to_dense(lat, eng.mps)
    e = T.join(M.build_mpo(lat), eng.mps, eng.mps)
    assert abs(e - np.vdot(v, H @ v)) < 1e-10
    hop = M.mpo_to_dense(lat, M.hopping_mpo(lat, 0))
    assert abs(T.join(M.hopping_mpo(lat, 0), eng.mps, eng.mps) - np.vdot(v, hop @ v)) < 1e-10
