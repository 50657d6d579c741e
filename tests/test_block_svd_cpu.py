"""CPU pin of the bond SVD's block subspace route, step for step as the device runs it.

The device route (randomised range finder with power steps and Rayleigh-Ritz) has no
counterpart in the reference, which calls LAPACK's full SVD per sector.  This restates its
steps in numpy — Cholesky QR with pivots at or below 1e-12 x the largest dropped
(chol_upper: the unblocked right-looking elimination; chol_left: the device's
left-looking 32-column panels, which also form the inverse factor T, checked against the
unblocked factor and its inverse by back substitution), the power steps
Q -> orth(A^H Q) -> orth(A P), the Ritz pairs, their residuals and the ceiling on the values
beyond the certified ones, and the check that settles a cut — and checks them against LAPACK
(numpy.linalg.svd): orthonormality of the Cholesky QR bases, the certified singular values,
the isometry and
the truncated product.  Planted slips (no second Cholesky QR pass, no power step) must fail.
"""
import numpy as np
import pytest

DROP, RESID, OVERSAMPLE = 1e-12, 1e-13, 32


def chol_upper(S, tau=DROP):
    """Upper factor R (S = R^H R) with dependent columns dropped; returns (R, live)."""
    S = np.array(S, dtype=complex)
    n = S.shape[0]
    dmax = max(np.real(np.diag(S)).max(), 0.0)
    R = np.zeros_like(S)
    live = np.zeros(n, dtype=bool)
    for k in range(n):
        piv = S[k, k].real
        lv = piv > tau * dmax and piv > 0
        live[k] = lv
        if not lv:
            continue
        r = np.sqrt(piv)
        R[k, k] = r
        R[k, k + 1:] = S[k, k + 1:] / r
        S[k + 1:, k + 1:] -= np.outer(R[k, k + 1:].conj(), R[k, k + 1:])
    return R, live


def tri_inv(R, live):
    n = R.shape[0]
    T = np.zeros_like(R)
    for c in range(n):
        if not live[c]:
            continue
        x = np.zeros(n, dtype=complex)
        for i in range(c, -1, -1):
            if not live[i]:
                continue
            acc = R[i, i + 1:c + 1] @ x[i + 1:c + 1]
            x[i] = 1.0 / R[i, i].real if i == c else -acc / R[i, i].real
        T[:c + 1, c] = x[:c + 1]
    return T


def chol_left(S, nb=32, tau=DROP):
    """chol_diag_kernel / chol_panel_kernel's order: per panel k of nb columns
    C = S_{k,k:} - R_{<k,k}^H R_{<k,k:}; its diagonal block factored right-looking with the same
    pivot drops and inverted by recursive doubling (Di);
    R_{k,c} = Di^H C_{k,c} right of it; T_{k,k} = Di and T_{<k,k} = -(T_{<k,<k} R_{<k,k}) Di.
    Returns (R, live, T)."""
    S = np.array(S, dtype=complex)
    n = S.shape[0]
    thr = tau * max(np.real(np.diag(S)).max(), 0.0)
    R = np.zeros_like(S)
    T = np.zeros_like(S)
    live = np.zeros(n, dtype=bool)
    for p in range(0, n, nb):
        w = min(nb, n - p)
        C = S[p:p + w, p:] - R[:p, p:p + w].conj().T @ R[:p, p:]
        Dg = np.triu(C[:, :w])
        Rf = np.zeros((w, w), dtype=complex)
        lv = np.zeros(w, dtype=bool)
        for k in range(w):
            # one phase per pivot: R's row is Dg's row over sqrt(piv); the trailing triangle
            # takes the update with Dg's unscaled row over piv
            piv = Dg[k, k].real
            lk = piv > thr and piv > 0
            lv[k] = lk
            sp = np.sqrt(piv) if lk else 0.0
            ri = 1 / sp if lk else 0.0
            rp = ri * ri
            Rf[k, k] = sp
            Rf[k, k + 1:] = Dg[k, k + 1:] * ri
            if lk:
                for i in range(k + 1, w):
                    Dg[i, i:] -= (np.conj(Dg[k, i]) * Dg[k, i:]) * rp
        # the inverse by recursive doubling over a 32-wide block (the kernel's padding beyond w
        # is zero and stays decoupled): T_AC = -T_A R_AC T_C for pairs of blocks of 1, 2, 4, ...
        Rb = np.zeros((nb, nb), dtype=complex)
        Rb[:w, :w] = Rf
        Tb = np.zeros((nb, nb), dtype=complex)
        for i in range(w):
            if lv[i]:
                Tb[i, i] = 1 / Rb[i, i].real
        h = 1
        while h < nb:
            for q in range(0, nb, 2 * h):
                ia, ic = slice(q, q + h), slice(q + h, q + 2 * h)
                Tb[ia, ic] = -Tb[ia, ia] @ (Rb[ia, ic] @ Tb[ic, ic])
            h *= 2
        Di = Tb[:w, :w]
        live[p:p + w] = lv
        R[p:p + w, p:p + w] = Rf
        R[p:p + w, p + w:] = Di.conj().T @ C[:, w:]
        T[p:p + w, p:p + w] = Di
        T[:p, p:p + w] = -(T[:p, :p] @ R[:p, p:p + w]) @ Di
    return R, live, T


def block_orth(Y, passes=2, refill=False, seed=99):
    """Cholesky QR passes; dependent columns come out zero, or with `refill` are replaced by
    random vectors at the largest column's scale and the first pass repeats (at most twice;
    later passes factor an orthonormal basis and are not checked)."""
    rng = np.random.default_rng(seed)
    for pas in range(passes):
        for attempt in range(3):
            S = Y.conj().T @ Y
            R, live, T = chol_left(S)
            if live.all() or not refill or pas > 0 or attempt == 2:
                break
            scale = np.sqrt(1.5 * np.real(np.diag(S)).max() / Y.shape[0])
            dead = ~live
            Y = Y.copy()
            Y[:, dead] = scale * (rng.uniform(-1, 1, (Y.shape[0], dead.sum()))
                                  + 1j * rng.uniform(-1, 1, (Y.shape[0], dead.sum())))
        Y = Y @ T
    return Y


def rand_block(rows, cols, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, (rows, cols)) + 1j * rng.uniform(-1, 1, (rows, cols))


def block_sector(M, maxdim, iso_rows=True, seed=0, power_max=3, b=None, need=None):
    """The device's block route for one sector: start block (one Cholesky QR pass), power
    steps (A^H Q one pass, A P two passes with refill), Rayleigh-Ritz, residuals and the
    ceiling.  Power steps stop once the leading min(b, maxdim + 1) pairs are certified (the
    device decides from the cut; `need` stands in for it).  Returns (certified singular
    values, Z isometric vectors, W = A^H Z, residuals, ceiling on every other value)."""
    A = M if iso_rows else M.conj().T
    b = b or maxdim + OVERSAMPLE
    Q = block_orth(A @ rand_block(A.shape[1], b, seed), passes=1)
    f2 = np.vdot(M, M).real
    for q in range(max(power_max, 1)):
        if power_max > 0:
            P = block_orth(A.conj().T @ Q, passes=1)
            Q = block_orth(A @ P, passes=2, refill=True, seed=seed + q)
        P = A.conj().T @ Q
        w, V = np.linalg.eigh((P.conj().T @ P) / f2)
        w, V = w[::-1], V[:, ::-1]
        kc = b
        theta = np.maximum(w[:kc], 0) * f2
        Z, W = Q @ V[:, :kc], P @ V[:, :kc]
        res = np.linalg.norm(A @ W - Z * theta[None, :], axis=0)
        cert = 0
        while cert < kc and res[cert] <= RESID * theta[0]:
            cert += 1
        if cert >= (need or kc):
            break
    ceil = np.sqrt(max(f2 - theta[:cert].sum(), 0.0) + 1e-13 * f2)
    return np.sqrt(theta[:cert]), Z, W, res, ceil


def cut_ok(sv, ceil, kept, cutoff, maxdim, total_kept, s_last, tail=1e-5):
    """svd_truncate's check on a block-route sector after the global cut: the kept values
    stay above the tail, and every value beyond the certified ones falls below the cutoff
    or, when maxdim binds, below the smallest kept value."""
    tail_ok = not (kept > 0 and sv[kept - 1] < tail * sv[0])
    return tail_ok and (ceil <= cutoff or (total_kept >= maxdim and ceil < s_last))


def spectrum_matrix(rng, nr, nc, s):
    u, _ = np.linalg.qr(rng.standard_normal((nr, len(s))) + 1j * rng.standard_normal((nr, len(s))))
    v, _ = np.linalg.qr(rng.standard_normal((nc, len(s))) + 1j * rng.standard_normal((nc, len(s))))
    return (u * s[None, :]) @ v.conj().T


def test_cholesky_qr_with_dropped_columns_matches_lapack_qr():
    rng = np.random.default_rng(1)
    Y = rng.standard_normal((300, 40)) + 1j * rng.standard_normal((300, 40))
    Y[:, 7] = Y[:, 2] * (1 - 2j) + Y[:, 5]  # dependent columns are dropped
    Y[:, 30] = 0.0
    Y[:, 11] *= 1e-4  # a small but independent column stays
    R, live = chol_upper(Y.conj().T @ Y)
    assert not live[7] and not live[30] and live[11] and live.sum() == 38
    T = tri_inv(R, live)
    np.testing.assert_allclose(R[np.ix_(live, live)] @ T[np.ix_(live, live)], np.eye(38), atol=1e-10)
    Q = block_orth(Y)  # dependent columns come out zero
    assert np.abs(Q[:, ~live]).max() == 0.0
    Ql = Q[:, live]
    np.testing.assert_allclose(Ql.conj().T @ Ql, np.eye(38), atol=1e-13)
    # the span is LAPACK's QR span of the live columns
    q2, _ = np.linalg.qr(Y[:, live])
    np.testing.assert_allclose(np.abs(np.linalg.svd(q2.conj().T @ Ql, compute_uv=False)), 1.0, atol=1e-12)
    Qr = block_orth(Y, refill=True)  # or are refilled: 40 orthonormal columns
    np.testing.assert_allclose(Qr.conj().T @ Qr, np.eye(40), atol=1e-13)


def test_left_looking_cholesky_and_its_inverse_match_the_unblocked_factor():
    """The left-looking panel factorisation gives the unblocked factor R and T = R^{-1} over
    the live columns, with dependent columns inside a panel (7, 40), on a panel edge (63) and
    in the last, partial panel (70), on a side (75) that is not a multiple of the panel; Y T
    has orthonormal live columns and zero dropped ones."""
    rng = np.random.default_rng(13)
    Y = rng.standard_normal((300, 75)) + 1j * rng.standard_normal((300, 75))
    Y[:, 7] = Y[:, 3] - 2j * Y[:, 5]
    Y[:, 40] = 0.5 * Y[:, 39]
    Y[:, 63] = Y[:, 0] + Y[:, 62]
    Y[:, 70] = 1j * Y[:, 65]
    S = Y.conj().T @ Y
    R1, live1 = chol_upper(S)
    R2, live2, T = chol_left(S)
    assert np.array_equal(live1, live2) and live1.sum() == 71
    np.testing.assert_allclose(R2, R1, atol=1e-12 * np.abs(R1).max())
    np.testing.assert_allclose(T, tri_inv(R1, live1), atol=1e-12 * np.abs(T).max())
    Q = Y @ T
    assert np.abs(Q[:, ~live1]).max() == 0.0
    Ql = Q[:, live1]
    np.testing.assert_allclose(Ql.conj().T @ Ql, np.eye(71), atol=1e-12)


def test_one_cholesky_qr_pass_is_not_enough_on_ill_conditioned_blocks():
    rng = np.random.default_rng(2)
    Y = spectrum_matrix(rng, 400, 60, np.logspace(0, -5.5, 60))
    one = block_orth(Y, passes=1)
    two = block_orth(Y, passes=2)
    assert np.abs(two.conj().T @ two - np.eye(two.shape[1])).max() < 1e-13
    assert np.abs(one.conj().T @ one - np.eye(one.shape[1])).max() > 1e-9


@pytest.mark.parametrize("iso_rows", [True, False])
def test_block_route_matches_lapack_on_a_bond_like_spectrum(iso_rows):
    """Sector of 800 x 1100 whose values fall like an evolved two-site tensor's: 60 at
    O(1..1e-2), then a cliff to 1e-7 and below.  Cut at maxdim 60 with a block of 60 + 32:
    the cliff values below the Cholesky QR's resolution leave the block and come back as
    random directions, the certified values and the ceiling (the mass beyond the certified
    pairs, ~1e-7, plus a rounding allowance of sqrt(1e-13) |M|_F) settle the cut, and values,
    isometry and product match LAPACK.  Cut at maxdim 64 with cutoff 1e-10: values of 1e-7
    must be kept, which the ceiling cannot settle, so the sector goes to the full route."""
    rng = np.random.default_rng(3)
    s = np.concatenate([np.logspace(0, -2, 60), 1e-7 * np.logspace(0, -6, 200)])
    M = spectrum_matrix(rng, 800, 1100, s)
    sv, Z, W, res, ceil = block_sector(M, 60, iso_rows)
    ref = np.linalg.svd(M, compute_uv=False)
    k = 60
    assert len(sv) >= k and ceil < 1e-6
    assert cut_ok(sv, ceil, k, 1e-10, 60, k, sv[k - 1])
    np.testing.assert_allclose(sv[:k], ref[:k], rtol=1e-9)
    np.testing.assert_allclose(Z.conj().T @ Z, np.eye(Z.shape[1]), atol=1e-13)
    rec = Z[:, :k] @ W[:, :k].conj().T if iso_rows else (Z[:, :k] @ W[:, :k].conj().T).conj().T
    u, s_, vh = np.linalg.svd(M, full_matrices=False)
    ref_rec = (u[:, :k] * s_[None, :k]) @ vh[:k]
    assert np.abs(rec - ref_rec).max() <= 1e-11 * np.abs(ref_rec).max()
    sv, Z, W, res, ceil = block_sector(M, 64, iso_rows)
    assert not cut_ok(sv, ceil, 64, 1e-10, 64, 64 if len(sv) >= 64 else len(sv), sv[min(63, len(sv) - 1)])


def test_block_route_power_steps_until_certified():
    """Seventy values from 1 to 0.7, then a tail from 1e-5: the range finder alone leaves
    Ritz residuals far above 1e-13 of the largest value, one power step brings every pair of
    the block of 96 under it, the flat top included, so the mass beyond them is the tail's
    and the cut at 64 is settled; the values match LAPACK."""
    rng = np.random.default_rng(4)
    s = np.concatenate([1.0 - 0.3 * np.arange(70) / 70, 1e-5 * np.exp(-np.arange(500) / 30.0)])
    M = spectrum_matrix(rng, 900, 1000, s)
    sv0, _, _, _, c0 = block_sector(M, 64, power_max=0)
    sv1, _, _, _, c1 = block_sector(M, 64, power_max=1, need=70)
    ref = np.linalg.svd(M, compute_uv=False)
    assert len(sv0) < 65 <= len(sv1)
    np.testing.assert_allclose(sv1[:65], ref[:65], rtol=1e-10)
    if len(sv0):  # with nothing certified the device goes to the full route at once
        assert not cut_ok(sv0, c0, min(64, len(sv0)), 1e-10, 64, 64, sv0[min(63, len(sv0) - 1)])
    assert cut_ok(sv1, c1, 64, 1e-10, 64, 64, sv1[63])


def test_block_route_rank_deficient_sector_refills_columns():
    """Rank 40 < block 96: the range finder's dependent columns drop out, the final basis
    refills them with random directions, so it keeps 96 orthonormal columns; the forty
    values match LAPACK, the rest are null (at rounding), and the ceiling sits at the rounding
    allowance, so a cut at any cutoff above it is settled."""
    rng = np.random.default_rng(5)
    M = spectrum_matrix(rng, 600, 700, np.logspace(0, -3, 40))
    sv, Z, W, res, ceil = block_sector(M, 64)
    assert Z.shape[1] == 96 and len(sv) >= 40
    np.testing.assert_allclose(Z.conj().T @ Z, np.eye(96), atol=1e-13)
    ref = np.linalg.svd(M, compute_uv=False)[:40]
    np.testing.assert_allclose(sv[:40], ref, rtol=1e-9)
    assert np.all(sv[40:] < 1e-6 * sv[0])
    assert ceil < 1e-6 and cut_ok(sv, ceil, 40, 1e-6, 64, 40, sv[39])
