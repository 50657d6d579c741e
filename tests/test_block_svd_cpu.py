"""CPU pin of the bond SVD's block subspace route, step for step as the device runs it.

The device route (randomised range finder with power steps and Rayleigh-Ritz) has no
counterpart in the reference, which calls LAPACK's full SVD per sector.  This restates its
steps in numpy — Cholesky QR with pivots at or below 1e-12 x the largest dropped
(right-looking, the order chol_upper_kernel eliminates in), the inverse of the factor by
back substitution (tri_inv_kernel), the compaction of live columns, the power steps
Q -> orth(A^H Q) -> orth(A P), the Ritz pairs, their residuals and the ceiling on the values
beyond the certified ones — and checks them against LAPACK (numpy.linalg.svd):
orthonormality of the Cholesky QR bases, the certified singular values, the isometry and
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


def block_orth(Y, passes=2):
    for _ in range(passes):
        R, live = chol_upper(Y.conj().T @ Y)
        Y = (Y @ tri_inv(R, live))[:, live]
    return Y


def rand_block(rows, cols, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, (rows, cols)) + 1j * rng.uniform(-1, 1, (rows, cols))


def block_sector(M, maxdim, iso_rows=True, seed=0, power_max=3, passes=2):
    """The device's block_sector: (certified singular values, Z isometric vectors,
    W = A^H Z, residuals, ceiling on every other value)."""
    A = M if iso_rows else M.conj().T
    b = maxdim + OVERSAMPLE
    Q = block_orth(A @ rand_block(A.shape[1], b, seed), passes)
    f2 = np.vdot(M, M).real
    for q in range(max(power_max, 1)):
        if power_max > 0:
            P = block_orth(A.conj().T @ Q, passes)
            Q = block_orth(A @ P, passes)
        P = A.conj().T @ Q
        w, V = np.linalg.eigh((P.conj().T @ P) / f2)
        w, V = w[::-1], V[:, ::-1]
        kc = min(Q.shape[1], maxdim + 1)
        theta = np.maximum(w[:kc], 0) * f2
        Z, W = Q @ V[:, :kc], P @ V[:, :kc]
        res = np.linalg.norm(A @ W - Z * theta[None, :], axis=0)
        cert = 0
        while cert < kc and res[cert] <= RESID * theta[0]:
            cert += 1
        if cert == kc:
            break
    ceil = np.sqrt(max(f2 - theta[:cert].sum(), 0.0) + 1e-13 * f2)
    return np.sqrt(theta[:cert]), Z, W, res, ceil


def cut_ok(sv, ceil, kept, cutoff, maxdim, total_kept, s_last, tail=1e-5):
    """svd_truncate's check on a block-route sector after the global cut."""
    ok = not (kept > 0 and sv[kept - 1] < tail * sv[0])
    if kept >= len(sv):
        ok = ok and (ceil <= cutoff or (total_kept >= maxdim and ceil < s_last))
    return ok


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
    Q = block_orth(Y)
    assert Q.shape == (300, 38)
    np.testing.assert_allclose(Q.conj().T @ Q, np.eye(38), atol=1e-13)
    # the span is LAPACK's QR span of the live columns
    q2, _ = np.linalg.qr(Y[:, live])
    np.testing.assert_allclose(np.abs(np.linalg.svd(q2.conj().T @ Q, compute_uv=False)), 1.0, atol=1e-12)


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
    O(1..1e-2), then a cliff to 1e-7 and below.  Cut at maxdim 60: the certified values
    stop at the cliff, the ceiling (the mass beyond it, ~1e-7, plus a rounding allowance of
    sqrt(1e-13) |M|_F) lies below the cut, and values, isometry and
    product match LAPACK.  Cut at maxdim 64 with cutoff 1e-10: values of 1e-7 must be kept,
    the ceiling is above the cutoff, and the sector has to go to the full route."""
    rng = np.random.default_rng(3)
    s = np.concatenate([np.logspace(0, -2, 60), 1e-7 * np.logspace(0, -6, 200)])
    M = spectrum_matrix(rng, 800, 1100, s)
    sv, Z, W, res, ceil = block_sector(M, 60, iso_rows)
    ref = np.linalg.svd(M, compute_uv=False)
    k = 60
    assert len(sv) == k and 1e-7 < ceil < 1e-6  # the mass beyond the cliff plus rounding slack
    assert cut_ok(sv, ceil, k, 1e-10, 60, k, sv[-1])
    np.testing.assert_allclose(sv, ref[:k], rtol=1e-9)
    np.testing.assert_allclose(Z[:, :k].conj().T @ Z[:, :k], np.eye(k), atol=1e-13)
    rec = Z[:, :k] @ W[:, :k].conj().T if iso_rows else (Z[:, :k] @ W[:, :k].conj().T).conj().T
    u, s_, vh = np.linalg.svd(M, full_matrices=False)
    ref_rec = (u[:, :k] * s_[None, :k]) @ vh[:k]
    assert np.abs(rec - ref_rec).max() <= 1e-11 * np.abs(ref_rec).max()
    sv, Z, W, res, ceil = block_sector(M, 64, iso_rows)
    assert len(sv) == k and not cut_ok(sv, ceil, k, 1e-10, 64, k, sv[-1])


def test_block_route_power_steps_until_certified():
    """Seventy values from 1 to 0.7, then a tail from 1e-5: the range finder alone leaves
    Ritz residuals far above 1e-13 of the largest value, one power step brings the leading
    maxdim + 1 pairs under it; the values then match LAPACK and the next value is certified."""
    rng = np.random.default_rng(4)
    s = np.concatenate([1.0 - 0.3 * np.arange(70) / 70, 1e-5 * np.exp(-np.arange(500) / 30.0)])
    M = spectrum_matrix(rng, 900, 1000, s)
    sv0, _, _, _, _ = block_sector(M, 64, power_max=0)
    sv1, _, _, _, _ = block_sector(M, 64, power_max=1)
    ref = np.linalg.svd(M, compute_uv=False)
    assert len(sv0) < 65 <= len(sv1)
    np.testing.assert_allclose(sv1[:65], ref[:65], rtol=1e-10)
    assert cut_ok(sv1, 0.0, 64, 1e-10, 64, 64, sv1[63])


def test_block_route_rank_deficient_sector_drops_columns():
    """Rank 40 < block 96: the start block loses its dependent columns and the certified
    values stop at the rank, with a ceiling at the rounding level, so a cut that keeps all
    forty is certified."""
    rng = np.random.default_rng(5)
    M = spectrum_matrix(rng, 600, 700, np.logspace(0, -3, 40))
    sv, Z, W, res, ceil = block_sector(M, 64)
    assert Z.shape[1] <= 41 and len(sv) >= 40
    ref = np.linalg.svd(M, compute_uv=False)[:40]
    np.testing.assert_allclose(sv[:40], ref, rtol=1e-9)
    assert ceil < 1e-6 and cut_ok(sv, ceil, len(sv), 1e-6, 64, len(sv), sv[-1])
